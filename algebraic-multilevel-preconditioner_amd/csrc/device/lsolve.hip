// lsolve.hip -- level-synchronous batched local solve: for every level of all subdomains'
// supernodal trees a pull launch and the unit launches of its large and small supernodes
// (forward), or a partial-product launch and the unit launches (backward).
#include "device.hpp"
#include "lsolve.hpp"

namespace s2l {

namespace {

template <typename T> __global__ __launch_bounds__(kLsThreads) void k_ls_pull(LsArgs<T> a, int32_t first, int32_t count)
{
    __shared__ T scr[kLsThreads];
    const int rr = blockIdx.y;
    if (a.skip && a.skip[rr]) return;
    const int t = threadIdx.x;
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kLsThreads + t;
    lsPullA(a, first, count, gid, t, rr, scr);
    __syncthreads();
    lsPullB(a, first, count, gid, t, rr, scr);
}

template <typename T> __global__ __launch_bounds__(kLsThreads) void k_ls_forward(LsArgs<T> a, int32_t first)
{
    extern __shared__ unsigned char lds_raw[];
    T* lds = reinterpret_cast<T*>(lds_raw);
    const int rr = blockIdx.y;
    if (a.skip && a.skip[rr]) return;
    const LsUnit u = a.units[first + blockIdx.x];
    const int q = threadIdx.x / kLsLanes, l = threadIdx.x % kLsLanes;
    lsFwdA(a, u, q, l, rr, lds);
    __syncthreads();
    lsFwdB(a, u, q, l, rr, lds);
    __syncthreads();
    lsFwdC(a, u, q, l, rr, lds);
    __syncthreads();
    lsFwdD(a, u, q, l, rr, lds);
}

// one small supernode per workgroup of one wavefront (the phases of wavefront 0)
template <typename T> __global__ __launch_bounds__(kLsLanes) void k_ls_forward_small(LsArgs<T> a, int32_t first)
{
    __shared__ T lds[2 * kLsLanes];
    const int rr = blockIdx.y;
    if (a.skip && a.skip[rr]) return;
    const LsUnit u = a.units[first + blockIdx.x];
    const int l = threadIdx.x;
    lsFwdA(a, u, 0, l, rr, lds);
    __syncthreads();
    lsFwdA2(a, u, 0, l, rr, lds);
    __syncthreads();
    lsFwdB(a, u, 0, l, rr, lds);
    __syncthreads();
    lsFwdD(a, u, 0, l, rr, lds);
}

template <typename T> __global__ __launch_bounds__(kLsLanes) void k_ls_backward_small(LsArgs<T> a, int32_t first)
{
    __shared__ T lds[2 * kLsLanes];
    const int rr = blockIdx.y;
    if (a.skip && a.skip[rr]) return;
    const LsUnit u = a.units[first + blockIdx.x];
    const int l = threadIdx.x;
    lsBwdA(a, u, 0, l, rr, lds);
    __syncthreads();
    lsBwdB(a, u, 0, l, rr, lds);
    __syncthreads();
    lsBwdC(a, u, 0, l, rr, lds);
}

template <typename T> __global__ __launch_bounds__(kLsLanes) void k_ls_back_partial(LsArgs<T> a, int32_t first)
{
    const int rr = blockIdx.y;
    if (a.skip && a.skip[rr]) return;
    const LsItem it = a.items[first + blockIdx.x];
    lsBackPartial(a, it, static_cast<int>(threadIdx.x), rr);
}

template <typename T> __global__ __launch_bounds__(kLsThreads) void k_ls_backward(LsArgs<T> a, int32_t first)
{
    extern __shared__ unsigned char lds_raw[];
    T* lds = reinterpret_cast<T*>(lds_raw);
    const int rr = blockIdx.y;
    if (a.skip && a.skip[rr]) return;
    const LsUnit u = a.units[first + blockIdx.x];
    const int q = threadIdx.x / kLsLanes, l = threadIdx.x % kLsLanes;
    lsBwdA(a, u, q, l, rr, lds);
    __syncthreads();
    lsBwdB(a, u, q, l, rr, lds);
    __syncthreads();
    lsBwdC(a, u, q, l, rr, lds);
    __syncthreads();
    lsBwdD(a, u, q, l, rr, lds);
}

} // namespace

template <typename T>
void launchLs(const LsArgs<T>& a, int kind, int32_t first, int32_t count, int32_t maxw, int nrhs, hipStream_t s)
{
    if (count <= 0) return;
    LsArgs<T> b = a;
    b.lds_w = maxw;
    const size_t lds = static_cast<size_t>(lsUnitLds(maxw)) * sizeof(T);
    if (kind == kLsPull) {
        const int64_t threads = static_cast<int64_t>(count) * kLsPullLanes;
        const dim3 grid(static_cast<unsigned>((threads + kLsThreads - 1) / kLsThreads), static_cast<unsigned>(nrhs));
        hipLaunchKernelGGL(k_ls_pull<T>, grid, dim3(kLsThreads), 0, s, b, first, count);
    } else if (kind == kLsForward) {
        hipLaunchKernelGGL(k_ls_forward<T>, dim3(static_cast<unsigned>(count), static_cast<unsigned>(nrhs)), dim3(kLsThreads),
                           lds, s, b, first);
    } else if (kind == kLsForwardSmall) {
        hipLaunchKernelGGL(k_ls_forward_small<T>, dim3(static_cast<unsigned>(count), static_cast<unsigned>(nrhs)),
                           dim3(kLsLanes), 0, s, b, first);
    } else if (kind == kLsBackwardSmall) {
        hipLaunchKernelGGL(k_ls_backward_small<T>, dim3(static_cast<unsigned>(count), static_cast<unsigned>(nrhs)),
                           dim3(kLsLanes), 0, s, b, first);
    } else if (kind == kLsBackPartial) {
        hipLaunchKernelGGL(k_ls_back_partial<T>, dim3(static_cast<unsigned>(count), static_cast<unsigned>(nrhs)),
                           dim3(kLsLanes), 0, s, b, first);
    } else {
        hipLaunchKernelGGL(k_ls_backward<T>, dim3(static_cast<unsigned>(count), static_cast<unsigned>(nrhs)), dim3(kLsThreads),
                           lds, s, b, first);
    }
}

template void launchLs<float>(const LsArgs<float>&, int, int32_t, int32_t, int32_t, int, hipStream_t);
template void launchLs<double>(const LsArgs<double>&, int, int32_t, int32_t, int32_t, int, hipStream_t);

} // namespace s2l
