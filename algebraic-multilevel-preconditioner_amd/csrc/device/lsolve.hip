// lsolve.hip -- level-synchronous batched local solve: one launch per level of the supernodal
// elimination trees of all subdomains, one workgroup of one wavefront per work item.
#include "device.hpp"
#include "lsolve.hpp"

namespace s2l {

namespace {

template <typename T> __global__ __launch_bounds__(kLsLanes) void k_ls_forward(LsArgs<T> a, int32_t first)
{
    extern __shared__ unsigned char lds_raw[];
    T* bl = reinterpret_cast<T*>(lds_raw);
    T* yl = bl + a.lds_w;
    const int rr = blockIdx.y;
    if (a.skip && a.skip[rr]) return;
    const LsItem it = a.items[first + blockIdx.x];
    const LsSn d = a.sn[it.sn];
    const int lane = threadIdx.x;
    lsForwardPull(a, d, rr, lane, bl);
    __syncthreads();
    lsForwardDiag(a, d, it.part, rr, lane, bl, yl);
    __syncthreads();
    lsForwardPush(a, d, it.part, rr, lane, yl);
}

template <typename T> __global__ __launch_bounds__(kLsLanes) void k_ls_back_partial(LsArgs<T> a, int32_t first)
{
    const int rr = blockIdx.y;
    if (a.skip && a.skip[rr]) return;
    const LsItem it = a.items[first + blockIdx.x];
    const LsSn d = a.sn[it.sn];
    lsBackPartial(a, d, it.part, rr, static_cast<int>(threadIdx.x));
}

template <typename T> __global__ __launch_bounds__(kLsLanes) void k_ls_backward(LsArgs<T> a, int32_t first)
{
    extern __shared__ unsigned char lds_raw[];
    T* bl = reinterpret_cast<T*>(lds_raw);
    const int rr = blockIdx.y;
    if (a.skip && a.skip[rr]) return;
    const LsItem it = a.items[first + blockIdx.x];
    const LsSn d = a.sn[it.sn];
    const int lane = threadIdx.x;
    lsBackwardPull(a, d, rr, lane, bl);
    __syncthreads();
    lsBackwardDiag(a, d, it.part, rr, lane, bl);
}

} // namespace

template <typename T>
void launchLs(const LsArgs<T>& a, int kind, int32_t first, int32_t count, int32_t maxw, int nrhs, hipStream_t s)
{
    if (count <= 0) return;
    const dim3 grid(static_cast<unsigned>(count), static_cast<unsigned>(nrhs)), block(kLsLanes);
    LsArgs<T> b = a;
    b.lds_w = maxw;
    const size_t lds = (static_cast<size_t>(maxw) + kLsLanes) * sizeof(T);
    if (kind == kLsForward)
        hipLaunchKernelGGL(k_ls_forward<T>, grid, block, lds, s, b, first);
    else if (kind == kLsBackPartial)
        hipLaunchKernelGGL(k_ls_back_partial<T>, grid, block, 0, s, b, first);
    else
        hipLaunchKernelGGL(k_ls_backward<T>, grid, block, lds, s, b, first);
}

template void launchLs<float>(const LsArgs<float>&, int, int32_t, int32_t, int32_t, int, hipStream_t);
template void launchLs<double>(const LsArgs<double>&, int, int32_t, int32_t, int32_t, int, hipStream_t);

} // namespace s2l
