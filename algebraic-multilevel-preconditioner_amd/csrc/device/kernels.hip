// kernels.hip -- HIP kernels of the two-level Schwarz apply path for CDNA4 (gfx950).
//
// Hot kernel: k_local_solve, the batched supernodal triangular solves of every subdomain in one
// launch (one 1024-thread workgroup per subdomain; the subdomain's vector lives in LDS). It
// fuses the restriction (gather x_loc = R_i r through the factor permutation), the forward and
// backward sweeps and the RAS prolongation (z = x0 + y on the owned rows). Supernodes are
// processed by levels of the supernodal elimination tree; a supernode up to 64 columns wide is
// solved by one wavefront (register-resident triangle, wave shuffles), a wider one by the whole
// workgroup in 64-column panels. The forward sweep pushes the products of each supernode's rows
// below the diagonal block into a T buffer which the ancestors pull in a fixed order, so every
// sum has one order: results are bitwise repeatable.
#include "device.hpp"

namespace s2l {

namespace {

constexpr int kWave = 64;

__device__ inline int pow2ceil(int w)
{
    int p = 1;
    while (p < w) p <<= 1;
    return p;
}

template <typename T> __device__ inline T waveSum(T v)
{
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// ------------------------------------------------------------------ local solve: forward sweep
template <typename T>
__device__ void fwdSmall(const LocalArgs<T>& a, const SubDesc& sd, const SnDesc s, T* x, T* tb, int lane)
{
    const int w = s.w, h = s.h, c0 = s.c0;
    T acc = T(0);
    if (lane < w) {
        const int row = c0 + lane;
        T v = x[row];
        const int cb = a.cptr[sd.cptr_off + row], ce = a.cptr[sd.cptr_off + row + 1];
        for (int q = cb; q < ce; ++q) v -= tb[a.cidx[q]];
        acc = v;
    }
    const T* F = a.diag + s.diag;
    for (int c = 0; c < w; ++c) {
        const T xc = __shfl(acc, c, kWave) / F[c * w + c];
        if (lane == c)
            acc = xc;
        else if (lane > c && lane < w)
            acc -= F[c * w + lane] * xc;
    }
    if (lane < w) x[c0 + lane] = acc;
    if (h > 0) {
        const int wp = pow2ceil(w), G = kWave / wp, g = lane / wp, c = lane % wp;
        const T xc = __shfl(acc, c < w ? c : 0, kWave);
        const T* B = a.below + s.below;
        for (int kb = 0; kb < h; kb += G) {
            const int k = kb + g;
            T v = (k < h && c < w) ? B[static_cast<int64_t>(k) * w + c] * xc : T(0);
            for (int off = wp >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
            if (c == 0 && k < h) tb[s.toff + k] = v;
        }
    }
}

template <typename T>
__device__ void fwdBig(const LocalArgs<T>& a, const SubDesc& sd, const SnDesc s, T* x, T* tb, int tid, int lane, int wave,
                       int nw)
{
    const int w = s.w, h = s.h, c0 = s.c0, nt = nw * kWave;
    for (int j = tid; j < w; j += nt) {
        const int row = c0 + j;
        T v = x[row];
        const int cb = a.cptr[sd.cptr_off + row], ce = a.cptr[sd.cptr_off + row + 1];
        for (int q = cb; q < ce; ++q) v -= tb[a.cidx[q]];
        x[row] = v;
    }
    __syncthreads();
    const T* F = a.diag + s.diag;
    for (int pc = 0; pc < w; pc += kWave) {
        const int pw = min(kWave, w - pc);
        if (wave == 0) {
            T acc = lane < pw ? x[c0 + pc + lane] : T(0);
            for (int c = 0; c < pw; ++c) {
                const T* Fc = F + static_cast<int64_t>(pc + c) * w + pc;
                const T xc = __shfl(acc, c, kWave) / Fc[c];
                if (lane == c)
                    acc = xc;
                else if (lane > c && lane < pw)
                    acc -= Fc[lane] * xc;
            }
            if (lane < pw) x[c0 + pc + lane] = acc;
        }
        __syncthreads();
        for (int j = pc + pw + tid; j < w; j += nt) {
            T sum = T(0);
            for (int c = 0; c < pw; ++c) sum += F[static_cast<int64_t>(pc + c) * w + j] * x[c0 + pc + c];
            x[c0 + j] -= sum;
        }
        __syncthreads();
    }
    const T* B = a.below + s.below;
    for (int k = wave; k < h; k += nw) {
        T v = T(0);
        for (int c = lane; c < w; c += kWave) v += B[static_cast<int64_t>(k) * w + c] * x[c0 + c];
        v = waveSum(v);
        if (lane == 0) tb[s.toff + k] = v;
    }
}

// ------------------------------------------------------------------ local solve: backward sweep
template <typename T> __device__ void bwdSmall(const LocalArgs<T>& a, const SnDesc s, T* x, int lane)
{
    const int w = s.w, h = s.h, c0 = s.c0;
    const int wp = pow2ceil(w), G = kWave / wp, g = lane / wp, c = lane % wp;
    T part = T(0);
    if (c < w && h > 0) {
        const T* B = a.below + s.below;
        const int32_t* R = a.ridx + s.ridx;
        for (int k = g; k < h; k += G) part += B[static_cast<int64_t>(k) * w + c] * x[R[k]];
    }
    for (int off = wp; off < kWave; off <<= 1) part += __shfl_xor(part, off, kWave);
    T acc = (c < w) ? x[c0 + c] - part : T(0);
    const T* Gd = a.diag + s.diag;
    for (int cc = w - 1; cc >= 0; --cc) {
        const T xc = __shfl(acc, cc, kWave) / Gd[cc * w + cc];
        if (c == cc)
            acc = xc;
        else if (c < cc)
            acc -= Gd[cc * w + c] * xc;
    }
    if (lane < w) x[c0 + lane] = acc;
}

template <typename T> __device__ void bwdBig(const LocalArgs<T>& a, const SnDesc s, T* x, int tid, int lane, int wave, int nw)
{
    const int w = s.w, h = s.h, c0 = s.c0, nt = nw * kWave;
    if (h > 0) {
        const T* B = a.below + s.below;
        const int32_t* R = a.ridx + s.ridx;
        for (int c = tid; c < w; c += nt) {
            T sum = T(0);
            for (int k = 0; k < h; ++k) sum += B[static_cast<int64_t>(k) * w + c] * x[R[k]];
            x[c0 + c] -= sum;
        }
    }
    __syncthreads();
    const T* Gd = a.diag + s.diag;
    const int np = (w + kWave - 1) / kWave;
    for (int pi = np - 1; pi >= 0; --pi) {
        const int pc = pi * kWave, pw = min(kWave, w - pc);
        if (wave == 0) {
            T acc = lane < pw ? x[c0 + pc + lane] : T(0);
            for (int cc = pw - 1; cc >= 0; --cc) {
                const T* Gc = Gd + static_cast<int64_t>(pc + cc) * w + pc;
                const T xc = __shfl(acc, cc, kWave) / Gc[cc];
                if (lane == cc)
                    acc = xc;
                else if (lane < cc)
                    acc -= Gc[lane] * xc;
            }
            if (lane < pw) x[c0 + pc + lane] = acc;
        }
        __syncthreads();
        for (int j = tid; j < pc; j += nt) {
            T sum = T(0);
            for (int cc = 0; cc < pw; ++cc) sum += Gd[static_cast<int64_t>(pc + cc) * w + j] * x[c0 + pc + cc];
            x[c0 + j] -= sum;
        }
        __syncthreads();
    }
}

template <typename T> __global__ __launch_bounds__(kLocalThreads) void k_local_solve(LocalArgs<T> a)
{
    const int r = blockIdx.y;
    if (a.skip && a.skip[r]) return;
    const SubDesc sd = a.sub[a.sub_first + blockIdx.x];
    extern __shared__ unsigned char smem_raw[];
    T* x = reinterpret_cast<T*>(smem_raw);
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave, nw = blockDim.x / kWave;
    const double* rhs = a.rhs + static_cast<int64_t>(r) * a.n_global;
    T* tb = a.tbuf + static_cast<int64_t>(r) * a.t_total + sd.t_off;
    for (int p = tid; p < sd.n; p += blockDim.x) x[p] = static_cast<T>(rhs[a.gmap[sd.map_off + p]]);
    __syncthreads();
    for (int L = 0; L < sd.f_nlev; ++L) {
        const int beg = a.lev[sd.f_lev + 2 * L], split = a.lev[sd.f_lev + 2 * L + 1], end = a.lev[sd.f_lev + 2 * L + 2];
        for (int q = beg + wave; q < split; q += nw) fwdSmall(a, sd, a.sn[a.lev_sn[q]], x, tb, lane);
        for (int q = split; q < end; ++q) fwdBig(a, sd, a.sn[a.lev_sn[q]], x, tb, tid, lane, wave, nw);
        __syncthreads();
    }
    for (int L = 0; L < sd.b_nlev; ++L) {
        const int beg = a.lev[sd.b_lev + 2 * L], split = a.lev[sd.b_lev + 2 * L + 1], end = a.lev[sd.b_lev + 2 * L + 2];
        for (int q = beg + wave; q < split; q += nw) bwdSmall(a, a.sn[a.lev_sn[q]], x, lane);
        for (int q = split; q < end; ++q) bwdBig(a, a.sn[a.lev_sn[q]], x, tid, lane, wave, nw);
        __syncthreads();
    }
    if (a.mode == 0) {
        double* out = a.out + static_cast<int64_t>(r) * a.n_global;
        const double* x0 = a.x0 ? a.x0 + static_cast<int64_t>(r) * a.n_global : nullptr;
        for (int p = tid; p < sd.n; p += blockDim.x) {
            const int l = a.lperm[sd.map_off + p];
            if (l < sd.n_owned) {
                const int64_t gi = sd.begin + l;
                const double y = static_cast<double>(x[p]);
                out[gi] = x0 ? x0[gi] + y : y;
            }
        }
    } else {
        double* out = a.out + static_cast<int64_t>(r) * a.n_local_total + sd.map_off;
        for (int p = tid; p < sd.n; p += blockDim.x) {
            const int l = a.lperm[sd.map_off + p];
            const double y = static_cast<double>(x[p]);
            out[l] = a.pou ? a.pou[sd.map_off + l] * y : y;
        }
    }
}

// ------------------------------------------------------------------ SpMV: one thread per row,
// sequential sum in the stored (local) column order with fma, as DistributedMatrix::apply.
__global__ void k_spmv(int32_t n, const int32_t* rp, const int32_t* col, const double* val, const double* x, double* y,
                       const double* b, const int32_t* skip)
{
    const int r = blockIdx.y;
    if (skip && skip[r]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t off = static_cast<int64_t>(r) * n;
    const double* xr = x + off;
    double s = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) s = fma(val[k], xr[col[k]], s);
    y[off + i] = b ? b[off + i] - s : s;
}

// ------------------------------------------------------------------ coarse space
constexpr int kCq = 8; // coarse columns per workgroup of the restriction

__global__ __launch_bounds__(256) void k_coarse_restrict(CoarseArgs c, const double* r, int64_t n, double* cvec,
                                                          const int32_t* skip)
{
    const int rr = blockIdx.z;
    if (skip && skip[rr]) return;
    const int i = blockIdx.x, cq = blockIdx.y * kCq;
    const int m = static_cast<int>(c.off[i + 1] - c.off[i]);
    if (cq >= m) return;
    const int rows = c.vrows[i];
    const double* V = c.V + c.voff[i];
    const int32_t* g = c.vrow_g + c.vrow_ptr[i];
    const double* rv = r + static_cast<int64_t>(rr) * n;
    double acc[kCq];
#pragma unroll
    for (int q = 0; q < kCq; ++q) acc[q] = 0.0;
    const int nq = min(kCq, m - cq);
    for (int l = threadIdx.x; l < rows; l += blockDim.x) {
        const double xv = rv[g[l]];
#pragma unroll
        for (int q = 0; q < kCq; ++q)
            if (q < nq) acc[q] += V[static_cast<int64_t>(cq + q) * rows + l] * xv;
    }
    __shared__ double red[256 / kWave][kCq];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < kCq; ++q) {
        const double v = waveSum(acc[q]);
        if (lane == 0) red[wave][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < nq) {
        double s = 0.0;
        for (int wv = 0; wv < static_cast<int>(blockDim.x / kWave); ++wv) s += red[wv][threadIdx.x];
        cvec[static_cast<int64_t>(rr) * c.n0 + c.off[i] + cq + threadIdx.x] = s;
    }
}

__global__ __launch_bounds__(256) void k_tri_gemv(int32_t n0, const double* Tm, int lower, const int32_t* perm,
                                                   const double* x, double* y, const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int i = blockIdx.x * (blockDim.x / kWave) + wave;
    if (i >= n0) return;
    const double* xv = x + static_cast<int64_t>(rr) * n0;
    const int lo = lower ? 0 : i, hi = lower ? i + 1 : n0;
    const double* row = Tm + static_cast<int64_t>(i) * n0;
    double s = 0.0;
    for (int j = lo + lane; j < hi; j += kWave) s += row[j] * xv[perm ? perm[j] : j];
    s = waveSum(s);
    if (lane == 0) y[static_cast<int64_t>(rr) * n0 + i] = s;
}

__global__ __launch_bounds__(256) void k_coarse_prolong(CoarseArgs c, const double* y, double* out, int64_t ld_out,
                                                         int boolean, const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int b = blockIdx.x;
    const int i = c.blk_sub[b];
    const int l = c.blk_row[b] + threadIdx.x;
    const int rows = c.vrows[i];
    if (l >= rows) return;
    const int m = static_cast<int>(c.off[i + 1] - c.off[i]);
    const double* V = c.V + c.voff[i];
    const double* yv = y + static_cast<int64_t>(rr) * c.n0 + c.off[i];
    double s = 0.0;
    for (int k = 0; k < m; ++k) s += V[static_cast<int64_t>(k) * rows + l] * yv[k];
    if (boolean)
        out[static_cast<int64_t>(rr) * ld_out + c.vrow_g[c.vrow_ptr[i] + l]] = s;
    else
        out[static_cast<int64_t>(rr) * ld_out + c.vrow_ptr[i] + l] = s;
}

// out[g] = x0[g] + sum of the listed local copies in order (owner first, then the neighbours in
// increasing order), HaloExchange::reverseAdd without messages.
__global__ void k_rev_gather(int32_t n, const int32_t* ptr, const int32_t* idx, const double* src, int64_t src_ld,
                             const double* x0, double* out, const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const double* sv = src + static_cast<int64_t>(rr) * src_ld;
    double s = 0.0;
    for (int q = ptr[g]; q < ptr[g + 1]; ++q) s += sv[idx[q]];
    const int64_t o = static_cast<int64_t>(rr) * n + g;
    out[o] = x0 ? x0[o] + s : s;
}

// ------------------------------------------------------------------ vectors
__global__ void k_axpby(int64_t n, double a, const double* x, double b, double* y, const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = static_cast<int64_t>(rr) * n + i;
    y[o] = a * x[o] + (b == 0.0 ? 0.0 : b * y[o]);
}

__global__ void k_copy(int64_t n, const double* x, double* y)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t o = static_cast<int64_t>(blockIdx.y) * n + i;
    if (i < n) y[o] = x[o];
}

constexpr int kRedThreads = 256;
constexpr int kRedComp = 3;

__device__ inline bool lastBlock(unsigned* counter, unsigned nblocks)
{
    __shared__ bool last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = (atomicAdd(counter, 1u) == nblocks - 1);
    __syncthreads();
    if (last) __threadfence();
    return last;
}

__global__ __launch_bounds__(kRedThreads) void k_reduce(int kind, int64_t n, const double* x, const double* y,
                                                         double* partial, unsigned* counter, double* out,
                                                         const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const double* xv = x + static_cast<int64_t>(rr) * n;
    const double* yv = y ? y + static_cast<int64_t>(rr) * n : nullptr;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = xv[i];
        if (kind == 3) {
            a0 += yv[i] * v;
        } else {
            a0 += v * v;
            if (kind == 1 || kind == 2) a1 += fabs(v);
            if (kind == 2) a2 += v * yv[i];
            if (kind == 4) a1 += v * yv[i];
        }
    }
    __shared__ double red[kRedThreads / kWave][kRedComp];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    a0 = waveSum(a0);
    a1 = waveSum(a1);
    a2 = waveSum(a2);
    if (lane == 0) {
        red[wave][0] = a0;
        red[wave][1] = a1;
        red[wave][2] = a2;
    }
    __syncthreads();
    double* pr = partial + (static_cast<int64_t>(rr) * gridDim.x + blockIdx.x) * kRedComp;
    if (threadIdx.x < kRedComp) {
        double s = 0.0;
        for (int wv = 0; wv < kRedThreads / kWave; ++wv) s += red[wv][threadIdx.x];
        pr[threadIdx.x] = s;
    }
    if (lastBlock(counter + rr, gridDim.x)) {
        if (threadIdx.x < kRedComp) {
            const double* base = partial + static_cast<int64_t>(rr) * gridDim.x * kRedComp;
            double s = 0.0;
            for (unsigned b = 0; b < gridDim.x; ++b) s += base[b * kRedComp + threadIdx.x];
            out[rr * kRedComp + threadIdx.x] = s;
        }
        if (threadIdx.x == 0) counter[rr] = 0;
    }
}

// ------------------------------------------------------------------ GMRES
constexpr int kGmThreads = 256;

// Basis layout: V[(k * nrhs + r) * n + i], so that column k of every right-hand side is one
// contiguous n x nrhs block (the input layout of the preconditioner and the SpMV).
__device__ inline const double* basis(const GmArgs& a, int r, int k) { return a.V + (static_cast<int64_t>(k) * a.nrhs + r) * a.n; }
__device__ inline double* basisW(const GmArgs& a, int r, int k) { return a.V + (static_cast<int64_t>(k) * a.nrhs + r) * a.n; }

// h[k] = V_k . w for k <= j; one partial per block, the last block sums them in order.
__device__ void gmDotsBody(const GmArgs& a, int j, int r, const double* w, double* hout)
{
    __shared__ double red[kGmThreads / kWave][kGmMaxK];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (int k = 0; k <= j; ++k) {
        const double* v = basis(a, r, k);
        double s = 0.0;
        for (int64_t i = i0; i < a.n; i += stride) s += v[i] * w[i];
        s = waveSum(s);
        if (lane == 0) red[wave][k] = s;
    }
    __syncthreads();
    double* pr = a.partial + (static_cast<int64_t>(r) * gridDim.x + blockIdx.x) * kGmMaxK;
    for (int k = threadIdx.x; k <= j; k += blockDim.x) {
        double s = 0.0;
        for (int wv = 0; wv < kGmThreads / kWave; ++wv) s += red[wv][k];
        pr[k] = s;
    }
    if (lastBlock(a.counter + r, gridDim.x)) {
        const double* base = a.partial + static_cast<int64_t>(r) * gridDim.x * kGmMaxK;
        for (int k = threadIdx.x; k <= j; k += blockDim.x) {
            double s = 0.0;
            for (unsigned b = 0; b < gridDim.x; ++b) s += base[b * kGmMaxK + k];
            hout[k] = s;
        }
        if (threadIdx.x == 0) a.counter[r] = 0;
    }
}

__global__ __launch_bounds__(kGmThreads) void k_gm_dots(GmArgs a, int j)
{
    const int r = blockIdx.y;
    if (a.flags[r]) return;
    gmDotsBody(a, j, r, a.w + static_cast<int64_t>(r) * a.n, a.h + r * (a.m + 1));
}

// w -= V h, then h2 = V^T w.
__global__ __launch_bounds__(kGmThreads) void k_gm_update_dots(GmArgs a, int j)
{
    const int r = blockIdx.y;
    if (a.flags[r]) return;
    __shared__ double hs[kGmMaxK];
    const double* h = a.h + r * (a.m + 1);
    for (int k = threadIdx.x; k <= j; k += blockDim.x) hs[k] = h[k];
    __syncthreads();
    double* w = a.w + static_cast<int64_t>(r) * a.n;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        double s = 0.0;
        for (int k = 0; k <= j; ++k) s += basis(a, r, k)[i] * hs[k];
        w[i] -= s;
    }
    gmDotsBody(a, j, r, w, a.h2 + r * (a.m + 1));
}

// w -= V h2, hn = ||w||; the last block finishes the Arnoldi step: Hessenberg column, Givens
// rotations, residual estimate and the stopping test (gmres.cpp:149-208).
__global__ __launch_bounds__(kGmThreads) void k_gm_update_norm(GmArgs a, int j)
{
    const int r = blockIdx.y;
    if (a.flags[r]) return;
    __shared__ double hs[kGmMaxK];
    const double* h2 = a.h2 + r * (a.m + 1);
    for (int k = threadIdx.x; k <= j; k += blockDim.x) hs[k] = h2[k];
    __syncthreads();
    double* w = a.w + static_cast<int64_t>(r) * a.n;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    double nrm = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        double s = 0.0;
        for (int k = 0; k <= j; ++k) s += basis(a, r, k)[i] * hs[k];
        const double v = w[i] - s;
        w[i] = v;
        nrm += v * v;
    }
    __shared__ double red[kGmThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    nrm = waveSum(nrm);
    if (lane == 0) red[wave] = nrm;
    __syncthreads();
    double* pr = a.partial + (static_cast<int64_t>(r) * gridDim.x + blockIdx.x) * kGmMaxK;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int wv = 0; wv < kGmThreads / kWave; ++wv) s += red[wv];
        pr[0] = s;
    }
    if (!lastBlock(a.counter + r, gridDim.x)) return;
    if (threadIdx.x != 0) return;
    a.counter[r] = 0;
    const double* base = a.partial + static_cast<int64_t>(r) * gridDim.x * kGmMaxK;
    double ss = 0.0;
    for (unsigned b = 0; b < gridDim.x; ++b) ss += base[b * kGmMaxK];
    const double hn = sqrt(ss);
    const int m = a.m;
    double* H = a.H + static_cast<int64_t>(r) * (m + 1) * m;
    double* Hc = H + static_cast<int64_t>(j) * (m + 1);
    double* cs = a.cs + r * m;
    double* sn = a.sn + r * m;
    double* g = a.g + r * (m + 1);
    const double* h = a.h + r * (m + 1);
    for (int i = 0; i <= j; ++i) Hc[i] = h[i] + hs[i];
    Hc[j + 1] = hn;
    GmState& st = a.st[r];
    st.hn = hn;
    for (int i = 0; i < j; ++i) {
        const double t = cs[i] * Hc[i] + sn[i] * Hc[i + 1];
        Hc[i + 1] = -sn[i] * Hc[i] + cs[i] * Hc[i + 1];
        Hc[i] = t;
    }
    const double den = hypot(Hc[j], Hc[j + 1]);
    cs[j] = (den > 0.0) ? Hc[j] / den : 1.0;
    sn[j] = (den > 0.0) ? Hc[j + 1] / den : 0.0;
    Hc[j] = den;
    Hc[j + 1] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    st.k += 1;
    st.iters += 1;
    st.rel = fabs(g[j + 1]) / st.bnorm;
    a.hist[static_cast<int64_t>(r) * a.hist_ld + st.iters] = st.rel;
    int done = 0;
    if (st.rel <= a.rtol) {
        st.conv = 1;
        done = 1;
    } else if (hn == 0.0) {
        st.breakdown = 1;
        done = 1;
    }
    if (st.iters >= st.maxit || j + 1 >= m) done = 1;
    st.done = done;
    a.flags[r] = done;
}

__global__ void k_gm_normalize(GmArgs a, int j)
{
    const int r = blockIdx.y;
    if (a.flags[r]) return;
    const double hn = a.st[r].hn;
    if (!(hn > 0.0)) return;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    basisW(a, r, j + 1)[i] = a.w[static_cast<int64_t>(r) * a.n + i] / hn;
}

// V_0 = r / beta, g = beta e_1, H = 0 for the right-hand sides that run this cycle.
__global__ void k_gm_start(GmArgs a, const double* rv)
{
    const int r = blockIdx.y;
    if (a.flags[r]) return;
    const double beta = a.st[r].beta;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < a.n) basisW(a, r, 0)[i] = rv[static_cast<int64_t>(r) * a.n + i] / beta;
    const int m = a.m;
    if (blockIdx.x == 0) {
        double* H = a.H + static_cast<int64_t>(r) * (m + 1) * m;
        for (int q = threadIdx.x; q < (m + 1) * m; q += blockDim.x) H[q] = 0.0;
        double* g = a.g + r * (m + 1);
        for (int q = threadIdx.x; q <= m; q += blockDim.x) g[q] = (q == 0) ? beta : 0.0;
    }
}

// y = H_k^{-1} g_k (upper triangular after the rotations).
__global__ void k_gm_lsq(GmArgs a, const int32_t* idle)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.nrhs || idle[r]) return;
    const int m = a.m, k = a.st[r].k;
    const double* H = a.H + static_cast<int64_t>(r) * (m + 1) * m;
    const double* g = a.g + r * (m + 1);
    double* y = a.y + r * m;
    for (int i = k - 1; i >= 0; --i) {
        double s = g[i];
        for (int l = i + 1; l < k; ++l) s -= H[static_cast<int64_t>(l) * (m + 1) + i] * y[l];
        y[i] = s / H[static_cast<int64_t>(i) * (m + 1) + i];
    }
}

// u = V_k y
__global__ void k_gm_combine(GmArgs a, double* u, const int32_t* idle)
{
    const int r = blockIdx.y;
    if (idle[r]) return;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int k = a.st[r].k;
    const double* y = a.y + r * a.m;
    double s = 0.0;
    for (int l = 0; l < k; ++l) s += basis(a, r, l)[i] * y[l];
    u[static_cast<int64_t>(r) * a.n + i] = s;
}

inline unsigned nblk(int64_t n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

} // namespace

// ------------------------------------------------------------------ launchers
template <typename T> void launchLocalSolve(const LocalArgs<T>& a, int nsub, int nrhs, size_t lds, hipStream_t s)
{
    hipLaunchKernelGGL(k_local_solve<T>, dim3(nsub, nrhs), dim3(kLocalThreads), lds, s, a);
}
template void launchLocalSolve<float>(const LocalArgs<float>&, int, int, size_t, hipStream_t);
template void launchLocalSolve<double>(const LocalArgs<double>&, int, int, size_t, hipStream_t);

template <typename T> const void* localSolveKernel() { return reinterpret_cast<const void*>(&k_local_solve<T>); }
template const void* localSolveKernel<float>();
template const void* localSolveKernel<double>();

void launchSpmv(int32_t n, const int32_t* rp, const int32_t* col, const double* val, const double* x, double* y,
                const double* b, int nrhs, const int32_t* skip, hipStream_t s)
{
    hipLaunchKernelGGL(k_spmv, dim3(nblk(n, 256), nrhs), dim3(256), 0, s, n, rp, col, val, x, y, b, skip);
}

void launchCoarseRestrict(const CoarseArgs& c, const double* r, int64_t n, double* cvec, int nrhs, const int32_t* skip,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_coarse_restrict, dim3(c.nsub, (c.mmax + kCq - 1) / kCq, nrhs), dim3(256), 0, s, c, r, n, cvec, skip);
}

void launchTriGemv(int32_t n0, const double* T, int lower, const int32_t* perm, const double* x, double* y, int nrhs,
                   const int32_t* skip, hipStream_t s)
{
    hipLaunchKernelGGL(k_tri_gemv, dim3(nblk(n0, 4), nrhs), dim3(256), 0, s, n0, T, lower, perm, x, y, skip);
}

void launchCoarseProlong(const CoarseArgs& c, const double* y, double* out, int64_t ld_out, int boolean, int nrhs,
                         const int32_t* skip, hipStream_t s)
{
    hipLaunchKernelGGL(k_coarse_prolong, dim3(c.nblk, nrhs), dim3(256), 0, s, c, y, out, ld_out, boolean, skip);
}

void launchRevGather(int32_t n, const int32_t* ptr, const int32_t* idx, const double* src, int64_t src_ld,
                     const double* x0, double* out, int nrhs, const int32_t* skip, hipStream_t s)
{
    hipLaunchKernelGGL(k_rev_gather, dim3(nblk(n, 256), nrhs), dim3(256), 0, s, n, ptr, idx, src, src_ld, x0, out, skip);
}

void launchAxpby(int64_t n, double a, const double* x, double b, double* y, int nrhs, const int32_t* skip, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpby, dim3(nblk(n, 256), nrhs), dim3(256), 0, s, n, a, x, b, y, skip);
}

void launchCopy(int64_t n, const double* x, double* y, int nrhs, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy, dim3(nblk(n, 256), nrhs), dim3(256), 0, s, n, x, y);
}

int reduceBlocks(int64_t n)
{
    const int64_t b = (n + kRedThreads * 4 - 1) / (kRedThreads * 4);
    return static_cast<int>(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

void launchReduce(int kind, int64_t n, const double* x, const double* y, double* partial, unsigned* counter, double* out,
                  int nrhs, const int32_t* skip, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce, dim3(reduceBlocks(n), nrhs), dim3(kRedThreads), 0, s, kind, n, x, y, partial, counter, out,
                       skip);
}

void gmDots(const GmArgs& a, int j, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_dots, dim3(a.nblocks, a.nrhs), dim3(kGmThreads), 0, s, a, j);
}
void gmUpdateDots(const GmArgs& a, int j, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_update_dots, dim3(a.nblocks, a.nrhs), dim3(kGmThreads), 0, s, a, j);
}
void gmUpdateNorm(const GmArgs& a, int j, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_update_norm, dim3(a.nblocks, a.nrhs), dim3(kGmThreads), 0, s, a, j);
}
void gmNormalize(const GmArgs& a, int j, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_normalize, dim3(nblk(a.n, 256), a.nrhs), dim3(256), 0, s, a, j);
}
void gmStartCycle(const GmArgs& a, const double* r, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_start, dim3(nblk(a.n, 256), a.nrhs), dim3(256), 0, s, a, r);
}
void gmLsq(const GmArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_lsq, dim3(nblk(a.nrhs, 64)), dim3(64), 0, s, a, a.flags + a.nrhs);
}
void gmCombine(const GmArgs& a, double* u, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_combine, dim3(nblk(a.n, 256), a.nrhs), dim3(256), 0, s, a, u, a.flags + a.nrhs);
}

} // namespace s2l
