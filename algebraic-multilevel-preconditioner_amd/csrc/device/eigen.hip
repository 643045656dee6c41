// eigen.hip -- kernels of the device local eigensolver (see eigen.hpp).
#include "eigen.hpp"

namespace s2l {

namespace {

constexpr int kEigThreads = 256;
constexpr int kEigLd = kEigMaxM + 1;  // per-chunk / per-subdomain stride of the dot products
constexpr int kRotRows = 64;          // rows per workgroup of the restart rotation

inline int blocks(int64_t n, int t) { return static_cast<int>((n + t - 1) / t); }

__device__ inline double waveSum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// adj = -s on the ghost rows (s = sum of |A(l, k)| over the columns k outside the subdomain,
// local_block_splitting.cpp:9-21), 0 on the owned rows; rowsum = sum_k |A~(l, k)|.
__global__ void k_eig_lump(EigArgs a, double* rowsum)
{
    const int64_t l = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= a.nl) return;
    double s = 0.0;
    if (a.ghost[l])
        for (int32_t q = a.orp[l]; q < a.orp[l + 1]; ++q) s += fabs(a.val[a.osrc[q]]);
    const double ad = -s;
    a.adj[l] = ad;
    double r = 0.0;
    bool diag = false;
    for (int32_t q = a.lrp[l]; q < a.lrp[l + 1]; ++q) {
        const double v = a.val[a.lsrc[q]];
        if (a.lcol[q] == l) {
            r += fabs(v + ad);
            diag = true;
        } else {
            r += fabs(v);
        }
    }
    if (!diag) r += fabs(ad);
    rowsum[l] = r;
}

// sigma_i = shift_rel * max_l rowsum (the norm estimate sqrt(|A~|_1 |A~|_inf) of a symmetric
// block, local_eigensolver.cpp:33-43 and :289), 1 in place of a zero norm
__global__ void k_eig_shift(EigArgs a, const double* rowsum, double shift_rel, double* sigma)
{
    const int i = blockIdx.x;
    __shared__ double red[kEigThreads / 64];
    double mx = 0.0;
    for (int64_t l = a.lptr[i] + threadIdx.x; l < a.lptr[i + 1]; l += blockDim.x) mx = fmax(mx, rowsum[l]);
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < kEigThreads / 64; ++w) m = fmax(m, red[w]);
        sigma[i] = shift_rel * (m > 0.0 ? m : 1.0);
    }
}

__global__ void k_eig_add_shift(EigArgs a, const double* sigma, double factor)
{
    const int32_t* ch = a.chunk + 3 * blockIdx.x;
    const double sg = factor * sigma[ch[0]];
    for (int64_t l = ch[1] + threadIdx.x; l < ch[2]; l += blockDim.x) a.adj[l] += sg;
}

__global__ void k_eig_spmv(EigArgs a, int b, const double* x, double* y)
{
    const int64_t l = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= a.nl) return;
    double acc = 0.0;
    if (b) {
        for (int32_t q = a.lrp[l]; q < a.lrp[l + 1]; ++q) {
            const int32_t k = a.lcol[q];
            acc += a.val[a.lsrc[q]] * (a.D[k] * x[k]);
        }
        y[l] = a.D[l] * acc;
    } else {
        for (int32_t q = a.lrp[l]; q < a.lrp[l + 1]; ++q) acc += a.val[a.lsrc[q]] * x[a.lcol[q]];
        y[l] = acc + a.adj[l] * x[l];
    }
}

// splitmix64 of (subdomain, row): a start vector that depends on neither the launch shape nor the
// other subdomains
__global__ void k_eig_start(EigArgs a, double* v)
{
    const int32_t* ch = a.chunk + 3 * blockIdx.x;
    const int sub = ch[0];
    for (int64_t l = ch[1] + threadIdx.x; l < ch[2]; l += blockDim.x) {
        uint64_t z = (static_cast<uint64_t>(sub) << 32) + static_cast<uint64_t>(l - a.lptr[sub]) + 0x5EEDull;
        z += 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        v[l] = 2.0 * (static_cast<double>(z >> 11) * 0x1.0p-53) - 1.0;
    }
}

// a wave per column (columns c = wave, wave + 4, ...): lanes over the chunk's rows
__global__ __launch_bounds__(kEigThreads) void k_eig_dots(EigArgs a, const double* X, int ncols, const double* w,
                                                         double* part)
{
    const int32_t* ch = a.chunk + 3 * blockIdx.x;
    if (a.done[ch[0]]) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = wave; c < ncols; c += kEigThreads / 64) {
        const double* xc = X + static_cast<int64_t>(c) * a.nl;
        double acc = 0.0;
        for (int64_t l = ch[1] + lane; l < ch[2]; l += 64) acc += xc[l] * w[l];
        acc = waveSum(acc);
        if (lane == 0) part[static_cast<int64_t>(blockIdx.x) * kEigLd + c] = acc;
    }
}

__global__ void k_eig_dots_final(EigArgs a, const double* part, int ncols, int mode, double* hb, double* hb2, double* H,
                                 int j, double* beta)
{
    const int i = blockIdx.x, c = threadIdx.x;
    if (a.done[i] || c >= ncols) return;
    double s = 0.0;
    for (int32_t k = a.chunk_ptr[i]; k < a.chunk_ptr[i + 1]; ++k) s += part[static_cast<int64_t>(k) * kEigLd + c];
    const int64_t o = static_cast<int64_t>(i) * kEigLd + c;
    if (mode == 0) {
        hb[o] = s;
    } else if (mode == 1) {
        hb2[o] = s;
        H[(static_cast<int64_t>(i) * a.m + j) * a.m + c] = hb[o] + s;
    } else if (c == 0) {
        beta[i] = sqrt(fmax(s, 0.0));
    }
}

__global__ __launch_bounds__(kEigThreads) void k_eig_update(EigArgs a, const double* V, int ncols, const double* h,
                                                           double* w)
{
    const int32_t* ch = a.chunk + 3 * blockIdx.x;
    const int sub = ch[0];
    if (a.done[sub]) return;
    __shared__ double hs[kEigLd];
    for (int c = threadIdx.x; c < ncols; c += blockDim.x) hs[c] = h[static_cast<int64_t>(sub) * kEigLd + c];
    __syncthreads();
    for (int64_t l = ch[1] + threadIdx.x; l < ch[2]; l += blockDim.x) {
        double acc = 0.0;
        for (int c = 0; c < ncols; ++c) acc += V[static_cast<int64_t>(c) * a.nl + l] * hs[c];
        w[l] -= acc;
    }
}

__global__ void k_eig_scale(EigArgs a, double* V, double* AV, int col, const double* w, const double* Aw, const double* beta)
{
    const int32_t* ch = a.chunk + 3 * blockIdx.x;
    const int sub = ch[0];
    if (a.done[sub]) return;
    const double b = beta[sub];
    const double inv = b > 0.0 ? 1.0 / b : 0.0;
    for (int64_t l = ch[1] + threadIdx.x; l < ch[2]; l += blockDim.x) {
        V[static_cast<int64_t>(col) * a.nl + l] = w[l] * inv;
        AV[static_cast<int64_t>(col) * a.nl + l] = Aw[l] * inv;
    }
}

__global__ void k_eig_check(EigArgs a, const double* S, const double* theta, const double* beta, int nev, double tol,
                            int32_t* conv, double* lam, int32_t* live)
{
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int m = a.m;
    for (int i = threadIdx.x; i < a.nsub; i += blockDim.x) {
        if (a.done[i]) continue;
        bool ok = true;
        for (int c = 0; c < nev; ++c) {
            const int k = m - 1 - c;
            const double th = theta[static_cast<int64_t>(i) * m + k];
            const double r = fabs(beta[i] * S[(static_cast<int64_t>(i) * m + k) * m + (m - 1)]);
            lam[static_cast<int64_t>(i) * nev + c] = th;
            ok = ok && r <= tol * fabs(th);
        }
        conv[i] = ok ? 1 : 0;
        if (!ok) atomicAdd(&cnt, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) live[0] = cnt;
}

// 64 rows of one chunk per workgroup: the rows of V (and then AV) in LDS, the kk leading Ritz
// vectors of the projected matrix in LDS, the rotated rows written back in place
__global__ __launch_bounds__(kEigThreads) void k_eig_rotate(EigArgs a, double* V, double* AV, const double* S, int kk)
{
    extern __shared__ double sm[];
    const int32_t* ch = a.chunk + 3 * blockIdx.x;
    const int sub = ch[0];
    if (a.done[sub]) return;
    const int64_t r0 = ch[1] + static_cast<int64_t>(blockIdx.y) * kRotRows;
    if (r0 >= ch[2]) return;
    const int nr = static_cast<int>(min<int64_t>(kRotRows, ch[2] - r0));
    const int m = a.m, ldv = m + 1;
    double* St = sm;                   // St[k * kk + c] = S(k, m - 1 - c)
    double* Vt = sm + m * kk;          // Vt[r * ldv + k]
    const double* Si = S + static_cast<int64_t>(sub) * m * m;
    for (int e = threadIdx.x; e < m * kk; e += blockDim.x) {
        const int k = e / kk, c = e - k * kk;
        St[e] = Si[static_cast<int64_t>(m - 1 - c) * m + k];
    }
    const int r = threadIdx.x & (kRotRows - 1), g = threadIdx.x / kRotRows;
    for (int pass = 0; pass < 2; ++pass) {
        double* X = pass == 0 ? V : AV;
        __syncthreads();
        for (int k = g; k < m; k += kEigThreads / kRotRows)
            if (r < nr) Vt[r * ldv + k] = X[static_cast<int64_t>(k) * a.nl + r0 + r];
        __syncthreads();
        if (r < nr) {
            for (int c = g; c < kk; c += kEigThreads / kRotRows) {
                double acc = 0.0;
                for (int k = 0; k < m; ++k) acc += Vt[r * ldv + k] * St[k * kk + c];
                X[static_cast<int64_t>(c) * a.nl + r0 + r] = acc;
            }
            if (g == 0) X[static_cast<int64_t>(kk) * a.nl + r0 + r] = X[static_cast<int64_t>(m) * a.nl + r0 + r];
        }
    }
}

__global__ void k_eig_reset(EigArgs a, double* H, const double* theta, int kk, const int32_t* conv, int32_t* done)
{
    const int i = blockIdx.x;
    if (done[i]) return;
    const int m = a.m;
    double* Hi = H + static_cast<int64_t>(i) * m * m;
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
        const int col = e / m, row = e - col * m;
        Hi[e] = (row == col && col < kk) ? theta[static_cast<int64_t>(i) * m + (m - 1 - col)] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x == 0 && conv[i]) done[i] = 1;
}

} // namespace

void launchEigLump(const EigArgs& a, double* rowsum, hipStream_t s)
{
    hipLaunchKernelGGL(k_eig_lump, dim3(blocks(a.nl, 256)), dim3(256), 0, s, a, rowsum);
}

void launchEigShift(const EigArgs& a, const double* rowsum, double shift_rel, double* sigma, hipStream_t s)
{
    hipLaunchKernelGGL(k_eig_shift, dim3(a.nsub), dim3(kEigThreads), 0, s, a, rowsum, shift_rel, sigma);
}

void launchEigAddShift(const EigArgs& a, const double* sigma, double factor, hipStream_t s)
{
    hipLaunchKernelGGL(k_eig_add_shift, dim3(a.nchunk), dim3(kEigThreads), 0, s, a, sigma, factor);
}

void launchEigSpmv(const EigArgs& a, int b, const double* x, double* y, hipStream_t s)
{
    hipLaunchKernelGGL(k_eig_spmv, dim3(blocks(a.nl, 256)), dim3(256), 0, s, a, b, x, y);
}

void launchEigStart(const EigArgs& a, double* v, hipStream_t s)
{
    hipLaunchKernelGGL(k_eig_start, dim3(a.nchunk), dim3(kEigThreads), 0, s, a, v);
}

void launchEigDots(const EigArgs& a, const double* X, int ncols, const double* w, double* part, hipStream_t s)
{
    hipLaunchKernelGGL(k_eig_dots, dim3(a.nchunk), dim3(kEigThreads), 0, s, a, X, ncols, w, part);
}

void launchEigDotsFinal(const EigArgs& a, const double* part, int ncols, int mode, double* hb, double* hb2, double* H, int j,
                        double* beta, hipStream_t s)
{
    hipLaunchKernelGGL(k_eig_dots_final, dim3(a.nsub), dim3(128), 0, s, a, part, ncols, mode, hb, hb2, H, j, beta);
}

void launchEigUpdate(const EigArgs& a, const double* V, int ncols, const double* h, double* w, hipStream_t s)
{
    hipLaunchKernelGGL(k_eig_update, dim3(a.nchunk), dim3(kEigThreads), 0, s, a, V, ncols, h, w);
}

void launchEigScale(const EigArgs& a, double* V, double* AV, int col, const double* w, const double* Aw, const double* beta,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_eig_scale, dim3(a.nchunk), dim3(kEigThreads), 0, s, a, V, AV, col, w, Aw, beta);
}

void launchEigCheck(const EigArgs& a, const double* S, const double* theta, const double* beta, int nev, double tol,
                    int32_t* conv, double* lam, int32_t* live, hipStream_t s)
{
    hipLaunchKernelGGL(k_eig_check, dim3(1), dim3(256), 0, s, a, S, theta, beta, nev, tol, conv, lam, live);
}

void launchEigRotate(const EigArgs& a, double* V, double* AV, const double* S, int kk, hipStream_t s)
{
    const size_t lds = (static_cast<size_t>(a.m) * kk + static_cast<size_t>(kRotRows) * (a.m + 1)) * sizeof(double);
    static bool attr = false;  // the largest basis needs more than the default 64 KiB of LDS
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_eig_rotate), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>((static_cast<size_t>(kEigMaxM) * kEigMaxM + static_cast<size_t>(kRotRows) * (kEigMaxM + 1)) * sizeof(double)));
        attr = true;
    }
    hipLaunchKernelGGL(k_eig_rotate, dim3(a.nchunk, kEigChunk / kRotRows), dim3(kEigThreads), lds, s, a, V, AV, S, kk);
}

void launchEigReset(const EigArgs& a, double* H, const double* theta, int kk, const int32_t* conv, int32_t* done,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_eig_reset, dim3(a.nsub), dim3(256), 0, s, a, H, theta, kk, conv, done);
}

} // namespace s2l
