// refactor.hip -- device refactorization kernels (see refactor.hpp).
#include "refactor.hpp"

namespace s2l {

namespace {

constexpr int kRfThreads = 256;

// Loads issued together per batch of the refactorization's sums (one memory round trip each).
constexpr int kRb = 8;

// One supernode's panel per workgroup (a height of the elimination tree per launch, so every
// descendant is final): A's entries, the descendants' updates in increasing descendant order
// (each a block of dot products over the descendant's columns), then a dense right-looking
// Cholesky of the panel [L_ss; L_rs]. Both blocks are column-major (L_ss w x w, L_rs h x w), so
// consecutive threads on consecutive rows read consecutive addresses.
__global__ __launch_bounds__(kRfThreads) void k_rf_level(RfArgs a, const int32_t* order)
{
    const int s = order[blockIdx.x];
    const RfSn r = a.sn[s];
    const int w = r.w, h = r.h, tid = threadIdx.x, nt = blockDim.x;
    double* D = a.work + r.wd;  // L_ss(q, c) at c w + q
    double* B = a.work + r.wb;  // L_rs(k, c) at c h + k
    const int64_t pan = static_cast<int64_t>(w) * w + static_cast<int64_t>(h) * w;
    for (int64_t q = tid; q < pan; q += nt) D[q] = 0.0;  // D and B are contiguous
    __syncthreads();
    for (int64_t e = tid; e < r.amap_count; e += nt) a.work[a.amap_pos[r.amap_first + e]] = a.val[a.amap_src[r.amap_first + e]];
    __syncthreads();
    if (a.diag_adj) {  // a shifted matrix with the same structure (the local eigenproblems' A~_ii + sigma I)
        const int32_t off = a.sub_off[r.sub];
        for (int c = tid; c < w; c += nt) D[static_cast<int64_t>(c) * w + c] += a.diag_adj[off + a.lperm[off + r.c0 + c]];
        __syncthreads();
    }
    for (int u = 0; u < r.upd_count; ++u) {
        const RfUpd U = a.upd[r.upd_first + u];
        const RfSn dn = a.sn[U.d];
        const double* Ld = a.work + dn.wb;  // rows below d: L_d(i, k) at k h_d + i
        const int32_t* Rd = a.ridx + dn.ridx;
        const int32_t* rel = a.rel + U.rel;
        const int wdd = dn.w, hd = dn.h;
        for (int j = U.i0; j < U.i1; ++j) {
            const int pc = Rd[j] - r.c0;
            for (int i = j + tid; i < hd; i += nt) {
                double v = 0.0;
                int k = 0;
                for (; k + kRb <= wdd; k += kRb) {
                    double x[kRb], y[kRb];
#pragma unroll
                    for (int q = 0; q < kRb; ++q) {
                        x[q] = Ld[static_cast<int64_t>(k + q) * hd + i];
                        y[q] = Ld[static_cast<int64_t>(k + q) * hd + j];
                    }
#pragma unroll
                    for (int q = 0; q < kRb; ++q) v += x[q] * y[q];
                }
                for (; k < wdd; ++k) v += Ld[static_cast<int64_t>(k) * hd + i] * Ld[static_cast<int64_t>(k) * hd + j];
                const int pr = rel[i - U.i0];
                if (pr < w)
                    D[static_cast<int64_t>(pc) * w + pr] -= v;
                else
                    B[static_cast<int64_t>(pc) * h + (pr - w)] -= v;
            }
        }
        __syncthreads();
    }
    __shared__ double piv;
    for (int j = 0; j < w; ++j) {
        if (tid == 0) {
            const double dj = D[static_cast<int64_t>(j) * w + j];
            const bool ok = dj > 0.0 && isfinite(dj);
            if (!ok) a.status[0] = 1 + s;
            piv = sqrt(ok ? dj : 1.0);
            D[static_cast<int64_t>(j) * w + j] = piv;
        }
        __syncthreads();
        const double p = piv;
        const int nd = w - j - 1;
        for (int t = tid; t < nd + h; t += nt) {
            if (t < nd)
                D[static_cast<int64_t>(j) * w + j + 1 + t] /= p;
            else
                B[static_cast<int64_t>(j) * h + (t - nd)] /= p;
        }
        __syncthreads();
        // trailing update, a thread per row q (rows j+1 .. w-1 of L_ss, then the rows of L_rs):
        // X(q, c) -= L(q, j) L(c, j) for j < c (c <= q in L_ss), eight columns per batch of loads
        for (int t = tid; t < nd + h; t += nt) {
            const bool inD = t < nd;
            const int q = inD ? j + 1 + t : t - nd;
            double* col0 = inD ? D + q : B + q;              // X(q, c) at col0[c * ld]
            const int64_t ld = inD ? w : h;
            const double lq = inD ? D[static_cast<int64_t>(j) * w + q] : B[static_cast<int64_t>(j) * h + q];
            const int cend = inD ? q + 1 : w;                // columns j+1 .. cend-1
            const double* Lj = D + static_cast<int64_t>(j) * w;  // L(c, j) at Lj[c]
            int c = j + 1;
            for (; c + kRb <= cend; c += kRb) {
                double x[kRb], l[kRb];
#pragma unroll
                for (int u = 0; u < kRb; ++u) {
                    x[u] = col0[static_cast<int64_t>(c + u) * ld];
                    l[u] = Lj[c + u];
                }
#pragma unroll
                for (int u = 0; u < kRb; ++u) col0[static_cast<int64_t>(c + u) * ld] = x[u] - lq * l[u];
            }
            for (; c < cend; ++c) col0[static_cast<int64_t>(c) * ld] -= lq * Lj[c];
        }
        __syncthreads();
    }
}

// Inverse X of L_ss row by row (row c of L X = I: X(c, j) = (delta_cj - sum_{j <= m < c} L(c, m)
// X(m, j)) / L(c, c), every column j at once, a barrier per row), then the solve layout: a small
// supernode's inverses column-major w x w (forward X, backward X^T), a wide one's row-major with the
// padded leading dimension; the block below row-major.
template <typename TF> __global__ __launch_bounds__(kRfThreads) void k_rf_pack(RfArgs a, TF* diag, TF* below)
{
    const int s = blockIdx.x;
    const RfSn r = a.sn[s];
    const int w = r.w, h = r.h, tid = threadIdx.x, nt = blockDim.x;
    const double* L = a.work + r.wd;   // L(c, m) at m w + c
    double* X = a.inv + a.inv_off[s];  // row-major: X(c, j) at c w + j
    for (int c = 0; c < w; ++c) {
        const double dc = L[static_cast<int64_t>(c) * w + c];
        for (int j = tid; j < w; j += nt) {
            double v = 0.0;
            if (j <= c) {
                double acc = j == c ? 1.0 : 0.0;
                for (int m = j; m < c; ++m) acc -= L[static_cast<int64_t>(m) * w + c] * X[static_cast<int64_t>(m) * w + j];
                v = acc / dc;
            }
            X[static_cast<int64_t>(c) * w + j] = v;
        }
        __syncthreads();
    }
    const bool wide = w > kSmallWidth;
    const int ld = wide ? wideLd(w) : w;
    for (int64_t e = tid; e < static_cast<int64_t>(w) * w; e += nt) {
        const int q = static_cast<int>(e / w), c = static_cast<int>(e - static_cast<int64_t>(q) * w);  // X(q, c) = X[q w + c]
        const double v = X[e];
        if (!wide) {
            diag[r.fdiag + static_cast<int64_t>(c) * w + q] = static_cast<TF>(v);   // forward: X(q, c) at c w + q
            diag[r.bdiag + static_cast<int64_t>(q) * w + c] = static_cast<TF>(v);   // backward: X^T(c, q) at q w + c
        } else {
            diag[r.fdiag + static_cast<int64_t>(q) * ld + c] = static_cast<TF>(v);  // row q, column c
            diag[r.bdiag + static_cast<int64_t>(c) * ld + q] = static_cast<TF>(v);  // X^T: row c, column q
        }
    }
    const double* Bw = a.work + r.wb;  // L_rs(k, c) at c h + k
    for (int64_t e = tid; e < static_cast<int64_t>(h) * w; e += nt) {
        const int k = static_cast<int>(e / w), c = static_cast<int>(e - static_cast<int64_t>(k) * w);
        below[r.below + static_cast<int64_t>(k) * ld + c] = static_cast<TF>(Bw[static_cast<int64_t>(c) * h + k]);
    }
}

// L_SS of dense top blockIdx.x: its supernodes' diagonal blocks and rows below (all inside the
// skeleton) at the skeleton rows dtop_q gives their factor positions.
__global__ __launch_bounds__(kRfThreads) void k_dt_assemble(RfArgs a, const DenseTop* dt, const int32_t* dt_rf, const int32_t* dtop_q,
                                                             const int64_t* off, double* Lall)
{
    const DenseTop t = dt[blockIdx.x];
    const int S = t.S, tid = threadIdx.x, nt = blockDim.x;
    double* L = Lall + off[blockIdx.x];  // L(q, c) at c S + q
    const int32_t* qOf = dtop_q + t.q_off;
    for (int64_t e = tid; e < static_cast<int64_t>(S) * S; e += nt) L[e] = 0.0;
    __syncthreads();
    for (int k = 0; k < t.sn_count; ++k) {
        const RfSn r = a.sn[dt_rf[t.sn_first + k]];
        const int w = r.w, h = r.h, q0 = qOf[r.c0];
        const double* D = a.work + r.wd;  // L_ss(q, c) at c w + q
        const double* B = a.work + r.wb;  // L_rs(k, c) at c h + k
        const int32_t* R = a.ridx + r.ridx;
        for (int64_t e = tid; e < static_cast<int64_t>(w) * (w + h); e += nt) {
            const int c = static_cast<int>(e / (w + h)), q = static_cast<int>(e - static_cast<int64_t>(c) * (w + h));
            if (q < w) {
                if (q >= c) L[static_cast<int64_t>(q0 + c) * S + q0 + q] = D[static_cast<int64_t>(c) * w + q];
            } else {
                const int kk = q - w;
                L[static_cast<int64_t>(q0 + c) * S + qOf[R[kk]]] = B[static_cast<int64_t>(c) * h + kk];
            }
        }
    }
}

// S^-1 (column-major, symmetric blocks of C) into the solve layout of every dense top.
template <typename TF> __global__ void k_dt_layout(const DenseTop* dt, const int64_t* off, const double* C, TF* diag)
{
    const DenseTop t = dt[blockIdx.y];
    const int S = t.S, ld = wideLd(S);
    const double* Ct = C + off[blockIdx.y];
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < static_cast<int64_t>(S) * ld;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int i = static_cast<int>(e / ld), j = static_cast<int>(e - static_cast<int64_t>(i) * ld);
        diag[t.sinv + e] = j < S ? static_cast<TF>(Ct[static_cast<int64_t>(j) * S + i]) : TF(0);
    }
}

__global__ void k_unit_columns(double* Y, int32_t n0, int32_t k0, int32_t cnt)
{
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<int64_t>(n0) * cnt) return;
    const int c = static_cast<int>(e / n0), q = static_cast<int>(e - static_cast<int64_t>(c) * n0);
    Y[e] = q == k0 + c ? 1.0 : 0.0;
}

__global__ void k_store_columns(const double* C, int32_t n0, int32_t k0, int32_t cnt, double* E, int32_t ld)
{
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<int64_t>(n0) * cnt) return;
    const int c = static_cast<int>(e / n0), q = static_cast<int>(e - static_cast<int64_t>(c) * n0);
    E[static_cast<int64_t>(k0 + c) * ld + q] = C[e];  // column-major E: column k0 + c
}

// T1 (row-major, ld) = the lower triangle of Linv (column-major, n x n), T2 = T1^T (upper).
__global__ void k_tri_layout(const double* Linv, int32_t n, int32_t ld, double* T1, double* T2)
{
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<int64_t>(n) * ld) return;
    const int row = static_cast<int>(e / ld), c = static_cast<int>(e - static_cast<int64_t>(row) * ld);
    T1[e] = (c < n && c <= row) ? Linv[static_cast<int64_t>(c) * n + row] : 0.0;  // Linv(row, c)
    T2[e] = (c < n && c >= row) ? Linv[static_cast<int64_t>(row) * n + c] : 0.0;  // Linv(c, row)
}

__global__ void k_gather(const double* src, const int32_t* idx, double* dst, int64_t n)
{
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = src[idx[q]];
}

inline unsigned blocks(int64_t n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

} // namespace

void launchRfLevel(const RfArgs& a, const int32_t* order, int32_t count, hipStream_t s)
{
    if (count > 0) hipLaunchKernelGGL(k_rf_level, dim3(count), dim3(kRfThreads), 0, s, a, order);
}

void launchRfPack(const RfArgs& a, int32_t nsn, void* diag, void* below, bool fp32, hipStream_t s)
{
    if (nsn <= 0) return;
    if (fp32)
        hipLaunchKernelGGL(k_rf_pack<float>, dim3(nsn), dim3(kRfThreads), 0, s, a, static_cast<float*>(diag), static_cast<float*>(below));
    else
        hipLaunchKernelGGL(k_rf_pack<double>, dim3(nsn), dim3(kRfThreads), 0, s, a, static_cast<double*>(diag),
                           static_cast<double*>(below));
}

void launchDtAssemble(const RfArgs& a, const DenseTop* dt, int32_t ndt, const int32_t* dt_rf, const int32_t* dtop_q,
                      const int64_t* off, double* L, hipStream_t s)
{
    if (ndt > 0) hipLaunchKernelGGL(k_dt_assemble, dim3(ndt), dim3(kRfThreads), 0, s, a, dt, dt_rf, dtop_q, off, L);
}

void launchDtLayout(const DenseTop* dt, int32_t ndt, const int64_t* off, const double* C, void* diag, bool fp32, hipStream_t s)
{
    if (ndt <= 0) return;
    const dim3 grid(64, static_cast<unsigned>(ndt));
    if (fp32)
        hipLaunchKernelGGL(k_dt_layout<float>, grid, dim3(256), 0, s, dt, off, C, static_cast<float*>(diag));
    else
        hipLaunchKernelGGL(k_dt_layout<double>, grid, dim3(256), 0, s, dt, off, C, static_cast<double*>(diag));
}

void launchGather(const double* src, const int32_t* idx, double* dst, int64_t n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_gather, dim3(blocks(n, 256)), dim3(256), 0, s, src, idx, dst, n);
}

void launchUnitColumns(double* Y, int32_t n0, int32_t k0, int32_t cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_unit_columns, dim3(blocks(static_cast<int64_t>(n0) * cnt, 256)), dim3(256), 0, s, Y, n0, k0, cnt);
}

void launchStoreColumns(const double* C, int32_t n0, int32_t k0, int32_t cnt, double* E, int32_t ld, hipStream_t s)
{
    hipLaunchKernelGGL(k_store_columns, dim3(blocks(static_cast<int64_t>(n0) * cnt, 256)), dim3(256), 0, s, C, n0, k0, cnt, E, ld);
}

void launchTriLayout(const double* Linv, int32_t n, int32_t ld, double* T1, double* T2, hipStream_t s)
{
    hipLaunchKernelGGL(k_tri_layout, dim3(blocks(static_cast<int64_t>(n) * ld, 256)), dim3(256), 0, s, Linv, n, ld, T1, T2);
}

} // namespace s2l
