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
constexpr int kPullSplit = 4;  // a backward chunk of at most blockDim / kPullSplit columns splits its rows below
constexpr int kMaxGroups = 32; // row groups of such a split

// A pair of right-hand sides solved together: T = float2 (FP32 arithmetic) or double2 (FP64
// arithmetic of FP32 entries) holds right-hand sides 2 r and 2 r + 1 in its two components.
template <typename T> struct PairOf {
    static constexpr bool value = false;
};
template <> struct PairOf<float2> {
    static constexpr bool value = true;
    using E = float;
};
template <> struct PairOf<double2> {
    static constexpr bool value = true;
    using E = double;
};

// Lane exchange of a value; a pair moves as two.
template <typename T> __device__ inline T shflXor(T v, int m) { return __shfl_xor(v, m, kWave); }
template <> __device__ inline float2 shflXor<float2>(float2 v, int m)
{
    return float2(__shfl_xor(v.x, m, kWave), __shfl_xor(v.y, m, kWave));
}
template <> __device__ inline double2 shflXor<double2>(double2 v, int m)
{
    return double2(__shfl_xor(v.x, m, kWave), __shfl_xor(v.y, m, kWave));
}

// A factor entry in the arithmetic type; for a pair of right-hand sides the same entry multiplies
// both (float2's one-argument constructor would leave the second component zero).
template <typename T, typename P> __device__ inline T widen(P v) { return static_cast<T>(v); }
template <> __device__ inline float2 widen<float2, float>(float v) { return float2(v, v); }
template <> __device__ inline double2 widen<double2, float>(float v)
{
    const double d = static_cast<double>(v);
    return double2(d, d);
}

template <typename T> __device__ inline T waveSum(T v)
{
    for (int off = kWave / 2; off > 0; off >>= 1) v += shflXor(v, off);
    return v;
}

// Global loads issued per batch by the dot products of the local solve: each batch is one memory
// round trip, so a thread's dot product of length n costs ceil(n / 8) dependent round trips. The
// remainder is one batch rather than a loop with a round trip per element: its loads past the end
// read the last element again (no branch around a load: the compiler would move a conversion of the
// loaded value into the branch and wait for each load there) and their products are skipped.
constexpr int kB = 8;

// sum over k < n of p[k * ps] * q[k * qs], in order k = 0, 1, ..., with the loads of a batch issued
// together ahead of their products: the scheduling barrier keeps the compiler from pulling the
// loads down next to their uses. p (factor storage, type P) is converted to the arithmetic type T.
template <typename T, typename P> __device__ inline T dot8(const P* p, int ps, const T* q, int qs, int n, T acc)
{
    int k = 0;
    for (; k + kB <= n; k += kB) {
        P a[kB];
        T b[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) a[u] = p[(k + u) * ps];
#pragma unroll
        for (int u = 0; u < kB; ++u) b[u] = q[(k + u) * qs];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kB; ++u) acc += widen<T>(a[u]) * b[u];
    }
    if (k < n) {
        P a[kB - 1];
        T b[kB - 1];
#pragma unroll
        for (int u = 0; u < kB - 1; ++u) a[u] = p[min(k + u, n - 1) * ps];
#pragma unroll
        for (int u = 0; u < kB - 1; ++u) b[u] = q[min(k + u, n - 1) * qs];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kB - 1; ++u)
            if (k + u < n) acc += widen<T>(a[u]) * b[u];
    }
    return acc;
}

// sum over k < n of p[k * ps] * x[r[k * rs]] (x in LDS), in order, with the loads of p and r of a
// batch issued together ahead of the products.
template <typename T, typename P> __device__ inline T gatherDot8(const P* p, int ps, const int32_t* r, int rs, const T* x, int n)
{
    T acc = T(0);
    int k = 0;
    for (; k + kB <= n; k += kB) {
        P a[kB];
        int32_t i[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            a[u] = p[(k + u) * ps];
            i[u] = r[(k + u) * rs];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kB; ++u) acc += widen<T>(a[u]) * x[i[u]];
    }
    if (k < n) {
        P a[kB - 1];
        int32_t i[kB - 1];
#pragma unroll
        for (int u = 0; u < kB - 1; ++u) {
            a[u] = p[min(k + u, n - 1) * ps];
            i[u] = r[min(k + u, n - 1) * rs];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kB - 1; ++u)
            if (k + u < n) acc += widen<T>(a[u]) * x[i[u]];
    }
    return acc;
}

// x[row] minus the products in the slots [cb, ce) pushed to it by its descendants, subtracted in
// slot order, eight loads per batch (the slot range comes with the chunk's rows).
template <typename T> __device__ inline T pullSlots(const T* tb, T v, int cb, int ce)
{
    int q = cb;
    for (; q + kB <= ce; q += kB) {
        T t[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) t[u] = tb[q + u];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kB; ++u) v -= t[u];
    }
    if (q < ce) {
        T t[kB - 1];
#pragma unroll
        for (int u = 0; u < kB - 1; ++u) t[u] = tb[min(q + u, ce - 1)];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kB - 1; ++u)
            if (q + u < ce) v -= t[u];
    }
    return v;
}

// Slots of a task without child tasks live in LDS (kLds: ls holds the first sc of the task's
// slots, [sb, sb + sc) of its subdomain's T buffer, the rest stay in global memory): the task is
// the only one pushing to its rows. Otherwise all slots are in global memory (tb). The choice is a
// template parameter, so each sweep's code sees one address space per access.
template <typename T, bool kLds> struct Slots {
    T* tb;   ///< the subdomain's T buffer (global)
    T* ls;   ///< kLds: the task's first sc slots in LDS
    int sb, sc;
    /// x minus the slots [cb, ce) of a row, in slot order (its LDS part first)
    __device__ T pull(T x, int cb, int ce) const
    {
        if (!kLds) return pullSlots(tb, x, cb, ce);
        const int e = sb + sc;
        if (ce <= e) return pullSlots(ls, x, cb - sb, ce - sb);
        if (cb >= e) return pullSlots(tb, x, cb, ce);
        return pullSlots(tb, pullSlots(ls, x, cb - sb, sc), e, ce);
    }
    /// a pushed product into its slot d
    __device__ void push(int d, T v) const
    {
        if (kLds && static_cast<unsigned>(d - sb) < static_cast<unsigned>(sc))
            ls[d - sb] = v;
        else
            tb[d] = v;
    }
};

// x[row] minus the products pushed to this row by its descendants: the row's slot range
// [cptr, cend), subtracted in slot order.
template <typename T, typename TF, bool kLds>
__device__ inline T pullRow(const LocalArgs<T, TF>& a, const SubDesc& sd, const Slots<T, kLds>& sl, const T* x, int row)
{
    return sl.pull(x[row], a.cptr[sd.cptr_off + row], a.cend[sd.cptr_off + row]);
}

// ------------------------------------------------------------------ wide supernodes (w > 64): whole workgroup
// The blocks of a wide supernode are row-major with a leading dimension padded to four entries
// and start 16-byte aligned, so a lane reads four consecutive columns with one vector load.

// Four consecutive entries from a 16-byte aligned address.
template <typename T> __device__ inline void load4(const T* p, T (&v)[4]);
template <> __device__ inline void load4<float>(const float* p, float (&v)[4])
{
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
}
template <> __device__ inline void load4<double>(const double* p, double (&v)[4])
{
    const double2 q0 = reinterpret_cast<const double2*>(p)[0], q1 = reinterpret_cast<const double2*>(p)[1];
    v[0] = q0.x;
    v[1] = q0.y;
    v[2] = q1.x;
    v[3] = q1.y;
}

// Rows of a dense block taken by one wavefront at a time: 4 in FP32, 2 in FP64 (registers).
template <typename T> constexpr int blockRows() { return sizeof(T) == 4 ? 4 : 2; }

// Sums of R rows spread over the 64 lanes (v[r]: this lane's part of row r), by a transposing
// butterfly: at the step over lane bit b every lane keeps half of its rows and adds the partner's
// part of them, so after log2(R) steps lane l holds row l / (64 / R) summed over R lanes, and
// plain steps over the remaining lane bits finish it (7 shuffles for 4 rows). The order is fixed.
template <typename T, int R> __device__ inline T transposeSum(T (&v)[R], int lane)
{
#pragma unroll
    for (int half = R / 2, bit = kWave / 2; half >= 1; half >>= 1, bit >>= 1) {
        const bool hi = (lane & bit) != 0;
#pragma unroll
        for (int j = 0; j < half; ++j) {
            const T keep = hi ? v[j + half] : v[j];
            const T send = hi ? v[j] : v[j + half];
            v[j] = keep + shflXor(send, bit);
        }
    }
    T s = v[0];
#pragma unroll
    for (int bit = kWave / R / 2; bit >= 1; bit >>= 1) s += shflXor(s, bit);
    return s;
}

// y_r = sum over c < cols of M[r * ld + c] * xs[c] for the rows r < rows of a wide supernode's
// block (xs in LDS): a wavefront takes blockRows rows at a time with each lane on four consecutive
// columns (one vector load per row and 256 columns), and out(r, y_r) is called by one lane per
// row. kTri 1 skips the column quads right of the rows (lower triangle), 2 those left of them
// (upper); the stored block is zero there. All lanes of a wavefront run the same rows.
template <typename T, typename TF, int kTri, int NB = 1, typename Out>
__device__ __forceinline__ void denseRows(const TF* M, int ld, int rows, int cols, const T* xs, int lane, int wave, int nw, Out out)
{
    // Two rows per group with an eight-byte factor or arithmetic type (FP64 factors, FP64 sums of
    // FP32 entries, a pair of FP32 right-hand sides): four rows of them spill registers. Those
    // passes take two groups of R rows, nw * R rows apart, whose loads are issued together; FP32
    // throughout takes one group of four rows (no registers are left for a second). The block's
    // entries (TF) are converted to the arithmetic type T. NB blocks of 4 * 64 columns have their
    // loads issued together (a caller with the registers of one workgroup per compute unit); the
    // products are summed in column order whatever NB is.
    constexpr int R = sizeof(T) >= 8 ? 2 : blockRows<TF>();
    constexpr int P = R == 2 ? 2 : 1;
    for (int r0 = wave * R; r0 < rows; r0 += P * nw * R) {
        const int rA = r0, rB = r0 + nw * R;
        const int nA = min(R, rows - rA), nB = P == 2 ? max(0, min(R, rows - rB)) : 0;
        T accA[R], accB[R];
#pragma unroll
        for (int r = 0; r < R; ++r) accA[r] = accB[r] = T(0);
        for (int cb0 = 0; cb0 < cols; cb0 += NB * 4 * kWave) {
            T xv[NB][4];
            TF mA[NB][R][4], mB[NB][R][4];
            bool okA[NB][R], okB[NB][R];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int c = cb0 + nb * 4 * kWave + 4 * lane;
                bool needA = c < cols, needB = c < cols && nB > 0;
                if (kTri == 1) {
                    needA = needA && c <= rA + nA - 1;
                    needB = needB && c <= rB + nB - 1;
                }
                if (kTri == 2) {
                    needA = needA && c + 3 >= rA;
                    needB = needB && c + 3 >= rB;
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) xv[nb][i] = c + i < cols ? xs[c + i] : T(0);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    okA[nb][r] = needA && r < nA;
                    okB[nb][r] = needB && r < nB;
                    if constexpr (std::is_same_v<T, TF>) {
                        // loads under a branch, the quads not needed set to zero
#pragma unroll
                        for (int i = 0; i < 4; ++i) mA[nb][r][i] = mB[nb][r][i] = TF(0);
                        if (okA[nb][r]) load4(M + static_cast<int64_t>(rA + r) * ld + c, mA[nb][r]);
                        if (okB[nb][r]) load4(M + static_cast<int64_t>(rB + r) * ld + c, mB[nb][r]);
                    } else {
                        // entries converted to T: every lane loads all its quads, those it does not
                        // need from the block's first (cached) quad, and zeroes them after the loads
                        // (under a branch the conversion joins the load, which then waits for each
                        // load before issuing the next)
                        load4(M + (okA[nb][r] ? static_cast<int64_t>(rA + r) * ld + c : 0), mA[nb][r]);
                        load4(M + (okB[nb][r] ? static_cast<int64_t>(rB + r) * ld + c : 0), mB[nb][r]);
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        if constexpr (std::is_same_v<T, TF>) {
                            accA[r] += mA[nb][r][i] * xv[nb][i];
                            accB[r] += mB[nb][r][i] * xv[nb][i];
                        } else {
                            accA[r] += (okA[nb][r] ? widen<T>(mA[nb][r][i]) : T(0)) * xv[nb][i];
                            accB[r] += (okB[nb][r] ? widen<T>(mB[nb][r][i]) : T(0)) * xv[nb][i];
                        }
                    }
        }
        const T yA = transposeSum<T, R>(accA, lane);
        const T yB = transposeSum<T, R>(accB, lane);
        const int rr = lane / (kWave / R);
        if (lane % (kWave / R) == 0) {
            if (rA + rr < rows) out(rA + rr, yA);
            if (rB + rr < rows && nB > 0) out(rB + rr, yB);
        }
    }
}

// y = X x_C for the inverse X of a wide supernode's diagonal block (lower or upper triangle),
// written over x_C.
template <typename T, typename TF>
__device__ __forceinline__ void wideDiag(const TF* X, int w, bool lower, int c0, T* x, T* ybuf, int tid, int lane, int wave, int nw, int nt)
{
    const int ld = wideLd(w);
    if (lower)
        denseRows<T, TF, 1>(X, ld, w, w, x + c0, lane, wave, nw, [&](int r, T v) { ybuf[r] = v; });
    else
        denseRows<T, TF, 2>(X, ld, w, w, x + c0, lane, wave, nw, [&](int r, T v) { ybuf[r] = v; });
    __syncthreads();
    for (int j = tid; j < w; j += nt) x[c0 + j] = ybuf[j];
    __syncthreads();
}

// Forward: every row pulls its products, y = X x_C, then the rows below push B y to their slots.
template <typename T, typename TF, bool kLds>
__device__ __forceinline__ void fwdWide(const LocalArgs<T, TF>& a, const SubDesc& sd, const SnDesc s, T* x, T* ybuf, const Slots<T, kLds>& sl, int tid, int lane,
                        int wave, int nt)
{
    const int w = s.w, h = s.h, c0 = s.c0, nw = nt / kWave;
    for (int j = tid; j < w; j += nt) x[c0 + j] = pullRow(a, sd, sl, x, c0 + j);
    __syncthreads();
    wideDiag(a.diag + s.diag, w, true, c0, x, ybuf, tid, lane, wave, nw, nt);
    const int32_t* dst = a.cidx + sd.dst_off + s.toff;
    denseRows<T, TF, 0>(a.below + s.below, wideLd(w), h, w, x + c0, lane, wave, nw, [&](int k, T v) { sl.push(dst[k], v); });
}

// Backward: x_C -= B^T x_R with G groups of w threads, group g summing the rows k = g, g + G, ...
// (consecutive threads on consecutive columns of a row), the partial sums combined in group
// order through LDS. Then x_C = X x_C.
template <typename T, typename TF>
__device__ __forceinline__ void bwdWide(const LocalArgs<T, TF>& a, const SnDesc s, T* x, T* ybuf, int tid, int lane, int wave, int nt)
{
    const int w = s.w, h = s.h, c0 = s.c0, nw = nt / kWave, ld = wideLd(w);
    if (h > 0) {
        const int G = max(1, min(min(nt / w, kChunkRows / w), min(h, kMaxGroups)));
        const TF* B = a.below + s.below;
        const int32_t* R = a.ridx + s.ridx;
        for (int u = tid; u < G * w; u += nt) {
            const int g = u / w, c = u - g * w;
            ybuf[u] = g < h ? gatherDot8(B + g * ld + c, G * ld, R + g, G, x, (h - g + G - 1) / G) : T(0);
        }
        __syncthreads();
        for (int c = tid; c < w; c += nt) {
            T sum = T(0);
            for (int g = 0; g < G; ++g) sum += ybuf[g * w + c];
            x[c0 + c] -= sum;
        }
        __syncthreads();
    }
    wideDiag(a.diag + s.diag, w, false, c0, x, ybuf, tid, lane, wave, nw, nt);
}

// Dense top, the rows [qlo, qhi) of part of a skeleton task: every row of the skeleton pulls the
// products of the subtrees (b_S - L_SR y_R, staged in order in ybuf), then x_q = sum over c of
// S^-1(q, c) b_c for the part's rows, four rows per wavefront pass with 16-byte loads of S^-1.
template <typename T, typename TF, bool kLds>
__device__ __forceinline__ void denseTop(const LocalArgs<T, TF>& a, const SubDesc& sd, const SnDesc s, const int32_t* rows, int qlo, int qhi, T* x,
                         T* ybuf, const Slots<T, kLds>& sl, int tid, int lane, int wave, int nt)
{
    const int S = s.w, ld = wideLd(S), nw = nt / kWave;
    for (int q = tid; q < S; q += nt) ybuf[q] = pullRow(a, sd, sl, x, rows[q]);
    __syncthreads();
    denseRows<T, TF, 0>(a.diag + s.diag + static_cast<int64_t>(qlo) * ld, ld, qhi - qlo, S, ybuf, lane, wave, nw,
                        [&](int r, T v) { x[rows[qlo + r]] = v; });
    __syncthreads();
}

// ------------------------------------------------------------------ small supernodes of one level, all at once
// Index of the last prefix entry <= t (prefix[0] = 0, nondecreasing).
__device__ inline int findItem(const int* prefix, int count, int t)
{
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// One chunk of small supernodes of one level (a step of the task's sweep), with every thread of the
// workgroup on one row (or column) at a time:
//   forward:  pull the slots of every row (its factor position and slot range come with the
//             chunk, so the pulls start together with the staging of the descriptors);
//             y = inverse(diagonal block) * row values; push the products of the rows below to
//             their slots;
//   backward: every column subtracts the products with the rows below; y = inverse(diagonal) * x.
// The step's fields are uniform over the workgroup.
template <typename T, typename TF, bool kLds>
__device__ __forceinline__ void runChunk(const LocalArgs<T, TF>& a, const SubDesc& sd, int q0, int cnt, int wbase, int hbase, int W, int H,
                         int roff, bool forward, T* x, T* ybuf, SnDesc* st, int* wp, int* hp, const Slots<T, kLds>& sl, int tid,
                         int lane, int wave, int nt)
{
    for (int i = tid; i < cnt; i += nt) {
        st[i] = a.lev_desc[q0 + i];
        wp[i] = a.lev_wpre[q0 + i] - wbase;
        hp[i] = a.lev_hpre[q0 + i] - hbase;
    }
    if (forward)
        for (int t = tid; t < W; t += nt) {
            const int row = a.lev_row[roff + t];
            x[row] = sl.pull(x[row], a.lev_cb[roff + t], a.lev_ce[roff + t]);
        }
    __syncthreads();
    // the supernode of each thread's first row (t = tid) and first row below, found once
    const int iw = tid < W ? findItem(wp, cnt, tid) : 0;
    const int ih = forward && tid < H ? findItem(hp, cnt, tid) : 0;
    if (!forward) {
        if (W * kPullSplit > nt) {
            for (int t = tid; t < W; t += nt) {
                const int i = t == tid ? iw : findItem(wp, cnt, t);
                const SnDesc& s = st[i];
                const int c = t - wp[i];
                if (s.h > 0) x[s.c0 + c] -= gatherDot8(a.below + s.below + c, s.w, a.ridx + s.ridx, 1, x, s.h);
            }
        } else {
            // a narrow chunk (few columns, long columns below): G groups of W threads, thread
            // (g, t) sums the rows g, g + G, ... below column t; partial sums combined in group order
            // (G at most the rows below and kMaxGroups, which bound the serial combine)
            const int G = max(1, min(min(nt / W, kChunkRows / W), min(H, kMaxGroups)));
            for (int u = tid; u < G * W; u += nt) {
                const int g = u / W, t = u - g * W;
                const int i = findItem(wp, cnt, t);
                const SnDesc& s = st[i];
                const int c = t - wp[i];
                ybuf[u] = g < s.h ? gatherDot8(a.below + s.below + g * s.w + c, G * s.w, a.ridx + s.ridx + g, G, x, (s.h - g + G - 1) / G)
                                  : T(0);
            }
            __syncthreads();
            for (int t = tid; t < W; t += nt) {
                const int i = t == tid ? iw : findItem(wp, cnt, t);
                T sum = T(0);
                for (int g = 0; g < G; ++g) sum += ybuf[g * W + t];
                x[st[i].c0 + (t - wp[i])] -= sum;
            }
        }
        __syncthreads();
    }
    for (int t = tid; t < W; t += nt) {
        const int i = t == tid ? iw : findItem(wp, cnt, t);
        const SnDesc& s = st[i];
        const int j = t - wp[i];
        const int lo = forward ? 0 : j, hi = forward ? j + 1 : s.w;
        const TF* D = a.diag + s.diag + lo * s.w + j;
        ybuf[t] = dot8(D, s.w, x + s.c0 + lo, 1, hi - lo, T(0));
    }
    __syncthreads();
    for (int t = tid; t < W; t += nt) {
        const int i = t == tid ? iw : findItem(wp, cnt, t);
        x[st[i].c0 + (t - wp[i])] = ybuf[t];
    }
    __syncthreads();
    if (forward) {
        for (int t = tid; t < H; t += nt) {
            const int i = t == tid ? ih : findItem(hp, cnt, t);
            const SnDesc& s = st[i];
            const int k = t - hp[i];
            const TF* B = a.below + s.below + static_cast<int64_t>(k) * s.w;
            sl.push(a.cidx[sd.dst_off + s.toff + k], dot8(B, 1, x + s.c0, 1, s.w, T(0)));
        }
        __syncthreads();
    }
}

__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int64_t uniform64(int64_t v)
{
    const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(v & 0xffffffff));
    const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(v >> 32));
    return (static_cast<int64_t>(hi) << 32) | static_cast<uint32_t>(lo);
}

// The steps of one sweep of a task, in order: chunks of the small supernodes of a level and the
// level's wide supernodes, level after level. The steps are staged in LDS kStepBlock at a time;
// every step ends with a barrier.
template <typename T, typename TF, bool kLds>
__device__ __forceinline__ void runSteps(const LocalArgs<T, TF>& a, const SubDesc& sd, int step0, int nsteps, bool forward, T* x, unsigned char* smem,
                         const Slots<T, kLds>& sl, const int32_t* rows, int qlo, int qhi, int tid, int lane, int wave, int nt)
{
    T* ybuf = reinterpret_cast<T*>(smem + a.y_off);
    SnDesc* st = reinterpret_cast<SnDesc*>(smem + a.st_off);
    int* wp = reinterpret_cast<int*>(smem + a.st_off + a.st_cap * sizeof(SnDesc));
    int* hp = wp + a.st_cap;
    StepDesc* sb = reinterpret_cast<StepDesc*>(smem + a.step_off);
    for (int k0 = 0; k0 < nsteps; k0 += kStepBlock) {
        const int kn = min(kStepBlock, nsteps - k0);
        for (int i = tid; i < kn; i += nt) sb[i] = a.steps[step0 + k0 + i];
        __syncthreads();
        for (int k = 0; k < kn; ++k) {
            const StepDesc& sp = sb[k];
            if (uniform(sp.kind) == 0) {
                runChunk(a, sd, uniform(sp.q0), uniform(sp.cnt), uniform(sp.wbase), uniform(sp.hbase), uniform(sp.W),
                         uniform(sp.H), uniform(sp.roff), forward, x, ybuf, st, wp, hp, sl, tid, lane, wave, nt);
            } else {
                SnDesc d;
                d.c0 = uniform(sp.wide.c0);
                d.w = uniform(sp.wide.w);
                d.h = uniform(sp.wide.h);
                d.toff = uniform(sp.wide.toff);
                d.diag = uniform64(sp.wide.diag);
                d.below = uniform64(sp.wide.below);
                d.ridx = uniform64(sp.wide.ridx);
                if (uniform(sp.kind) == 2)
                    denseTop(a, sd, d, rows, qlo, qhi, x, ybuf, sl, tid, lane, wave, nt);
                else if (forward)
                    fwdWide(a, sd, d, x, ybuf, sl, tid, lane, wave, nt);
                else
                    bwdWide(a, d, x, ybuf, tid, lane, wave, nt);
                __syncthreads();
            }
        }
    }
}

// One entry of the level-ordered queue: a task's forward sweep, its backward sweep, or both for a
// task without parent. Entries of one launch are independent (one depth group of the queue, or
// one subdomain's entries run in order by a single workgroup); a later launch on the same stream
// sees the T buffer and the shared results of earlier launches.
template <typename T, typename TF>
__device__ __forceinline__ void runEntry(const LocalArgs<T, TF>& a, long long e, T* x, unsigned char* smem, int tid, int lane, int wave, int nt)
{
    // kPair: T holds right-hand sides 2 r and 2 r + 1 (the second absent when nrhs is odd)
    constexpr bool kPair = PairOf<T>::value;
    const int r = static_cast<int>(e / a.qcount);
    const int rA = kPair ? 2 * r : r, rB = rA + 1;
    const bool onA = !(a.skip && a.skip[rA]);
    const bool onB = kPair && rB < a.nrhs && !(a.skip && a.skip[rB]);
    if (!onA && !onB) return;
    const int code = a.queue[e % a.qcount];
    const int tk = (code & ((1 << kPartShift) - 1)) >> 1;
    const int part = code >> kPartShift;
    const bool backward = (code & 1) != 0;
    const TaskDesc td = a.task[tk];
    const bool combined = td.parent < 0;
    const SubDesc sd = a.sub[td.sub];
    T* tb = a.tbuf + static_cast<int64_t>(r) * a.t_total + sd.t_off;
    T* xg = a.xg + static_cast<int64_t>(r) * a.n_local_total + sd.map_off;
    const int32_t* rows = a.task_rows + td.rows_first;
    // a forward sweep of a task without child tasks keeps its rows' slots in LDS, as many as fit
    const bool ldsSlots = !backward && td.slot_count > 0 && a.slot_cap > 0;
    const Slots<T, false> slg{tb, nullptr, 0, 0};
    const Slots<T, true> sll{tb, reinterpret_cast<T*>(smem + a.slot_off), td.slot_base, min(td.slot_count, a.slot_cap)};
    // the rows this entry computes: all of the task's, or a dense top part's range
    const int qlo = td.nparts > 0 ? static_cast<int>(static_cast<int64_t>(td.rows_count) * part / td.nparts) : 0;
    const int qhi = td.nparts > 0 ? static_cast<int>(static_cast<int64_t>(td.rows_count) * (part + 1) / td.nparts) : td.rows_count;
    if (!backward) {
        const int64_t ld = a.local_in ? a.n_local_total : a.n_global;
        const double* rhs = a.rhs + static_cast<int64_t>(rA) * ld;
        for (int q = tid; q < td.rows_count; q += blockDim.x) {
            const int row = rows[q];
            const int g = a.local_in ? sd.map_off + a.lperm[sd.map_off + row] : a.gmap[sd.map_off + row];
            if constexpr (kPair) {
                using E = typename PairOf<T>::E;
                x[row] = T(static_cast<E>(rhs[g]), onB ? static_cast<E>(rhs[ld + g]) : E(0));
            } else
                x[row] = static_cast<T>(rhs[g]);
        }
        __syncthreads();
        if (ldsSlots)
            runSteps(a, sd, td.f_step, td.f_nsteps, true, x, smem, sll, rows, qlo, qhi, tid, lane, wave, nt);
        else
            runSteps(a, sd, td.f_step, td.f_nsteps, true, x, smem, slg, rows, qlo, qhi, tid, lane, wave, nt);
        if (!combined) {
            for (int q = tid; q < td.rows_count; q += blockDim.x) xg[rows[q]] = x[rows[q]];
            __syncthreads();
            return;
        }
    } else {
        for (int q = tid; q < td.rows_count; q += blockDim.x) x[rows[q]] = xg[rows[q]];
        const int32_t* ext = a.ext_rows + td.ext_first;
        for (int q = tid; q < td.ext_count; q += blockDim.x) x[ext[q]] = xg[ext[q]];
        __syncthreads();
    }
    runSteps(a, sd, td.b_step, td.b_nsteps, false, x, smem, slg, rows, qlo, qhi, tid, lane, wave, nt);
    if (td.nchild > 0)
        for (int q = qlo + tid; q < qhi; q += blockDim.x) xg[rows[q]] = x[rows[q]];
    // the owned rows of the result (z = x0 + y), or the local buffer (scaled by D_i) for a reverse add
    auto store = [&](int rr, int row, double y) {
        const int l = a.lperm[sd.map_off + row];
        if (a.mode == 0) {
            if (l < sd.n_owned) {
                const int64_t gi = static_cast<int64_t>(rr) * a.n_global + sd.begin + l;
                a.out[gi] = a.x0 ? a.x0[gi] + y : y;
            }
        } else {
            a.out[static_cast<int64_t>(rr) * a.n_local_total + sd.map_off + l] = a.pou ? a.pou[sd.map_off + l] * y : y;
        }
    };
    for (int q = qlo + tid; q < qhi; q += blockDim.x) {
        const int row = rows[q];
        if constexpr (kPair) {
            if (onA) store(rA, row, static_cast<double>(x[row].x));
            if (onB) store(rB, row, static_cast<double>(x[row].y));
        } else {
            store(rA, row, static_cast<double>(x[row]));
        }
    }
    __syncthreads();
}

// Batched local solve. Grid mode: workgroup b runs entry b of the launch (all subdomains' tasks
// of one depth group, for every right-hand side). Serial mode: one workgroup runs the entries of
// one subdomain in queue order (the one-launch-per-subdomain comparison and single solves).
//
// kThreads 1024: one workgroup per compute unit (16 wavefronts, 128 VGPRs each); 512: two
// workgroups share a compute unit (8 wavefronts each, 4 per SIMD together: at most 128 VGPRs), each
// hiding the other's memory round trips and barriers.
template <typename T, typename TF, int kThreads>
__global__ __launch_bounds__(kThreads, kThreads == 1024 ? 1 : 4) void k_local_solve(LocalArgs<T, TF> a)
{
    extern __shared__ unsigned char smem_raw[];
    T* x = reinterpret_cast<T*>(smem_raw);
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave, nt = blockDim.x;
    if (a.serial) {
        const long long total = static_cast<long long>(a.qcount) * a.nrhs;
        for (long long e = 0; e < total; ++e) runEntry(a, e, x, smem_raw, tid, lane, wave, nt);
    } else {
        runEntry(a, static_cast<long long>(blockIdx.x), x, smem_raw, tid, lane, wave, nt);
    }
}

// ------------------------------------------------------------------ dense tops in a launch of their own
// A launch whose entries are all dense tops (tasks without parent whose one step is x_S = S^-1 b_S):
// workgroup b runs entry b. The part's workgroup forms b_S (the right-hand side minus the products
// pulled from the subtree tasks) for every row of the skeleton in LDS, then x_q = S^-1(q, :) b_S for
// its rows [qlo, qhi), written to xg for the subtree tasks' backward sweeps and stored as results:
// the operations of the dense top step of k_local_solve in the same order (the same bits), with
// under half its registers: two workgroups of 16 wavefronts per compute unit keep more loads of
// S^-1 in flight (3D Poisson 64^3, 64 skeletons of about 820 rows: 48.5 -> 43.7 us per launch).
// dt[dtb[task] + q] = (factor position, global row, local index, first slot) of the skeleton's row
// q, dt_ce its end of slots.
// kNB > 1 (a launch of at most one workgroup per compute unit): the registers of one workgroup per
// compute unit, with the loads of kNB blocks of 256 columns of S^-1 issued together (the same sums
// in the same order: the same bits).
constexpr int kDtThreads = 1024;

template <typename T, typename TF, int kNB>
__global__ __launch_bounds__(kDtThreads, kNB == 1 ? 2 : 1) void k_dense_top(LocalArgs<T, TF> a, const int4* dt,
                                                                          const int32_t* dt_ce, const int32_t* dtb)
{
    extern __shared__ unsigned char smem_raw[];
    T* bS = reinterpret_cast<T*>(smem_raw);
    constexpr bool kPair = PairOf<T>::value;
    const long long e = blockIdx.x;
    const int r = static_cast<int>(e / a.qcount);
    const int rA = kPair ? 2 * r : r, rB = rA + 1;
    const bool onA = !(a.skip && a.skip[rA]);
    const bool onB = kPair && rB < a.nrhs && !(a.skip && a.skip[rB]);
    if (!onA && !onB) return;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave, nt = blockDim.x;
    const int code = a.queue[e % a.qcount];
    const int tk = (code & ((1 << kPartShift) - 1)) >> 1;
    const int part = code >> kPartShift;
    const TaskDesc td = a.task[tk];
    const SubDesc sd = a.sub[td.sub];
    const StepDesc sp = a.steps[td.f_step];
    const int S = sp.wide.w, ld = wideLd(S);
    const int4* rowq = dt + dtb[tk];
    const int32_t* ceq = dt_ce + dtb[tk];
    const T* tb = a.tbuf + static_cast<int64_t>(r) * a.t_total + sd.t_off;
    const int64_t ldr = a.local_in ? a.n_local_total : a.n_global;
    const double* rhs = a.rhs + static_cast<int64_t>(rA) * ldr;
    for (int q = tid; q < S; q += nt) {
        const int4 w = rowq[q];
        const int g = a.local_in ? sd.map_off + w.z : w.y;
        T v;
        if constexpr (kPair) {
            using E = typename PairOf<T>::E;
            v = T(static_cast<E>(rhs[g]), onB ? static_cast<E>(rhs[ldr + g]) : E(0));
        } else
            v = static_cast<T>(rhs[g]);
        bS[q] = pullSlots(tb, v, w.w, ceq[q]);
    }
    __syncthreads();
    const int qlo = static_cast<int>(static_cast<int64_t>(S) * part / td.nparts);
    const int qhi = static_cast<int>(static_cast<int64_t>(S) * (part + 1) / td.nparts);
    T* xg = a.xg + static_cast<int64_t>(r) * a.n_local_total + sd.map_off;
    auto store = [&](int rr, int l, double y) {
        if (a.mode == 0) {
            if (l < sd.n_owned) {
                const int64_t gi = static_cast<int64_t>(rr) * a.n_global + sd.begin + l;
                a.out[gi] = a.x0 ? a.x0[gi] + y : y;
            }
        } else {
            a.out[static_cast<int64_t>(rr) * a.n_local_total + sd.map_off + l] = a.pou ? a.pou[sd.map_off + l] * y : y;
        }
    };
    auto out = [&](int k, T v) {
        const int4 w = rowq[qlo + k];
        if (td.nchild > 0) xg[w.x] = v;
        if constexpr (kPair) {
            if (onA) store(rA, w.z, static_cast<double>(v.x));
            if (onB) store(rB, w.z, static_cast<double>(v.y));
        } else {
            store(rA, w.z, static_cast<double>(v));
        }
    };
    denseRows<T, TF, 0, kNB>(a.diag + sp.wide.diag + static_cast<int64_t>(qlo) * ld, ld, qhi - qlo, S, bS, lane, wave, nt / kWave,
                             out);
}

// ------------------------------------------------------------------ SpMV: one thread per row,
// sequential sum in the stored (local) column order with fma, as DistributedMatrix::apply.
// Row i of A x: the row's entries eight at a time, columns and values in one round trip, x in the
// next.
__device__ __forceinline__ double spmvRow(const int32_t* rp, const int32_t* col, const double* val, const double* xr, int i)
{
    double s = 0.0;
    const int ke = rp[i + 1];
    for (int k0 = rp[i]; k0 < ke; k0 += 8) {
        int c[8];
        double v[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            c[u] = k0 + u < ke ? col[k0 + u] : 0;
            v[u] = k0 + u < ke ? val[k0 + u] : 0.0;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = k0 + u < ke ? xr[c[u]] : 0.0;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k0 + u < ke) s = fma(v[u], xv[u], s);
    }
    return s;
}

__global__ void k_spmv(int32_t n, const int32_t* rp, const int32_t* col, const double* val, const double* x, double* y,
                       const double* b, const int32_t* skip)
{
    const int r = blockIdx.y;
    if (skip && skip[r]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t off = static_cast<int64_t>(r) * n;
    const double s = spmvRow(rp, col, val, x + off, i);
    y[off + i] = b ? b[off + i] - s : s;
}

// The same products for a group of up to kSpmvGroup right-hand sides per thread: the row's column
// indices and values are read once for the group, and every right-hand side's sum keeps the order
// and the fma of k_spmv (bitwise the same results).
constexpr int kSpmvGroup = 8;
__global__ __launch_bounds__(256) void k_spmv_group(int32_t n, const int32_t* rp, const int32_t* col, const double* val,
                                                     const double* x, double* y, const double* b, int nrhs,
                                                     const int32_t* skip)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r0 = blockIdx.y * kSpmvGroup;
    bool on[kSpmvGroup];
    double s[kSpmvGroup];
    bool any = false;
#pragma unroll
    for (int u = 0; u < kSpmvGroup; ++u) {
        on[u] = r0 + u < nrhs && !(skip && skip[r0 + u]);
        any = any || on[u];
        s[u] = 0.0;
    }
    if (!any) return;
    const double* xr = x + static_cast<int64_t>(r0) * n;
    const int ke = rp[i + 1];
    for (int k = rp[i]; k < ke; ++k) {
        const double v = val[k];
        const int c = col[k];
        double xv[kSpmvGroup];
#pragma unroll
        for (int u = 0; u < kSpmvGroup; ++u) xv[u] = on[u] ? xr[static_cast<int64_t>(u) * n + c] : 0.0;
#pragma unroll
        for (int u = 0; u < kSpmvGroup; ++u) s[u] = fma(v, xv[u], s[u]);
    }
#pragma unroll
    for (int u = 0; u < kSpmvGroup; ++u) {
        if (!on[u]) continue;
        const int64_t o = static_cast<int64_t>(r0 + u) * n + i;
        y[o] = b ? b[o] - s[u] : s[u];
    }
}

// ------------------------------------------------------------------ coarse space
constexpr int kCrows = 256;  // rows of V_i per workgroup of the restriction (one per thread)


// c_i = V_i^T r for fewer right-hand sides than the matrix-core path takes, in one stage:
// workgroup (i, group, rhs) owns kCg coarse columns of subdomain i over all its rows (no partial
// sums across workgroups). Each thread sums its rows (tid, tid + blockDim, ...) in order, the
// wavefronts' sums are combined in wavefront order.
constexpr int kCg = 4;           // coarse columns per workgroup
constexpr int kCgThreads = 512;
constexpr int kCgRows = 4;       // rows per thread whose loads are issued together
// c_q = sum over the rows l of block i of V_i(l, q) r(g(l)), for the kCg columns q of the workgroup:
// each thread takes kCgRows rows at a time (their V entries, their rows of r: through the row map,
// or directly when the block's rows are consecutive global rows), then wavefront and workgroup sums.
__global__ __launch_bounds__(kCgThreads) void k_coarse_restrict_cols(CoarseArgs c, const double* r, int64_t n, double* cvec,
                                                                   const int32_t* skip)
{
    const int rr = blockIdx.z;
    if (skip && skip[rr]) return;
    const int i = blockIdx.x, q0 = blockIdx.y * kCg;
    const int m = static_cast<int>(c.off[i + 1] - c.off[i]);
    if (q0 >= m) return;
    const int rows = c.vrows[i];
    const double* V = c.V + c.voff[i] + static_cast<int64_t>(q0) * rows;
    const int32_t* g = c.vrow_g + c.vrow_ptr[i];
    const int base = c.vbase[i];
    const double* rv = r + static_cast<int64_t>(rr) * n;
    const int nq = min(kCg, m - q0);
    double acc[kCg];
#pragma unroll
    for (int q = 0; q < kCg; ++q) acc[q] = 0.0;
    for (int l0 = threadIdx.x; l0 < rows; l0 += kCgRows * kCgThreads) {
        // loads past the block's rows or columns read its last row or column, then count zero
        int gl[kCgRows];
        double va[kCgRows][kCg];
#pragma unroll
        for (int u = 0; u < kCgRows; ++u) {
            const int l = min(l0 + u * kCgThreads, rows - 1);
            gl[u] = base >= 0 ? base + l : g[l];
#pragma unroll
            for (int q = 0; q < kCg; ++q) va[u][q] = V[static_cast<int64_t>(min(q, nq - 1)) * rows + l];
        }
        double xv[kCgRows];
#pragma unroll
        for (int u = 0; u < kCgRows; ++u) xv[u] = rv[gl[u]];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kCgRows; ++u) {
            if (l0 + u * kCgThreads >= rows) continue;
#pragma unroll
            for (int q = 0; q < kCg; ++q) acc[q] += (q < nq ? va[u][q] : 0.0) * xv[u];
        }
    }
    __shared__ double red[kCgThreads / kWave][kCg];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < kCg; ++q) {
        const double v = waveSum(acc[q]);
        if (lane == 0) red[wave][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < nq) {
        double t = 0.0;
        for (int w = 0; w < kCgThreads / kWave; ++w) t += red[w][threadIdx.x];
        cvec[static_cast<int64_t>(rr) * c.n0 + c.off[i] + q0 + threadIdx.x] = t;
    }
}

// ------------------------------------------------------------------ coarse products of several
// right-hand sides on the matrix cores (v_mfma_f64_16x16x4f64: D(16x16) += A(16x4) B(4x16) in
// FP64). Lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[l / 16 + 4 j][l % 16]
// for j = 0..3. A group of up to 16 right-hand sides shares every read of V and of A00's inverted
// factors, which the one-right-hand-side kernels read once per right-hand side.
typedef double mfmaAcc __attribute__((ext_vector_type(4)));
constexpr int kMfmaTiles = 4;  // coarse columns of one subdomain covered by the restriction: 64

__device__ inline mfmaAcc mfma4(double a, double b, mfmaAcc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// First stage of c_i = V_i^T r for the right-hand sides 16 g .. 16 g + 15: workgroup (i, chunk, g),
// wavefront w on rows [64 w, 64 w + 64) of the chunk's 256, the four wavefronts' tiles summed in
// wavefront order; partial sums in the layout of k_coarse_restrict.
__global__ __launch_bounds__(kCrows) void k_coarse_restrict_mfma(CoarseArgs c, const double* r, int64_t n, double* part,
                                                                  int nrhs, const int32_t* skip)
{
    const int i = blockIdx.x, ch = blockIdx.y, g0 = blockIdx.z * 16;
    const int rows = c.vrows[i];
    const int nch = (rows + kCrows - 1) / kCrows;
    if (ch >= nch) return;
    const int m = static_cast<int>(c.off[i + 1] - c.off[i]);
    const int ntile = (m + 15) / 16;
    const double* V = c.V + c.voff[i];
    const int32_t* gr = c.vrow_g + c.vrow_ptr[i];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int rhs = g0 + (lane & 15);
    const bool rok = rhs < nrhs;
    const double* rv = r + static_cast<int64_t>(rok ? rhs : 0) * n;
    mfmaAcc acc[kMfmaTiles];
#pragma unroll
    for (int t = 0; t < kMfmaTiles; ++t) acc[t] = mfmaAcc{0.0, 0.0, 0.0, 0.0};
    const int k0 = ch * kCrows + wave * 64, k1 = min(rows, k0 + 64);
    for (int kb = k0; kb < k1; kb += 4) {
        const int k = kb + (lane >> 4);
        const bool kok = k < k1;
        const double b = (kok && rok) ? rv[gr[k]] : 0.0;
#pragma unroll
        for (int t = 0; t < kMfmaTiles; ++t) {
            if (t < ntile) {
                const int q = 16 * t + (lane & 15);
                const double a = (kok && q < m) ? V[static_cast<int64_t>(q) * rows + k] : 0.0;
                acc[t] = mfma4(a, b, acc[t]);
            }
        }
    }
    __shared__ double red[kCrows / kWave][kMfmaTiles][256];
#pragma unroll
    for (int t = 0; t < kMfmaTiles; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) red[wave][t][((lane >> 4) + 4 * j) * 16 + (lane & 15)] = acc[t][j];
    __syncthreads();
    // thread u: coarse column q = 16 t + u / 16, right-hand side g0 + u % 16
    for (int t = 0; t < ntile; ++t) {
        const int u = threadIdx.x, q = 16 * t + u / 16, rr = g0 + (u & 15);
        if (q < m && rr < nrhs && !(skip && skip[rr])) {
            double sum = 0.0;
            for (int w = 0; w < kCrows / kWave; ++w) sum += red[w][t][u];
            part[((static_cast<int64_t>(rr) * c.nsub + i) * c.rchunks + ch) * c.mmax + q] = sum;
        }
    }
}

// y = T x for 16 right-hand sides at a time (T an inverted triangular factor of A00, lower 1 or
// upper 0, or the full inverse, lower 2: row-major n0 x n0 with explicit zeros and rows padded to
// ld, a multiple of eight; x rows permuted by perm when given). Workgroup (t, g) computes rows [16 t, 16 t + 16) with its kTriWaves wavefronts
// splitting the columns the triangle holds into blocks of eight (wavefront w takes blocks w,
// w + kTriWaves, ...). A lane reads two consecutive entries of its row with one 16-byte load; they
// feed the two products over the columns k = 8 b + 2 q and 8 b + 2 q + 1 (q = lane / 16), and the
// matching entries of x go to the same lanes. The wavefronts' 16 x 16 results are summed in
// wavefront order.
constexpr int kTriWaves = 16;
constexpr int kTriBatch = 4;  // blocks of eight columns per batch of loads
template <bool kPerm>
__global__ __launch_bounds__(kTriWaves * kWave) void k_tri_gemm_mfma(int32_t n0, int32_t ld, const double* T, int lower,
                                                                      const int32_t* perm, const double* x, double* y,
                                                                      int nrhs, const int32_t* skip)
{
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int r0 = blockIdx.x * 16, g0 = blockIdx.y * 16;
    const int rhs = g0 + (lane & 15), q = lane >> 4;
    const bool rok = rhs < nrhs;
    const double* xv = x + static_cast<int64_t>(rok ? rhs : 0) * n0;
    const int row = min(r0 + (lane & 15), n0 - 1);  // rows past n0 repeat the last one; their sums are dropped
    const double* Trow = T + static_cast<int64_t>(row) * ld;
    const int b0 = lower == 0 ? (r0 >> 3) : 0;
    const int b1 = lower == 1 ? ((min(n0, r0 + 16) + 7) >> 3) : (ld >> 3);
    mfmaAcc acc = {0.0, 0.0, 0.0, 0.0};
    for (int bb = b0 + wave; bb < b1; bb += kTriWaves * kTriBatch) {
        double2 a[kTriBatch];
        double xa[kTriBatch], xb[kTriBatch];
#pragma unroll
        for (int u = 0; u < kTriBatch; ++u) {
            const int b = bb + u * kTriWaves;
            const int k = 8 * b + 2 * q;
            const bool ok = b < b1;
            a[u] = ok ? *reinterpret_cast<const double2*>(Trow + k) : double2{0.0, 0.0};
            const bool oka = ok && rok && k < n0, okb = ok && rok && k + 1 < n0;
            if (kPerm) {
                xa[u] = oka ? xv[perm[k]] : 0.0;
                xb[u] = okb ? xv[perm[k + 1]] : 0.0;
            } else {
                xa[u] = oka ? xv[k] : 0.0;
                xb[u] = okb ? xv[k + 1] : 0.0;
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kTriBatch; ++u) {
            acc = mfma4(a[u].x, xa[u], acc);
            acc = mfma4(a[u].y, xb[u], acc);
        }
    }
    __shared__ double red[kTriWaves][256];
#pragma unroll
    for (int j = 0; j < 4; ++j) red[wave][(q + 4 * j) * 16 + (lane & 15)] = acc[j];
    __syncthreads();
    if (threadIdx.x < 256) {
        const int u = threadIdx.x, rw = r0 + u / 16, rr = g0 + (u & 15);
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < kTriWaves; ++w) sum += red[w][u];
        if (rw < n0 && rr < nrhs && !(skip && skip[rr])) y[static_cast<int64_t>(rr) * n0 + rw] = sum;
    }
}

// x0 rows of V_i = V_i y_i for 16 right-hand sides at a time: workgroup (prolongation block b, g),
// wavefront w on rows [64 w, 64 w + 64) of the block's 256, 16 rows per tile, K = m_i.
__global__ __launch_bounds__(256) void k_coarse_prolong_mfma(CoarseArgs c, const double* y, double* out, int64_t ld_out,
                                                              int boolean, int nrhs, const int32_t* skip)
{
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int b = blockIdx.x, g0 = blockIdx.y * 16;
    const int i = c.blk_sub[b];
    const int rows = c.vrows[i];
    const int m = static_cast<int>(c.off[i + 1] - c.off[i]);
    const double* V = c.V + c.voff[i];
    const int rhs = g0 + (lane & 15);
    const bool rok = rhs < nrhs;
    const double* yv = y + static_cast<int64_t>(rok ? rhs : 0) * c.n0 + c.off[i];
    const bool write = rok && !(skip && skip[rhs]);
    // K = m <= 64 in at most 16 steps of four: the entries of y first, then V eight steps at a time
    double bv[16];
#pragma unroll
    for (int st = 0; st < 16; ++st) {
        const int q = 4 * st + (lane >> 4);
        bv[st] = (q < m && rok) ? yv[q] : 0.0;
    }
    const int nst = (m + 3) / 4;
    for (int t = 0; t < 4; ++t) {
        const int l0 = c.blk_row[b] + wave * 64 + t * 16;
        if (l0 >= rows) break;
        const int row = l0 + (lane & 15);
        mfmaAcc acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s0 = 0; s0 < 16; s0 += 8) {
            if (s0 >= nst) break;
            double av[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = 4 * (s0 + u) + (lane >> 4);
                av[u] = (q < m && row < rows) ? V[static_cast<int64_t>(q) * rows + row] : 0.0;
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = mfma4(av[u], bv[s0 + u], acc);
        }
        if (!write) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int l = l0 + (lane >> 4) + 4 * j;
            if (l < rows) {
                const int64_t dst = boolean ? c.vrow_g[c.vrow_ptr[i] + l] : c.vrow_ptr[i] + l;
                out[static_cast<int64_t>(rhs) * ld_out + dst] = acc[j];
            }
        }
    }
}

// c_i[q] = sum over the chunks of subdomain i, in chunk order.
__global__ void k_coarse_final(CoarseArgs c, const double* part, double* cvec, const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int i = blockIdx.x;
    const int m = static_cast<int>(c.off[i + 1] - c.off[i]);
    const int nch = (c.vrows[i] + kCrows - 1) / kCrows;
    const double* base = part + (static_cast<int64_t>(rr) * c.nsub + i) * c.rchunks * c.mmax;
    for (int q = threadIdx.x; q < m; q += blockDim.x) {
        double t = 0.0;
        for (int b = 0; b < nch; ++b) t += base[static_cast<int64_t>(b) * c.mmax + q];
        cvec[static_cast<int64_t>(rr) * c.n0 + c.off[i] + q] = t;
    }
}

// 16-byte loads per lane and batch of k_tri_gemv: a row of up to 1536 entries in one round trip
constexpr int kGemvBatch = 12;

// y = T x for one right-hand side per grid row: one wavefront per row of the triangle (lower 1,
// upper 0) or of a full matrix (lower 2), a lane reading two consecutive entries with one 16-byte
// load (rows padded to ld, a multiple of eight, with zeros), kGemvBatch such loads per batch; lane
// partial sums (entries 2 lane + 128 k in increasing k), then the wavefront sum.

__global__ __launch_bounds__(256) void k_tri_gemv(int32_t n0, int32_t ld, const double* Tm, int lower, const int32_t* perm,
                                                   const double* x, double* y, const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int i = blockIdx.x * (blockDim.x / kWave) + wave;
    if (i >= n0) return;
    const double* xv = x + static_cast<int64_t>(rr) * n0;
    const int lo = lower == 0 ? (i & ~1) : 0, hi = lower == 1 ? i + 1 : n0;  // even start: 16-byte aligned
    const double* row = Tm + static_cast<int64_t>(i) * ld;
    double s = 0.0;
    for (int j0 = lo; j0 < hi; j0 += 2 * kGemvBatch * kWave) {
        double2 t[kGemvBatch];
        double xa[kGemvBatch], xb[kGemvBatch];
#pragma unroll
        for (int u = 0; u < kGemvBatch; ++u) {
            const int j = j0 + 2 * lane + 2 * kWave * u;
            const bool ok = j < hi;
            t[u] = ok ? *reinterpret_cast<const double2*>(row + j) : double2{0.0, 0.0};
            xa[u] = ok ? xv[perm ? perm[j] : j] : 0.0;
            xb[u] = ok && j + 1 < n0 ? xv[perm ? perm[j + 1] : j + 1] : 0.0;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kGemvBatch; ++u) s += t[u].x * xa[u] + t[u].y * xb[u];
    }
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
    for (int k0 = 0; k0 < m; k0 += 8) {
        double v[8], w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v[u] = k0 + u < m ? V[static_cast<int64_t>(k0 + u) * rows + l] : 0.0;
            w[u] = k0 + u < m ? yv[k0 + u] : 0.0;
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u] * w[u];
    }
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


// Two-stage sums without atomics: every block of a reduction kernel writes its partial sums
// (a wave tree, then the waves in order), and a second launch adds the partials of all blocks, one
// block per sum: thread t takes blocks t, t + blockDim.x, ... in order, then a wave tree and the
// waves in order. The order is fixed, so results are bitwise repeatable.
constexpr int kBlockThreads = 256;
constexpr int kBlockWaves = kBlockThreads / kWave;

// Partial sums (a0, a1, a2) of this block for right-hand side rr: partial[(rr * gridDim.x + b) * 3 + k].
__device__ void writePartial3(double a0, double a1, double a2, double* partial, int rr)
{
    __shared__ double red[kBlockWaves][kRedComp];
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
    if (threadIdx.x < kRedComp) {
        double t = 0.0;
        for (int wv = 0; wv < kBlockWaves; ++wv) t += red[wv][threadIdx.x];
        partial[(static_cast<int64_t>(rr) * gridDim.x + blockIdx.x) * kRedComp + threadIdx.x] = t;
    }
}

// Block-wide sum of one value per thread (fixed order); the result is valid in thread 0.
__device__ double blockSum(double v)
{
    __shared__ double red[kBlockWaves];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    v = waveSum(v);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int wv = 0; wv < kBlockWaves; ++wv) t += red[wv];
    return t;
}

// out[rr * out_ld + k] = sum over b < nb of partial[(rr * nb + b) * ld + k]; grid (K, nrhs).
__global__ __launch_bounds__(kBlockThreads) void k_final_sum(const double* partial, int ld, int nb, double* out, int out_ld,
                                                             const int32_t* skip)
{
    const int k = blockIdx.x, rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const double* base = partial + static_cast<int64_t>(rr) * nb * ld + k;
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += base[static_cast<int64_t>(b) * ld];
    const double t = blockSum(v);
    if (threadIdx.x == 0) out[static_cast<int64_t>(rr) * out_ld + k] = t;
}

__global__ __launch_bounds__(kRedThreads) void k_reduce(int kind, int64_t n, const double* x, const double* y,
                                                         double* partial, const int32_t* skip)
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
            if (kind == 5) a1 += v;
        }
    }
    writePartial3(a0, a1, a2, partial, rr);
}

// ------------------------------------------------------------------ BiCGStab (bicgstab.cpp:81-137)
// Per right-hand side coefficients coef[4 r + k] are set by the host: the scalar recurrences of
// BiCGStab stay on the host in FP64, the vector updates and their reductions are fused here.

// p = r (first step of a cycle, coef[2] != 0) or p = r + beta (p - omega v)
__global__ void k_bicg_p(int64_t n, const double* r, double* p, const double* v, const double* coef, const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = static_cast<int64_t>(rr) * n + i;
    const double beta = coef[4 * rr], omega = coef[4 * rr + 1];
    p[o] = coef[4 * rr + 2] != 0.0 ? r[o] : r[o] + beta * (p[o] - omega * v[o]);
}

// s = r - alpha v; sums (s.s, |s|_1)
__global__ __launch_bounds__(kRedThreads) void k_bicg_s(int64_t n, const double* r, const double* v, double* sv,
                                                        const double* coef, double* partial, const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int64_t base = static_cast<int64_t>(rr) * n;
    const double alpha = coef[4 * rr];
    double a0 = 0.0, a1 = 0.0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double s = r[base + i] - alpha * v[base + i];
        sv[base + i] = s;
        a0 += s * s;
        a1 += fabs(s);
    }
    writePartial3(a0, a1, 0.0, partial, rr);
}

// x += alpha y + omega z; r = s - omega t; sums (r.r, |r|_1, rh.r)
__global__ __launch_bounds__(kRedThreads) void k_bicg_xr(int64_t n, double* x, const double* y, const double* z,
                                                         const double* sv, const double* t, double* r, const double* rh,
                                                         const double* coef, double* partial, const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int64_t base = static_cast<int64_t>(rr) * n;
    const double alpha = coef[4 * rr], omega = coef[4 * rr + 1];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t o = base + i;
        x[o] = x[o] + (alpha * y[o] + omega * z[o]);
        const double rv = sv[o] - omega * t[o];
        r[o] = rv;
        a0 += rv * rv;
        a1 += fabs(rv);
        a2 += rh[o] * rv;
    }
    writePartial3(a0, a1, a2, partial, rr);
}

// OpenFOAM normalisation (krylov.cpp:47-57): sums (|A x0 - xm A 1|_1, |b - xm A 1|_1), xm = coef[4 r]
__global__ __launch_bounds__(kRedThreads) void k_of_norm(int64_t n, const double* ax, const double* b, const double* a1,
                                                         const double* coef, double* partial, const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int64_t base = static_cast<int64_t>(rr) * n;
    const double xm = coef[4 * rr];
    double a0 = 0.0, b0 = 0.0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double m = xm * a1[i];
        a0 += fabs(ax[base + i] - m);
        b0 += fabs(b[base + i] - m);
    }
    writePartial3(a0, b0, 0.0, partial, rr);
}

// The scalar recurrences and decisions of BiCGStab (bicgstab.cpp:15-154) for every right-hand
// side, one thread each, in FP64 with the reference's order of operations; stage by stage they
// read the sums of the preceding launch and write the masks and coefficients of the next ones.
__device__ inline bool biSatisfied(const BiArgs& a, const BiState& st, double slack)
{
    return a.of ? (st.l1 <= slack * st.thr) : (st.rn <= slack * a.rtol * st.bnorm);
}
__device__ inline void biPush(const BiArgs& a, BiState& st, int q, double v)
{
    if (st.nhist < a.maxit + 1) a.hist[static_cast<int64_t>(q) * (a.maxit + 1) + st.nhist] = v;
    ++st.nhist;
}

__global__ __launch_bounds__(64) void k_bicg_control(BiArgs a, int stage)
{
    __shared__ int live;
    if (threadIdx.x == 0) live = 0;
    __syncthreads();
    const int R = a.nrhs;
    auto M = [&](int m, int q) -> int32_t& { return a.mask[m * R + q]; };
    auto C = [&](int m, int q, int k) -> double& { return a.coef[static_cast<int64_t>(m) * 4 * R + 4 * q + k]; };
    for (int q = threadIdx.x; q < R; q += blockDim.x) {
        BiState st = a.st[q];
        const double* red = a.red + 3 * q;
        switch (stage) {
        case kBiBnorm: {
            st = BiState{};
            st.bnorm = sqrt(red[0]);
            st.normF = 1.0;
            st.rho_old = st.alpha = st.omega = 1.0;
            st.first = 1;
            const bool zero = st.bnorm == 0.0;  // x = 0, converged (bicgstab.cpp:24-29)
            // a non-finite right-hand side ends the solve as a breakdown (the reference's tests
            // against rtol * ||b|| would pass for ||b|| = inf and report convergence)
            const bool bad = !isfinite(st.bnorm);
            if (zero || bad) {
                st.done = 1;
                st.conv = zero ? 1 : 0;
                st.brk = bad ? 1 : 0;
                biPush(a, st, q, zero ? 0.0 : st.bnorm);
            }
            M(kBiZero, q) = zero ? 0 : 1;
            M(kBiLive, q) = (zero || bad) ? 1 : 0;
            for (int m = kBiNom; m < kBiMasks; ++m) M(m, q) = 1;
            for (int m = 0; m < kBiMasks; ++m)
                for (int k = 0; k < 4; ++k) C(m, q, k) = 0.0;
            break;
        }
        case kBiXmean: {  // OpenFOAM normalisation: the mean of x0 decides the form of F
            const bool on = !M(kBiLive, q);
            const double xm = a.n ? red[1] / static_cast<double>(a.n) : 0.0;
            M(kBiNom, q) = on && xm == 0.0 ? 0 : 1;
            M(kBiWithm, q) = on && xm != 0.0 ? 0 : 1;
            C(kBiWithm, q, 0) = on ? xm : 0.0;
            break;
        }
        case kBiNomAx:
            if (!M(kBiNom, q)) st.normF = red[1];
            break;
        case kBiNomB:
            if (!M(kBiNom, q)) st.normF += red[1] + 1e-20;
            break;
        case kBiWithmNorm:
            if (!M(kBiWithm, q)) st.normF = red[0] + red[1] + 1e-20;
            break;
        case kBiR0:
            if (!M(kBiLive, q)) {
                st.rn = sqrt(red[0]);
                st.l1 = red[1];
                st.l1_0 = st.l1;
                if (a.of) st.thr = fmax(a.rtol * st.l1, a.atol * st.normF);
                biPush(a, st, q, st.rn / st.bnorm);
            }
            break;
        case kBiCycle: {  // cycle start (bicgstab.cpp:73-82), then the first half's participants
            bool start = false;
            if (!st.done && st.phase == 0) {
                if (biSatisfied(a, st, 1.0)) {
                    st.conv = st.done = 1;
                } else if (st.its >= a.maxit) {
                    st.done = 1;
                } else {
                    st.rhn = st.rn;
                    st.rho = st.rn * st.rn;
                    st.rho_old = st.alpha = st.omega = 1.0;
                    st.first = 1;
                    st.candidate = 0;
                    st.phase = 1;
                    start = true;
                }
            }
            bool step = false;
            if (!st.done && st.phase == 1) {
                if (st.its >= a.maxit)
                    st.phase = 2;
                else
                    step = true;
            }
            M(kBiStart, q) = start ? 0 : 1;
            M(kBiP, q) = step ? 0 : 1;
            C(kBiP, q, 0) = step ? (st.first ? 0.0 : (st.rho / st.rho_old) * (st.alpha / st.omega)) : 0.0;
            C(kBiP, q, 1) = step ? st.omega : 0.0;
            C(kBiP, q, 2) = step ? (st.first ? 1.0 : 0.0) : 0.0;
            break;
        }
        case kBiAlpha: {  // alpha = rho / (r^, v)
            bool half = false;
            if (!M(kBiP, q)) {
                st.first = 0;
                ++st.its;
                const double rhv = red[0];
                if (rhv == 0.0 || !isfinite(rhv)) {
                    st.brk = 1;
                    st.phase = 2;
                } else {
                    st.alpha = st.rho / rhv;
                    half = true;
                }
            }
            M(kBiHalf, q) = half ? 0 : 1;
            C(kBiHalf, q, 0) = half ? st.alpha : 0.0;
            break;
        }
        case kBiHalfStep: {  // norms of s: the half step may end the cycle
            bool hx = false, full = false;
            if (!M(kBiHalf, q)) {
                st.rn = sqrt(red[0]);
                st.l1 = red[1];
                biPush(a, st, q, st.rn / st.bnorm);
                const bool sat = biSatisfied(a, st, 1.0);
                if (sat || st.its >= a.maxit) {
                    st.candidate = sat ? 1 : 0;
                    st.phase = 2;
                    hx = true;  // x += alpha y; r = s
                } else {
                    full = true;
                }
            }
            M(kBiHx, q) = hx ? 0 : 1;
            C(kBiHx, q, 0) = hx ? st.alpha : 0.0;
            M(kBiFull, q) = full ? 0 : 1;
            break;
        }
        case kBiOmega:  // omega = (t, s) / (t, t)
            if (!M(kBiFull, q)) {
                ++st.its;
                const double d0 = red[0], d1 = red[1];
                st.omega = d0 > 0.0 ? d1 / d0 : 0.0;
                C(kBiFull, q, 0) = st.alpha;
                C(kBiFull, q, 1) = st.omega;
            }
            break;
        case kBiXr:
            if (!M(kBiFull, q)) {
                st.rho_old = st.rho;
                st.rho = red[2];
                st.rn = sqrt(red[0]);
                st.l1 = red[1];
                biPush(a, st, q, st.rn / st.bnorm);
                if (biSatisfied(a, st, 1.0)) {
                    st.candidate = 1;
                    st.phase = 2;
                } else if (st.omega == 0.0 || fabs(st.rho) < 1e-14 * st.rhn * st.rn) {
                    st.phase = 2;  // restart with r^ = r
                } else if (st.its >= a.maxit) {
                    st.phase = 2;
                }
            }
            M(kBiEnd, q) = (!st.done && st.phase == 2) ? 0 : 1;
            break;
        case kBiEndCycle:  // true residual at the end of a cycle (bicgstab.cpp:138-151)
            if (!M(kBiEnd, q)) {
                st.rn = sqrt(red[0]);
                st.l1 = red[1];
                st.phase = 0;
                if (st.brk)
                    st.done = 1;
                else if (st.candidate && biSatisfied(a, st, 1.1))
                    st.conv = st.done = 1;
                else if (st.its >= a.maxit)
                    st.done = 1;
                else
                    ++st.restarts;
            }
            if (!st.done) atomicAdd(&live, 1);
            break;
        default:
            break;
        }
        a.st[q] = st;
    }
    __syncthreads();
    if (stage == kBiEndCycle && threadIdx.x == 0) a.live[0] = live;
}

__global__ void k_fill_masked(int64_t n, double v, double* y, const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[static_cast<int64_t>(rr) * n + i] = v;
}

__global__ void k_fill(int64_t n, double v, double* y)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

// stop at the half step: x += alpha y; r = s
__global__ void k_bicg_half(int64_t n, double* x, const double* y, const double* sv, double* r, const double* coef,
                            const int32_t* skip)
{
    const int rr = blockIdx.y;
    if (skip && skip[rr]) return;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = static_cast<int64_t>(rr) * n + i;
    x[o] = x[o] + coef[4 * rr] * y[o];
    r[o] = sv[o];
}

// ------------------------------------------------------------------ GMRES

// Basis layout: V[(k * nrhs + r) * n + i], so that column k of every right-hand side is one
// contiguous n x nrhs block (the input layout of the preconditioner and the SpMV).
__device__ inline const double* basis(const GmArgs& a, int r, int k) { return a.V + (static_cast<int64_t>(k) * a.nrhs + r) * a.n; }
__device__ inline double* basisW(const GmArgs& a, int r, int k) { return a.V + (static_cast<int64_t>(k) * a.nrhs + r) * a.n; }

// Per-block partials of h[k] = V_k . w for k <= j (summed by k_gm_final).
__device__ void gmDotsBody(const GmArgs& a, int j, int r, const double* w)
{
    __shared__ double red[kGmThreads / kWave * kGmMaxK];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (int k0 = 0; k0 <= j; k0 += 8) {
        double s[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) s[q] = 0.0;
        for (int64_t i = i0; i < a.n; i += stride) {
            const double wi = w[i];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (k0 + q <= j) s[q] += basis(a, r, k0 + q)[i] * wi;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double v = waveSum(s[q]);
            if (lane == 0 && k0 + q <= j) red[wave * kGmMaxK + k0 + q] = v;
        }
    }
    __syncthreads();
    double* pr = a.partial + (static_cast<int64_t>(r) * gridDim.x + blockIdx.x) * kGmMaxK;
    for (int k = threadIdx.x; k <= j; k += blockDim.x) {
        double s = 0.0;
        for (int wv = 0; wv < kGmThreads / kWave; ++wv) s += red[wv * kGmMaxK + k];
        pr[k] = s;
    }
}

__global__ __launch_bounds__(kGmThreads) void k_gm_dots(GmArgs a, int j)
{
    const int r = blockIdx.y;
    if (a.flags[r]) return;
    gmDotsBody(a, j, r, a.w + static_cast<int64_t>(r) * a.n);
}

// w = A z with the rows of k_spmv (the same fma order: the same bits), then the partials of
// k_gm_dots over the same grid (the same bits): one launch for the SpMV of the Arnoldi step and the
// first pass of CGS2. A thread's rows are the ones it reads back for the dots.
__global__ __launch_bounds__(kGmThreads) void k_gm_spmv_dots(GmArgs a, int j, const int32_t* rp, const int32_t* col,
                                                             const double* val, const double* z)
{
    const int r = blockIdx.y;
    if (a.flags[r]) return;
    const int64_t off = static_cast<int64_t>(r) * a.n;
    double* w = a.w + off;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n; i += stride)
        w[i] = spmvRow(rp, col, val, z + off, static_cast<int>(i));
    gmDotsBody(a, j, r, w);
}

// w -= V h, then partials of h2 = V^T w.
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
#pragma unroll 8
        for (int k = 0; k <= j; ++k) s += basis(a, r, k)[i] * hs[k];
        w[i] -= s;
    }
    gmDotsBody(a, j, r, w);
}

// w -= V h2 and the partials of ||w||^2 (summed by k_gm_givens).
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
#pragma unroll 8
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
    // the partials of ||w||^2 one after another in this right-hand side's region (its h2 partials
    // there were summed by the previous launch), so each workgroup of k_gm_givens_normalize reads
    // them as one contiguous block
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int wv = 0; wv < kGmThreads / kWave; ++wv) s += red[wv];
        a.partial[static_cast<int64_t>(r) * gridDim.x * kGmMaxK + blockIdx.x] = s;
    }
}

// h (second == 0) or h2 (second == 1) = sums of the partials of k_gm_dots / k_gm_update_dots;
// grid (j + 1, nrhs).
__global__ __launch_bounds__(kBlockThreads) void k_gm_final(GmArgs a, int j, int second)
{
    const int k = blockIdx.x, r = blockIdx.y;
    if (a.flags[r]) return;
    const double* base = a.partial + static_cast<int64_t>(r) * a.nblocks * kGmMaxK + k;
    double v = 0.0;
#pragma unroll 4
    for (int b = threadIdx.x; b < a.nblocks; b += blockDim.x) v += base[static_cast<int64_t>(b) * kGmMaxK];
    const double t = blockSum(v);
    if (threadIdx.x == 0) (second ? a.h2 : a.h)[r * (a.m + 1) + k] = t;
}

// The end of the Arnoldi step for right-hand side r with hn = ||w||: Hessenberg column, Givens
// rotations, residual estimate and the stopping test (gmres.cpp:149-208); one thread.
__device__ void gmGivensStep(const GmArgs& a, int j, int r, double hn)
{
    const int m = a.m;
    double* H = a.H + static_cast<int64_t>(r) * (m + 1) * m;
    double* Hc = H + static_cast<int64_t>(j) * (m + 1);
    double* cs = a.cs + r * m;
    double* sn = a.sn + r * m;
    double* g = a.g + r * (m + 1);
    const double* h = a.h + r * (m + 1);
    const double* h2 = a.h2 + r * (m + 1);
    for (int i = 0; i <= j; ++i) Hc[i] = h[i] + h2[i];
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
    if (a.of) {
        // the OpenFOAM test needs v_{j+1}: k_gm_of_final decides and sets the flag
    } else if (st.rel <= a.rtol) {
        st.conv = 1;
        done = 1;
    } else if (hn == 0.0) {
        st.breakdown = 1;
        done = 1;
    }
    if (!isfinite(st.rel) || !isfinite(hn)) done = 1;  // the host ends the solve (non-finite residual)
    if (st.iters >= st.maxit || j + 1 >= m) done = 1;
    st.done = done;
    if (!a.of) a.flags[r] = done;
}

// hn = ||w|| from the partials of k_gm_update_norm (every workgroup sums them in the same order,
// so all hold the same bits), the Givens step (workgroup 0) and v_{j+1} = w / hn. Workgroup 0 may
// set the right-hand side's flag while others start: one that reads it set skips its rows of
// v_{j+1}, which the finished cycle does not read.
__global__ __launch_bounds__(kBlockThreads) void k_gm_givens_normalize(GmArgs a, int j)
{
    const int r = blockIdx.y;
    if (a.flags[r]) return;
    const double* base = a.partial + static_cast<int64_t>(r) * a.nblocks * kGmMaxK;
    double v = 0.0;
#pragma unroll 4
    for (int b = threadIdx.x; b < a.nblocks; b += blockDim.x) v += base[b];
    const double ss = blockSum(v);
    __shared__ double hs;
    if (threadIdx.x == 0) hs = sqrt(ss);
    __syncthreads();
    const double hn = hs;
    if (blockIdx.x == 0 && threadIdx.x == 0) gmGivensStep(a, j, r, hn);
    if (!(hn > 0.0)) return;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    basisW(a, r, j + 1)[i] = a.w[static_cast<int64_t>(r) * a.n + i] / hn;
}

// OpenFOAM stopping test of iteration j (gmres.cpp:106-122): the GMRES residual is
// r_j = g_{j+1} V_{j+2} Q^T e_{j+1}, Q^T = G_0^T ... G_j^T; its 1-norm is compared with st.thr.
// Partials of sum_i |sum_k V_k[i] q_k| per block; every block forms q from the rotations in the
// same order.
__global__ __launch_bounds__(kGmThreads) void k_gm_of_partial(GmArgs a, int j)
{
    const int r = blockIdx.y;
    if (a.flags[r]) return;
    __shared__ double q[kGmMaxK + 1];
    const double* cs = a.cs + r * a.m;
    const double* sn = a.sn + r * a.m;
    if (threadIdx.x == 0) {
        for (int i = 0; i <= j; ++i) q[i] = 0.0;
        q[j + 1] = 1.0;
        for (int i = j; i >= 0; --i) {
            const double u = q[i], v = q[i + 1];
            q[i] = cs[i] * u - sn[i] * v;
            q[i + 1] = sn[i] * u + cs[i] * v;
        }
    }
    __syncthreads();
    double s1 = 0.0;
    if (a.st[r].hn > 0.0) {  // hn == 0: v_{j+1} was not formed and the residual is zero
        const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
        for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n; i += stride) {
            double v = 0.0;
            for (int k = 0; k <= j + 1; ++k) v += basis(a, r, k)[i] * q[k];
            s1 += fabs(v);
        }
    }
    const double t = blockSum(s1);
    if (threadIdx.x == 0) a.partial[static_cast<int64_t>(r) * gridDim.x * kGmMaxK + blockIdx.x] = t;
}

// sum|r_j| = |g_{j+1}| x (the partials in block order); convergence and the right-hand side's flag.
__global__ __launch_bounds__(kBlockThreads) void k_gm_of_final(GmArgs a, int j)
{
    const int r = blockIdx.y;
    if (a.flags[r]) return;
    const double* base = a.partial + static_cast<int64_t>(r) * a.nblocks * kGmMaxK;
    double v = 0.0;
#pragma unroll 4
    for (int b = threadIdx.x; b < a.nblocks; b += blockDim.x) v += base[b];
    const double t = blockSum(v);
    if (threadIdx.x == 0) {
        GmState& st = a.st[r];
        const double gj = a.g[r * (a.m + 1) + j + 1];
        const double l1 = (st.hn == 0.0 || gj == 0.0) ? 0.0 : fabs(gj) * t;
        st.l1 = l1;
        if (l1 <= st.thr) {
            st.conv = 1;
            st.done = 1;
        }
        a.flags[r] = st.done;
    }
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

// u = B_k y (B = V, or Z = M^{-1} V for flexible GMRES)
__global__ void k_gm_combine(GmArgs a, const double* B, double* u, const int32_t* idle)
{
    const int r = blockIdx.y;
    if (idle[r]) return;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int k = a.st[r].k;
    const double* y = a.y + r * a.m;
    double s = 0.0;
#pragma unroll 8
    for (int l = 0; l < k; ++l) s += B[(static_cast<int64_t>(l) * a.nrhs + r) * a.n + i] * y[l];
    u[static_cast<int64_t>(r) * a.n + i] = s;
}

inline unsigned nblk(int64_t n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

} // namespace

// ------------------------------------------------------------------ launchers
template <typename T, typename TF> void launchLocalSolve(const LocalArgs<T, TF>& a, int grid, size_t lds, int threads, hipStream_t s)
{
    if (threads == 512)
        hipLaunchKernelGGL((k_local_solve<T, TF, 512>), dim3(grid), dim3(512), lds, s, a);
    else
        hipLaunchKernelGGL((k_local_solve<T, TF, 1024>), dim3(grid), dim3(1024), lds, s, a);
}
template void launchLocalSolve<float, float>(const LocalArgs<float, float>&, int, size_t, int, hipStream_t);
template void launchLocalSolve<double, double>(const LocalArgs<double, double>&, int, size_t, int, hipStream_t);
template void launchLocalSolve<double, float>(const LocalArgs<double, float>&, int, size_t, int, hipStream_t);
template void launchLocalSolve<float2, float>(const LocalArgs<float2, float>&, int, size_t, int, hipStream_t);
template void launchLocalSolve<double2, float>(const LocalArgs<double2, float>&, int, size_t, int, hipStream_t);

template <typename T, typename TF> const void* localSolveKernel(int threads)
{
    return threads == 512 ? reinterpret_cast<const void*>(&k_local_solve<T, TF, 512>)
                          : reinterpret_cast<const void*>(&k_local_solve<T, TF, 1024>);
}
template const void* localSolveKernel<float, float>(int);
template const void* localSolveKernel<double, double>(int);
template const void* localSolveKernel<double, float>(int);
template const void* localSolveKernel<float2, float>(int);
template const void* localSolveKernel<double2, float>(int);

template <typename T, typename TF>
void launchDenseTop(const LocalArgs<T, TF>& a, int grid, const int4* dt, const int32_t* dt_ce, const int32_t* dtb, size_t lds,
                    int cus, hipStream_t s)
{
    // FP64 entries or pairs of FP64 right-hand sides: two blocks of columns per batch (three need
    // more than the 128 registers)
    constexpr int kWideNB = (sizeof(TF) == 8 || sizeof(T) == 16) ? 2 : 3;
    if (grid <= cus)
        hipLaunchKernelGGL((k_dense_top<T, TF, kWideNB>), dim3(grid), dim3(kDtThreads), lds, s, a, dt, dt_ce, dtb);
    else
        hipLaunchKernelGGL((k_dense_top<T, TF, 1>), dim3(grid), dim3(kDtThreads), lds, s, a, dt, dt_ce, dtb);
}
template void launchDenseTop<float, float>(const LocalArgs<float, float>&, int, const int4*, const int32_t*, const int32_t*, size_t,
                                           int, hipStream_t);
template void launchDenseTop<double, double>(const LocalArgs<double, double>&, int, const int4*, const int32_t*, const int32_t*,
                                             size_t, int, hipStream_t);
template void launchDenseTop<double, float>(const LocalArgs<double, float>&, int, const int4*, const int32_t*, const int32_t*,
                                            size_t, int, hipStream_t);
template void launchDenseTop<float2, float>(const LocalArgs<float2, float>&, int, const int4*, const int32_t*, const int32_t*,
                                            size_t, int, hipStream_t);
template void launchDenseTop<double2, float>(const LocalArgs<double2, float>&, int, const int4*, const int32_t*, const int32_t*,
                                             size_t, int, hipStream_t);

void launchSpmv(int32_t n, const int32_t* rp, const int32_t* col, const double* val, const double* x, double* y,
                const double* b, int nrhs, const int32_t* skip, hipStream_t s)
{
    if (nrhs >= 2)
        hipLaunchKernelGGL(k_spmv_group, dim3(nblk(n, 256), (nrhs + kSpmvGroup - 1) / kSpmvGroup), dim3(256), 0, s, n, rp, col,
                           val, x, y, b, nrhs, skip);
    else
        hipLaunchKernelGGL(k_spmv, dim3(nblk(n, 256), nrhs), dim3(256), 0, s, n, rp, col, val, x, y, b, skip);
}

void launchCoarseRestrict(const CoarseArgs& c, const double* r, int64_t n, double* cvec, double* part, int nrhs,
                          const int32_t* skip, hipStream_t s)
{
    if (nrhs >= kMfmaMinRhs && c.mmax <= 16 * kMfmaTiles)
        hipLaunchKernelGGL(k_coarse_restrict_mfma, dim3(c.nsub, c.rchunks, (nrhs + 15) / 16), dim3(kCrows), 0, s, c, r, n, part,
                           nrhs, skip);
    else {
        hipLaunchKernelGGL(k_coarse_restrict_cols, dim3(c.nsub, (c.mmax + kCg - 1) / kCg, nrhs), dim3(kCgThreads), 0, s, c, r,
                           n, cvec, skip);
        return;
    }
    hipLaunchKernelGGL(k_coarse_final, dim3(c.nsub, nrhs), dim3(64), 0, s, c, part, cvec, skip);
}

void launchTriGemv(int32_t n0, int32_t ld, const double* T, int lower, const int32_t* perm, const double* x, double* y, int nrhs,
                   const int32_t* skip, hipStream_t s)
{
    if (nrhs >= kMfmaMinRhs) {
        const dim3 grid(nblk(n0, 16), (nrhs + 15) / 16), block(kTriWaves * kWave);
        if (perm)
            hipLaunchKernelGGL(k_tri_gemm_mfma<true>, grid, block, 0, s, n0, ld, T, lower, perm, x, y, nrhs, skip);
        else
            hipLaunchKernelGGL(k_tri_gemm_mfma<false>, grid, block, 0, s, n0, ld, T, lower, perm, x, y, nrhs, skip);
    } else {
        hipLaunchKernelGGL(k_tri_gemv, dim3(nblk(n0, 4), nrhs), dim3(256), 0, s, n0, ld, T, lower, perm, x, y, skip);
    }
}

void launchCoarseProlong(const CoarseArgs& c, const double* y, double* out, int64_t ld_out, int boolean, int nrhs,
                         const int32_t* skip, hipStream_t s)
{
    if (nrhs >= kMfmaMinRhs)
        hipLaunchKernelGGL(k_coarse_prolong_mfma, dim3(c.nblk, (nrhs + 15) / 16), dim3(256), 0, s, c, y, out, ld_out, boolean, nrhs,
                           skip);
    else
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

// Workgroups of the GMRES vector kernels: one element per thread up to 2048 workgroups of
// kGmThreads (a thread's elements are dependent round trips: four per thread cost a dot product
// about three times one), then a grid stride.
int gmBlocks(int64_t n)
{
    const int64_t b = (n + kGmThreads - 1) / kGmThreads;
    return static_cast<int>(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

int reduceBlocks(int64_t n)
{
    const int64_t b = (n + kRedThreads * 4 - 1) / (kRedThreads * 4);
    return static_cast<int>(b < 1 ? 1 : (b > 256 ? 256 : b));
}

namespace {
void finalSum3(const double* partial, int nb, double* out, int nrhs, const int32_t* skip, hipStream_t s)
{
    hipLaunchKernelGGL(k_final_sum, dim3(kRedComp, nrhs), dim3(kBlockThreads), 0, s, partial, kRedComp, nb, out, kRedComp,
                       skip);
}
} // namespace

void launchReduce(int kind, int64_t n, const double* x, const double* y, double* partial, double* out, int nrhs,
                  const int32_t* skip, hipStream_t s)
{
    const int nb = reduceBlocks(n);
    hipLaunchKernelGGL(k_reduce, dim3(nb, nrhs), dim3(kRedThreads), 0, s, kind, n, x, y, partial, skip);
    finalSum3(partial, nb, out, nrhs, skip, s);
}

void bicgP(int64_t n, const double* r, double* p, const double* v, const double* coef, int nrhs, const int32_t* skip,
           hipStream_t s)
{
    hipLaunchKernelGGL(k_bicg_p, dim3(nblk(n, 256), nrhs), dim3(256), 0, s, n, r, p, v, coef, skip);
}
void bicgS(int64_t n, const double* r, const double* v, double* sv, const double* coef, double* partial, double* out,
           int nrhs, const int32_t* skip, hipStream_t s)
{
    const int nb = reduceBlocks(n);
    hipLaunchKernelGGL(k_bicg_s, dim3(nb, nrhs), dim3(kRedThreads), 0, s, n, r, v, sv, coef, partial, skip);
    finalSum3(partial, nb, out, nrhs, skip, s);
}
void bicgXR(int64_t n, double* x, const double* y, const double* z, const double* sv, const double* t, double* r,
            const double* rh, const double* coef, double* partial, double* out, int nrhs, const int32_t* skip,
            hipStream_t s)
{
    const int nb = reduceBlocks(n);
    hipLaunchKernelGGL(k_bicg_xr, dim3(nb, nrhs), dim3(kRedThreads), 0, s, n, x, y, z, sv, t, r, rh, coef, partial, skip);
    finalSum3(partial, nb, out, nrhs, skip, s);
}
void bicgHalf(int64_t n, double* x, const double* y, const double* sv, double* r, const double* coef, int nrhs,
              const int32_t* skip, hipStream_t s)
{
    hipLaunchKernelGGL(k_bicg_half, dim3(nblk(n, 256), nrhs), dim3(256), 0, s, n, x, y, sv, r, coef, skip);
}
void ofNorm(int64_t n, const double* ax, const double* b, const double* a1, const double* coef, double* partial,
            double* out, int nrhs, const int32_t* skip, hipStream_t s)
{
    const int nb = reduceBlocks(n);
    hipLaunchKernelGGL(k_of_norm, dim3(nb, nrhs), dim3(kRedThreads), 0, s, n, ax, b, a1, coef, partial, skip);
    finalSum3(partial, nb, out, nrhs, skip, s);
}
void launchBicgControl(const BiArgs& a, int stage, hipStream_t s)
{
    hipLaunchKernelGGL(k_bicg_control, dim3(1), dim3(64), 0, s, a, stage);
}

void launchFillMasked(int64_t n, double v, double* y, int nrhs, const int32_t* skip, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_masked, dim3(nblk(n, 256), nrhs), dim3(256), 0, s, n, v, y, skip);
}

void launchFill(int64_t n, double v, double* y, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, s, n, v, y);
}

void gmSpmvDots(const GmArgs& a, int j, const int32_t* rp, const int32_t* col, const double* val, const double* z,
                hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_spmv_dots, dim3(a.nblocks, a.nrhs), dim3(kGmThreads), 0, s, a, j, rp, col, val, z);
}
void gmFinalH(const GmArgs& a, int j, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_final, dim3(j + 1, a.nrhs), dim3(kBlockThreads), 0, s, a, j, 0);
}
void gmDots(const GmArgs& a, int j, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_dots, dim3(a.nblocks, a.nrhs), dim3(kGmThreads), 0, s, a, j);
    hipLaunchKernelGGL(k_gm_final, dim3(j + 1, a.nrhs), dim3(kBlockThreads), 0, s, a, j, 0);
}
void gmUpdateDots(const GmArgs& a, int j, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_update_dots, dim3(a.nblocks, a.nrhs), dim3(kGmThreads), 0, s, a, j);
    hipLaunchKernelGGL(k_gm_final, dim3(j + 1, a.nrhs), dim3(kBlockThreads), 0, s, a, j, 1);
}
void gmUpdateNorm(const GmArgs& a, int j, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_update_norm, dim3(a.nblocks, a.nrhs), dim3(kGmThreads), 0, s, a, j);
}
void gmNormalize(const GmArgs& a, int j, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_givens_normalize, dim3(nblk(a.n, kBlockThreads), a.nrhs), dim3(kBlockThreads), 0, s, a, j);
}
void gmOpenFoamTest(const GmArgs& a, int j, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_of_partial, dim3(a.nblocks, a.nrhs), dim3(kGmThreads), 0, s, a, j);
    hipLaunchKernelGGL(k_gm_of_final, dim3(1, a.nrhs), dim3(kBlockThreads), 0, s, a, j);
}
void gmStartCycle(const GmArgs& a, const double* r, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_start, dim3(nblk(a.n, 256), a.nrhs), dim3(256), 0, s, a, r);
}
void gmLsq(const GmArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_lsq, dim3(nblk(a.nrhs, 64)), dim3(64), 0, s, a, a.flags + a.nrhs);
}
void gmCombine(const GmArgs& a, const double* B, double* u, hipStream_t s)
{
    hipLaunchKernelGGL(k_gm_combine, dim3(nblk(a.n, 256), a.nrhs), dim3(256), 0, s, a, B, u, a.flags + a.nrhs);
}

} // namespace s2l
