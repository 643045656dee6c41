// lsolve.hpp -- the work of the level-synchronous local solve, split into the phases between which a
// workgroup synchronizes. The kernels in lsolve.hip call these phases with the thread's wavefront
// and lane and __syncthreads() in between; every phase only reads LDS written before the last sync.
#ifndef S2L_LSOLVE_HPP
#define S2L_LSOLVE_HPP

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lsched.hpp"

namespace s2l {

constexpr int kLsLanes = 64;
constexpr int kLsThreads = kLsLanes * kLsWaves;

/// Arguments of the level-synchronous local-solve launches (device pointers).
template <typename T> struct LsArgs {
    const LsSn* sn;
    const LsUnit* units;
    const LsItem* items;
    const int32_t* pull_rows;
    const T* diag;          ///< packed column-major triangles of the inverted diagonal blocks
    const T* belowF;        ///< forward blocks below the diagonal, column-major
    const T* belowB;        ///< backward blocks, row-major
    const int32_t* ridx;    ///< rows below every block (factor positions)
    const int32_t* cptr;    ///< slot range of every concatenated local row
    const int32_t* cidx;    ///< slot of every push
    const int32_t* gmap;    ///< concatenated factor position -> global row
    const int32_t* lperm;   ///< concatenated factor position -> local index
    const int32_t* owned;   ///< n_I of every subdomain
    const int32_t* begin;   ///< first owned global row of every subdomain
    const double* rhs;      ///< n_global x nrhs
    const double* x0;       ///< mode 0: added to the owned rows (may be null)
    double* out;            ///< mode 0: n_global x nrhs; mode 1: local buffer n_local_total x nrhs
    const double* pou;      ///< mode 1: D_i by concatenated local index (null = 1)
    const int32_t* skip;    ///< per right-hand side, nonzero = nothing to do (may be null)
    T* slots;               ///< nslots per right-hand side
    T* part;                ///< backward partial dot products, npart per right-hand side
    T* bbuf;                ///< forward pulled right-hand sides, n_local_total per right-hand side
    T* yf;                  ///< forward results, n_local_total per right-hand side
    T* xb;                  ///< backward results, n_local_total per right-hand side
    int64_t n_global, n_local_total, nslots, npart;
    int32_t mode;
    int32_t lds_w;          ///< values of the pulled-row buffer of a large unit (widest supernode of the launch)
};

/// LDS of a unit workgroup, in values: large: rows [0, lds_w), wave partials [lds_w, lds_w + 256),
/// block result [lds_w + 256, lds_w + 320); small: wave q owns [128 q, 128 q + 128).
__host__ __device__ inline int lsUnitLds(int lds_w)
{
    const int large = lds_w + kLsThreads + kLsLanes;
    const int small = kLsWaves * 2 * kLsLanes;
    return large > small ? large : small;
}

/// Offset of column c in a packed lower (rows c..w-1) or upper (rows 0..c) triangle.
__host__ __device__ inline int64_t triLowerCol(int64_t c, int64_t w) { return c * w - c * (c - 1) / 2; }
__host__ __device__ inline int64_t triUpperCol(int64_t c) { return c * (c + 1) / 2; }

/// Columns [c0, c1) of wavefront q when four wavefronts split w columns.
__host__ __device__ inline void lsWaveCols(int w, int q, int& c0, int& c1)
{
    const int cw = (w + kLsWaves - 1) / kLsWaves;
    c0 = q * cw;
    c1 = (c0 + cw < w) ? c0 + cw : w;
}

// ------------------------------------------------------------------ forward pull

/// Phase A (thread t of the block, global thread gid): four lanes per row sum strided parts of the
/// row's slots into scr[t].
template <typename T>
__host__ __device__ inline void lsPullA(const LsArgs<T>& a, int32_t first, int32_t count, int64_t gid, int t, int rr, T* scr)
{
    const int64_t ri = gid / kLsPullLanes;
    const int g = static_cast<int>(gid % kLsPullLanes);
    T v = 0;
    if (ri < count) {
        const int32_t row = a.pull_rows[first + ri];
        const T* tb = a.slots + static_cast<int64_t>(rr) * a.nslots;
        const int32_t q1 = a.cptr[row + 1];
#pragma unroll 4
        for (int32_t q = a.cptr[row] + g; q < q1; q += kLsPullLanes) v += tb[q];
    }
    scr[t] = v;
}

/// Phase B: the first lane of a row writes b = float(rhs) - its slots.
template <typename T>
__host__ __device__ inline void lsPullB(const LsArgs<T>& a, int32_t first, int32_t count, int64_t gid, int t, int rr, const T* scr)
{
    const int64_t ri = gid / kLsPullLanes;
    if (gid % kLsPullLanes != 0 || ri >= count) return;
    const int32_t row = a.pull_rows[first + ri];
    const T s = (scr[t] + scr[t + 1]) + (scr[t + 2] + scr[t + 3]);
    const T b = static_cast<T>(a.rhs[static_cast<int64_t>(rr) * a.n_global + a.gmap[row]]);
    a.bbuf[static_cast<int64_t>(rr) * a.n_local_total + row] = b - s;
}

// ------------------------------------------------------------------ forward unit

/// Lanes per row of a small supernode's pulls: 64 / (w rounded up to a power of two).
__host__ __device__ inline int lsLanesPerRow(int w)
{
    int p = 1;
    while (p < w) p <<= 1;
    return kLsLanes / p;
}

/// Phase A: a large unit loads its rows' pulled right-hand side (from the pull launch) into LDS; a
/// small supernode pulls its rows itself, lsLanesPerRow lanes per row summing strided parts of
/// the row's slots into LDS.
template <typename T>
__host__ __device__ inline void lsFwdA(const LsArgs<T>& a, const LsUnit& u, int q, int l, int rr, T* lds)
{
    if (u.large) {
        const T* b = a.bbuf + static_cast<int64_t>(rr) * a.n_local_total;
        const LsSn& d = a.sn[u.sn[0]];
        for (int r = q * kLsLanes + l; r < d.w; r += kLsThreads) lds[r] = b[d.row0 + d.c0 + r];
    } else if (u.sn[q] >= 0) {
        const LsSn& d = a.sn[u.sn[q]];
        const int G = lsLanesPerRow(d.w);
        const int r = l / G, g = l % G;
        T v = 0;
        if (r < d.w) {
            const int32_t row = d.row0 + d.c0 + r;
            const T* tb = a.slots + static_cast<int64_t>(rr) * a.nslots;
            const int32_t q1 = a.cptr[row + 1];
#pragma unroll 4
            for (int32_t s = a.cptr[row] + g; s < q1; s += G) v += tb[s];
        }
        lds[q * 2 * kLsLanes + kLsLanes + l] = v;
    }
}

/// Phase A2 (small supernodes): b of row l = float(rhs) - the sum of its lanes' parts.
template <typename T>
__host__ __device__ inline void lsFwdA2(const LsArgs<T>& a, const LsUnit& u, int q, int l, int rr, T* lds)
{
    if (u.large || u.sn[q] < 0) return;
    const LsSn& d = a.sn[u.sn[q]];
    if (l >= d.w) return;
    const int G = lsLanesPerRow(d.w);
    const T* sp = lds + q * 2 * kLsLanes + kLsLanes + l * G;
    T acc = 0;
    for (int g = 0; g < G; ++g) acc += sp[g];
    const int32_t row = d.row0 + d.c0 + l;
    const T b = static_cast<T>(a.rhs[static_cast<int64_t>(rr) * a.n_global + a.gmap[row]]);
    lds[q * 2 * kLsLanes + l] = b - acc;
}

/// Phase B: y = inv(L_ss) b; a large unit: partial products of wavefront q's columns.
template <typename T>
__host__ __device__ inline void lsFwdB(const LsArgs<T>& a, const LsUnit& u, int q, int l, int rr, T* lds)
{
    if (u.large) {
        const LsSn& d = a.sn[u.sn[0]];
        const int r = u.part * kLsLanes + l;
        int c0, c1;
        lsWaveCols(d.w, q, c0, c1);
        if (c1 > r + 1) c1 = r + 1;
        const T* D = a.diag + d.diag;
        T p = 0;
        if (r < d.w) {
#pragma unroll 8
            for (int c = c0; c < c1; ++c) p += D[triLowerCol(c, d.w) + (r - c)] * lds[c];
        }
        lds[a.lds_w + q * kLsLanes + l] = p;
    } else if (u.sn[q] >= 0) {
        const LsSn& d = a.sn[u.sn[q]];
        const T* bq = lds + q * 2 * kLsLanes;
        T y = 0;
        if (l < d.w) {
            const T* D = a.diag + d.diag;
#pragma unroll 8
            for (int c = 0; c <= l; ++c) y += D[triLowerCol(c, d.w) + (l - c)] * bq[c];
            a.yf[static_cast<int64_t>(rr) * a.n_local_total + d.row0 + d.c0 + l] = y;
        }
        lds[q * 2 * kLsLanes + kLsLanes + l] = y;
    }
}

/// Phase C (large units): wavefront 0 adds the four partials in order.
template <typename T>
__host__ __device__ inline void lsFwdC(const LsArgs<T>& a, const LsUnit& u, int q, int l, int rr, T* lds)
{
    if (!u.large || q != 0) return;
    const LsSn& d = a.sn[u.sn[0]];
    const T* p = lds + a.lds_w;
    const int r = u.part * kLsLanes + l;
    const T y = (p[l] + p[kLsLanes + l]) + (p[2 * kLsLanes + l] + p[3 * kLsLanes + l]);
    if (r < d.w) a.yf[static_cast<int64_t>(rr) * a.n_local_total + d.row0 + d.c0 + r] = y;
    lds[a.lds_w + kLsThreads + l] = (r < d.w) ? y : T(0);
}

/// Phase D: push the products of the rows below with the unit's columns, one slot per (row, block).
template <typename T>
__host__ __device__ inline void lsFwdD(const LsArgs<T>& a, const LsUnit& u, int q, int l, int rr, const T* lds)
{
    T* tb = a.slots + static_cast<int64_t>(rr) * a.nslots;
    if (u.large) {
        const LsSn& d = a.sn[u.sn[0]];
        const int cb = u.part * kLsLanes;
        const int ce = (d.w < cb + kLsLanes) ? d.w : cb + kLsLanes;
        const T* Bm = a.belowF + d.below;
        const T* y = lds + a.lds_w + kLsThreads;
        const int32_t* dst = a.cidx + d.toff + static_cast<int64_t>(u.part) * d.h;
        for (int k = q * kLsLanes + l; k < d.h; k += kLsThreads) {
            T acc = 0;
#pragma unroll 8
            for (int c = cb; c < ce; ++c) acc += Bm[static_cast<int64_t>(c) * d.h + k] * y[c - cb];
            tb[dst[k]] = acc;
        }
    } else if (u.sn[q] >= 0) {
        const LsSn& d = a.sn[u.sn[q]];
        const T* Bm = a.belowF + d.below;
        const T* y = lds + q * 2 * kLsLanes + kLsLanes;
        const int32_t* dst = a.cidx + d.toff;
        for (int k = l; k < d.h; k += kLsLanes) {
            T acc = 0;
#pragma unroll 8
            for (int c = 0; c < d.w; ++c) acc += Bm[static_cast<int64_t>(c) * d.h + k] * y[c];
            tb[dst[k]] = acc;
        }
    }
}

// ------------------------------------------------------------------ backward

/// Partial dot products of a large supernode: rows [64 rb, 64 rb + 64) with ancestor rows
/// [64 kc, 64 kc + 64) of the block below (one wavefront per item).
template <typename T>
__host__ __device__ inline void lsBackPartial(const LsArgs<T>& a, const LsItem& it, int l, int rr)
{
    const LsSn& d = a.sn[it.sn];
    const int rb = it.part / d.nkc, kc = it.part % d.nkc;
    const int r = rb * kLsLanes + l;
    if (r >= d.w) return;
    const int k0 = kc * kLsKChunk;
    const int k1 = (d.h < k0 + kLsKChunk) ? d.h : k0 + kLsKChunk;
    const T* Bm = a.belowB + d.below;
    const int32_t* ri = a.ridx + d.ridx;
    const T* x = a.xb + static_cast<int64_t>(rr) * a.n_local_total + d.row0;
    T acc = 0;
#pragma unroll 8
    for (int k = k0; k < k1; ++k) acc += Bm[static_cast<int64_t>(k) * d.w + r] * x[ri[k]];
    a.part[static_cast<int64_t>(rr) * a.npart + d.poff + static_cast<int64_t>(kc) * d.w + r] = acc;
}

/// Phase A: large units: b = forward result - the partials, all rows into LDS; small supernodes:
/// partial dot products of the rows with the ancestors' values, lsLanesPerRow lanes per row.
template <typename T>
__host__ __device__ inline void lsBwdA(const LsArgs<T>& a, const LsUnit& u, int q, int l, int rr, T* lds)
{
    const int64_t vo = static_cast<int64_t>(rr) * a.n_local_total;
    if (u.large) {
        const LsSn& d = a.sn[u.sn[0]];
        const T* yf = a.yf + vo + d.row0 + d.c0;
        const T* P = a.part + static_cast<int64_t>(rr) * a.npart + d.poff;
        for (int r = q * kLsLanes + l; r < d.w; r += kLsThreads) {
            T b = yf[r];
#pragma unroll 4
            for (int kc = 0; kc < d.nkc; ++kc) b -= P[static_cast<int64_t>(kc) * d.w + r];
            lds[r] = b;
        }
    } else if (u.sn[q] >= 0) {
        const LsSn& d = a.sn[u.sn[q]];
        const int G = lsLanesPerRow(d.w);
        const int r = l / G, g = l % G;
        T acc = 0;
        if (r < d.w) {
            const T* Bm = a.belowB + d.below;
            const int32_t* ri = a.ridx + d.ridx;
            const T* x = a.xb + vo + d.row0;
#pragma unroll 8
            for (int k = g; k < d.h; k += G) acc += Bm[static_cast<int64_t>(k) * d.w + r] * x[ri[k]];
        }
        lds[q * 2 * kLsLanes + kLsLanes + l] = acc;
    }
}

/// Phase B (small supernodes): b of row l = forward result - the sum of its lanes' partials.
template <typename T>
__host__ __device__ inline void lsBwdB(const LsArgs<T>& a, const LsUnit& u, int q, int l, int rr, T* lds)
{
    if (u.large || u.sn[q] < 0) return;
    const LsSn& d = a.sn[u.sn[q]];
    if (l >= d.w) return;
    const int G = lsLanesPerRow(d.w);
    const T* s = lds + q * 2 * kLsLanes + kLsLanes + l * G;
    T acc = 0;
    for (int g = 0; g < G; ++g) acc += s[g];
    lds[q * 2 * kLsLanes + l] = a.yf[static_cast<int64_t>(rr) * a.n_local_total + d.row0 + d.c0 + l] - acc;
}

/// Result of row r of supernode d: to the backward results and the output.
template <typename T>
__host__ __device__ inline void lsBwdOut(const LsArgs<T>& a, const LsSn& d, int r, int rr, T x)
{
    const int32_t row = d.row0 + d.c0 + r;
    a.xb[static_cast<int64_t>(rr) * a.n_local_total + row] = x;
    const int32_t lidx = a.lperm[row];
    const double y = static_cast<double>(x);
    if (a.mode == 0) {
        if (lidx < a.owned[d.sub]) {
            const int64_t gi = static_cast<int64_t>(rr) * a.n_global + a.begin[d.sub] + lidx;
            a.out[gi] = a.x0 ? a.x0[gi] + y : y;
        }
    } else {
        const int64_t li = static_cast<int64_t>(d.row0) + lidx;
        a.out[static_cast<int64_t>(rr) * a.n_local_total + li] = a.pou ? a.pou[li] * y : y;
    }
}

/// Phase C: x = inv(U_ss) b; a large unit: partial products of wavefront q's columns.
template <typename T>
__host__ __device__ inline void lsBwdC(const LsArgs<T>& a, const LsUnit& u, int q, int l, int rr, T* lds)
{
    if (u.large) {
        const LsSn& d = a.sn[u.sn[0]];
        const int r = u.part * kLsLanes + l;
        int c0, c1;
        lsWaveCols(d.w, q, c0, c1);
        if (c0 < r) c0 = r;
        const T* D = a.diag + d.diag;
        T p = 0;
#pragma unroll 8
        for (int c = c0; c < c1; ++c) p += D[triUpperCol(c) + r] * lds[c];
        lds[a.lds_w + q * kLsLanes + l] = p;
    } else if (u.sn[q] >= 0) {
        const LsSn& d = a.sn[u.sn[q]];
        if (l >= d.w) return;
        const T* bq = lds + q * 2 * kLsLanes;
        const T* D = a.diag + d.diag;
        T x = 0;
#pragma unroll 8
        for (int c = l; c < d.w; ++c) x += D[triUpperCol(c) + l] * bq[c];
        lsBwdOut(a, d, l, rr, x);
    }
}

/// Phase D (large units): wavefront 0 adds the four partials in order and writes the rows.
template <typename T>
__host__ __device__ inline void lsBwdD(const LsArgs<T>& a, const LsUnit& u, int q, int l, int rr, const T* lds)
{
    if (!u.large || q != 0) return;
    const LsSn& d = a.sn[u.sn[0]];
    const int r = u.part * kLsLanes + l;
    if (r >= d.w) return;
    const T* p = lds + a.lds_w;
    const T x = (p[l] + p[kLsLanes + l]) + (p[2 * kLsLanes + l] + p[3 * kLsLanes + l]);
    lsBwdOut(a, d, r, rr, x);
}

} // namespace s2l

#endif
