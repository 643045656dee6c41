// lsolve.hpp -- the work of one item of the level-synchronous local solve, split into the phases
// between which a workgroup of one wavefront (64 lanes) synchronizes. The kernels in lsolve.hip call
// these phases with lane = threadIdx.x and __syncthreads() in between.
#ifndef S2L_LSOLVE_HPP
#define S2L_LSOLVE_HPP

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lsched.hpp"

namespace s2l {

constexpr int kLsLanes = 64;

/// Arguments of the level-synchronous local-solve launches (device pointers).
template <typename T> struct LsArgs {
    const LsSn* sn;
    const LsItem* items;
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
    T* yf;                  ///< forward results, n_local_total per right-hand side
    T* xb;                  ///< backward results, n_local_total per right-hand side
    int64_t n_global, n_local_total, nslots, npart;
    int32_t mode;
    int32_t lds_w;          ///< values of the pulled-row buffer in LDS (widest supernode of the launch)
};

/// Offset of column c in a packed lower (rows c..w-1) or upper (rows 0..c) triangle.
__host__ __device__ inline int64_t triLowerCol(int64_t c, int64_t w) { return c * w - c * (c - 1) / 2; }
__host__ __device__ inline int64_t triUpperCol(int64_t c) { return c * (c + 1) / 2; }

// ---- forward item (supernode d, row block j)

/// Phase 1: b of every row of the supernode = float(rhs) - the products its descendants pushed.
template <typename T>
__host__ __device__ inline void lsForwardPull(const LsArgs<T>& a, const LsSn& d, int rr, int lane, T* bl)
{
    const int32_t rowc = d.row0 + d.c0;
    const double* rhs = a.rhs + static_cast<int64_t>(rr) * a.n_global;
    const T* tb = a.slots + static_cast<int64_t>(rr) * a.nslots;
    for (int r = lane; r < d.w; r += kLsLanes) {
        const int32_t row = rowc + r;
        T b = static_cast<T>(rhs[a.gmap[row]]);
        const int32_t q1 = a.cptr[row + 1];
#pragma unroll 8
        for (int32_t q = a.cptr[row]; q < q1; ++q) b -= tb[q];
        bl[r] = b;
    }
}

/// Phase 2: rows [64 j, 64 j + 64) of y = inv(L_ss) b; y to the forward results and to yl.
template <typename T>
__host__ __device__ inline void lsForwardDiag(const LsArgs<T>& a, const LsSn& d, int j, int rr, int lane, const T* bl, T* yl)
{
    const int r = j * kLsLanes + lane;
    T y = 0;
    if (r < d.w) {
        const T* D = a.diag + d.diag;
#pragma unroll 8
        for (int c = 0; c <= r; ++c) y += D[triLowerCol(c, d.w) + (r - c)] * bl[c];
        a.yf[static_cast<int64_t>(rr) * a.n_local_total + d.row0 + d.c0 + r] = y;
    }
    yl[lane] = y;
}

/// Phase 3: push the products of the rows below with the block's columns, one slot per (row, block).
template <typename T>
__host__ __device__ inline void lsForwardPush(const LsArgs<T>& a, const LsSn& d, int j, int rr, int lane, const T* yl)
{
    const int cb = j * kLsLanes;
    const int ce = (d.w < cb + kLsLanes) ? d.w : cb + kLsLanes;
    const T* Bm = a.belowF + d.below;
    T* tb = a.slots + static_cast<int64_t>(rr) * a.nslots;
    const int32_t* dst = a.cidx + d.toff + static_cast<int64_t>(j) * d.h;
    for (int k = lane; k < d.h; k += kLsLanes) {
        T acc = 0;
#pragma unroll 8
        for (int c = cb; c < ce; ++c) acc += Bm[static_cast<int64_t>(c) * d.h + k] * yl[c - cb];
        tb[dst[k]] = acc;
    }
}

// ---- backward items

/// Partial dot products of a wide supernode: rows [64 rb, 64 rb + 64) with ancestor rows
/// [64 kc, 64 kc + 64) of the block below.
template <typename T>
__host__ __device__ inline void lsBackPartial(const LsArgs<T>& a, const LsSn& d, int part, int rr, int lane)
{
    const int rb = part / d.nkc, kc = part % d.nkc;
    const int r = rb * kLsLanes + lane;
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

/// Phase 1: b of every row = forward result - (block below) x(ancestors); from the partials when
/// the supernode has them, else summed here.
template <typename T>
__host__ __device__ inline void lsBackwardPull(const LsArgs<T>& a, const LsSn& d, int rr, int lane, T* bl)
{
    const int64_t vo = static_cast<int64_t>(rr) * a.n_local_total;
    const T* yf = a.yf + vo + d.row0 + d.c0;
    if (d.nkc > 0) {
        const T* P = a.part + static_cast<int64_t>(rr) * a.npart + d.poff;
        for (int r = lane; r < d.w; r += kLsLanes) {
            T b = yf[r];
#pragma unroll 4
            for (int kc = 0; kc < d.nkc; ++kc) b -= P[static_cast<int64_t>(kc) * d.w + r];
            bl[r] = b;
        }
    } else {
        const T* Bm = a.belowB + d.below;
        const int32_t* ri = a.ridx + d.ridx;
        const T* x = a.xb + vo + d.row0;
        for (int r = lane; r < d.w; r += kLsLanes) {
            T acc = 0;
#pragma unroll 8
            for (int k = 0; k < d.h; ++k) acc += Bm[static_cast<int64_t>(k) * d.w + r] * x[ri[k]];
            bl[r] = yf[r] - acc;
        }
    }
}

/// Phase 2: rows [64 j, 64 j + 64) of x = inv(U_ss) b; to the backward results and the output.
template <typename T>
__host__ __device__ inline void lsBackwardDiag(const LsArgs<T>& a, const LsSn& d, int j, int rr, int lane, const T* bl)
{
    const int r = j * kLsLanes + lane;
    if (r >= d.w) return;
    const T* D = a.diag + d.diag;
    T x = 0;
#pragma unroll 8
    for (int c = r; c < d.w; ++c) x += D[triUpperCol(c) + r] * bl[c];
    const int32_t row = d.row0 + d.c0 + r;
    a.xb[static_cast<int64_t>(rr) * a.n_local_total + row] = x;
    const int32_t l = a.lperm[row];
    const double y = static_cast<double>(x);
    if (a.mode == 0) {
        if (l < a.owned[d.sub]) {
            const int64_t gi = static_cast<int64_t>(rr) * a.n_global + a.begin[d.sub] + l;
            a.out[gi] = a.x0 ? a.x0[gi] + y : y;
        }
    } else {
        const int64_t li = static_cast<int64_t>(d.row0) + l;
        a.out[static_cast<int64_t>(rr) * a.n_local_total + li] = a.pou ? a.pou[li] * y : y;
    }
}

} // namespace s2l

#endif
