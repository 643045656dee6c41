// refactor.hpp -- numeric refactorization on the device: the new values of A, the same symbolic
// structure. Host side: the plan (which entries of A start each supernode's panel, which
// descendants update it and where their rows land). Device side: a left-looking supernodal
// Cholesky one elimination-tree height at a time, the repack into the layout of the local solve,
// and the coarse matrix A_00 = W^T A W with its Cholesky factor and triangular inverses.
#ifndef S2L_REFACTOR_HPP
#define S2L_REFACTOR_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "pack.hpp"

namespace s2l {

/// One forward supernode of a local Cholesky factor in the refactorization.
struct RfSn {
    int32_t c0, w, h, sub;
    int64_t wd;        ///< work offset of L_ss (w x w, column-major)
    int64_t wb;        ///< work offset of the block below (h x w, column-major)
    int64_t ridx;      ///< offset of its h row indices (factor positions, sorted) in the solve's ridx
    int32_t upd_first, upd_count;  ///< updates from descendants, in increasing descendant order
    int64_t amap_first;
    int32_t amap_count, pad;
    int64_t fdiag, bdiag, below;   ///< offsets of the forward / backward diagonal inverses and of the block below in the solve layout
};

/// An update of supernode s by descendant d: the rows of d at positions [i0, i1) of its row list are
/// columns of s; the rows at [i0, h_d) land at panel rows rel[0 .. h_d - i0) of s (a row < w of the
/// diagonal block, or w + its index among the rows below s).
struct RfUpd {
    int32_t d, i0, i1, pad;
    int64_t rel;
};

struct RfPlan {
    std::vector<RfSn> sn;
    std::vector<RfUpd> upd;
    std::vector<int32_t> rel;
    std::vector<int64_t> amap_pos;   ///< work position of every entry of A_ii(perm, perm)'s lower triangle
    std::vector<int32_t> amap_src;   ///< its index in the stored values of A
    std::vector<int32_t> order;      ///< supernodes by height (all subdomains)
    std::vector<int32_t> level_ptr;  ///< start of every height in order, + end
    int64_t work = 0;                ///< FP64 entries of the work array (panels)
    int64_t inv = 0;                 ///< FP64 entries of the inverse scratch (sum of w^2)
    std::vector<int64_t> inv_off;    ///< per supernode
    // dense tops (Packed::dtop): their supernodes' plan indices and the offsets of their S x S blocks
    std::vector<int32_t> dtop_rf;    ///< plan index of every entry of Packed::dtop_sn
    std::vector<int64_t> dtop_off;   ///< per dense top: offset of its S x S scratch blocks
    int64_t dtop_total = 0;          ///< FP64 entries of one scratch array of all dense tops
};

/// The plan of all subdomains' factors (Cholesky only). local_ptr / gid: the subdomains' maps;
/// rp / col: A in the stored order (rows in the owners' local column order); P: the packed factors
/// (their solve-layout offsets); fwd_base: the first forward supernode of every subdomain in P.sn
/// (backward ones follow at fwd_base + count).
RfPlan buildRefactorPlan(const std::vector<HostFactor>& f, const std::vector<int32_t>& local_ptr,
                         const std::vector<int32_t>& gid, const std::vector<int32_t>& rp, const std::vector<int32_t>& col,
                         const Packed& P);

/// Device arguments.
struct RfArgs {
    const RfSn* sn;
    const RfUpd* upd;
    const int32_t* rel;
    const int32_t* ridx;      ///< the solve's row indices
    const int64_t* amap_pos;
    const int32_t* amap_src;
    const double* val;        ///< A, stored order
    double* work;
    double* inv;
    const int64_t* inv_off;
    int32_t* status;          ///< 0 ok; 1 + the first failing supernode: a pivot not > 0
    const double* diag_adj;   ///< null, or added to the diagonal of A_ii (by local index)
    const int32_t* lperm;     ///< local index of every factor position (with diag_adj)
    const int32_t* sub_off;   ///< first local position of every subdomain (with diag_adj)
};

/// The panels of supernodes order[first .. first + count) (one height): A entries, descendants'
/// updates, dense Cholesky of L_ss and L_rs = A_rs L_ss^-T.
void launchRfLevel(const RfArgs& a, const int32_t* order, int32_t count, hipStream_t s);
/// Inverse of every L_ss and the solve layout of all supernodes: forward inverse (lower),
/// backward inverse (its transpose), block below; @p fp32: float storage.
void launchRfPack(const RfArgs& a, int32_t nsn, void* diag, void* below, bool fp32, hipStream_t s);

/// Dense tops: L_SS of every dense top assembled from the refactorized panels into the column-major
/// S x S blocks of @p L (zero outside the factor's entries).
void launchDtAssemble(const RfArgs& a, const DenseTop* dt, int32_t ndt, const int32_t* dt_rf, const int32_t* dtop_q,
                      const int64_t* off, double* L, hipStream_t s);
/// S^-1 (column-major blocks of @p C) into the solve layout: row-major with leading dimension
/// wideLd(S) at each top's offset of @p diag (FP32 when @p fp32).
void launchDtLayout(const DenseTop* dt, int32_t ndt, const int64_t* off, const double* C, void* diag, bool fp32, hipStream_t s);

/// dst[q] = src[idx[q]], q < n.
void launchGather(const double* src, const int32_t* idx, double* dst, int64_t n, hipStream_t s);
/// Columns [k0, k0 + cnt) of the n0 x n0 identity, n0 x cnt (column-major).
void launchUnitColumns(double* Y, int32_t n0, int32_t k0, int32_t cnt, hipStream_t s);
/// E(:, k0 .. k0 + cnt) = C (n0 x cnt, column-major) into column-major E with leading dimension ld.
void launchStoreColumns(const double* C, int32_t n0, int32_t k0, int32_t cnt, double* E, int32_t ld, hipStream_t s);
/// The layout of the coarse solve from Linv = L^-1 (column-major n x n): T1 = Linv, T2 = Linv^T,
/// both row-major with leading dimension ld, zero outside their triangle and in the padding.
void launchTriLayout(const double* Linv, int32_t n, int32_t ld, double* T1, double* T2, hipStream_t s);

} // namespace s2l

#endif
