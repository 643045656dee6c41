// pack.hpp -- host-side packing of the local factors into the layout of the batched
// supernodal triangular-solve kernel.
#ifndef S2L_PACK_HPP
#define S2L_PACK_HPP

#include <cstdint>
#include <vector>

namespace s2l {

constexpr int32_t kSmallWidth = 64;   ///< supernodes up to one wavefront wide are "small"
constexpr int32_t kChunkSn = 512;     ///< small supernodes staged in LDS at a time
constexpr int32_t kChunkRows = 4096;  ///< most rows of one chunk (and widest supernode)
constexpr int32_t kMinStepRows = 1024; ///< least rows of the LDS result buffer (the largest workgroup)
constexpr int32_t kMaxSkeleton = 16;       ///< most supernodes cut from the subtrees into the top task
constexpr int32_t kDenseTopMax = 2048;     ///< widest skeleton replaced by a dense inverse (default limit)
constexpr int32_t kMaxParts = 32;          ///< most workgroups sharing the rows of a dense top

/// Queue entry codes: bit 0 the phase (0 forward or combined, 1 backward), bits 1-24 the task,
/// bits 25-30 the part of a dense top task (each part computes a range of its rows).
constexpr int kPartShift = 25;
constexpr int32_t kMaxTasks = 1 << (kPartShift - 1);
inline int32_t queueCode(int32_t task, int phase, int part) { return (part << kPartShift) | (task << 1) | phase; }
inline int32_t codeTask(int32_t code) { return (code & ((1 << kPartShift) - 1)) >> 1; }

/// Leading dimension of a wide supernode's blocks: w rounded up to a multiple of four entries.
constexpr int32_t wideLd(int32_t w) { return (w + 3) & ~3; }

/// One supernode of a sweep: columns [c0, c0 + w) of the factor, h rows below the diagonal block.
/// Blocks of a small supernode (w <= kSmallWidth): the inverse of the diagonal block column-major
/// w x w, the block below row-major h x w. Blocks of a wide one start at a multiple of four
/// entries and are row-major with leading dimension wideLd(w) (zero padding).
struct SnDesc {
    int32_t c0, w, h, toff; ///< toff: forward only, offset of its h products in the T buffer
    int64_t diag;           ///< offset of the inverse of the diagonal block
    int64_t below;          ///< offset of the h x w block below
    int64_t ridx;           ///< offset of the h row indices (factor positions)
};

/// One step of a task's sweep, in execution order: a chunk of small supernodes of one level, or one
/// wide supernode, or (kind 2) the dense top: x_S = S^-1 (b_S - pulled products) over the S rows
/// of a skeleton task, S^-1 the inverse of the skeleton's Schur complement (wide.w = S, wide.diag
/// its row-major offset with leading dimension wideLd(S)).
/// A chunk lists its supernodes' descriptors in level order (lev_desc[q0 .. q0 + cnt)), the prefix
/// sums of their widths and heights (lev_wpre / lev_hpre minus wbase / hbase), and, in the forward
/// sweep, the factor position and the pulled slot range [cb, ce) of each of its W rows at lev_row /
/// lev_cb / lev_ce [roff, roff + W): everything a chunk reads before its first product is one load
/// away from the step. A task's steps are staged in LDS.
struct StepDesc {
    int32_t kind;                           ///< 0 chunk of small supernodes, 1 wide supernode, 2 dense top
    int32_t q0, cnt, wbase, hbase, W, H, roff;
    SnDesc wide;                            ///< kind 1: the wide supernode; kind 2: w = S, diag = S^-1
};

/// Per-subdomain descriptor.
struct SubDesc {
    int32_t n;        ///< local size n_i
    int32_t map_off;  ///< offset of the n_i entries of gmap / lperm (= offset of Omega_i in the local buffers)
    int32_t cptr_off; ///< offset of the n_i slot ranges (row r pulls T slots [cptr[r], cend[r]))
    int32_t n_owned;  ///< n_I (set by the context)
    int64_t t_off;    ///< offset of the subdomain's T buffer region
    int32_t begin;    ///< first owned global row (set by the context)
    int32_t dst_off;  ///< offset of the slot of every pushed product (indexed by toff + k)
};

/// A task of the work queue: one or more sibling subtrees of the supernodal elimination tree, or
/// the skeleton above the subtrees (with the root). Forward entries run bottom-up, backward entries top-down; a
/// task without parent runs both sweeps in one entry.
struct TaskDesc {
    int32_t sub;
    int32_t f_lev, f_nlev;          ///< forward level table of the task's supernodes (none for a dense top)
    int32_t b_lev, b_nlev;          ///< backward level table
    int32_t sn_first, sn_count;     ///< the task's supernodes (indices into task_sn)
    int32_t ext_first, ext_count;   ///< rows of ancestor tasks read by the backward sweep
    int32_t parent;                 ///< parent task (global index) or -1
    int32_t child_first, nchild;    ///< children tasks (indices into task_child)
    int32_t depth;                  ///< depth of the task's root supernode
    int32_t rows_first, rows_count; ///< the task's rows (factor positions, indices into task_rows)
    int32_t f_step, f_nsteps;       ///< forward steps (indices into steps)
    int32_t b_step, b_nsteps;       ///< backward steps
    int32_t nparts;                 ///< dense top: workgroups sharing its rows (one queue entry each); else 0
    int32_t slot_base, slot_count;  ///< the slots of the task's rows (T buffer of its subdomain); count 0
                                    ///< for a task with child tasks (they push to its rows too)
};

/// A dense top: the skeleton task of one elimination tree whose supernodes (all above the subtree
/// tasks) are replaced by the dense inverse of their Schur complement. The inverse is
/// U_SS^-1 L_SS^-1 (L_SS^-T L_SS^-1 for Cholesky), computed in FP64 and stored row-major.
struct DenseTop {
    int32_t sub, task, S;
    int32_t sn_first, sn_count;     ///< its forward supernodes (global indices into sn), in dtop_sn
    int32_t q_off;                  ///< offset of its subdomain's position -> row map in dtop_q
    int64_t sinv;                   ///< offset of S^-1 in diag
};

/// Host copy of one uploaded factor (FP64 values).
struct HostFactor {
    int32_t n = 0;
    bool cholesky = true;
    std::vector<int32_t> perm;
    std::vector<int64_t> lcp, ucp;
    std::vector<int32_t> lr, ur;
    std::vector<double> lv, uv;
};

/// Packed factors of all subdomains (values still FP64).
struct Packed {
    std::vector<SnDesc> sn;           ///< forward and backward supernodes of all subdomains
    std::vector<TaskDesc> task;
    std::vector<int32_t> task_sn;     ///< forward supernode ids of every task
    std::vector<int32_t> task_child;
    std::vector<int32_t> ext_rows;    ///< factor positions
    std::vector<int32_t> task_rows;   ///< factor positions of every task's columns
    std::vector<int32_t> queue;       ///< entries (queueCode: task, phase 0 forward/combined or 1 backward, part)
    std::vector<int32_t> qgroup;      ///< start of every depth group in the queue (+ end)
    std::vector<int32_t> qsub;        ///< start of every subdomain's slice in queue_sub (+ end)
    std::vector<int32_t> queue_sub;   ///< the same entries grouped by subdomain
    std::vector<int32_t> lev;         ///< level tables (7 ints per level, see appendLevels)
    std::vector<int32_t> lev_sn;      ///< supernode indices listed by level
    std::vector<int32_t> lev_wpre, lev_hpre; ///< prefix sums of small widths / heights per level
    std::vector<int32_t> lev_chunk;   ///< chunk starts (positions in lev_sn) of every level, + end
    std::vector<SnDesc> lev_desc;     ///< the descriptors in lev_sn order
    std::vector<int32_t> lev_row, lev_cb, lev_ce; ///< forward: factor position and slot range of each small row, level order
    std::vector<StepDesc> steps;      ///< the steps of every task's sweeps
    std::vector<double> diag, below;  ///< values
    std::vector<int32_t> ridx;        ///< below-row indices
    std::vector<int32_t> cptr, cend, cidx;  ///< forward contributions: slot range per row; slot of every product
    std::vector<SubDesc> sub;
    std::vector<int32_t> fwd_base, fwd_count, bwd_base; ///< per subdomain: its forward and backward supernodes in sn
    std::vector<DenseTop> dtop;
    std::vector<int32_t> dtop_sn;     ///< forward supernodes of the dense tops
    std::vector<int32_t> dtop_q;      ///< per subdomain with a dense top: row of every factor position in its top (-1 outside)
    int64_t entries_dense = 0;        ///< entries of the dense tops' inverses (S^2 each, read once per solve)
    int32_t max_task_slots = 0;       ///< most slots of the rows of a task without child tasks
    int64_t entries_skel_fwd = 0, entries_skel_bwd = 0; ///< factor entries of the supernodes they replace
    int64_t t_total = 0;
    int64_t entries_fwd = 0, entries_bwd = 0;
    int32_t max_levels_fwd = 0, max_levels_bwd = 0, max_n = 0;
    int32_t max_step_rows = 0;        ///< most rows of a step (a chunk's rows, a wide supernode's or a dense top's width)
    int32_t max_step_sn = 0;          ///< most small supernodes of a chunk
    int64_t bytes_index = 0;          ///< bytes of index data read per solve (ridx + contributions)
};

/// Packs factors; map_off[i] gives the offset of subdomain i in the gmap/lperm arrays.
/// Supernode partition of a factor in CSC form (diagonal first, rows sorted): columns with the
/// same structure below the diagonal form one supernode. Returns the first column of every
/// supernode and n.
std::vector<int32_t> detectSupernodes(int32_t n, const std::vector<int64_t>& cp, const std::vector<int32_t>& row);
/// In-place FP64 inverses of a lower / upper triangular m x m block, column-major with leading dimension ld.
void invertLower(double* B, int ld, int m);
void invertUpper(double* B, int ld, int m);

/// @p tasks_per_sub: wanted subtree tasks per subdomain (1 = one task per subdomain).
/// @p dense_top_max: widest skeleton (sum of its supernode widths) replaced by a dense top; 0 = none.
Packed packFactors(const std::vector<HostFactor>& f, const std::vector<int32_t>& map_off, int tasks_per_sub,
                   int dense_top_max = kDenseTopMax);

} // namespace s2l

#endif
