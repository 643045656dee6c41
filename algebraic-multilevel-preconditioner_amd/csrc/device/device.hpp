// device.hpp -- kernel argument blocks and launchers of the MI355X apply path.
#ifndef S2L_DEVICE_HPP
#define S2L_DEVICE_HPP

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lsolve.hpp"
#include "pack.hpp"

namespace s2l {

constexpr int kLocalThreads = 1024; ///< threads of a local-solve workgroup (one per queue entry); 512 is the other choice
constexpr int kGmMaxK = 65;         ///< GMRES restart length + 1 handled by the device loop
constexpr int kGmThreads = 256;     ///< threads per workgroup of the GMRES vector kernels
constexpr int kStepBlock = 32;      ///< steps of a task's sweep staged in LDS at a time

/// Arguments of the batched local solve (all pointers are device pointers).
/// T: arithmetic type of the local solve (LDS vector, pushed products); TF: storage type of the
/// factor entries (TF = float with T = double: FP32 factors applied in FP64 arithmetic).
template <typename T, typename TF = T> struct LocalArgs {
    const SubDesc* sub;
    const int32_t* lev_wpre;  ///< per small supernode: widths of the small supernodes before it in its level
    const int32_t* lev_hpre;  ///< same for the heights
    int32_t y_off;            ///< byte offset of the result buffer in the dynamic LDS
    int32_t st_off;           ///< byte offset of the staged supernode descriptors
    int32_t step_off;         ///< byte offset of the staged steps (kStepBlock StepDesc)
    int32_t st_cap;           ///< supernode descriptors the staging area holds (the most of a chunk)
    const SnDesc* lev_desc;   ///< descriptors in level order
    const int32_t* lev_row;   ///< forward: factor position of each small row, level order
    const int32_t* lev_cb;    ///< forward: first pulled slot of each small row
    const int32_t* lev_ce;    ///< forward: end of its pulled slots
    const StepDesc* steps;    ///< the steps of every task's sweeps
    const TF* diag;
    const TF* below;
    const int32_t* ridx;
    const int32_t* cptr;  ///< first slot of every row
    const int32_t* cend;  ///< end of every row's slots
    const int32_t* cidx;  ///< slot of every pushed product
    int32_t slot_off;     ///< byte offset in the dynamic LDS of the slots of a task without child tasks
    int32_t slot_cap;     ///< slots that fit there (a task with more keeps the rest in global memory)
    T* tbuf;              ///< T buffer, t_total entries per right-hand side
    const int32_t* gmap;  ///< factor position -> global row (gather)
    const int32_t* lperm; ///< factor position -> local index
    const double* rhs;    ///< global right-hand sides (n x nrhs)
    const double* x0;     ///< added to the owned output rows (may be null)
    double* out;          ///< mode 0: global output (n x nrhs); mode 1: local buffer (sum n_i x nrhs)
    const double* pou;    ///< mode 1, RAS: D_i by local index (null = 1)
    const int32_t* skip;  ///< per right-hand side: nonzero = nothing to do (may be null)
    int64_t n_global;
    int64_t n_local_total;
    int64_t t_total;
    int32_t mode;         ///< 0 RAS with Boolean PoU (direct write), 1 local buffer for a reverse add
    int32_t local_in;     ///< 1: rhs is a local buffer (sum n_i x nrhs, by local index) instead of global
    // ---- level-ordered work queue
    const TaskDesc* task;
    const int32_t* task_child;
    const int32_t* task_rows;
    const int32_t* ext_rows;
    const int32_t* queue;  ///< entries (task << 1 | phase) of this launch
    int32_t qcount;        ///< entries per right-hand side in this launch
    int32_t serial;        ///< 1: one workgroup runs all entries in order
    int32_t nrhs;
    T* xg;                 ///< per-task results shared between entries (n_local_total per rhs)
};

/// Coarse space arguments.
struct CoarseArgs {
    int32_t nsub, n0, mmax;
    int32_t rchunks;       ///< chunks of 256 rows of the largest V_i block (restriction grid)
    const int64_t* off;    ///< coarse offsets (nsub + 1)
    const int64_t* voff;   ///< offset of each V_i block
    const int32_t* vrows;  ///< rows of each V_i block
    const int32_t* vrow_ptr; ///< offset of each block's rows in vrow_g
    const int32_t* vrow_g; ///< global row of each V row
    const int32_t* vbase;  ///< per block: its global rows are vbase + 0, 1, ... (contiguous), or -1
    const double* V;
    const int32_t* blk_sub; ///< prolongation blocks: subdomain and first row
    const int32_t* blk_row;
    int32_t nblk;
};

/// Per right-hand side state of the device GMRES loop.
struct GmState {
    double bnorm, beta, rel, hn;
    double thr;         ///< OpenFOAM test: threshold on sum|r| (max(rtol sum|r0|, atol F))
    double l1;          ///< OpenFOAM test: sum|r_j| of the last iteration's GMRES residual
    int32_t k;          ///< columns of the current cycle
    int32_t iters;      ///< total iterations
    int32_t done;       ///< cycle finished (converged estimate, breakdown or maxit)
    int32_t conv;       ///< estimate below the tolerance
    int32_t breakdown;  ///< hn == 0 (happy breakdown)
    int32_t maxit;
    int32_t pad[2];
};

/// Per right-hand side state of the device BiCGStab loop (bicgstab.cpp:15-154).
struct BiState {
    double bnorm, rn, l1, l1_0, thr, normF, rho, rho_old, alpha, omega, rhn, pad0;
    int32_t its, restarts, conv, brk, candidate, first, done, phase, nhist, pad1;
};
/// Participation masks of the BiCGStab launches (1 = skip that right-hand side), each with its
/// own four coefficients per right-hand side.
enum BiMask { kBiZero, kBiLive, kBiNom, kBiWithm, kBiStart, kBiP, kBiHalf, kBiHx, kBiFull, kBiEnd, kBiMasks };
/// Stages of the BiCGStab control kernel: each reads the sums of the launch before it.
enum BiStage { kBiBnorm, kBiXmean, kBiNomAx, kBiNomB, kBiWithmNorm, kBiR0, kBiCycle, kBiAlpha, kBiHalfStep, kBiOmega, kBiXr, kBiEndCycle };
struct BiArgs {
    BiState* st;
    const double* red;  ///< sums of the last reduction, 3 per right-hand side
    int32_t* mask;      ///< [kBiMasks][nrhs]
    double* coef;       ///< [kBiMasks][4 nrhs]
    double* hist;       ///< [nrhs][maxit + 1] relative residual norms
    int32_t* live;      ///< right-hand sides not finished (written by kBiEndCycle)
    int64_t n;
    double rtol, atol;
    int32_t nrhs, maxit, of, pad;
};
void launchBicgControl(const BiArgs& a, int stage, hipStream_t s);
/// y_r = v for the right-hand sides r that skip[r] does not exclude
void launchFillMasked(int64_t n, double v, double* y, int nrhs, const int32_t* skip, hipStream_t s);

template <typename T, typename TF> void launchLocalSolve(const LocalArgs<T, TF>& a, int grid, size_t lds, int threads, hipStream_t s);
template <typename T, typename TF> const void* localSolveKernel(int threads);
/// A launch of @p grid dense tops (every entry of a.queue a dense top): dt / dt_ce / dtb the row
/// table of the skeletons (factor position, global row, local index, first slot; end of slots;
/// offset of a task's rows in the table), lds the bytes of the widest skeleton's b_S; cus the compute
/// units of the device (a launch of at most that many workgroups takes the variant with more loads
/// in flight).
template <typename T, typename TF>
void launchDenseTop(const LocalArgs<T, TF>& a, int grid, const int4* dt, const int32_t* dt_ce, const int32_t* dtb, size_t lds,
                    int cus, hipStream_t s);

/// One launch of the level-synchronous local solve: items [first, first + count) of kind LsKind
/// (widest supernode maxw) for nrhs right-hand sides.
template <typename T>
void launchLs(const LsArgs<T>& a, int kind, int32_t first, int32_t count, int32_t maxw, int nrhs, hipStream_t s);

void launchSpmv(int32_t n, const int32_t* rp, const int32_t* col, const double* val, const double* x, double* y,
                const double* b, int nrhs, const int32_t* skip, hipStream_t s);
/// Right-hand sides from which the coarse products run on the matrix cores (16 per group).
constexpr int kMfmaMinRhs = 4;
/// c = R_0 r; part: nrhs x nsub x rchunks x mmax partials of the first stage
void launchCoarseRestrict(const CoarseArgs& c, const double* r, int64_t n, double* cvec, double* part, int nrhs,
                          const int32_t* skip, hipStream_t s);
/// Leading dimension of the inverted A00 factors: n0 rounded up to a multiple of 8 (zero padding).
inline int32_t triLd(int32_t n0) { return (n0 + 7) & ~7; }
/// y = T x (perm: x[perm[k]]) for the lower or upper triangle of T (n0 x n0, leading dimension ld)
void launchTriGemv(int32_t n0, int32_t ld, const double* T, int lower, const int32_t* perm, const double* x, double* y, int nrhs,
                   const int32_t* skip, hipStream_t s);
void launchCoarseProlong(const CoarseArgs& c, const double* y, double* out, int64_t ld_out, int boolean, int nrhs,
                         const int32_t* skip, hipStream_t s);
void launchRevGather(int32_t n, const int32_t* ptr, const int32_t* idx, const double* src, int64_t src_ld,
                     const double* x0, double* out, int nrhs, const int32_t* skip, hipStream_t s);

// ---- vector kernels
void launchAxpby(int64_t n, double a, const double* x, double b, double* y, int nrhs, const int32_t* skip, hipStream_t s);
void launchCopy(int64_t n, const double* x, double* y, int nrhs, hipStream_t s);

/// Deterministic sum reductions of nrhs columns: out[r * ncomp + c] for c < ncomp;
/// kind 0: x.x ; 1: x.x, sum|x| ; 2: x.x, sum|x|, x.y ; 3: y.x (single) ; 4: x.x, x.y
void bicgP(int64_t n, const double* r, double* p, const double* v, const double* coef, int nrhs, const int32_t* skip,
           hipStream_t s);
void bicgS(int64_t n, const double* r, const double* v, double* sv, const double* coef, double* partial, double* out,
           int nrhs, const int32_t* skip, hipStream_t s);
void bicgXR(int64_t n, double* x, const double* y, const double* z, const double* sv, const double* t, double* r,
            const double* rh, const double* coef, double* partial, double* out, int nrhs, const int32_t* skip,
            hipStream_t s);
void bicgHalf(int64_t n, double* x, const double* y, const double* sv, double* r, const double* coef, int nrhs,
              const int32_t* skip, hipStream_t s);
void ofNorm(int64_t n, const double* ax, const double* b, const double* a1, const double* coef, double* partial,
            double* out, int nrhs, const int32_t* skip, hipStream_t s);
void launchFill(int64_t n, double v, double* y, hipStream_t s);
/// kind 0: (x.x); 1: (x.x, |x|_1); 2: (x.x, |x|_1, x.y); 3: (y.x); 4: (x.x, x.y); 5: (x.x, sum x)
void launchReduce(int kind, int64_t n, const double* x, const double* y, double* partial, double* out, int nrhs,
                  const int32_t* skip, hipStream_t s);
int reduceBlocks(int64_t n);
int gmBlocks(int64_t n);

// ---- GMRES
struct GmArgs {
    int64_t n;
    int32_t m;          ///< restart
    int32_t nrhs;
    double* V;          ///< basis, V[(k * nrhs + r) * n + i]
    double* Z;          ///< flexible GMRES: M^{-1} v_k in the same layout (null otherwise)
    double* w;          ///< n per rhs
    double* H;          ///< (m + 1) x m per rhs, column-major
    double* cs;
    double* sn;
    double* g;          ///< m + 1 per rhs
    double* h;          ///< m + 1 per rhs
    double* h2;
    double* y;          ///< m per rhs
    GmState* st;
    int32_t* flags;     ///< [0, nrhs): 1 once the rhs has finished its cycle; [nrhs, 2 nrhs): 1 if idle this cycle
    double* hist;       ///< (maxit + 1) per rhs
    int32_t hist_ld;
    double rtol;
    double* partial;    ///< block partial sums
    int32_t nblocks;
    int32_t of;         ///< 1: OpenFOAM stopping test (gmres.cpp:106-122), run by gmOpenFoamTest
};
void gmDots(const GmArgs& a, int j, hipStream_t s);
/// w = A z (one right-hand side per grid row) and the first CGS2 pass's partials in one launch;
/// gmFinalH sums them into h (together: gmDots after an SpMV into w)
void gmSpmvDots(const GmArgs& a, int j, const int32_t* rp, const int32_t* col, const double* val, const double* z,
                hipStream_t s);
void gmFinalH(const GmArgs& a, int j, hipStream_t s);
void gmUpdateDots(const GmArgs& a, int j, hipStream_t s);
void gmUpdateNorm(const GmArgs& a, int j, hipStream_t s);
void gmNormalize(const GmArgs& a, int j, hipStream_t s);
/// OpenFOAM stopping test of iteration j: sum|g_{j+1} V_{j+2} Q^T e_{j+1}| against st.thr; sets
/// the convergence and the flags the l2 test sets in gmNormalize.
void gmOpenFoamTest(const GmArgs& a, int j, hipStream_t s);
void gmStartCycle(const GmArgs& a, const double* r, hipStream_t s);
void gmLsq(const GmArgs& a, hipStream_t s);
/// u = B_k y for the basis B = V (or Z for flexible GMRES)
void gmCombine(const GmArgs& a, const double* B, double* u, hipStream_t s);

} // namespace s2l

#endif
