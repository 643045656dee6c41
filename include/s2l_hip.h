/*
 * s2l_hip.h -- C-ABI of the MI355X apply path of schwarz2lvl's two-level overlapping Schwarz
 * preconditioner inside right-preconditioned restarted GMRES / BiCGStab.
 *
 * One context holds every subdomain of the problem on one GPU. The host uploads once per setup
 * (or refactorization) what the reference's setup produced; the apply, the SpMV and the Krylov
 * iterations then run as HIP kernels. No torch types, plain pointers and sizes.
 *
 * Each entry point names the reference interface it replaces (paths relative to the reference
 * repository). Conventions:
 *   - return value: 0 on success, a negative S2L_E* code on failure; s2l_last_error(ctx) gives
 *     the message (ctx may be NULL for failures of s2l_ctx_create). No exception crosses the ABI.
 *   - ownership: the context copies every uploaded host array; the caller keeps its buffers.
 *     Device buffers given to s2l_apply / s2l_spmv / s2l_krylov_solve belong to the caller.
 *   - vectors are in the renumbered global order (subdomain p owns rows
 *     [ownership[p], ownership[p+1]), src/partition/domain_decomposer.cpp:254-269); several
 *     right-hand sides are stored one after the other (column-major n x nrhs).
 *   - threading: a context is used by one host thread at a time (like the reference's
 *     Preconditioner objects, whose apply() mutates work vectors); different contexts are
 *     independent. All work of a context is ordered on the stream given (NULL = the context's).
 */
#ifndef S2L_HIP_H
#define S2L_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    S2L_OK = 0,
    S2L_EINVAL = -1,   /* invalid argument / shape */
    S2L_ESTATE = -2,   /* call out of order (e.g. apply before upload) */
    S2L_EHIP = -3,     /* HIP runtime error */
    S2L_ELIMIT = -4,   /* a documented size limit is exceeded */
    S2L_ENUMERIC = -5  /* singular coarse matrix, non-finite values */
};

typedef struct s2l_ctx s2l_ctx;

/* Creates a context on HIP device @p device. Replaces the construction of the preconditioner
 * objects (src/main.cpp:143-168); the process model of one rank per subdomain
 * (README.md:281) becomes one context for all subdomains. */
int s2l_ctx_create(int device, s2l_ctx** out);
int s2l_ctx_destroy(s2l_ctx* ctx);
const char* s2l_last_error(const s2l_ctx* ctx);

/* Global operator A (n x n, CSR, int32 indices, FP64 values) in the renumbered order, together
 * with the subdomain ownership (N + 1 offsets). Every row is reordered to the local column
 * order of its owner (owned columns, then ghosts by increasing global id), which is the order
 * of A(Omega_I, Omega_i) in DistributedMatrix::apply (include/linalg/distributed_matrix.hpp:33-38,
 * include/partition/subdomain.hpp:112). Uploading again with the same n, row pointers, columns and
 * ownership refreshes the values only; any other upload is a new system, and the subdomains,
 * factors and coarse space must be uploaded again before the next apply (S2L_ESTATE otherwise). */
int s2l_upload_system(s2l_ctx* ctx, int32_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                      int32_t nsub, const int32_t* ownership);

/* Overlapping subdomains: local -> global maps Omega_i = [owned | ghosts sorted] concatenated
 * (local_ptr: nsub + 1 offsets, gid: global ids), as Subdomain::globalId
 * (src/partition/subdomain.cpp:150-161). @p pou: NULL for the Boolean partition of unity (1 on
 * owned, 0 on ghosts), else D_i concatenated (src/partition/subdomain.cpp:281-302).
 * @p one_level: 0 RAS, 1 ASM (src/preconditioner/one_level_schwarz.cpp:80-99). Replaces the
 * HaloExchange pattern (src/linalg/halo_exchange.cpp): the forward exchange becomes the
 * on-device gather x_loc[l] = x[gid[l]], the reverse add a deterministic gather-sum. */
int s2l_upload_subdomains(s2l_ctx* ctx, int32_t nsub, const int32_t* local_ptr, const int32_t* gid, const double* pou,
                          int one_level);

/* Local factor of subdomain @p i: A_ii(perm, perm) = L U with perm[k] = local index of factor
 * position k; L lower triangular in CSC (colptr n + 1, rows sorted, diagonal first), U given by
 * U^T in the same form; ut_* = NULL means U = L^T (Cholesky). Replaces LocalSolver::factorize's
 * result (include/linalg/local_solver.hpp:61-138; e.g. Eigen SimplicialLLT matrixL() with its
 * permutation). The values are stored on the device in FP64 or FP32 according to
 * s2l_finalize_local(). */
int s2l_upload_local_factor(s2l_ctx* ctx, int32_t i, int32_t n, const int32_t* perm, const int64_t* l_colptr,
                            const int32_t* l_row, const double* l_val, const int64_t* ut_colptr, const int32_t* ut_row,
                            const double* ut_val);

/* Packs every uploaded factor into the device layout (supernodes, level schedule) with values in
 * @p precision: 0 FP64, 1 FP32 (--local-precision single, include/core/config.hpp:72: the
 * right-hand side is rounded to float, both sweeps run in float, the result is widened, as
 * solveIn<float> in src/linalg/local_solver.cpp:77-84), 2 FP32 factor entries applied in FP64
 * arithmetic (no reference counterpart: the bytes of FP32 factors, the sums and the vector of
 * FP64; the local solve is then a linear operator up to FP64 rounding; task-queue launch modes 0
 * and 1 only). */
int s2l_finalize_local(s2l_ctx* ctx, int precision);

/* Coarse space: V_i blocks (rows x m_i, column-major, concatenated; rows = owned rows with the
 * Boolean PoU, all local rows otherwise), offsets (nsub + 1) and A_00 dense row-major (n0 x n0);
 * @p E = NULL forms A_00 = W^T A W on the device from the uploaded operator and V
 * (coarse_space.cpp:95-105; needs s2l_upload_system first), so that a caller need not assemble it.
 * @p symmetric: Cholesky of A_00, else LU with partial pivoting (both FP64, factored on the host
 * at upload). @p correction: 0 deflated, 1 additive (two_level_schwarz.hpp:69). n0 = 0 gives the
 * one-level preconditioner (two_level_schwarz.cpp:136-139). Replaces CoarseSpace
 * (src/preconditioner/coarse_space.cpp:208-246) and RootCoarseSolver
 * (src/preconditioner/coarse_solver.cpp:94-179). */
int s2l_upload_coarse(s2l_ctx* ctx, int32_t n0, const int64_t* offsets, const double* V, const double* E, int symmetric,
                      int correction);

/* Refactorize on the device (UpdateLevel::Refactorize, src/preconditioner/two_level_schwarz.cpp:
 * 35-41 with coarse_space.cpp:186-206 and subdomain.cpp:129-148): @p val holds new values of A in the
 * order of the arrays last given to s2l_upload_system (same pattern; host memory if @p host, else
 * device memory). The local matrices A_ii are gathered from them, every local Cholesky factor is
 * computed again on the device over its symbolic structure (a left-looking supernodal factorization,
 * one height of the elimination trees per launch) and repacked into the layout of the local solve,
 * and A_00 = W^T A W is formed on the device with the same coarse basis, factored and inverted
 * (dense FP64). Needs Cholesky local factors, a symmetric A_00 and the task-queue schedule; the
 * level-synchronous layout is not refreshed (launch mode 2 is refused afterwards). S2L_ENUMERIC when
 * a factor loses positive definiteness. The first call builds the symbolic plan. */
int s2l_refactorize(s2l_ctx* ctx, const double* val, int host);

/* Local eigenproblems of the coarse space on the device (LocalEigensolver::compute and
 * computeSparse, src/preconditioner/local_eigensolver.cpp:215-226 and :277-421, the operator of
 * ProjectedOperator :174-211): for every subdomain i not marked in @p skip (may be NULL), the @p nev
 * largest eigenpairs of B_i u = lambda (A~_ii + sigma_i I) u with B_i = D_i A_ii D_i, A~_ii the
 * lumped block (src/preconditioner/local_block_splitting.cpp:9-21) and sigma_i = shift_rel *
 * sqrt(|A~_ii|_1 |A~_ii|_inf), multiplied by 1e3 while a factorization fails (four attempts, :285-303).
 * The values of A are the context's current ones (upload or refactorize). A~_ii + sigma_i I is
 * factored on the device over the structure of A_ii's Cholesky factor; every subdomain runs a
 * thick-restart Lanczos process (basis @p ncv, restarted with (nev + ncv) / 2 Ritz vectors, at most
 * @p max_restarts restarts, full reorthogonalization in the A~ + sigma I inner product) until the
 * Ritz estimates |beta_m s_mk| of the nev largest are <= tol |theta_k|. The kernel of A~_ii is not
 * detected separately: its directions with B q != 0 get eigenvalues of order 1/sigma and come first.
 * Outputs (host memory): @p lambda nsub x nev (decreasing), @p Z the Ritz vectors (nev columns of
 * local_total rows: column c of subdomain i at Z[c * local_total + local_ptr[i] + l], normalized in
 * the A~ + sigma I norm), @p converged per subdomain (may be NULL), @p info [restarts, Lanczos steps,
 * factorization attempts, subdomains converged] (may be NULL). Needs symmetric A (Cholesky local
 * factors), FP64 local factors (precision 0), the task-queue schedule, 1 <= nev, nev + 2 <= ncv <= 96
 * and ncv + 1 rows in every subdomain solved here. */
int s2l_local_eigensolve(s2l_ctx* ctx, int32_t nev, int32_t ncv, double shift_rel, double tol, int32_t max_restarts,
                         const int32_t* skip, double* lambda, double* Z, int32_t* converged, int32_t* info);

/* z = M^{-1} r for @p nrhs right-hand sides (device pointers, n x nrhs). Replaces
 * Preconditioner::apply (include/preconditioner/preconditioner.hpp:73) as implemented by
 * TwoLevelSchwarz::apply (src/preconditioner/two_level_schwarz.cpp:134-163). @p stream: a
 * hipStream_t or NULL. */
int s2l_apply(s2l_ctx* ctx, const double* d_r, double* d_z, int32_t nrhs, void* stream);

/* z = M^{-1} r like s2l_apply, allocating nothing: no reference counterpart; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The apply's scratch (the coarse vectors, x0 and w, the local buffers of the
 * local solves) lives in a static device array of 128 MiB per GPU, held until the process ends;
 * S2L_ELIMIT when @p nrhs right-hand sides need more (about (2 n + 2 sum n_i) x 8 bytes per
 * right-hand side). Every call on one GPU uses that GPU's array: calls that overlap on one GPU
 * (from any contexts or streams) race. Nothing in it needs to start at zero. Event timing is not
 * recorded for these calls. Same kernels and results as s2l_apply. */
int s2l_apply_noalloc(s2l_ctx* ctx, const double* d_r, double* d_z, int32_t nrhs, void* stream);

/* y = A x (device pointers). Replaces DistributedMatrix::apply
 * (include/linalg/distributed_matrix.hpp:33-38). */
int s2l_spmv(s2l_ctx* ctx, const double* d_x, double* d_y, int32_t nrhs, void* stream);

/* Krylov solve with the uploaded operator and preconditioner. Replaces KrylovSolver::solve
 * (include/solver/krylov.hpp:70-71) of Gmres (src/solver/gmres.cpp:18-161) and BiCgStab
 * (src/solver/bicgstab.cpp:15-154), made by makeKrylovSolver (src/solver/krylov.cpp:23-36).
 * d_x: initial guess in, solution out. Every right-hand side is an independent solve whose
 * kernels are shared (batched launches); results[r] and history + r * (maxit + 1) hold its
 * KrylovResult (include/solver/krylov.hpp:43-54). */
typedef struct s2l_krylov_opts {
    int method;       /* 0 GMRES(restart), 1 BiCGStab */
    int restart;      /* GMRES restart length m (<= 64) */
    double rtol;      /* relative tolerance */
    double atol;      /* OpenFOAM norm only: absolute tolerance (0 = off) */
    int maxit;        /* maximum applications of M^{-1} */
    int stop_norm;    /* 0 l2, 1 openfoam (include/solver/krylov.hpp:29-45) */
    int flexible;     /* GMRES only: 0 the reference's x = x0 + M^{-1}(V y) at the end of a cycle
                         (gmres.cpp:143-156); 1 flexible GMRES (Saad, SIAM J. Sci. Comput. 14,
                         1993): the basis Z_j = M^{-1} v_j is stored and x = x0 + Z y, which keeps
                         the Arnoldi relation exact for an inexact (FP32) preconditioner and saves
                         the apply at the end of a cycle. No reference counterpart. */
} s2l_krylov_opts;

typedef struct s2l_krylov_result {
    int converged;
    int breakdown;    /* BiCGStab: (r^, A y) = 0 or not finite (bicgstab.cpp:97-100); GMRES: a residual
                         was not finite and the solve stopped unconverged (the reference's GMRES would
                         restart until maxit) */
    int iterations;
    int restarts;
    double rel_residual;
    double true_rel_residual;
    double initial_residual_of;
    double final_residual_of;
    int history_len;
} s2l_krylov_result;

int s2l_krylov_solve(s2l_ctx* ctx, const s2l_krylov_opts* opts, const double* d_b, double* d_x, int32_t nrhs,
                     s2l_krylov_result* results, double* history, void* stream);

/* ---- Continuous batching: a GMRES run over a fixed number of right-hand-side slots.
 * No single reference counterpart: it runs Gmres::solve (src/solver/gmres.cpp:18-161) for every
 * slot, and its callers are the request loops around the solve (src/solver/time_loop.cpp:57-195
 * solves one right-hand side after another). Every slot is an independent solve whose kernels are
 * shared with the other slots (one launch per kernel for all of them); a slot's results do not
 * depend on which other slots are busy, so a request gets the same bits alone and in a mix. A
 * request enters a free slot with s2l_solver_admit and starts at the next s2l_solver_step, which
 * runs one restart cycle of every busy slot; a finished slot's result is taken with
 * s2l_solver_take, which frees the slot. The solver owns its right-hand sides, solutions and
 * Krylov workspace (counted in s2l_device_footprint of its context) and uses the context's stream
 * and apply buffers: the solvers of one context run one at a time (one host thread); solvers of
 * different contexts run concurrently and their launches overlap on the device. */
typedef struct s2l_solver s2l_solver;
int s2l_solver_create(s2l_ctx* ctx, const s2l_krylov_opts* opts, int32_t slots, s2l_solver** out);
int s2l_solver_destroy(s2l_solver* solver);
/* b (and x0, NULL = 0): n values, host memory if @p host, else device memory. */
int s2l_solver_admit(s2l_solver* solver, int32_t slot, const double* b, const double* x0, int host);
/* One restart cycle of every busy slot (admitted ones start first). finished[slot] (slots
 * entries, may be NULL) = 1 for a slot that holds a finished solve not yet taken. */
int s2l_solver_step(s2l_solver* solver, int32_t* finished);
/* KrylovResult of a finished slot, its history (maxit + 1 entries, may be NULL) and its solution
 * (n values to host memory if @p host, else device memory; may be NULL); frees the slot. */
int s2l_solver_take(s2l_solver* solver, int32_t slot, s2l_krylov_result* result, double* history, double* x, int host);

/* ---- Host-memory conveniences (copy in, run on the context's stream, copy out, synchronize). */
int s2l_apply_host(s2l_ctx* ctx, const double* r, double* z, int32_t nrhs);
int s2l_spmv_host(s2l_ctx* ctx, const double* x, double* y, int32_t nrhs);
int s2l_krylov_solve_host(s2l_ctx* ctx, const s2l_krylov_opts* opts, const double* b, double* x, int32_t nrhs,
                          s2l_krylov_result* results, double* history);
/* One local solve of subdomain i alone on the device, y = A_ii^{-1} b (local numbering); for
 * parity tests of the batched triangular-solve kernel. */
int s2l_local_solve_host(s2l_ctx* ctx, int32_t i, const double* b, double* y);

/* ---- Device memory helpers (so that callers need no HIP headers). */
int s2l_device_alloc(s2l_ctx* ctx, int64_t bytes, void** out);
int s2l_device_free(s2l_ctx* ctx, void* p);
int s2l_memcpy_h2d(s2l_ctx* ctx, void* dst, const void* src, int64_t bytes);
int s2l_memcpy_d2h(s2l_ctx* ctx, void* dst, const void* src, int64_t bytes);
/* dst[k] = src[k], k < count, by a plain streaming kernel on @p stream (NULL = the context's):
 * 16 bytes of HBM traffic per element, the calibration point of the traffic counters. */
int s2l_copy_d2d(s2l_ctx* ctx, double* dst, const double* src, int64_t count, void* stream);
int s2l_synchronize(s2l_ctx* ctx);

/* ---- Introspection and measurement. */
typedef struct s2l_stats {
    int32_t n, nsub, n0, precision;
    int64_t nnz_a;
    int64_t factor_entries_fwd;   /* stored entries read by the forward sweeps */
    int64_t factor_entries_bwd;   /* stored entries read by the backward sweeps */
    int64_t local_rows;           /* sum of n_i */
    int64_t coarse_rows;          /* rows of the V_i blocks */
    int32_t levels_fwd_max, levels_bwd_max, supernodes;
    double bytes_local_solve;     /* algorithmic HBM bytes of one batched local-solve launch (1 rhs) */
    double bytes_apply;           /* algorithmic HBM bytes of one preconditioner application (1 rhs) */
    double bytes_spmv;            /* algorithmic HBM bytes of one SpMV (1 rhs) */
    double refactor_ms[4];        /* the last s2l_refactorize: new values of A, local factors, A_00, total [ms] */
    int32_t lds_slot_cap;         /* task queue: slots of a task without child tasks kept in LDS (its others in HBM) */
    int32_t max_task_slots;       /* task queue: most slots of a task without child tasks */
    int64_t lds_slots;            /* task queue: slots kept in LDS, summed over tasks (not counted as HBM bytes) */
} s2l_stats;
int s2l_get_stats(s2l_ctx* ctx, s2l_stats* out);
/* Bytes of device memory the context holds (its uploaded system, subdomains, factors, coarse
 * space, graphs' buffers and the Krylov workspace grown so far); buffers of s2l_device_alloc are the
 * caller's and not counted. No reference counterpart: the footprint a server admits resident
 * matrices by. */
int s2l_device_footprint(s2l_ctx* ctx, int64_t* bytes);

/* Event timing of the batched local solve: when enabled, the launches of every local solve (one
 * apply's: all depth groups of the task queue, or all levels, or all subdomains) are bracketed by
 * HIP events on their stream; s2l_timing_read returns the number of local solves and their summed
 * milliseconds. */
int s2l_timing_enable(s2l_ctx* ctx, int on);
int s2l_timing_read(s2l_ctx* ctx, int64_t* solves, double* ms_total, int reset);

/* Per-phase timing with HIP events on the work's stream, under the timer names of the reference
 * (utils::ScopedTimer in src/solver/gmres.cpp:67,71,77,108,
 * src/preconditioner/two_level_schwarz.cpp:144,149,155,159 and
 * src/preconditioner/coarse_space.cpp:237,241,244; s2l_phase_name gives them). The parts nest as in
 * the reference: solve.preconditioner contains the coarse, spmv and local parts of each apply.
 * The halo exchange of the one-level part is fused into the local solve (its gather), so
 * solve.preconditioner.local covers solve.preconditioner.local.{halo,lu}. Enabling it launches the
 * kernels one by one (no HIP graphs); s2l_phase_timing_read synchronizes with the recorded events
 * and returns the summed milliseconds and the number of timed intervals per phase. */
enum {
    S2L_PHASE_PRECONDITIONER = 0,
    S2L_PHASE_PREC_COARSE = 1,
    S2L_PHASE_PREC_COARSE_RESTRICT = 2,
    S2L_PHASE_PREC_COARSE_SOLVE = 3,
    S2L_PHASE_PREC_COARSE_PROLONG = 4,
    S2L_PHASE_PREC_SPMV = 5,
    S2L_PHASE_PREC_LOCAL = 6,
    S2L_PHASE_SPMV = 7,
    S2L_PHASE_ORTHOGONALIZATION = 8,
    S2L_PHASE_STOPPING_TEST = 9,
    S2L_NPHASES = 10
};
const char* s2l_phase_name(int phase);
int s2l_phase_timing_enable(s2l_ctx* ctx, int on);
int s2l_phase_timing_read(s2l_ctx* ctx, double* ms, int64_t* counts, int reset);

/* Schedule of the batched local solve: 0 = task queue, one launch per depth group of all
 * subdomains' task trees, a subtree per workgroup with its vector in LDS (default); 1 = the same
 * tasks with one launch per subdomain (the comparison point); 2 = level-synchronous, one launch per
 * level of all subdomains' supernodal trees with one wavefront per supernode or row block (no LDS
 * limit on the subdomain size; chosen automatically when a subdomain does not fit mode 0). */
int s2l_set_local_launch_mode(s2l_ctx* ctx, int mode);

/* Work split of the batched local solve: the supernodal elimination tree of every subdomain is cut
 * into subtrees of at most about 1/@p tasks of its dense entries below a skeleton of at most 16
 * supernodes (one task with the root); sibling subtrees are packed into shared tasks. The subtree
 * tasks of all subdomains run in one forward and one backward launch, the top tasks (both sweeps)
 * in between (1 = one task per subdomain, one launch). Splits agree to rounding; each is bitwise
 * repeatable. 0 (default) = compute units / subdomains, at least 1, so that the subtree launch
 * holds about one workgroup per compute unit. Takes effect at the next s2l_finalize_local(). */
int s2l_set_tasks_per_subdomain(s2l_ctx* ctx, int tasks);

/* Dense tops (no reference counterpart; a layout of LocalSolver::solve,
 * src/linalg/local_solver.cpp:119-126): when a subdomain is split into tasks, the skeleton above its
 * subtree tasks (the top separators, a chain of dependent dense supernodes in both sweeps) is
 * replaced by the dense inverse of its Schur complement S^-1 = U_SS^-1 L_SS^-1, computed in FP64 and
 * stored in the factors' precision, whenever the skeleton holds at most @p max_width columns: the
 * top task pulls the subtrees' products into its S rows and multiplies by S^-1, its rows split over
 * as many workgroups as there are tasks per subdomain. Results agree with the sparse skeleton to
 * rounding. 0 turns it off; at most 2048; default 2048. Takes effect at the next
 * s2l_finalize_local(). */
int s2l_set_dense_top(s2l_ctx* ctx, int32_t max_width);

/* Threads of a local-solve workgroup (no reference counterpart): 1024 (default, one workgroup per
 * compute unit) or 512 (two per compute unit, each hiding the other's memory round trips; falls back
 * to 1024 when half the LDS does not hold a subdomain's layout). Pair it with twice the tasks per
 * subdomain. Takes effect at the next s2l_finalize_local(). */
int s2l_set_local_threads(s2l_ctx* ctx, int32_t threads);

/* Autotuner of the local-solve schedule (no reference counterpart; it picks the layout of the
 * batched LocalSolver::solve, src/linalg/local_solver.cpp:119-126). Repacks the uploaded factors
 * (in the context's precision) for every candidate schedule: 1024-thread workgroups with tasks per
 * subdomain around one workgroup per compute unit (1, b/2, b-1, b, b+1, 2b with b = compute units /
 * subdomains), 512-thread workgroups with b and 2b tasks, each with and without dense tops, and the
 * level-synchronous schedule (launch mode 2; not with precision 2); times @p applies batched local
 * solves of @p nrhs right-hand sides of each with HIP events after 3 untimed ones, and keeps the
 * fastest (its tasks, dense-top width, workgroup size and launch mode stay set for later
 * s2l_finalize_local calls). Writes the candidates to @p entries (at most @p max_entries: 16 always
 * suffice), their number to @p count and the index of the one kept to @p best. It repacks from the
 * uploaded host factors: call it before s2l_refactorize, whose new values it would replace. */
typedef struct s2l_tune_entry {
    int32_t tasks, dense_top, launch_mode;
    int32_t threads;  /* threads per local-solve workgroup (s2l_set_local_threads) */
    double us;        /* microseconds per batched local solve of nrhs right-hand sides */
} s2l_tune_entry;
int s2l_autotune_local(s2l_ctx* ctx, int32_t nrhs, int32_t applies, s2l_tune_entry* entries, int32_t max_entries,
                       int32_t* count, int32_t* best);

/* HIP graphs in the device Krylov loops (no counterpart in the reference, whose apply is a host
 * call): on, the kernels of one preconditioner apply and the following SpMV of a GMRES iteration
 * are captured once per input/output buffer pair and replayed as one graph launch; every upload,
 * schedule change or work-buffer growth drops the captured graphs. Off (default) launches every
 * kernel by itself: on MI355X the replays measured within noise of it (the launch gaps left are
 * the host polls of the convergence flags). Results are identical either way. */
int s2l_set_graphs(s2l_ctx* ctx, int on);

#ifdef __cplusplus
}
#endif

#endif /* S2L_HIP_H */
