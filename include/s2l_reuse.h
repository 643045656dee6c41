/*
 * s2l_reuse.h -- C-ABI of the reuse policy of the C++ host classes (prec::ReusePolicy in
 * s2l_backend.hpp), for hosts that are not C++ (the Python time loop binds it with ctypes).
 *
 * Replaces the reference's prec::ReusePolicy (include/preconditioner/reuse_policy.hpp:20-136,
 * src/preconditioner/reuse_policy.cpp:14-84): the update level to perform before each solve of a
 * sequence of matrices with one sparsity pattern. There is one implementation, in
 * libs2l_backend.so; these functions forward to it.
 *
 * Levels: 0 reuse, 1 refactorize, 2 rebuild (UpdateLevel of s2l_backend.hpp).
 */
#ifndef S2L_REUSE_H
#define S2L_REUSE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct s2l_reuse s2l_reuse;

/* New policy; mode "off" | "refactor" | "auto", growth the safety net on q (<= 0: none), max_age the
 * most refactorizations between rebuilds (<= 0: none). NULL for an unknown mode (message from
 * s2l_reuse_last_error()). Replaces ReusePolicy::ReusePolicy (reuse_policy.hpp:66-73). */
s2l_reuse* s2l_reuse_create(const char* mode, double growth, int max_age);
void s2l_reuse_destroy(s2l_reuse* p);
/* Message of the last failing call on this thread. */
const char* s2l_reuse_last_error(void);

/* Level to perform before the next solve (reuse_policy.cpp:41-52); -1 for a NULL policy. */
int s2l_reuse_decide(const s2l_reuse* p, int matrix_changed);
/* Outcome of a solve performed after level `done` (reuse_policy.cpp:54-84): its iterations, residual
 * reduction, convergence, the update's and the solve's times [s] and the kind of solve (each kind
 * keeps its own reference rate). Returns 0, or -1 for a NULL policy or a level outside 0..2. */
int s2l_reuse_record(s2l_reuse* p, int done, int iterations, double reduction, int converged, double t_update,
                     double t_solve, int kind);

/* State (0 for a NULL policy): q of the last solve, the reference q of a kind, refactorizations since
 * the last rebuild, seconds lost to the old basis since then, the measured extra cost of a rebuild. */
double s2l_reuse_last(const s2l_reuse* p);
double s2l_reuse_reference(const s2l_reuse* p, int kind);
int s2l_reuse_age(const s2l_reuse* p);
double s2l_reuse_wasted(const s2l_reuse* p);
double s2l_reuse_rebuild_premium(const s2l_reuse* p);

/* iterations / log10(1 / reduction); 0 when undefined. */
double s2l_reuse_per_decade(int iterations, double reduction);
/* Element of rank floor((n - 1) / 2) in increasing order of v[0 .. n); 0 for n <= 0. */
double s2l_reuse_lower_median(const double* v, int n);

#ifdef __cplusplus
}
#endif

#endif
