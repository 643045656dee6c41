/*
 * oracle.c -- CPU restatement (plain C) of the kernels of the two-level Schwarz apply path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker of the HIP path: only tests/, the
 * smoke check and the CPU-baseline leg of bench.py load it. The product path never calls it.
 *
 * Contents (each follows the reference function named beside it):
 *   orc_local_solve_{d,f}  LocalSolver::solve with a factor P A P^T = L U
 *                          (src/linalg/local_solver.cpp:119-126; the float variant is the
 *                          single-precision round trip solveIn<float>, :77-84: b rounded to
 *                          float, both sweeps in float, x widened). The sweeps follow Eigen's
 *                          sparse triangular solves: the lower one column-oriented (axpy form),
 *                          the upper one row-oriented (dot form) over the columns of U^T.
 *   orc_spmv_owned         DistributedMatrix::apply (include/linalg/distributed_matrix.hpp:33-38):
 *                          x_loc = R_i x (halo forward), y_I = A(Omega_I, Omega_i) x_loc, each
 *                          row summed sequentially in local column order with fma.
 *   orc_dense_chol/_lu     the root coarse solver's factorization of A_00
 *                          (src/preconditioner/coarse_solver.cpp:99-138; dense textbook
 *                          Cholesky / LU with partial pivoting instead of the sparse chain)
 *   orc_dense_*_solve      RootCoarseSolver::solve (coarse_solver.cpp:166-179).
 *
 * Parity pins of these functions live in tests/ (dense references, backward errors).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* x = (L U)^{-1} b in the factor ordering: rp[k] = b[perm[k]], x[perm[k]] = rp[k]. */
void orc_local_solve_d(int32_t n, const int32_t* perm, const int64_t* lcp, const int32_t* lr, const double* lv,
                       const int64_t* ucp, const int32_t* ur, const double* uv, const double* b, double* x,
                       double* work)
{
    double* y = work;
    for (int32_t k = 0; k < n; ++k) y[k] = b[perm[k]];
    /* L y = b: column j, divide by the diagonal (first entry), then axpy below. */
    for (int32_t j = 0; j < n; ++j) {
        const int64_t p0 = lcp[j], p1 = lcp[j + 1];
        const double t = y[j] / lv[p0];
        y[j] = t;
        for (int64_t p = p0 + 1; p < p1; ++p) y[lr[p]] -= lv[p] * t;
    }
    /* U x = y, U^T given by columns: row j of U = column j of U^T (diagonal first). */
    for (int32_t j = n - 1; j >= 0; --j) {
        const int64_t p0 = ucp[j], p1 = ucp[j + 1];
        double t = y[j];
        for (int64_t p = p0 + 1; p < p1; ++p) t -= uv[p] * y[ur[p]];
        y[j] = t / uv[p0];
    }
    for (int32_t k = 0; k < n; ++k) x[perm[k]] = y[k];
}

void orc_local_solve_f(int32_t n, const int32_t* perm, const int64_t* lcp, const int32_t* lr, const float* lv,
                       const int64_t* ucp, const int32_t* ur, const float* uv, const double* b, double* x,
                       float* work)
{
    float* y = work;
    for (int32_t k = 0; k < n; ++k) y[k] = (float)b[perm[k]];
    for (int32_t j = 0; j < n; ++j) {
        const int64_t p0 = lcp[j], p1 = lcp[j + 1];
        const float t = y[j] / lv[p0];
        y[j] = t;
        for (int64_t p = p0 + 1; p < p1; ++p) y[lr[p]] -= lv[p] * t;
    }
    for (int32_t j = n - 1; j >= 0; --j) {
        const int64_t p0 = ucp[j], p1 = ucp[j + 1];
        float t = y[j];
        for (int64_t p = p0 + 1; p < p1; ++p) t -= uv[p] * y[ur[p]];
        y[j] = t / uv[p0];
    }
    for (int32_t k = 0; k < n; ++k) x[perm[k]] = (double)y[k];
}

/* y[r] = sum_k A(r, col_k) x[gid[col_k]] for the n_owned rows, in storage order. */
void orc_spmv_owned(int32_t n_owned, const int64_t* rp, const int32_t* col, const double* val, const int64_t* gid,
                    const double* x, double* y)
{
    for (int32_t r = 0; r < n_owned; ++r) {
        double s = 0.0;
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) s = fma(val[k], x[gid[col[k]]], s);
        y[r] = s;
    }
}

/* In-place lower Cholesky of the row-major n x n matrix (upper part untouched). 0 on success. */
int orc_dense_chol(int32_t n, double* A)
{
    for (int32_t j = 0; j < n; ++j) {
        double d = A[(int64_t)j * n + j];
        for (int32_t k = 0; k < j; ++k) d -= A[(int64_t)j * n + k] * A[(int64_t)j * n + k];
        if (!(d > 0.0)) return -1;
        const double ljj = sqrt(d);
        A[(int64_t)j * n + j] = ljj;
        for (int32_t i = j + 1; i < n; ++i) {
            double s = A[(int64_t)i * n + j];
            for (int32_t k = 0; k < j; ++k) s -= A[(int64_t)i * n + k] * A[(int64_t)j * n + k];
            A[(int64_t)i * n + j] = s / ljj;
        }
    }
    return 0;
}

/* b <- (L L^T)^{-1} b with the lower factor of orc_dense_chol. */
void orc_dense_chol_solve(int32_t n, const double* L, double* b)
{
    for (int32_t i = 0; i < n; ++i) {
        double s = b[i];
        for (int32_t k = 0; k < i; ++k) s -= L[(int64_t)i * n + k] * b[k];
        b[i] = s / L[(int64_t)i * n + i];
    }
    for (int32_t i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int32_t k = i + 1; k < n; ++k) s -= L[(int64_t)k * n + i] * b[k];
        b[i] = s / L[(int64_t)i * n + i];
    }
}

/* In-place LU with partial pivoting of the row-major matrix: P A = L U, piv[i] = row swapped with i. */
int orc_dense_lu(int32_t n, double* A, int32_t* piv)
{
    for (int32_t j = 0; j < n; ++j) {
        int32_t p = j;
        double amax = fabs(A[(int64_t)j * n + j]);
        for (int32_t i = j + 1; i < n; ++i)
            if (fabs(A[(int64_t)i * n + j]) > amax) {
                amax = fabs(A[(int64_t)i * n + j]);
                p = i;
            }
        piv[j] = p;
        if (amax == 0.0) return -1;
        if (p != j)
            for (int32_t c = 0; c < n; ++c) {
                const double t = A[(int64_t)j * n + c];
                A[(int64_t)j * n + c] = A[(int64_t)p * n + c];
                A[(int64_t)p * n + c] = t;
            }
        const double d = A[(int64_t)j * n + j];
        for (int32_t i = j + 1; i < n; ++i) {
            const double l = A[(int64_t)i * n + j] / d;
            A[(int64_t)i * n + j] = l;
            if (l != 0.0)
                for (int32_t c = j + 1; c < n; ++c) A[(int64_t)i * n + c] -= l * A[(int64_t)j * n + c];
        }
    }
    return 0;
}

void orc_dense_lu_solve(int32_t n, const double* LU, const int32_t* piv, double* b)
{
    for (int32_t j = 0; j < n; ++j)
        if (piv[j] != j) {
            const double t = b[j];
            b[j] = b[piv[j]];
            b[piv[j]] = t;
        }
    for (int32_t i = 0; i < n; ++i) {
        double s = b[i];
        for (int32_t k = 0; k < i; ++k) s -= LU[(int64_t)i * n + k] * b[k];
        b[i] = s;
    }
    for (int32_t i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int32_t k = i + 1; k < n; ++k) s -= LU[(int64_t)i * n + k] * b[k];
        b[i] = s / LU[(int64_t)i * n + i];
    }
}
