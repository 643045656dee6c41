/*
 * s2l_setup.h -- host-side setup stand-in of the two-level Schwarz apply path.
 *
 * In the reference, the setup that feeds the apply path is its own host code:
 * the matrix generators (src/utils/matrix_generators.cpp:57-306), the METIS
 * partition and renumbering (src/partition/graph_partitioner.cpp:17-94,
 * src/partition/domain_decomposer.cpp:254-269) and the Eigen/UMFPACK local
 * factorizations (src/linalg/local_solver.cpp:436-451). That code cannot ship
 * with this library, so this header exposes a from-scratch replacement that
 * produces the same kind of data (CSR matrix, partition vector, local factors
 * in compressed-column form with a permutation). Everything here runs on the
 * host; none of it is on the timed apply path.
 *
 * Conventions: 0-based indices; CSR/CSC arrays are caller-owned; every
 * function returns 0 on success and a negative code on failure, with a message
 * available from s2l_setup_last_error().
 */
#ifndef S2L_SETUP_H
#define S2L_SETUP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failing call on this thread. */
const char* s2l_setup_last_error(void);

/* ---------------------------------------------------------------- generators */

/* Options of the generators. Mirrors utils::GeneratorOptions
 * (include/utils/matrix_generators.hpp) plus the fields of the extra families. */
typedef struct s2l_gen_opts {
    int nx, ny, nz;          /* cells per direction (ny, nz = 0 -> nx)                  */
    double nu;               /* convdiff/coupled viscosity                               */
    int ncomp;               /* coupled: unknowns per cell                                */
    double time;             /* pseudo-time (bubbles rise, convdiff pulse)                */
    double pulse;            /* convdiff/coupled pulsation amplitude                      */
    int pulse_period;        /* pulsation period [steps]                                  */
    double rho_ratio;        /* bubbles: liquid/gas density ratio                         */
    int nbubbles;            /* bubbles: number of bubbles                                */
    double radius;           /* bubbles: radius (unit domain)                             */
    double cfl;              /* bubbles: rise per step [cells]                            */
    int buoyancy;            /* bubbles: 1 = buoyancy source for the generated rhs        */
    unsigned seed;           /* bubbles / channels: seed of the random geometry           */
    double rmin_cells;       /* bubbles extension: per-bubble radii uniform in            */
    double rmax_cells;       /*   [rmin, rmax] cells when rmax_cells > 0                   */
    double contrast;         /* highcontrast3d: coefficient inside the channels           */
    double channel_fraction; /* highcontrast3d: volume fraction of the channels           */
    double aniso;            /* aniso2d: coefficient of d2/dx2 (d2/dy2 has 1)             */
} s2l_gen_opts;

/* Fills @p o with the reference's defaults (include/core/config.hpp:19-117). */
void s2l_gen_defaults(s2l_gen_opts* o);

/* Opaque generated problem: global CSR (int64 indices) + generated rhs + grid coordinates. */
typedef struct s2l_problem s2l_problem;

/* Generates a matrix. Families:
 *   laplace2d/3d, convdiff2d/3d, coupled2d/3d, bubbles2d/3d  (the reference's generators)
 *   poisson2d/3d    5/7-point finite differences, Dirichlet on every face
 *   aniso2d         5-point -aniso u_xx - u_yy, Dirichlet on every face
 *   highcontrast3d  7-point finite volumes, coefficient 1 with seeded axis-aligned channels
 * Returns NULL on error. */
s2l_problem* s2l_generate(const char* name, const s2l_gen_opts* o);
void s2l_problem_free(s2l_problem* p);
int64_t s2l_problem_n(const s2l_problem* p);
int64_t s2l_problem_nnz(const s2l_problem* p);
int s2l_problem_symmetric(const s2l_problem* p);
int s2l_problem_block_size(const s2l_problem* p);
int s2l_problem_dim(const s2l_problem* p);
/* Copies the CSR arrays (rowptr n+1, col nnz, val nnz), the generated rhs (n) and the cell
 * coordinates (n*3, cell centres, block unknowns share their cell). Any pointer may be NULL. */
void s2l_problem_copy(const s2l_problem* p, int64_t* rowptr, int64_t* col, double* val, double* rhs, double* xyz);

/* Random right-hand side: std::mt19937_64(seed), std::uniform_real_distribution<double>(-1, 1),
 * n draws in the original order (src/partition/domain_decomposer.cpp:158-164). */
void s2l_random_rhs(int64_t n, uint64_t seed, double* out);

/* Right-hand side of one solve of a sequence: seed_seq{seed, step, corrector, 0x5eed}
 * (src/solver/time_loop.cpp:31-43). */
void s2l_random_rhs_seq(int64_t n, uint64_t seed, int step, int corrector, double* out);

/* ---------------------------------------------------------------- Matrix Market files */

/* Reads a sparse matrix (MatrixMarketIO::readMatrix, src/partition/matrix_market_io.cpp:92-140):
 * `matrix coordinate real|integer|pattern general|symmetric|skew-symmetric`; symmetric storage is
 * expanded (the mirror entry negated for skew-symmetric), explicit zeros dropped and duplicates
 * summed. The matrix must be square (src/partition/domain_decomposer.cpp:153). The problem's
 * symmetric flag is the header's (symmetric), its rhs and coordinates are empty and its block size
 * 1. Returns NULL on error (complex / Hermitian / array files, bad sizes or indices, short file). */
s2l_problem* s2l_read_matrix(const char* path);
/* Reads a dense vector (MatrixMarketIO::readVector, matrix_market_io.cpp:142-172): array format
 * with one column, or coordinate format (entries added into their rows). Stores its length in *n
 * and, when out is not NULL, its first min(n, cap) entries into out. Returns 0 or -1. */
int s2l_read_vector(const char* path, double* out, int64_t cap, int64_t* n);
/* Writes A (CSR, int64 indices, n x n) as `coordinate real general` with 17 significant digits
 * (MatrixMarketIO::writeMatrix, matrix_market_io.cpp:174-186). Returns 0 or -1. */
int s2l_write_matrix(const char* path, int64_t n, const int64_t* rowptr, const int64_t* col, const double* val);
/* Writes v as `array real general` with one column (MatrixMarketIO::writeVector, :188-194). */
int s2l_write_vector(const char* path, int64_t n, const double* v);
/* File name of step @p step of a sequence (stepFileName, src/partition/domain_decomposer.cpp:80-97):
 * a pattern with one %d / %Nd / %0Nd conversion is formatted with the step; a name without '%' is
 * returned unchanged. Writes at most cap bytes including the NUL. Returns the length, or -1 on a
 * bad pattern or a buffer that is too short. */
int64_t s2l_step_file_name(const char* pattern, int step, char* out, int64_t cap);
/* Hash of a sparsity pattern (hashPattern, domain_decomposer.cpp:116-130): FNV-1a steps over the
 * 8 little-endian bytes of n, of every row pointer and of every column index. */
uint64_t s2l_hash_pattern(int64_t n, const int64_t* rowptr, const int64_t* col);

/* ---------------------------------------------------------------- OpenFOAM LDU matrices */

/* The global problem of an OpenFOAM lduMatrix, built in memory as the reference's OpenFOAM solver
 * writes it to a Matrix Market file (openfoam/bubblyPressureFoam/bubblyPressureFoam.C:120-170):
 * A(gid[c], gid[c]) = diag[c] (D(), the diagonal with the boundary contributions), for every face
 * f: A(gid[l[f]], gid[u[f]]) = upper[f] and A(gid[u[f]], gid[l[f]]) = lower[f] (lower NULL: a
 * symmetric matrix, lower = upper), b[gid[c]] = source[c]. gid NULL is the identity; source NULL
 * leaves the problem's rhs empty. Exact zeros are dropped and rows sorted, so the result equals
 * the file route bitwise. The problem's symmetric flag is (lower == NULL). Returns NULL on error
 * (a face with equal or out-of-range cells, gid not a permutation of 0..n-1). */
s2l_problem* s2l_from_ldu(int64_t n, int64_t nfaces, const int64_t* lower_addr, const int64_t* upper_addr,
                          const double* diag, const double* upper, const double* lower, const int64_t* gid,
                          const double* source);
/* Length of the problem's right-hand side (n for a generated problem or an LDU one with a source,
 * 0 for a read matrix). */
int64_t s2l_problem_rhs_size(const s2l_problem* p);

/* ---------------------------------------------------------------- partition */

/* Partition of G(A + A^T) into @p nparts parts. With @p xyz (n*3 coordinates) a recursive
 * coordinate bisection along the longest extent; without, a recursive bisection of the graph
 * in breadth-first order from a pseudo-peripheral vertex. Parts are never empty. */
int s2l_partition(int64_t n, const int64_t* rowptr, const int64_t* col, int nparts, const double* xyz,
                  int block_size, int32_t* part);

/* ---------------------------------------------------------------- local factorization */

/* Opaque factor of one local matrix. */
typedef struct s2l_factor s2l_factor;

/* Factorizes the n x n matrix given in CSR (int32 indices) with a nested-dissection ordering
 * and a supernodal left-looking algorithm in FP64.
 *   kind 0: Cholesky L L^T (falls back to LU when a pivot is not positive, as the reference's
 *           chain Cholesky -> LDL^T -> LU, src/linalg/local_solver.cpp:328-378);
 *   kind 1: LU without pivoting on the symmetrized pattern (L unit lower, U upper).
 * @p leaf: maximum size of a nested-dissection leaf (0 = default). Returns NULL on error. */
s2l_factor* s2l_factorize(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, int kind,
                          int leaf);
void s2l_factor_free(s2l_factor* f);
/* Sizes: n, nnz(L) and nnz(U^T) as stored (explicit zeros of the dense supernode blocks
 * included), supernode count, resulting kind (0 Cholesky, 1 LU), flops of the factorization. */
void s2l_factor_info(const s2l_factor* f, int32_t* n, int64_t* nnz_l, int64_t* nnz_ut, int32_t* nsuper,
                     int* kind, double* flops);
/* Copies the factor: perm[k] = local index of factor position k (A(perm, perm) = L U);
 * L in CSC (colptr n+1, rowidx, val; rows sorted, diagonal first); U^T in CSC with the same
 * layout (for Cholesky U^T = L and the same arrays are returned). Pointers may be NULL. */
void s2l_factor_copy(const s2l_factor* f, int32_t* perm, int64_t* l_colptr, int32_t* l_row, double* l_val,
                     int64_t* ut_colptr, int32_t* ut_row, double* ut_val);
/* Supernode partition used by the factorization (first column of every supernode + n). */
void s2l_factor_supernodes(const s2l_factor* f, int32_t* sn_begin);

#ifdef __cplusplus
}
#endif

#endif /* S2L_SETUP_H */
