// eigen.hpp -- the local eigenproblems of the coarse space on the device: for every subdomain the
// largest eigenpairs of B_i u = lambda (A~_ii + sigma_i I) u, B_i = D_i A_ii D_i, A~_ii the lumped
// block (local_eigensolver.cpp:215-226, 277-421). All subdomains run one thick-restart Lanczos
// process each, in lockstep: the operator (A~_ii + sigma_i I)^{-1} B_i is a batched sparse
// product and a batched local solve with the factors of the shifted blocks (refactorized on the
// device from the structure of A_ii's factors), inner products are in the (A~_ii + sigma_i I) norm,
// and the projected matrices are diagonalized by a batched symmetric eigensolver.
#ifndef S2L_EIGEN_HPP
#define S2L_EIGEN_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace s2l {

/// Rows of a subdomain handled by one workgroup of the vector kernels.
constexpr int kEigChunk = 1024;
/// Largest Krylov basis (ncv) of the device eigensolver.
constexpr int kEigMaxM = 96;

/// Host-built tables of the eigensolver (pattern only; the values come from A at every call).
struct EigPlan {
    std::vector<int32_t> chunk;      ///< per chunk: subdomain, first row, end row (local positions)
    std::vector<int32_t> chunk_ptr;  ///< first chunk of every subdomain, + end
    std::vector<int32_t> lrp;        ///< A_ii of all subdomains: CSR over the local positions
    std::vector<int32_t> lcol;       ///< local position of every entry's column
    std::vector<int32_t> lsrc;       ///< index of every entry in the stored values of A
    std::vector<int32_t> orp;        ///< per local row: entries of its global row outside the subdomain
    std::vector<int32_t> osrc;
    std::vector<int32_t> ghost;      ///< 1 on the ghost rows
    std::vector<double> D;           ///< D_i by local position
};

EigPlan buildEigPlan(const std::vector<int32_t>& local_ptr, const std::vector<int32_t>& gid,
                     const std::vector<int32_t>& ownership, const std::vector<int32_t>& rp, const std::vector<int32_t>& col,
                     const std::vector<double>& pou, bool boolean);

struct EigArgs {
    int32_t nsub, nchunk, m;
    int64_t nl;                 ///< local positions of all subdomains (the column stride of every block)
    const int32_t* lptr;        ///< nsub + 1
    const int32_t* chunk;       ///< 3 per chunk
    const int32_t* chunk_ptr;   ///< nsub + 1
    const int32_t* lrp;
    const int32_t* lcol;
    const int32_t* lsrc;
    const int32_t* orp;
    const int32_t* osrc;
    const int32_t* ghost;
    const double* D;
    const double* val;          ///< A, stored order
    double* adj;                ///< diagonal of A~_ii + sigma_i I minus that of A_ii (by local position)
    const int32_t* done;        ///< per subdomain: nonzero = frozen (converged or not solved here)
};

/// adj = -lump on the ghost rows, 0 elsewhere; rowsum = sum |A~_ii(l, :)| (for sigma_i)
void launchEigLump(const EigArgs& a, double* rowsum, hipStream_t s);
/// sigma_i = shift_rel * max_l rowsum (per subdomain), adj += scale * sigma_i
void launchEigShift(const EigArgs& a, const double* rowsum, double shift_rel, double* sigma, hipStream_t s);
void launchEigAddShift(const EigArgs& a, const double* sigma, double factor, hipStream_t s);
/// y = D A_ii D x (b = 1) or y = (A~_ii + sigma_i I) x (b = 0), one vector of all subdomains
void launchEigSpmv(const EigArgs& a, int b, const double* x, double* y, hipStream_t s);
/// a seeded pseudo-random vector in [-1, 1) per subdomain
void launchEigStart(const EigArgs& a, double* v, hipStream_t s);
/// partial dot products X[:, c]^T w over every chunk, c < ncols (columns stride nl)
void launchEigDots(const EigArgs& a, const double* X, int ncols, const double* w, double* part, hipStream_t s);
/// per subdomain: the sums of the chunks; mode 0 hb = s; 1 hb2 = s and H(:, j) = hb + s; 2 beta = sqrt(s)
void launchEigDotsFinal(const EigArgs& a, const double* part, int ncols, int mode, double* hb, double* hb2, double* H, int j,
                        double* beta, hipStream_t s);
/// w -= V[:, 0..ncols) h (h per subdomain, kEigMaxM + 1 apart)
void launchEigUpdate(const EigArgs& a, const double* V, int ncols, const double* h, double* w, hipStream_t s);
/// V[:, col] = w / beta, AV[:, col] = Aw / beta
void launchEigScale(const EigArgs& a, double* V, double* AV, int col, const double* w, const double* Aw, const double* beta,
                    hipStream_t s);
/// convergence of the nev largest Ritz pairs from the eigenvectors S (column-major, ascending) and
/// eigenvalues theta of H: conv = |beta S(m-1, k)| <= tol |theta_k|; lam = the nev largest; live = count
void launchEigCheck(const EigArgs& a, const double* S, const double* theta, const double* beta, int nev, double tol,
                    int32_t* conv, double* lam, int32_t* live, hipStream_t s);
/// thick restart: V[:, c] = V[:, 0..m) S[:, m-1-c] for c < kk (and AV), V[:, kk] = V[:, m]
void launchEigRotate(const EigArgs& a, double* V, double* AV, const double* S, int kk, hipStream_t s);
/// H = diag(theta_{m-1}, ..., theta_{m-kk}) of the restarted basis; newly converged subdomains freeze
void launchEigReset(const EigArgs& a, double* H, const double* theta, int kk, const int32_t* conv, int32_t* done,
                    hipStream_t s);

} // namespace s2l

#endif
