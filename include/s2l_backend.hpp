// s2l_backend.hpp -- C++17 host classes of the MI355X apply path behind the reference's
// preconditioner and Krylov interfaces, over the C-ABI of s2l_hip.h.
//
// They mirror, with the same names, argument meaning and error behaviour:
//   prec::Preconditioner   include/preconditioner/preconditioner.hpp:64-96 (UpdateLevel :30)
//   prec::TwoLevelSchwarz  src/preconditioner/two_level_schwarz.cpp:129-163 (here HipTwoLevelSchwarz)
//   solv::KrylovResult     include/solver/krylov.hpp:43-54
//   solv::KrylovSolver     include/solver/krylov.hpp:60-76, Gmres / BiCgStab (gmres.hpp, bicgstab.hpp)
//   solv::makeKrylovSolver src/solver/krylov.cpp:23-36 (from core::SolverConfig, include/core/config.hpp:19-117)
//   linalg::DistributedMatrix include/linalg/distributed_matrix.hpp (the operator argument of solve)
// Vectors are std::vector<double> with the owned entries of every subdomain in the renumbered
// global order (all subdomains live in one process and on one GPU). Errors of the C-ABI become
// std::runtime_error with the library's message; invalid arguments std::invalid_argument.
#ifndef S2L_BACKEND_HPP
#define S2L_BACKEND_HPP

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "s2l_hip.h"

namespace s2l {

namespace core {
using Vec = std::vector<double>;

/// The run-time parameters the device path reads, with the names, meanings and defaults of the
/// reference's SolverConfig (include/core/config.hpp:19-117); the setup parameters (partitioning,
/// eigenproblems, local factorization) are the host setup's.
struct SolverConfig {
    // preconditioner
    std::string pou = "boolean";            ///< "boolean" | "multiplicity"
    std::string one_level = "ras";          ///< "ras" | "asm"
    std::string coarse = "deflated";        ///< "deflated" | "additive" | "none"
    std::string local_precision = "double"; ///< "double" | "single" (FP32 factors, solveIn<float>)
    // Krylov solver
    std::string krylov = "gmres";           ///< "gmres" (restarted, right preconditioning) | "bicgstab"
    int restart = 30;                       ///< GMRES restart length
    double rtol = 1e-8;                     ///< relative tolerance
    double atol = 0.0;                      ///< OpenFOAM norm only: absolute tolerance on the normalised residual
    std::string stop_norm = "l2";           ///< "l2" | "openfoam"
    int maxit = 1000;                       ///< maximum applications of the preconditioner
    // added for the device path (the switch at src/main.cpp:143-168 and src/solver/time_loop.cpp:128-151)
    std::string backend = "hip";            ///< "hip": this library
    int device = 0;                         ///< HIP device of the context
};
} // namespace core

namespace linalg {

/// The operator of a Krylov solve (linalg::DistributedMatrix, include/linalg/distributed_matrix.hpp:
/// y = A x on the owned entries). The device Krylov solvers take the operator uploaded with their
/// preconditioner (solv::HipOperator).
class DistributedMatrix {
public:
    virtual ~DistributedMatrix() = default;
    virtual void apply(const core::Vec& x, core::Vec& y) const = 0;
    /// owned rows (all rows: every subdomain lives in this process)
    virtual int32_t nOwned() const = 0;
};

} // namespace linalg

/// What the host setup produced, in the form the device layout is built from (the reference's
/// setup objects: LocalSystem rows, Subdomain index maps and partition of unity, LocalSolver
/// factors, CoarseSpace basis and A_00). Arrays are concatenated over the subdomains.
struct SetupData {
    // operator A in the renumbered order, every row in its owner's local column order
    int32_t n = 0;
    std::vector<int32_t> rowptr, col;
    std::vector<double> val;
    std::vector<int32_t> ownership;     ///< nsub + 1 offsets (subdomain p owns [ownership[p], ownership[p+1]))
    // overlapping subdomains Omega_i = [owned | ghosts sorted]
    std::vector<int32_t> local_ptr;     ///< nsub + 1 offsets into gid
    std::vector<int32_t> gid;           ///< global ids of the local unknowns
    std::vector<double> pou;            ///< D_i concatenated; empty = Boolean partition of unity
    // local factors A_ii(perm, perm) = L U (U^T empty = Cholesky, U = L^T)
    struct Factor {
        int32_t n = 0;
        std::vector<int32_t> perm;
        std::vector<int64_t> l_colptr, ut_colptr;
        std::vector<int32_t> l_row, ut_row;
        std::vector<double> l_val, ut_val;
    };
    std::vector<Factor> factors;
    // coarse space: V_i blocks (rows x m_i column-major; rows = owned rows for the Boolean PoU, all
    // local rows otherwise) and A_00 (n0 x n0 row-major; empty: formed on the device as W^T A W)
    int32_t n0 = 0;
    std::vector<int64_t> coarse_offsets; ///< nsub + 1
    std::vector<double> V, E;
    bool symmetric = true;
    std::string one_level = "ras";       ///< "ras" | "asm"
    std::string correction = "deflated"; ///< "deflated" | "additive"
    std::string precision = "double";    ///< local factors: "double" | "single" | "mixed" (FP32 entries, FP64 arithmetic)
    int nsub() const { return static_cast<int>(local_ptr.size()) - 1; }
    /// Same size and sparsity pattern of A (rowptr and columns): the condition of a refactorization
    /// (DomainDecomposer::updateValues, src/partition/domain_decomposer.cpp:345-352).
    static bool samePattern(const SetupData& a, const SetupData& b);
};

namespace prec {

/// Update levels of a preconditioner for new values of A with the same pattern.
enum class UpdateLevel : int { Reuse = 0, Refactorize = 1, Rebuild = 2 };
const char* toString(UpdateLevel l);

/// Application of a (right) preconditioner z = M^{-1} r (owned entries).
class Preconditioner {
public:
    virtual ~Preconditioner() = default;
    virtual void apply(const core::Vec& r, core::Vec& z) const = 0;
    virtual std::string name() const = 0;
    virtual UpdateLevel update(UpdateLevel /*requested*/) { return UpdateLevel::Reuse; }
    virtual bool upToDate() const { return true; }
    virtual long localFallbacks() const { return 0; }
    virtual long coarseFallbacks() const { return 0; }
};

/// The host setup a device preconditioner is built from and brought up to date by. The
/// reference's TwoLevelSchwarz runs this setup itself from its Subdomain (partition, overlap, local
/// factorizations, eigenproblems, A_00: two_level_schwarz.cpp:14-47); here it stays host code
/// behind this interface and the device side uploads what it returns.
class SetupSource {
public:
    virtual ~SetupSource() = default;
    /// Full setup from the current values of A (Rebuild: factors, new coarse basis and A_00).
    virtual SetupData build() = 0;
    /// New local factors and A_00 = V^T A V from the current values, with the subdomains and the
    /// coarse basis of the last build() (Refactorize, two_level_schwarz.cpp:35-41).
    virtual SetupData refactorize() = 0;
    /// Changes whenever the values of A change (Subdomain::valuesVersion).
    virtual uint64_t valuesVersion() const = 0;
};

/// The two-level Schwarz preconditioner with every subdomain resident on one GPU.
class HipTwoLevelSchwarz final : public Preconditioner {
public:
    /// Creates a context on @p device and uploads the setup (copies it; @p s may be freed).
    explicit HipTwoLevelSchwarz(const SetupData& s, int device = 0);
    /// Builds from @p source (source->build()) and keeps it for update(): the drop-in for
    /// TwoLevelSchwarz(sd, A, cfg) with its update levels.
    explicit HipTwoLevelSchwarz(std::shared_ptr<SetupSource> source, int device = 0);
    ~HipTwoLevelSchwarz() override;
    HipTwoLevelSchwarz(const HipTwoLevelSchwarz&) = delete;
    HipTwoLevelSchwarz& operator=(const HipTwoLevelSchwarz&) = delete;

    /// z = M^{-1} r with host vectors of n owned entries (z is resized).
    void apply(const core::Vec& r, core::Vec& z) const override;
    /// z = M^{-1} r for nrhs device vectors (n x nrhs) on @p stream (a hipStream_t or null).
    void applyDevice(const double* d_r, double* d_z, int nrhs = 1, void* stream = nullptr) const;
    /// "RAS,deflated" and the like (two_level_schwarz.cpp:129-132).
    std::string name() const override;
    /// The preconditioner was built for other values of A (Subdomain::updateValues in the reference).
    void markStale() { up_to_date_ = false; }
    /// Refactorize on the device side: uploads the new values of A, of the local factors and of
    /// A_00 computed by the host (same size, pattern and coarse basis: std::invalid_argument
    /// otherwise). Returns Refactorize.
    UpdateLevel refresh(const SetupData& s);
    /// The semantics of Preconditioner::update (preconditioner.hpp:75-87, two_level_schwarz.cpp:
    /// 30-47): Reuse, or Refactorize of an up-to-date preconditioner, does nothing and returns
    /// Reuse; Refactorize uploads source->refactorize(); Rebuild is always performed and uploads
    /// source->build(). Without a source (built from SetupData) a stale preconditioner is brought
    /// up to date by refresh(), and Refactorize / Rebuild throw std::logic_error.
    UpdateLevel update(UpdateLevel requested) override;
    bool upToDate() const override;

    int32_t n() const { return n_; }
    s2l_ctx* context() const { return ctx_; }
    /// y = A x with host vectors (the operator of the uploaded system).
    void spmv(const core::Vec& x, core::Vec& y) const;

private:
    void create(const SetupData& s, int device);
    void upload(const SetupData& s);
    s2l_ctx* ctx_ = nullptr;
    int32_t n_ = 0;
    std::string name_;
    bool up_to_date_ = true;
    std::shared_ptr<SetupSource> source_;
    uint64_t version_ = 0;
    std::vector<int32_t> rowptr_, col_;  ///< pattern of the uploaded system (refresh checks it)
};

/// Chooses the update level before each solve of a sequence of matrices with one sparsity pattern
/// (prec::ReusePolicy, include/preconditioner/reuse_policy.hpp:20-136). Off rebuilds before every
/// solve; Refactor never rebuilds after the first setup; Auto reuses an unchanged matrix,
/// refactorizes a changed one and rebuilds once the time lost to the aging coarse basis reaches
/// the measured extra cost of a rebuild (the lower medians of the rebuild and refactorization
/// times), or when a safety net fires: no convergence, q above growth times its reference, more
/// than max_age refactorizations. q is iterations per decade of residual reduction; every kind of
/// solve (e.g. a corrector index) has its own reference, the q of its first solve after a rebuild.
class ReusePolicy {
public:
    enum class Mode { Off, Refactor, Auto };
    ReusePolicy(Mode mode, double growth, int max_age) : mode_(mode), growth_(growth), max_age_(max_age) {}
    /// "off" | "refactor" | "auto"; std::invalid_argument otherwise.
    static Mode parseMode(const std::string& s);
    /// Level to perform before the next solve.
    UpdateLevel decide(bool matrix_changed) const;
    /// Outcome of a solve performed after @p done (t_update: its time, t_solve: the solve's) [s].
    void record(UpdateLevel done, int iterations, double reduction, bool converged, double t_update, double t_solve,
                int kind = 0);
    /// iterations / log10(1 / reduction); 0 when undefined (no iteration, reduction outside (0, 1)).
    static double perDecade(int iterations, double reduction);
    /// Element of rank floor((n - 1) / 2) in increasing order; 0 for none.
    static double lowerMedian(std::vector<double> v);
    /// Lower median of the rebuild times minus that of the refactorization times (>= 0; 0 = unknown).
    double rebuildPremium() const;
    double reference(int kind = 0) const;
    double last() const { return last_; }
    int age() const { return age_; }
    double wasted() const { return wasted_; }
    Mode mode() const { return mode_; }

private:
    Mode mode_;
    double growth_;
    int max_age_;
    std::vector<double> ref_;       ///< per kind: q of the first solve after the last rebuild (0 = none yet)
    double last_ = 0.0, last_ref_ = 0.0;
    int age_ = 0;
    bool converged_ = true;
    double wasted_ = 0.0;           ///< seconds lost to the old basis since the last rebuild
    std::vector<double> t_rebuild_, t_refactor_;
};

} // namespace prec

namespace solv {

enum class StopNorm { L2, OpenFOAM };
StopNorm parseStopNorm(const std::string& s);
const char* toString(StopNorm s);

/// Outcome of a Krylov solve (include/solver/krylov.hpp:43-54).
struct KrylovResult {
    bool converged = false;
    bool breakdown = false;
    int iterations = 0;
    int restarts = 0;
    double rel_residual = 0.0;
    double true_rel_residual = 0.0;
    double initial_residual_of = 0.0;
    double final_residual_of = 0.0;
    std::vector<double> history;
};

/// The operator A of a HipTwoLevelSchwarz (the system uploaded with its setup), as the
/// linalg::DistributedMatrix built from the same Subdomain as the preconditioner in the reference.
class HipOperator final : public linalg::DistributedMatrix {
public:
    explicit HipOperator(const prec::HipTwoLevelSchwarz& M) : M_(M) {}
    void apply(const core::Vec& x, core::Vec& y) const override { M_.spmv(x, y); }
    int32_t nOwned() const override { return M_.n(); }
    const prec::HipTwoLevelSchwarz& owner() const { return M_; }

private:
    const prec::HipTwoLevelSchwarz& M_;
};

/// Solves A x = b with a preconditioner M; x is the initial guess in and the solution out
/// (KrylovSolver::solve, include/solver/krylov.hpp:60-76). The device solvers run the whole loop
/// on the GPU: A must be the HipOperator of M's context and M a HipTwoLevelSchwarz, else
/// std::invalid_argument (there is no host loop to fall back to).
class KrylovSolver {
public:
    virtual ~KrylovSolver() = default;
    virtual KrylovResult solve(const linalg::DistributedMatrix& A, const prec::Preconditioner& M, const core::Vec& b,
                               core::Vec& x) const = 0;
    virtual std::string name() const = 0;
};

/// Right-preconditioned restarted GMRES(m) with the whole loop on the device (gmres.cpp:18-161).
/// @p flexible: store M^{-1} v_j (flexible GMRES; no reference counterpart, see s2l_krylov_opts).
class HipGmres final : public KrylovSolver {
public:
    HipGmres(int restart, double rtol, int maxit, StopNorm stop = StopNorm::L2, double atol = 0.0, bool flexible = false);
    KrylovResult solve(const linalg::DistributedMatrix& A, const prec::Preconditioner& M, const core::Vec& b,
                       core::Vec& x) const override;
    std::string name() const override { return "GMRES(" + std::to_string(m_) + ")"; }

private:
    int m_, maxit_;
    double rtol_, atol_;
    StopNorm stop_;
    bool flexible_;
};

/// Right-preconditioned BiCGStab, OpenFOAM's PBiCGStab variant (bicgstab.cpp:15-154).
class HipBiCgStab final : public KrylovSolver {
public:
    HipBiCgStab(double rtol, int maxit, StopNorm stop = StopNorm::L2, double atol = 0.0);
    KrylovResult solve(const linalg::DistributedMatrix& A, const prec::Preconditioner& M, const core::Vec& b,
                       core::Vec& x) const override;
    std::string name() const override { return "BiCGStab"; }

private:
    int maxit_;
    double rtol_, atol_;
    StopNorm stop_;
};

/// Krylov solver of a configuration (src/solver/krylov.cpp:23-36): --krylov gmres (with --restart)
/// or bicgstab; --rtol, --atol, --maxit, --stop-norm. std::invalid_argument for an unknown solver
/// or stopping norm, an absolute tolerance without the OpenFOAM norm, or another backend than "hip".
std::unique_ptr<KrylovSolver> makeKrylovSolver(const core::SolverConfig& cfg);
/// Same, with the tolerances @p rtol and @p atol instead of those of @p cfg (e.g. the last corrector).
std::unique_ptr<KrylovSolver> makeKrylovSolver(const core::SolverConfig& cfg, double rtol, double atol);
/// The same factory from the individual parameters.
std::unique_ptr<KrylovSolver> makeKrylovSolver(const std::string& krylov, int restart, double rtol, double atol,
                                               int maxit, const std::string& stop_norm = "l2");

} // namespace solv

} // namespace s2l

#endif
