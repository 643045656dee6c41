// backend.cpp -- C++17 host classes over the C-ABI (include/s2l_backend.hpp).
#include "s2l_backend.hpp"
#include "s2l_reuse.h"
#include <cmath>
#include <algorithm>

#include <stdexcept>

namespace s2l {

namespace {

void check(s2l_ctx* ctx, int rc)
{
    if (rc == S2L_OK) return;
    const std::string msg = s2l_last_error(ctx);
    if (rc == S2L_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

} // namespace

bool SetupData::samePattern(const SetupData& a, const SetupData& b)
{
    return a.n == b.n && a.rowptr == b.rowptr && a.col == b.col;
}

namespace prec {

const char* toString(UpdateLevel l)
{
    switch (l) {
    case UpdateLevel::Reuse: return "reuse";
    case UpdateLevel::Refactorize: return "refactorize";
    case UpdateLevel::Rebuild: return "rebuild";
    }
    return "?";
}

HipTwoLevelSchwarz::HipTwoLevelSchwarz(const SetupData& s, int device) { create(s, device); }

HipTwoLevelSchwarz::HipTwoLevelSchwarz(std::shared_ptr<SetupSource> source, int device) : source_(std::move(source))
{
    if (!source_) throw std::invalid_argument("HipTwoLevelSchwarz: null setup source");
    create(source_->build(), device);
    version_ = source_->valuesVersion();
}

void HipTwoLevelSchwarz::create(const SetupData& s, int device)
{
    if (s.one_level != "ras" && s.one_level != "asm") throw std::invalid_argument("one-level variant must be ras or asm");
    if (s.correction != "deflated" && s.correction != "additive")
        throw std::invalid_argument("coarse correction must be deflated or additive");
    if (s.precision != "double" && s.precision != "single" && s.precision != "mixed")
        throw std::invalid_argument("precision must be double, single or mixed");
    check(nullptr, s2l_ctx_create(device, &ctx_));
    try {
        upload(s);
    } catch (...) {
        s2l_ctx_destroy(ctx_);
        ctx_ = nullptr;
        throw;
    }
    name_ = std::string(s.one_level == "ras" ? "RAS" : "ASM") + (s.correction == "deflated" ? ",deflated" : ",additive");
}

HipTwoLevelSchwarz::~HipTwoLevelSchwarz()
{
    if (ctx_) s2l_ctx_destroy(ctx_);
}

void HipTwoLevelSchwarz::upload(const SetupData& s)
{
    const int nsub = s.nsub();
    if (nsub < 1 || static_cast<int>(s.ownership.size()) != nsub + 1 || static_cast<int>(s.factors.size()) != nsub)
        throw std::invalid_argument("setup: ownership, local_ptr and factors must describe the same subdomains");
    if (static_cast<int32_t>(s.rowptr.size()) != s.n + 1) throw std::invalid_argument("setup: rowptr must have n + 1 entries");
    check(ctx_, s2l_upload_system(ctx_, s.n, s.rowptr.data(), s.col.data(), s.val.data(), nsub, s.ownership.data()));
    check(ctx_, s2l_upload_subdomains(ctx_, nsub, s.local_ptr.data(), s.gid.data(), s.pou.empty() ? nullptr : s.pou.data(),
                                      s.one_level == "asm" ? 1 : 0));
    for (int i = 0; i < nsub; ++i) {
        const SetupData::Factor& f = s.factors[static_cast<size_t>(i)];
        const bool lu = !f.ut_colptr.empty();
        check(ctx_, s2l_upload_local_factor(ctx_, i, f.n, f.perm.data(), f.l_colptr.data(), f.l_row.data(), f.l_val.data(),
                                            lu ? f.ut_colptr.data() : nullptr, lu ? f.ut_row.data() : nullptr,
                                            lu ? f.ut_val.data() : nullptr));
    }
    check(ctx_, s2l_finalize_local(ctx_, s.precision == "single" ? 1 : (s.precision == "mixed" ? 2 : 0)));
    check(ctx_, s2l_upload_coarse(ctx_, s.n0, s.n0 > 0 ? s.coarse_offsets.data() : nullptr, s.n0 > 0 ? s.V.data() : nullptr,
                                  s.n0 > 0 && !s.E.empty() ? s.E.data() : nullptr, s.symmetric ? 1 : 0,
                                  s.correction == "additive" ? 1 : 0));
    n_ = s.n;
    rowptr_ = s.rowptr;
    col_ = s.col;
}

void HipTwoLevelSchwarz::apply(const core::Vec& r, core::Vec& z) const
{
    if (static_cast<int32_t>(r.size()) != n_) throw std::invalid_argument("apply: r must have n entries");
    z.resize(static_cast<size_t>(n_));
    check(ctx_, s2l_apply_host(ctx_, r.data(), z.data(), 1));
}

void HipTwoLevelSchwarz::applyDevice(const double* d_r, double* d_z, int nrhs, void* stream) const
{
    check(ctx_, s2l_apply(ctx_, d_r, d_z, nrhs, stream));
}

void HipTwoLevelSchwarz::spmv(const core::Vec& x, core::Vec& y) const
{
    if (static_cast<int32_t>(x.size()) != n_) throw std::invalid_argument("spmv: x must have n entries");
    y.resize(static_cast<size_t>(n_));
    check(ctx_, s2l_spmv_host(ctx_, x.data(), y.data(), 1));
}

std::string HipTwoLevelSchwarz::name() const { return name_; }

UpdateLevel HipTwoLevelSchwarz::refresh(const SetupData& s)
{
    if (s.n != n_) throw std::invalid_argument("refresh: the system size changed; rebuild instead");
    if (s.rowptr != rowptr_ || s.col != col_)
        throw std::invalid_argument("refresh: the sparsity pattern changed; rebuild instead");
    upload(s);
    up_to_date_ = true;
    return UpdateLevel::Refactorize;
}

bool HipTwoLevelSchwarz::upToDate() const { return source_ ? version_ == source_->valuesVersion() : up_to_date_; }

UpdateLevel HipTwoLevelSchwarz::update(UpdateLevel requested)
{
    // two_level_schwarz.cpp:30-47: Reuse, or Refactorize when up to date, does nothing
    if (requested == UpdateLevel::Reuse || (requested == UpdateLevel::Refactorize && upToDate())) return UpdateLevel::Reuse;
    if (!source_)
        throw std::logic_error(std::string(toString(requested)) +
                               " needs a setup source: build the preconditioner from a SetupSource, or refresh() it");
    if (requested == UpdateLevel::Refactorize) {
        refresh(source_->refactorize());
        version_ = source_->valuesVersion();
        return UpdateLevel::Refactorize;
    }
    // Rebuild is always performed: new partition-dependent data, factors and coarse basis
    const SetupData s = source_->build();
    if (s.one_level != "ras" && s.one_level != "asm") throw std::invalid_argument("one-level variant must be ras or asm");
    upload(s);
    name_ = std::string(s.one_level == "ras" ? "RAS" : "ASM") + (s.correction == "deflated" ? ",deflated" : ",additive");
    version_ = source_->valuesVersion();
    up_to_date_ = true;
    return UpdateLevel::Rebuild;
}

// ---- ReusePolicy (src/preconditioner/reuse_policy.cpp:14-84): the same rules, in this file's terms
ReusePolicy::Mode ReusePolicy::parseMode(const std::string& s)
{
    if (s == "off") return Mode::Off;
    if (s == "refactor") return Mode::Refactor;
    if (s == "auto") return Mode::Auto;
    throw std::invalid_argument("unknown reuse mode '" + s + "' (off|refactor|auto)");
}

double ReusePolicy::perDecade(int iterations, double reduction)
{
    const bool defined = iterations > 0 && reduction > 0.0 && reduction < 1.0;
    return defined ? iterations / std::log10(1.0 / reduction) : 0.0;
}

double ReusePolicy::lowerMedian(std::vector<double> v)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    return v[(v.size() - 1) / 2];
}

double ReusePolicy::rebuildPremium() const
{
    const double rebuild = lowerMedian(t_rebuild_);
    if (!(rebuild > 0.0)) return 0.0;
    return std::max(0.0, rebuild - lowerMedian(t_refactor_));
}

double ReusePolicy::reference(int kind) const
{
    return (kind >= 0 && static_cast<std::size_t>(kind) < ref_.size()) ? ref_[static_cast<std::size_t>(kind)] : 0.0;
}

UpdateLevel ReusePolicy::decide(bool matrix_changed) const
{
    switch (mode_) {
    case Mode::Off:
        return UpdateLevel::Rebuild;
    case Mode::Refactor:
        return matrix_changed ? UpdateLevel::Refactorize : UpdateLevel::Reuse;
    case Mode::Auto:
        break;
    }
    if (!matrix_changed) return UpdateLevel::Reuse;
    const bool not_converged = !converged_;
    const bool too_old = max_age_ > 0 && age_ >= max_age_;
    const bool too_slow = growth_ > 0.0 && last_ref_ > 0.0 && last_ > growth_ * last_ref_;
    const double premium = rebuildPremium();
    const bool break_even = premium > 0.0 && wasted_ >= premium;
    return (not_converged || too_old || too_slow || break_even) ? UpdateLevel::Rebuild : UpdateLevel::Refactorize;
}

void ReusePolicy::record(UpdateLevel done, int iterations, double reduction, bool converged, double t_update,
                         double t_solve, int kind)
{
    if (done == UpdateLevel::Rebuild) {
        // a fresh basis: every kind measures a new reference, nothing is wasted yet
        age_ = 0;
        wasted_ = 0.0;
        last_ref_ = 0.0;
        std::fill(ref_.begin(), ref_.end(), 0.0);
        if (t_update > 0.0) t_rebuild_.push_back(t_update);
    } else if (done == UpdateLevel::Refactorize) {
        age_ += 1;
        if (t_update > 0.0) t_refactor_.push_back(t_update);
    }
    converged_ = converged;
    const double q = perDecade(iterations, reduction);
    if (!(q > 0.0)) return;
    last_ = q;
    const std::size_t k = static_cast<std::size_t>(std::max(kind, 0));
    if (ref_.size() <= k) ref_.resize(k + 1, 0.0);
    if (!(ref_[k] > 0.0)) {
        ref_[k] = q;
        last_ref_ = q;
        return;
    }
    last_ref_ = ref_[k];
    // iterations beyond those the reference rate needs for this reduction, at this solve's cost each
    const double extra = iterations - ref_[k] * std::log10(1.0 / reduction);
    if (extra > 0.0) wasted_ += extra * (t_solve / iterations);
}

} // namespace prec

namespace solv {

StopNorm parseStopNorm(const std::string& s)
{
    if (s == "l2") return StopNorm::L2;
    if (s == "openfoam") return StopNorm::OpenFOAM;
    throw std::invalid_argument("unknown stopping norm '" + s + "' (l2|openfoam)");
}

const char* toString(StopNorm s) { return s == StopNorm::L2 ? "l2" : "openfoam"; }

namespace {

const prec::HipTwoLevelSchwarz& deviceSide(const linalg::DistributedMatrix& A, const prec::Preconditioner& M)
{
    const auto* hm = dynamic_cast<const prec::HipTwoLevelSchwarz*>(&M);
    if (!hm) throw std::invalid_argument("the device Krylov solvers need a HipTwoLevelSchwarz preconditioner");
    const auto* op = dynamic_cast<const HipOperator*>(&A);
    if (!op) throw std::invalid_argument("the device Krylov solvers need the HipOperator of the preconditioner");
    if (op->owner().context() != hm->context())
        throw std::invalid_argument("the operator and the preconditioner must share their context");
    return *hm;
}

KrylovResult run(const linalg::DistributedMatrix& A, const prec::Preconditioner& M, const core::Vec& b, core::Vec& x, int method,
                 int restart, double rtol, double atol, int maxit, StopNorm stop, bool flexible)
{
    const prec::HipTwoLevelSchwarz& hm = deviceSide(A, M);
    const size_t n = static_cast<size_t>(hm.n());
    if (b.size() != n) throw std::invalid_argument("solve: b must have n entries");
    if (x.size() != n) x.assign(n, 0.0);
    s2l_krylov_opts o{};
    o.method = method;
    o.restart = restart;
    o.rtol = rtol;
    o.atol = atol;
    o.maxit = maxit;
    o.stop_norm = stop == StopNorm::OpenFOAM ? 1 : 0;
    o.flexible = flexible ? 1 : 0;
    s2l_krylov_result r{};
    std::vector<double> hist(static_cast<size_t>(maxit) + 1);
    check(hm.context(), s2l_krylov_solve_host(hm.context(), &o, b.data(), x.data(), 1, &r, hist.data()));
    KrylovResult out;
    out.converged = r.converged != 0;
    out.breakdown = r.breakdown != 0;
    out.iterations = r.iterations;
    out.restarts = r.restarts;
    out.rel_residual = r.rel_residual;
    out.true_rel_residual = r.true_rel_residual;
    out.initial_residual_of = r.initial_residual_of;
    out.final_residual_of = r.final_residual_of;
    out.history.assign(hist.begin(), hist.begin() + r.history_len);
    return out;
}

} // namespace

HipGmres::HipGmres(int restart, double rtol, int maxit, StopNorm stop, double atol, bool flexible)
    : m_(restart), maxit_(maxit), rtol_(rtol), atol_(atol), stop_(stop), flexible_(flexible)
{
    if (restart < 1) throw std::invalid_argument("GMRES restart must be >= 1");
    if (atol > 0.0 && stop != StopNorm::OpenFOAM) throw std::invalid_argument("an absolute tolerance needs the OpenFOAM stopping norm");
}

KrylovResult HipGmres::solve(const linalg::DistributedMatrix& A, const prec::Preconditioner& M, const core::Vec& b, core::Vec& x) const
{
    return run(A, M, b, x, 0, m_, rtol_, atol_, maxit_, stop_, flexible_);
}

HipBiCgStab::HipBiCgStab(double rtol, int maxit, StopNorm stop, double atol)
    : maxit_(maxit), rtol_(rtol), atol_(atol), stop_(stop)
{
}

KrylovResult HipBiCgStab::solve(const linalg::DistributedMatrix& A, const prec::Preconditioner& M, const core::Vec& b,
                                core::Vec& x) const
{
    return run(A, M, b, x, 1, 0, rtol_, atol_, maxit_, stop_, false);
}

std::unique_ptr<KrylovSolver> makeKrylovSolver(const std::string& krylov, int restart, double rtol, double atol, int maxit,
                                               const std::string& stop_norm)
{
    const StopNorm stop = parseStopNorm(stop_norm);
    if (atol > 0.0 && stop != StopNorm::OpenFOAM) throw std::invalid_argument("an absolute tolerance needs the OpenFOAM stopping norm");
    if (krylov == "gmres") return std::make_unique<HipGmres>(restart, rtol, maxit, stop, atol);
    if (krylov == "bicgstab") return std::make_unique<HipBiCgStab>(rtol, maxit, stop, atol);
    throw std::invalid_argument("unknown Krylov solver '" + krylov + "' (gmres|bicgstab)");
}

std::unique_ptr<KrylovSolver> makeKrylovSolver(const core::SolverConfig& cfg, double rtol, double atol)
{
    if (cfg.backend != "hip") throw std::invalid_argument("unknown backend '" + cfg.backend + "' (this library: hip)");
    return makeKrylovSolver(cfg.krylov, cfg.restart, rtol, atol, cfg.maxit, cfg.stop_norm);
}

std::unique_ptr<KrylovSolver> makeKrylovSolver(const core::SolverConfig& cfg) { return makeKrylovSolver(cfg, cfg.rtol, cfg.atol); }

} // namespace solv

} // namespace s2l

// ---- C-ABI of the reuse policy (include/s2l_reuse.h): forwards to prec::ReusePolicy
struct s2l_reuse {
    s2l::prec::ReusePolicy policy;
};

namespace {
thread_local std::string g_reuse_error;
}

extern "C" {

s2l_reuse* s2l_reuse_create(const char* mode, double growth, int max_age)
{
    try {
        if (mode == nullptr) throw std::invalid_argument("no reuse mode (off|refactor|auto)");
        return new s2l_reuse{s2l::prec::ReusePolicy(s2l::prec::ReusePolicy::parseMode(mode), growth, max_age)};
    } catch (const std::exception& e) {
        g_reuse_error = e.what();
        return nullptr;
    }
}

void s2l_reuse_destroy(s2l_reuse* p) { delete p; }

const char* s2l_reuse_last_error(void) { return g_reuse_error.c_str(); }

int s2l_reuse_decide(const s2l_reuse* p, int matrix_changed)
{
    if (p == nullptr) return -1;
    return static_cast<int>(p->policy.decide(matrix_changed != 0));
}

int s2l_reuse_record(s2l_reuse* p, int done, int iterations, double reduction, int converged, double t_update,
                     double t_solve, int kind)
{
    if (p == nullptr || done < 0 || done > 2) return -1;
    p->policy.record(static_cast<s2l::prec::UpdateLevel>(done), iterations, reduction, converged != 0, t_update, t_solve,
                     kind);
    return 0;
}

double s2l_reuse_last(const s2l_reuse* p) { return p ? p->policy.last() : 0.0; }
double s2l_reuse_reference(const s2l_reuse* p, int kind) { return p ? p->policy.reference(kind) : 0.0; }
int s2l_reuse_age(const s2l_reuse* p) { return p ? p->policy.age() : 0; }
double s2l_reuse_wasted(const s2l_reuse* p) { return p ? p->policy.wasted() : 0.0; }
double s2l_reuse_rebuild_premium(const s2l_reuse* p) { return p ? p->policy.rebuildPremium() : 0.0; }

double s2l_reuse_per_decade(int iterations, double reduction)
{
    return s2l::prec::ReusePolicy::perDecade(iterations, reduction);
}

double s2l_reuse_lower_median(const double* v, int n)
{
    if (v == nullptr || n <= 0) return 0.0;
    return s2l::prec::ReusePolicy::lowerMedian(std::vector<double>(v, v + n));
}

} // extern "C"
