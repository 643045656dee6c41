// test_backend.cpp -- the C++ host classes (include/s2l_backend.hpp) on the GPU, checked like the
// reference's own tests: apply against a sequential reference (test/test_preconditioner.cpp:
// 104-184), GMRES convergence and true residual (:187-191, test/test_stopping.cpp:110-117),
// BiCGStab accounting (test/test_bicgstab.cpp:82-109), solver names and configuration
// (test/test_bicgstab.cpp:247-261), update semantics.
//
// Usage: test_backend <setup file> [<refactorized setup> <rebuilt setup>]   (written by
// tests/test_gpu_backend.py: the setup arrays and the oracle's expected results; the optional
// files drive the update levels through a SetupSource). Prints one line per check; exit status 1
// if any failed.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "s2l_backend.hpp"
#include "setup_io.hpp"

using namespace s2l;

namespace {

int failures = 0;

void check(bool ok, const std::string& what)
{
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", what.c_str());
    std::fflush(stdout);
    if (!ok) ++failures;
}

double norm2(const core::Vec& v)
{
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

double relDiff(const core::Vec& a, const core::Vec& b)
{
    core::Vec d(a.size());
    for (size_t i = 0; i < a.size(); ++i) d[i] = a[i] - b[i];
    return norm2(d) / norm2(b);
}

std::string sci(double v)
{
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%.2e", v);
    return buf;
}

SetupData readSetup(const s2ltest::ArrayFile& f)
{
    SetupData s;
    s.n = static_cast<int32_t>(f.scalar("n"));
    s.rowptr = f.i32("rowptr");
    s.col = f.i32("col");
    s.val = f.f64("val");
    s.ownership = f.i32("ownership");
    s.local_ptr = f.i32("local_ptr");
    s.gid = f.i32("gid");
    if (f.has("pou")) s.pou = f.f64("pou");
    for (int i = 0; i < s.nsub(); ++i) {
        const std::string p = "f" + std::to_string(i) + "_";
        SetupData::Factor F;
        F.perm = f.i32(p + "perm");
        F.n = static_cast<int32_t>(F.perm.size());
        F.l_colptr = f.i64(p + "lcp");
        F.l_row = f.i32(p + "lrow");
        F.l_val = f.f64(p + "lval");
        if (f.has(p + "ucp")) {
            F.ut_colptr = f.i64(p + "ucp");
            F.ut_row = f.i32(p + "urow");
            F.ut_val = f.f64(p + "uval");
        }
        s.factors.push_back(std::move(F));
    }
    s.n0 = static_cast<int32_t>(f.scalar("n0"));
    s.coarse_offsets = f.i64("coarse_offsets");
    if (s.n0 > 0) {
        s.V = f.f64("V");
        s.E = f.f64("E");
    }
    s.symmetric = f.scalar("symmetric") != 0;
    s.one_level = f.text("one_level");
    s.correction = f.text("correction");
    s.precision = f.text("precision");
    return s;
}

// The host setup behind the update levels: build() gives the first setup, then the rebuilt one;
// refactorize() the refactorized one; bump() stands for Subdomain::updateValues.
class FileSource final : public prec::SetupSource {
public:
    FileSource(const std::string& first, const std::string& refac, const std::string& rebuilt)
        : first_(first), refac_(refac), rebuilt_(rebuilt)
    {
    }
    SetupData build() override
    {
        ++builds;
        return readSetup(s2ltest::ArrayFile(builds == 1 ? first_ : rebuilt_));
    }
    SetupData refactorize() override
    {
        ++refactorizations;
        return readSetup(s2ltest::ArrayFile(refac_));
    }
    uint64_t valuesVersion() const override { return version_; }
    void bump() { ++version_; }
    int builds = 0, refactorizations = 0;

private:
    std::string first_, refac_, rebuilt_;
    uint64_t version_ = 1;
};

// An operator that is not the device one: the device solvers refuse it (no host loop).
class HostOperator final : public linalg::DistributedMatrix {
public:
    explicit HostOperator(int32_t n) : n_(n) {}
    void apply(const core::Vec& x, core::Vec& y) const override { y = x; }
    int32_t nOwned() const override { return n_; }

private:
    int32_t n_;
};

} // namespace

int main(int argc, char** argv)
{
    if (argc != 2 && argc != 4) {
        std::fprintf(stderr, "usage: %s <setup file> [<refactorized setup> <rebuilt setup>]\n", argv[0]);
        return 2;
    }
    try {
        const s2ltest::ArrayFile f(argv[1]);
        const SetupData s = readSetup(f);
        prec::HipTwoLevelSchwarz M(s);
        const solv::HipOperator A(M);
        const double tol_apply = s.precision == "single" ? 1e-4 : 1e-12;

        // 1. apply against the oracle's result for r = cos(linspace(0, 7, n))
        const core::Vec r = f.f64("r");
        const core::Vec zref = f.f64("z_expected");
        core::Vec z;
        M.apply(r, z);
        const double e = relDiff(z, zref);
        check(z.size() == r.size() && e < tol_apply, M.name() + ": apply vs the oracle (rel. err " + sci(e) + ")");
        check(M.name() == f.text("name"), "name() = " + M.name());

        // 2. GMRES(30) to 1e-8: converged, true residual, history, same iterations as the oracle
        const core::Vec b = f.f64("b");
        auto gm = solv::makeKrylovSolver("gmres", 30, 1e-8, 0.0, 500);
        core::Vec x;
        const solv::KrylovResult g = gm->solve(A, M, b, x);
        core::Vec Ax;
        A.apply(x, Ax);
        core::Vec res(b.size());
        for (size_t i = 0; i < b.size(); ++i) res[i] = b[i] - Ax[i];
        const double t2 = norm2(res) / norm2(b);
        check(g.converged && t2 <= 1.1e-8 && std::abs(t2 - g.true_rel_residual) < 1e-12 &&
                  g.history.size() == static_cast<size_t>(g.iterations) + 1,
              gm->name() + " converged in " + std::to_string(g.iterations) + " it., true residual " + sci(t2));
        if (s.precision == "double") {
            const int want = static_cast<int>(f.scalar("gmres_iterations"));
            check(g.iterations == want, "GMRES iterations equal the oracle's (" + std::to_string(want) + ")");
            const double ex = relDiff(x, f.f64("x_expected"));
            check(ex < 1e-8, "GMRES solution vs the oracle (rel. diff " + sci(ex) + ")");
        }

        // 3. BiCGStab: applications of M = iterations, history of one entry per application
        auto bi = solv::makeKrylovSolver("bicgstab", 30, 1e-8, 0.0, 1000);
        core::Vec y;
        const solv::KrylovResult br = bi->solve(A, M, b, y);
        check(br.converged && !br.breakdown && br.true_rel_residual <= 1.1e-8 &&
                  br.history.size() == static_cast<size_t>(br.iterations) + 1 && bi->name() == "BiCGStab",
              "BiCGStab converged in " + std::to_string(br.iterations) + " applications of M, " + std::to_string(br.restarts) +
                  " restarts");
        if (s.precision == "double") {
            const int want = static_cast<int>(f.scalar("bicgstab_iterations"));
            check(br.iterations == want, "BiCGStab iterations equal the oracle's (" + std::to_string(want) + ")");
        }

        // 4. configuration: names, unknown values and an absolute tolerance without the OpenFOAM norm rejected
        auto rejected = [](auto fn) {
            try {
                fn();
            } catch (const std::invalid_argument&) {
                return true;
            }
            return false;
        };
        check(solv::makeKrylovSolver("gmres", 30, 1e-8, 0.0, 10)->name() == "GMRES(30)" &&
                  rejected([] { solv::makeKrylovSolver("cg", 30, 1e-8, 0.0, 10); }) &&
                  rejected([] { solv::makeKrylovSolver("bicgstab", 30, 1e-8, 1e-3, 10, "l2"); }) &&
                  rejected([] { solv::parseStopNorm("max"); }),
              "configuration: GMRES(30), unknown solver / stopping norm and atol without openfoam rejected");
        check(rejected([&] {
                  core::Vec small(3), out;
                  M.apply(small, out);
              }),
              "apply of a vector of the wrong size rejected");

        // 5. update semantics: up to date -> Reuse; stale -> refresh (Refactorize), Rebuild -> new object
        bool ok = M.update(prec::UpdateLevel::Refactorize) == prec::UpdateLevel::Reuse && M.upToDate();
        M.markStale();
        bool threw = false;
        try {
            M.update(prec::UpdateLevel::Refactorize);
        } catch (const std::logic_error&) {
            threw = true;
        }
        ok = ok && threw && !M.upToDate();
        ok = ok && M.refresh(s) == prec::UpdateLevel::Refactorize && M.upToDate();
        core::Vec z2;
        M.apply(r, z2);
        ok = ok && relDiff(z2, z) == 0.0;
        check(ok, "update: Reuse when up to date, refresh after new values, same apply bits after refresh");

        // 6. a refresh whose pattern differs (two columns of a row swapped, same nnz) is refused
        {
            SetupData p = s;
            size_t row = 0;
            while (row + 1 < p.rowptr.size() && p.rowptr[row + 1] - p.rowptr[row] < 2) ++row;
            std::swap(p.col[static_cast<size_t>(p.rowptr[row])], p.col[static_cast<size_t>(p.rowptr[row]) + 1]);
            check(!SetupData::samePattern(p, s) && rejected([&] { M.refresh(p); }) && M.upToDate(),
                  "refresh with another sparsity pattern of the same size rejected");
        }

        // 7. the solvers of a configuration (makeKrylovSolver(cfg[, rtol, atol])); the device loop
        //    refuses an operator that is not the HIP operator of the preconditioner
        {
            core::SolverConfig cfg;
            cfg.krylov = "gmres";
            cfg.restart = 30;
            cfg.maxit = 500;
            auto k1 = solv::makeKrylovSolver(cfg);
            cfg.krylov = "bicgstab";
            auto k2 = solv::makeKrylovSolver(cfg, 1e-6, 0.0);
            core::SolverConfig bad = cfg;
            bad.backend = "cpu";
            core::SolverConfig bad2 = cfg;
            bad2.atol = 1e-3;
            const HostOperator H(M.n());
            core::Vec xx;
            check(k1->name() == "GMRES(30)" && k2->name() == "BiCGStab" && rejected([&] { solv::makeKrylovSolver(bad); }) &&
                      rejected([&] { solv::makeKrylovSolver(bad2); }) && rejected([&] { k1->solve(H, M, b, xx); }),
                  "makeKrylovSolver(SolverConfig): names, backend and atol checks; a host operator is rejected");
        }

        // 8. GMRES with the OpenFOAM stopping norm (gmres.cpp:106-122): same iterations as the
        //    oracle, sum|b - A x| <= rtol sum|b| from x0 = 0, normalised residuals reported
        if (f.has("gmres_of_iterations")) {
            core::SolverConfig cfg;
            cfg.restart = 30;
            cfg.maxit = 500;
            cfg.stop_norm = "openfoam";
            const double rtol = 1e-6;
            auto gof = solv::makeKrylovSolver(cfg, rtol, 0.0);
            core::Vec xo;
            const solv::KrylovResult ro = gof->solve(A, M, b, xo);
            core::Vec Axo;
            A.apply(xo, Axo);
            double l1 = 0.0, b1 = 0.0;
            for (size_t i = 0; i < b.size(); ++i) {
                l1 += std::abs(b[i] - Axo[i]);
                b1 += std::abs(b[i]);
            }
            const int want = static_cast<int>(f.scalar("gmres_of_iterations"));
            check(ro.converged && (s.precision != "double" || ro.iterations == want) && l1 <= 1.1 * rtol * b1 &&
                      std::abs(ro.initial_residual_of - 1.0) < 1e-12 && std::abs(ro.final_residual_of - l1 / b1) < 1e-9 * rtol,
                  "GMRES, OpenFOAM norm: " + std::to_string(ro.iterations) + " it. (oracle " + std::to_string(want) +
                      "), sum|r| / sum|b| = " + sci(l1 / b1));
        }

        // 9. update levels through a SetupSource (two_level_schwarz.cpp:30-47)
        if (argc == 4) {
            auto src = std::make_shared<FileSource>(argv[1], argv[2], argv[3]);
            prec::HipTwoLevelSchwarz P(src);
            const s2ltest::ArrayFile f1(argv[2]), f2(argv[3]);
            core::Vec za, zb, zc, zd;
            P.apply(r, za);
            const bool fresh = P.upToDate() && P.update(prec::UpdateLevel::Refactorize) == prec::UpdateLevel::Reuse &&
                               relDiff(za, zref) < tol_apply;
            src->bump();
            const bool stale = !P.upToDate();
            const bool refac = P.update(prec::UpdateLevel::Refactorize) == prec::UpdateLevel::Refactorize && P.upToDate() &&
                               src->refactorizations == 1;
            P.apply(r, zb);
            const double e1 = relDiff(zb, f1.f64("z_expected"));
            const bool rebuilt = P.update(prec::UpdateLevel::Rebuild) == prec::UpdateLevel::Rebuild && src->builds == 2;
            P.apply(r, zc);
            const double e2 = relDiff(zc, f2.f64("z_expected"));
            const bool reuse = P.update(prec::UpdateLevel::Reuse) == prec::UpdateLevel::Reuse;
            P.apply(r, zd);
            check(fresh && stale && refac && e1 < tol_apply && rebuilt && e2 < tol_apply && reuse && relDiff(zd, zc) == 0.0,
                  "SetupSource: Refactorize after new values (apply vs oracle " + sci(e1) + "), Rebuild always performed (" +
                      sci(e2) + "), Reuse keeps the bits");
        }
    } catch (const std::exception& ex) {
        std::printf("FAIL exception: %s\n", ex.what());
        return 1;
    }
    std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
    return failures ? 1 : 0;
}
