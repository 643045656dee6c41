// Host-side checks of the C++ backend that need no GPU: the pattern condition of a refactorization
// (DomainDecomposer::updateValues, src/partition/domain_decomposer.cpp:345-352: the size and the
// sparsity pattern must not change) and the Krylov factory from a SolverConfig
// (src/solver/krylov.cpp:23-36).
#include "s2l_backend.hpp"

#include <cstdio>
#include <stdexcept>
#include <utility>

using namespace s2l;

static int failures = 0;
static void check(bool ok, const char* what)
{
    std::printf("%s: %s\n", ok ? "PASS" : "FAIL", what);
    if (!ok) ++failures;
}

template <typename F> static bool rejected(F f)
{
    try {
        f();
    } catch (const std::invalid_argument&) {
        return true;
    }
    return false;
}

int main()
{
    // 1D Laplacian on 4 rows
    SetupData a;
    a.n = 4;
    a.rowptr = {0, 2, 5, 8, 10};
    a.col = {0, 1, 0, 1, 2, 1, 2, 3, 2, 3};
    a.val = {2, -1, -1, 2, -1, -1, 2, -1, -1, 2};
    SetupData b = a;
    for (double& v : b.val) v *= 3.0;
    check(SetupData::samePattern(a, b), "new values, same pattern: a refactorization");
    SetupData c = a;
    std::swap(c.col[2], c.col[3]);  // same nnz per row, columns permuted within row 1
    check(!SetupData::samePattern(a, c), "columns permuted within a row: another pattern");
    SetupData d = a;
    d.rowptr = {0, 2, 4, 8, 10};  // an entry moved to the next row
    d.col = {0, 1, 0, 1, 1, 2, 2, 3, 2, 3};
    check(!SetupData::samePattern(a, d), "same nnz, an entry in another row: another pattern");
    SetupData e = a;
    e.n = 5;
    check(!SetupData::samePattern(a, e), "another size");

    core::SolverConfig cfg;
    check(solv::makeKrylovSolver(cfg)->name() == "GMRES(30)", "default configuration: GMRES(30)");
    cfg.krylov = "bicgstab";
    check(solv::makeKrylovSolver(cfg, 1e-4, 0.0)->name() == "BiCGStab", "bicgstab");
    cfg.krylov = "cg";
    check(rejected([&] { solv::makeKrylovSolver(cfg); }), "unknown solver rejected");
    cfg.krylov = "gmres";
    cfg.atol = 1e-6;
    check(rejected([&] { solv::makeKrylovSolver(cfg); }), "atol with the l2 norm rejected");
    cfg.stop_norm = "openfoam";
    check(solv::makeKrylovSolver(cfg)->name() == "GMRES(30)", "atol with the OpenFOAM norm");
    cfg.backend = "cpu";
    check(rejected([&] { solv::makeKrylovSolver(cfg); }), "another backend rejected");
    std::printf("%d failure(s)\n", failures);
    return failures ? 1 : 0;
}
