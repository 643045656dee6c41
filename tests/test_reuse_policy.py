"""The reuse policy (prec::ReusePolicy, include/preconditioner/reuse_policy.hpp:20-136) against the
cases of the reference's own test (test/test_reuse.cpp:50-110), through the Python binding of its
C-ABI (include/s2l_reuse.h) and in C++ (a test program built with g++ and run on the CPU): one
implementation, two hosts. A planted mistake (the break-even comparison without the premium) must
fail one of the cases."""
import os
import shutil
import subprocess

import pytest

from amp_amd.reuse import REBUILD, REFACTORIZE, REUSE, ReusePolicy as P

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "algebraic-multilevel-preconditioner_amd")


def _near(a, b):
    return abs(a - b) < 1e-12


def policy_cases(P):
    """The reference's policy checks; returns the names of the failing ones."""
    bad = []
    if not (_near(P.per_decade(10, 1e-5), 2.0) and P.per_decade(0, 1e-5) == 0.0 and P.per_decade(5, 1.0) == 0.0):
        bad.append("per decade")
    a = P("auto", 0.0, 0)
    a.record(REBUILD, 10, 1e-8, True, 2.0, 1.0)
    ok = a.decide(False) == REUSE and a.decide(True) == REFACTORIZE and a.wasted() == 0.0
    a.record(REFACTORIZE, 15, 1e-8, True, 0.2, 1.5)
    ok = ok and _near(a.wasted(), 0.5) and a.decide(True) == REFACTORIZE
    a.record(REUSE, 15, 1e-8, True, 0.0, 1.5)
    a.record(REFACTORIZE, 19, 1e-8, True, 0.2, 1.9)
    ok = ok and _near(a.wasted(), 1.9) and _near(a.rebuild_premium(), 1.8)
    ok = ok and a.decide(True) == REBUILD and a.decide(False) == REUSE
    a.record(REBUILD, 12, 1e-8, True, 2.5, 1.2)
    ok = ok and a.decide(True) == REFACTORIZE and a.wasted() == 0.0 and _near(a.reference(), 1.5)
    if not ok:
        bad.append("break-even")
    g = P("auto", 1.5, 0)
    g.record(REBUILD, 10, 1e-8, True, 1e9, 1.0)
    g.record(REFACTORIZE, 16, 1e-8, True, 0.1, 1.6)
    ok = g.decide(True) == REBUILD
    n = P("auto", 0.0, 0)
    n.record(REBUILD, 10, 1e-8, True, 1e9, 1.0)
    n.record(REFACTORIZE, 500, 1e-3, False, 0.1, 50.0)
    ok = ok and n.decide(True) == REBUILD
    b = P("auto", 0.0, 2)
    b.record(REBUILD, 10, 1e-8, True, 1e9, 1.0)
    ok = ok and b.decide(True) == REFACTORIZE
    b.record(REFACTORIZE, 10, 1e-8, True, 0.1, 1.0)
    ok = ok and b.decide(True) == REFACTORIZE
    b.record(REFACTORIZE, 10, 1e-8, True, 0.1, 1.0)
    ok = ok and b.decide(True) == REBUILD
    if not ok:
        bad.append("safety nets")
    ok = (P.lower_median([]) == 0.0 and P.lower_median([3.0]) == 3.0 and P.lower_median([5.0, 1.0]) == 1.0
          and P.lower_median([4.0, 1.0, 2.0]) == 2.0 and P.lower_median([4.0, 1.0, 2.0, 3.0]) == 2.0)
    s = P("auto", 0.0, 0)
    s.record(REBUILD, 10, 1e-8, True, 2.0, 1.0)
    s.record(REFACTORIZE, 10, 1e-8, True, 0.2, 1.0)
    s.record(REBUILD, 10, 1e-8, True, 30.0, 1.0)
    ok = ok and _near(s.rebuild_premium(), 1.8)
    s.record(REFACTORIZE, 10, 1e-8, True, 9.0, 1.0)
    ok = ok and _near(s.rebuild_premium(), 1.8)
    if not ok:
        bad.append("lower median")
    off = P("off", 0.0, 0)
    if not (off.decide(False) == REBUILD and off.decide(True) == REBUILD):
        bad.append("off")
    r = P("refactor", 0.0, 0)
    r.record(REBUILD, 10, 1e-8, True, 2.0, 1.0)
    r.record(REFACTORIZE, 900, 1e-3, False, 0.1, 90.0)
    if not (r.decide(False) == REUSE and r.decide(True) == REFACTORIZE):
        bad.append("refactor")
    return bad


def test_python_policy_matches_reference_cases():
    assert policy_cases(P) == []
    with pytest.raises(ValueError):
        P("sometimes")


def test_per_kind_references():
    """Solves of another kind (e.g. the second corrector) are measured against their own first solve
    after a rebuild, so a kind that is slower by nature does not count as aging."""
    p = P("auto", 0.0, 0)
    p.record(REBUILD, 10, 1e-8, True, 2.0, 1.0, kind=0)
    p.record(REUSE, 20, 1e-8, True, 0.0, 2.0, kind=1)   # first solve of kind 1: its reference
    assert p.wasted() == 0.0 and _near(p.reference(1), 2.5)
    p.record(REFACTORIZE, 20, 1e-8, True, 0.2, 2.0, kind=1)
    assert p.wasted() == 0.0
    p.record(REFACTORIZE, 12, 1e-8, True, 0.2, 1.2, kind=0)   # 2 iterations over kind 0's reference
    assert _near(p.wasted(), 0.2)


def test_planted_break_even_mistake_is_caught(monkeypatch):
    """Rebuilding as soon as anything is wasted (the premium left out) fails the reference's cases."""
    def decide(self, matrix_changed):
        if self.mode == "off":
            return REBUILD
        if not matrix_changed:
            return REUSE
        return REBUILD if self.wasted() > 0.0 else REFACTORIZE
    monkeypatch.setattr(P, "decide", decide)
    assert "break-even" in policy_cases(P)


def test_c_abi_rejects_bad_arguments():
    """Unknown modes and levels are errors; a NULL policy answers -1 / 0 and is not dereferenced."""
    from amp_amd import _lib
    L = _lib.backend_lib()
    assert not L.s2l_reuse_create(b"sometimes", 0.0, 0)
    assert "sometimes" in L.s2l_reuse_last_error().decode()
    assert not L.s2l_reuse_create(None, 0.0, 0)
    assert L.s2l_reuse_decide(None, 1) == -1
    assert L.s2l_reuse_record(None, 0, 10, 1e-8, 1, 0.0, 1.0, 0) == -1
    assert L.s2l_reuse_age(None) == 0 and L.s2l_reuse_wasted(None) == 0.0
    h = L.s2l_reuse_create(b"auto", 0.0, 0)
    try:
        assert L.s2l_reuse_record(h, 3, 10, 1e-8, 1, 0.0, 1.0, 0) == -1
        assert L.s2l_reuse_record(h, 2, 10, 1e-8, 1, 2.0, 1.0, 0) == 0
        assert _near(L.s2l_reuse_reference(h, 0), 1.25)
    finally:
        L.s2l_reuse_destroy(h)
    with pytest.raises(ValueError):
        P("auto").record("sometimes", 1, 0.1, True, 0.0, 1.0)


def test_cpp_policy_matches_reference_cases(tmp_path):
    """The C++ ReusePolicy of the backend (include/s2l_backend.hpp) on the same cases, on the CPU."""
    lib = os.path.join(PKG, "libs2l_backend.so")
    if not os.path.exists(lib) or shutil.which("g++") is None:
        pytest.skip("the backend library or g++ is missing (build with build.py all)")
    exe = str(tmp_path / "test_reuse_policy")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(REPO, "include"), "-o", exe,
                           os.path.join(REPO, "tests", "cpp", "test_reuse_policy.cpp"), "-L", PKG, "-ls2l_backend",
                           "-ls2l_hip", f"-Wl,-rpath,{PKG}"])
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failure(s)" in p.stdout


def _bubbles(time):
    from amp_amd import hostsetup as hs
    return hs.generate("bubbles3d", nx=12, nbubbles=5, radius=0.14, time=time)


def test_host_refactorize_keeps_basis_and_refreshes_a00():
    """Refactorize on the host (same partition, subdomains and coarse basis; new local factors and
    A_00 = W^T A W): the deflation consistency R_0 (r - A x0) = 0 of the reference's reuse test
    (test/test_reuse.cpp:229,336-346) holds with the new values, which it does not with the stale
    A_00 (planted), and the new factors solve the new local matrices (backward error < 1e-12)."""
    import numpy as np
    from amp_amd import hostsetup as hs
    from oracle import oracle as orc
    A0, _, xyz, sym, bs = _bubbles(0.0)
    st0 = hs.build_setup(A0, hs.random_rhs(A0.shape[0], 1), 4, xyz=xyz, symmetric=sym, block_size=bs, nev=5)
    A1 = _bubbles(4.0)[0]
    assert (A1 != A0).nnz > 0
    st1 = hs.refactorize(st0, A1, hs.random_rhs(A1.shape[0], 2))
    assert all(np.array_equal(v0, v1) for v0, v1 in zip(st0.coarse.V, st1.coarse.V))
    r = np.cos(np.linspace(0.0, 7.0, st1.n))

    def residual_of_deflation(st_values, st_coarse):
        O = orc.OracleSchwarz(st_coarse)
        w = r - st_values.dec.A @ O.coarse(r)
        return max(np.abs(V.T @ w[sd.global_ids()]).max() for sd, V in zip(st_values.subs, st_values.coarse.V))

    assert residual_of_deflation(st1, st1) < 1e-10 * np.abs(r).max()
    assert residual_of_deflation(st1, st0) > 1e-6 * np.abs(r).max()   # stale A_00
    rng = np.random.default_rng(5)
    for sd, f in zip(st1.subs, st1.factors):
        b = rng.standard_normal(sd.n_local)
        x = orc.OracleLocalFactor(f, "double").solve(b)
        err = np.abs(sd.Aii @ x - b).max() / (abs(sd.Aii).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
        assert err < 1e-12
