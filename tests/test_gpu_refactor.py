"""Refactorization on the device (s2l_refactorize): new values of A with the same pattern, the local
Cholesky factors recomputed on the GPU over their symbolic structure and repacked, A_00 = W^T A W
formed, factored and inverted on the GPU, with the subdomains and the coarse basis kept
(two_level_schwarz.cpp:35-41). Checked against the host refactorization uploaded into a new context
and against the oracle on the refactorized setup."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _pair(name, kw, N, nev, t1, **opts):
    from amp_amd import hostsetup as hs
    A0, _, xyz, sym, bs = hs.generate(name, **kw, time=0.0)
    st0 = hs.build_setup(A0, hs.random_rhs(A0.shape[0], 1), N, xyz=xyz, symmetric=sym, block_size=bs, nev=nev, **opts)
    A1 = hs.generate(name, **kw, time=t1)[0]
    assert (A1 != A0).nnz > 0
    st1 = hs.refactorize(st0, A1, hs.random_rhs(A1.shape[0], 2), overlap=opts.get("overlap", 1))
    return st0, A1, st1


@pytest.mark.parametrize("name,kw,N,nev,overlap", [("bubbles3d", dict(nx=16, nbubbles=6, radius=0.12), 4, 6, 1),
                                                   ("bubbles3d", dict(nx=24, nbubbles=8, radius=0.1), 4, 8, 2)])
def test_device_refactorization_matches_host_and_oracle(name, kw, N, nev, overlap):
    """FP64: the apply after the device refactorization equals the apply of a context built from the
    host refactorization and the oracle on it (1e-12); a second refactorization back to the first
    values gives the first apply (1e-12); the SpMV uses the new values bitwise."""
    from amp_amd import device
    print(f"case device refactorization {name} {kw} N={N} overlap {overlap}", flush=True)
    st0, A1, st1 = _pair(name, kw, N, nev, 3.0, overlap=overlap)
    M = device.HipTwoLevelSchwarz(st0, "double")
    r = np.cos(np.linspace(0.0, 7.0, st0.n))
    z0 = M.apply(r)
    assert M.refactorize_device(A1) == "refactorize"
    z1 = M.apply(r)
    zh = device.HipTwoLevelSchwarz(st1, "double").apply(r)
    zo = orc.OracleSchwarz(st1).apply(r)
    print(f"device vs host refactorization {_rel(z1, zh):.2e}, vs oracle {_rel(z1, zo):.2e}", flush=True)
    assert _rel(z1, zh) < 1e-12 and _rel(z1, zo) < 1e-12
    x = np.sin(np.linspace(-1.0, 2.0, st0.n))
    assert np.array_equal(M.ctx.spmv(x), orc.OracleSchwarz(st1).spmv(x))
    from amp_amd import hostsetup as hs
    A0 = hs.generate(name, **kw, time=0.0)[0]
    M.refactorize_device(A0)
    assert _rel(M.apply(r), z0) < 1e-12


@pytest.mark.parametrize("precision", ["single", "mixed"])
def test_device_refactorization_reduced_precision(precision):
    """FP32 factor storage: the device refactorization agrees with the host refactorization uploaded
    in the same precision to the FP32 rounding of the factors (1e-5), and GMRES converges to a true
    residual <= 1.1e-8 with it."""
    from amp_amd import device
    st0, A1, st1 = _pair("bubbles3d", dict(nx=16, nbubbles=6, radius=0.12), 4, 6, 3.0)
    M = device.HipTwoLevelSchwarz(st0, precision)
    M.refactorize_device(A1)
    r = np.cos(np.linspace(0.0, 7.0, st0.n))
    assert _rel(M.apply(r), device.HipTwoLevelSchwarz(st1, precision).apply(r)) < 1e-5
    x, res = device.HipGmres(30, 1e-8, 500).solve(M, st1.dec.b)
    assert res.converged and res.true_rel_residual <= 1.1e-8


def test_device_refactorization_refuses_a_new_pattern():
    from amp_amd import device, hostsetup as hs
    st0, A1, _ = _pair("bubbles3d", dict(nx=12, nbubbles=4, radius=0.14), 4, 4, 2.0)
    M = device.HipTwoLevelSchwarz(st0, "double")
    A2 = A1.tolil()
    A2[0, A1.shape[0] - 1] = 1e-3
    with pytest.raises(ValueError):
        M.refactorize_device(A2.tocsr())


@pytest.mark.parametrize("krylov", ["gmres", "bicgstab"])
def test_time_loop_device_refactorization(krylov):
    """The setup-reuse sequence with the refactorizations on the device: every solve converges to a
    true residual <= 1.1e-8 and takes the iterations of the same sequence refactorized on the host
    (FP64 local factors: the two refactorizations agree to rounding)."""
    from amp_amd.timeloop import run_time_loop
    print(f"case time loop device refactorization {krylov}", flush=True)
    kw = dict(nx=20, nbubbles=6, rho_ratio=1e-3, rmin_cells=2.0, rmax_cells=4.0, cfl=1.0, seed=2024)
    common = dict(nparts=4, steps=3, correctors=1, reuse="refactor", precision="double", restart=30, nev=6,
                  krylov=krylov)
    rd = run_time_loop("bubbles3d", kw, device_refactor=True, **common)
    rh = run_time_loop("bubbles3d", kw, device_refactor=False, **common)
    assert [r["level"] for r in rd] == ["rebuild", "refactorize", "refactorize"]
    for a, b in zip(rd, rh):
        assert a["converged"] and a["true_rel_residual"] <= 1.1e-8
        assert a["iterations"] == b["iterations"]
