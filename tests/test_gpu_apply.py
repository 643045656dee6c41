"""GPU parity of the apply path against the CPU oracle (through the C-ABI)."""
import numpy as np
import pytest

from amp_amd import hostsetup as hs

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

CASES = [("laplace2d", dict(nx=20), 4), ("convdiff2d", dict(nx=24), 4), ("coupled2d", dict(nx=12, ncomp=2), 4),
         ("convdiff3d", dict(nx=8), 4), ("poisson3d", dict(nx=16), 8)]


def _setup(name, kw, N, **opts):
    A, _, xyz, sym, bs = hs.generate(name, **kw)
    b = hs.random_rhs(A.shape[0], 12345)
    return hs.build_setup(A, b, N, xyz=xyz, symmetric=sym, block_size=bs, **opts)


@pytest.mark.parametrize("name,kw,N", CASES)
def test_apply_fp64_matches_oracle(name, kw, N):
    from amp_amd import device
    st = _setup(name, kw, N, nev=8)
    print(f"case {name} {kw} N={N}", flush=True)
    M = device.HipTwoLevelSchwarz(st, "double")
    r = np.cos(np.linspace(0.0, 7.0, st.n))
    z = M.apply(r)
    zo = orc.OracleSchwarz(st).apply(r)
    assert np.linalg.norm(z - zo) / np.linalg.norm(zo) < 1e-12


@pytest.mark.parametrize("name,kw,N", CASES[:2])
def test_spmv_bitwise(name, kw, N):
    from amp_amd import device
    st = _setup(name, kw, N, nev=4)
    M = device.HipTwoLevelSchwarz(st, "double")
    x = np.sin(np.linspace(-1.0, 2.0, st.n))
    y = M.ctx.spmv(x)
    yo = orc.OracleSchwarz(st).spmv(x)
    assert np.array_equal(y, yo)


@pytest.mark.parametrize("name,kw,N", CASES)
def test_gmres_fp64_matches_oracle(name, kw, N):
    from amp_amd import device
    st = _setup(name, kw, N, nev=8)
    print(f"case {name} {kw} N={N}", flush=True)
    M = device.HipTwoLevelSchwarz(st, "double")
    x, res = device.HipGmres(30, 1e-8, 500).solve(M, st.dec.b)
    O = orc.OracleSchwarz(st)
    xo, ro = orc.gmres(O.spmv, O.apply, st.dec.b, restart=30, rtol=1e-8, maxit=500)
    assert res.converged and ro.converged
    assert res.iterations == ro.iterations
    assert res.true_rel_residual < 1e-7
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
    assert len(res.history) == res.iterations + 1 + 0 * len(ro.history)
