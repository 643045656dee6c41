"""Pairs of right-hand sides in FP64 arithmetic of FP32 factor entries (precision "mixed"): two
right-hand sides per local-solve workgroup, a double pair per row in LDS, one read of the factors
for both. Each component takes the FP64 operations of one right-hand side in the same order (two
rows per wavefront pass of a dense block either way, columns summed in order whatever the batch of
loads), so a batch gives the bits of its right-hand sides solved one at a time."""
import numpy as np
import pytest

from amp_amd import hostsetup as hs

pytestmark = pytest.mark.gpu

CASES = [("convdiff2d", dict(nx=24), 4), ("poisson3d", dict(nx=16), 8)]


def _setup(name, kw, N, **opts):
    A, _, xyz, sym, bs = hs.generate(name, **kw)
    b = hs.random_rhs(A.shape[0], 12345)
    return hs.build_setup(A, b, N, xyz=xyz, symmetric=sym, block_size=bs, **opts)


@pytest.mark.parametrize("nrhs", [2, 3])
@pytest.mark.parametrize("name,kw,N", CASES)
def test_mixed_pairs_equal_single_applies(name, kw, N, nrhs):
    """Two and three right-hand sides (an odd count leaves the last pair half empty): bitwise the
    applies of one right-hand side at a time, and bitwise repeatable."""
    from amp_amd import device
    st = _setup(name, kw, N, nev=8, overlap=2)
    print(f"case {name} {kw} N={N} mixed pairs nrhs={nrhs}", flush=True)
    M = device.HipTwoLevelSchwarz(st, "mixed")
    rng = np.random.default_rng(17)
    R = rng.standard_normal((nrhs, st.n))
    Z = M.ctx.apply(R)
    Zs = np.stack([M.ctx.apply(R[k]) for k in range(nrhs)])
    print(f"max abs diff {np.abs(Z - Zs).max():.2e}", flush=True)
    assert np.array_equal(Z, Zs)
    assert np.array_equal(Z, M.ctx.apply(R))


@pytest.mark.parametrize("name,kw,N", CASES)
def test_mixed_pairs_with_matrix_core_coarse_products(name, kw, N):
    """Five right-hand sides: the local solves in pairs and the coarse products on the FP64 matrix
    cores (from four right-hand sides, a different summation order than one at a time): every
    right-hand side within 1e-12 of its apply alone."""
    from amp_amd import device
    st = _setup(name, kw, N, nev=8, overlap=2)
    print(f"case {name} {kw} N={N} mixed pairs nrhs=5", flush=True)
    M = device.HipTwoLevelSchwarz(st, "mixed")
    R = np.random.default_rng(19).standard_normal((5, st.n))
    Z = M.ctx.apply(R)
    Zs = np.stack([M.ctx.apply(R[k]) for k in range(5)])
    rel = np.linalg.norm(Z - Zs, axis=1) / np.linalg.norm(Zs, axis=1)
    print(f"max rel diff {rel.max():.2e}", flush=True)
    assert rel.max() < 1e-12


def test_mixed_batched_gmres_equals_solves_alone():
    """A GMRES solve of three right-hand sides (local solves in pairs) gives each right-hand side
    the iterations and bits of its solve alone, and the FP64 iteration count."""
    from amp_amd import device
    st = _setup("convdiff2d", dict(nx=24), 4, nev=8, overlap=2)
    M = device.HipTwoLevelSchwarz(st, "mixed")
    M64 = device.HipTwoLevelSchwarz(st, "double")
    B = np.stack([hs.random_rhs(st.n, 300 + k) for k in range(3)])
    B[2] *= 1e-3
    X, R = M.ctx.krylov(B, None, "gmres", 30, 1e-8, 1000)
    for k in range(3):
        x, r = M.ctx.krylov(B[k], None, "gmres", 30, 1e-8, 1000)
        x64, r64 = M64.ctx.krylov(B[k], None, "gmres", 30, 1e-8, 1000)
        print(f"rhs {k}: batched {R[k].iterations} its, alone {r.iterations}, FP64 {r64.iterations}", flush=True)
        assert R[k].converged and R[k].iterations == r.iterations and np.array_equal(X[k], x)
        assert R[k].iterations <= 1.1 * r64.iterations
        assert np.linalg.norm(X[k] - x64) / np.linalg.norm(x64) < 1e-6
