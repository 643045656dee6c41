"""GPU parity of the apply path against the CPU oracle (through the C-ABI)."""
import numpy as np
import pytest

from amp_amd import hostsetup as hs

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

CASES = [("laplace2d", dict(nx=20), 4), ("convdiff2d", dict(nx=24), 4), ("coupled2d", dict(nx=12, ncomp=2), 4),
         ("convdiff3d", dict(nx=8), 4), ("poisson3d", dict(nx=16), 8)]


def _setup(name, kw, N, **opts):
    """Host setup of a generated problem on N subdomains."""
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
    assert len(res.history) == len(ro.history) == res.iterations + 1
    assert np.allclose(res.history, ro.history, rtol=1e-6, atol=1e-14)


@pytest.mark.parametrize("name,kw,N", [CASES[0], CASES[1], CASES[4]])
def test_local_solve_fp64_matches_oracle(name, kw, N):
    from amp_amd import device
    st = _setup(name, kw, N, nev=4)
    M = device.HipTwoLevelSchwarz(st, "double")
    rng = np.random.default_rng(7)
    for i in range(N):
        f = st.factors[i]
        b = rng.standard_normal(f.n)
        y = M.ctx.local_solve(i, b)
        yo = orc.OracleLocalFactor(f, "double").solve(b)
        assert np.linalg.norm(y - yo) / np.linalg.norm(yo) < 1e-12, i


@pytest.mark.parametrize("name,kw,N", CASES)
def test_apply_fp32_as_accurate_as_oracle_float_path(name, kw, N):
    """FP32 local factors: the device result is as close to the FP64 apply as the reference's
    single-precision round trip (solveIn<float>) restated in the oracle."""
    from amp_amd import device
    st = _setup(name, kw, N, nev=8)
    M = device.HipTwoLevelSchwarz(st, "single")
    r = np.cos(np.linspace(0.0, 7.0, st.n))
    z = M.apply(r)
    z64 = orc.OracleSchwarz(st, "double").apply(r)
    z32o = orc.OracleSchwarz(st, "single").apply(r)
    e_dev = np.linalg.norm(z - z64) / np.linalg.norm(z64)
    e_orc = np.linalg.norm(z32o - z64) / np.linalg.norm(z64)
    assert e_dev < 1e-4
    assert e_dev <= 3.0 * e_orc + 1e-7


@pytest.mark.parametrize("name,kw,N", [CASES[0], CASES[1], CASES[4]])
def test_apply_mixed_precision_linear_and_accurate(name, kw, N):
    """FP32 factor entries in FP64 arithmetic (precision "mixed", no reference counterpart): the
    device apply is linear to FP64 rounding (1e-13), its error against the FP64 oracle is that of
    rounding the factors (within 3x the oracle's FP32 path, like the FP32 mode), it agrees with the
    oracle's mixed restatement to the difference of the two factor roundings, and the local solves
    agree with the oracle's."""
    from amp_amd import device
    st = _setup(name, kw, N, nev=8)
    print(f"case {name} {kw} N={N} mixed", flush=True)
    M = device.HipTwoLevelSchwarz(st, "mixed")
    r1 = np.cos(np.linspace(0.0, 7.0, st.n))
    r2 = np.sin(np.linspace(0.0, 3.0, st.n))
    z1, z2, z12 = M.apply(r1), M.apply(r2), M.apply(2.0 * r1 - 3.0 * r2)
    assert np.linalg.norm(z12 - (2.0 * z1 - 3.0 * z2)) / np.linalg.norm(z12) < 1e-13
    assert np.array_equal(z1, M.apply(r1))
    z64 = orc.OracleSchwarz(st, "double").apply(r1)
    z32o = orc.OracleSchwarz(st, "single").apply(r1)
    zmo = orc.OracleSchwarz(st, "mixed").apply(r1)
    e_dev = np.linalg.norm(z1 - z64) / np.linalg.norm(z64)
    assert e_dev <= 3.0 * np.linalg.norm(z32o - z64) / np.linalg.norm(z64) + 1e-9
    assert np.linalg.norm(z1 - zmo) / np.linalg.norm(zmo) < 1e-5
    rng = np.random.default_rng(3)
    for i in range(N):
        f = st.factors[i]
        b = rng.standard_normal(f.n)
        y = M.ctx.local_solve(i, b)
        yo = orc.OracleLocalFactor(f, "double").solve(b)
        assert np.linalg.norm(y - yo) / np.linalg.norm(yo) < 1e-5, i


def test_mixed_precision_refuses_level_synchronous_mode():
    """The level-synchronous schedule stores factors in its arithmetic type: mode 2 is refused for
    the mixed precision, and the context keeps working in the task queue."""
    from amp_amd import device
    st = _setup("laplace2d", dict(nx=20), 4, nev=4)
    M = device.HipTwoLevelSchwarz(st, "mixed")
    with pytest.raises(Exception):
        M.ctx.set_launch_mode(2)
    r = np.cos(np.linspace(0.0, 7.0, st.n))
    assert np.all(np.isfinite(M.apply(r)))


@pytest.mark.parametrize("precision", ["double", "single"])
def test_apply_bitwise_repeatable_and_launch_modes_agree(precision):
    """Unchanged matrix -> bitwise identical apply (test/test_reuse.cpp:309-316) in every launch
    mode; the task queue in one launch per depth group and in one launch per subdomain give the
    same bits; the level-synchronous launches agree with them to rounding."""
    from amp_amd import device
    st = _setup("poisson3d", dict(nx=16), 8, nev=8, overlap=2)
    M = device.HipTwoLevelSchwarz(st, precision)
    r = np.cos(np.linspace(0.0, 7.0, st.n))
    out = {}
    for mode in (2, 0, 1):
        M.ctx.set_launch_mode(mode)
        z0 = M.apply(r)
        z1 = M.apply(r)
        assert np.array_equal(z0, z1), mode
        out[mode] = z0
    M.ctx.set_launch_mode(2)
    assert np.array_equal(out[0], out[1])
    tol = 1e-13 if precision == "double" else 1e-5
    assert np.linalg.norm(out[2] - out[0]) / np.linalg.norm(out[0]) < tol


@pytest.mark.parametrize("precision", ["double", "single", "mixed"])
def test_subtree_tasks_agree(precision):
    """Splitting every subdomain's elimination tree into subtree tasks (several workgroups per
    subdomain, several launches) changes the schedule, and with it only the rounding of the sums
    below narrow chunks and the operator of the dense top (the skeleton's S^-1, whose rows depend
    on the split; stored in FP32 by the single and mixed precisions): 1 to 4 tasks per subdomain
    agree to rounding and with the oracle, on Cholesky (Poisson) and LU (convection) factors, and
    every split is bitwise repeatable."""
    from amp_amd import device
    tol = 1e-13 if precision == "double" else 2e-6
    for name, kw, N in (("poisson3d", dict(nx=16), 8), ("convdiff2d", dict(nx=24), 4)):
        st = _setup(name, kw, N, nev=8, overlap=2)
        print(f"case {name} {kw} N={N} {precision}", flush=True)
        r = np.cos(np.linspace(0.0, 7.0, st.n))
        out = []
        for t in (1, 2, 3, 4):
            M = device.HipTwoLevelSchwarz(st, precision, tasks_per_subdomain=t)
            z = M.apply(r)
            assert np.array_equal(z, M.apply(r)), t
            out.append(z)
        for z in out[1:]:
            assert np.linalg.norm(z - out[0]) / np.linalg.norm(out[0]) < tol
        zo = orc.OracleSchwarz(st).apply(r)
        err = np.linalg.norm(out[0] - zo) / np.linalg.norm(zo)
        assert err < (1e-12 if precision == "double" else 1e-5)


@pytest.mark.parametrize("mode", [2, 0])
@pytest.mark.parametrize("precision", ["double", "single"])
def test_wide_supernodes_and_partials(mode, precision):
    """Subdomains whose separators exceed a wavefront (row blocks, backward partial products)."""
    from amp_amd import device
    st = _setup("poisson3d", dict(nx=24), 4, nev=8, overlap=2)
    print(f"case poisson3d 24^3 N=4 mode {mode} {precision}", flush=True)
    M = device.HipTwoLevelSchwarz(st, precision)
    M.ctx.set_launch_mode(mode)
    r = np.cos(np.linspace(0.0, 7.0, st.n))
    z = M.apply(r)
    z64 = orc.OracleSchwarz(st, "double").apply(r)
    err = np.linalg.norm(z - z64) / np.linalg.norm(z64)
    if precision == "double":
        assert err < 1e-12
    else:
        z32o = orc.OracleSchwarz(st, "single").apply(r)
        assert err <= 3.0 * np.linalg.norm(z32o - z64) / np.linalg.norm(z64) + 1e-7
    rng = np.random.default_rng(11)
    for i in range(4):
        f = st.factors[i]
        b = rng.standard_normal(f.n)
        y = M.ctx.local_solve(i, b)
        yo = orc.OracleLocalFactor(f, "double").solve(b)
        assert np.linalg.norm(y - yo) / np.linalg.norm(yo) < (1e-12 if precision == "double" else 1e-5), i


def test_gmres_mixed_precision_matches_fp64_iterations():
    """FP32 factor entries in FP64 arithmetic with the reference's GMRES: the preconditioner is a
    linear operator to FP64 rounding, so the solution formed at the end of a cycle agrees with the
    Arnoldi estimate and the solve takes the FP64 iteration count (+1 at most), with a true residual
    <= 1.1e-8 and a solution within 1e-6 of the FP64 one."""
    from amp_amd import device
    st = _setup("poisson3d", dict(nx=20), 8, nev=12, overlap=2)
    x64, r64 = device.HipGmres(50, 1e-8, 500).solve(device.HipTwoLevelSchwarz(st, "double"), st.dec.b)
    xm, rm = device.HipGmres(50, 1e-8, 500).solve(device.HipTwoLevelSchwarz(st, "mixed"), st.dec.b)
    assert rm.converged and rm.true_rel_residual <= 1.1e-8
    assert rm.iterations <= r64.iterations + 1
    assert np.linalg.norm(xm - x64) / np.linalg.norm(x64) < 1e-6


def test_gmres_fp32_local_within_target():
    """FP32 local factors / FP64 Krylov: converges to 1e-8 (true residual) within +10% (+1) of the
    FP64 iteration count (test/test_factorization.cpp:296-306 asks <= 1.1e-8 true residual)."""
    from amp_amd import device
    st = _setup("poisson3d", dict(nx=20), 8, nev=12, overlap=2)
    x64, r64 = device.HipGmres(50, 1e-8, 500).solve(device.HipTwoLevelSchwarz(st, "double"), st.dec.b)
    x32, r32 = device.HipGmres(50, 1e-8, 500).solve(device.HipTwoLevelSchwarz(st, "single"), st.dec.b)
    assert r32.converged and r32.true_rel_residual <= 1.1e-8
    assert r32.iterations <= int(1.1 * r64.iterations) + 1
    assert np.linalg.norm(x32 - x64) / np.linalg.norm(x64) < 1e-6


@pytest.mark.parametrize("precision", ["double", "single"])
def test_many_rhs_coarse_products_on_matrix_cores(precision):
    """16 right-hand sides (coarse restriction, A00 products and prolongation on the FP64 matrix
    cores, one read of V and of the inverted factors for the group) agree with 16 single applies to
    rounding (FP64 local factors: 1e-12; FP32 ones: the float rounding of the local right-hand
    sides, 1e-6), and with the oracle. High-contrast channels, 8 subdomains, 12 coarse vectors each."""
    from amp_amd import device
    st = _setup("highcontrast3d", dict(nx=14, seed=5, contrast=1e6, channel_fraction=0.1), 8, nev=12, overlap=1)
    print(f"case highcontrast3d 14^3 N=8 16 rhs {precision}", flush=True)
    M = device.HipTwoLevelSchwarz(st, precision)
    rng = np.random.default_rng(7)
    R = rng.standard_normal((16, st.n))
    Z = M.ctx.apply(R)
    Zs = np.stack([M.ctx.apply(R[k]) for k in range(16)])
    diff = np.linalg.norm(Z - Zs, axis=1) / np.linalg.norm(Zs, axis=1)
    assert diff.max() < (1e-12 if precision == "double" else 1e-6), diff.max()
    if precision == "double":
        O = orc.OracleSchwarz(st)
        for k in (0, 7, 15):
            zo = O.apply(R[k])
            assert np.linalg.norm(Z[k] - zo) / np.linalg.norm(zo) < 1e-12


@pytest.mark.parametrize("nrhs", [2, 3])
@pytest.mark.parametrize("name,kw,N", [CASES[1], CASES[4]])
def test_fp32_pairs_of_rhs_equal_single_applies(name, kw, N, nrhs):
    """FP32 local solves of two right-hand sides in one workgroup (a float pair per row in LDS, one
    read of the factors for both; an odd count leaves the last pair half empty) give the results of
    one right-hand side at a time to 1e-6 (the same float operations per component, except that the
    dense blocks of wide supernodes and dense tops sum two rows per wavefront pass for a pair and four
    for one right-hand side), and every batch is bitwise repeatable."""
    from amp_amd import device
    st = _setup(name, kw, N, nev=8, overlap=2)
    print(f"case {name} {kw} N={N} fp32 pairs nrhs={nrhs}", flush=True)
    M = device.HipTwoLevelSchwarz(st, "single")
    rng = np.random.default_rng(13)
    R = rng.standard_normal((nrhs, st.n))
    Z = M.ctx.apply(R)
    Zs = np.stack([M.ctx.apply(R[k]) for k in range(nrhs)])
    diff = np.linalg.norm(Z - Zs, axis=1) / np.linalg.norm(Zs, axis=1)
    print(f"pairs bitwise equal to single applies: {np.array_equal(Z, Zs)}, max rel diff {diff.max():.2e}", flush=True)
    assert diff.max() < 1e-6, diff.max()
    assert np.array_equal(Z, M.ctx.apply(R))


@pytest.mark.parametrize("name,kw,N,nev,nrhs", [("convdiff2d", dict(nx=32), 8, 9, 16), ("convdiff2d", dict(nx=32), 8, 9, 5),
                                                ("poisson3d", dict(nx=12), 6, 13, 16)])
def test_many_rhs_coarse_odd_sizes(name, kw, N, nev, nrhs):
    """Coarse dimensions that are not multiples of 8 or 16 (zero-padded rows of the inverted A00
    factors, a last partial tile of rows), a partial group of right-hand sides, and a nonsymmetric
    A00 (LU with row pivoting: the first product gathers x in pivot order): the batched FP64 apply
    against single applies (1e-12) and the oracle (1e-12)."""
    from amp_amd import device
    st = _setup(name, kw, N, nev=nev)
    print(f"case {name} {kw} N={N} nev={nev} nrhs={nrhs}", flush=True)
    M = device.HipTwoLevelSchwarz(st, "double")
    assert M.ctx.stats().n0 % 8 != 0
    rng = np.random.default_rng(11)
    R = rng.standard_normal((nrhs, st.n))
    Z = M.ctx.apply(R)
    Zs = np.stack([M.ctx.apply(R[k]) for k in range(nrhs)])
    diff = np.linalg.norm(Z - Zs, axis=1) / np.linalg.norm(Zs, axis=1)
    assert diff.max() < 1e-12, diff.max()
    O = orc.OracleSchwarz(st)
    for k in (0, nrhs - 1):
        zo = O.apply(R[k])
        assert np.linalg.norm(Z[k] - zo) / np.linalg.norm(zo) < 1e-12


@pytest.mark.parametrize("nrhs", [2, 8, 13])
def test_spmv_groups_bitwise(nrhs):
    """Several right-hand sides share each row's reads of A (groups of eight per thread) and give the
    same bits as one product at a time and as the oracle."""
    from amp_amd import device
    st = _setup("convdiff2d", dict(nx=24), 4, nev=4)
    M = device.HipTwoLevelSchwarz(st, "double")
    X = np.stack([np.sin(np.linspace(-1.0, 2.0 + k, st.n)) for k in range(nrhs)])
    Y = M.ctx.spmv(X)
    O = orc.OracleSchwarz(st)
    for k in range(nrhs):
        assert np.array_equal(Y[k], M.ctx.spmv(X[k]))
        assert np.array_equal(Y[k], O.spmv(X[k]))


@pytest.mark.parametrize("name,kw,N,correction", [("poisson3d", dict(nx=16), 8, "deflated"),
                                                   ("convdiff2d", dict(nx=24), 4, "additive")])
def test_coarse_matrix_formed_on_the_device(name, kw, N, correction):
    """s2l_upload_coarse with E = NULL: A_00 = W^T A W formed on the device from the uploaded operator
    and basis (coarse_space.cpp:95-105) gives the apply of the host's A_00 (Cholesky for Poisson,
    LU with pivoting for convection-diffusion) to 1e-12, and the oracle's."""
    from amp_amd import device
    st = _setup(name, kw, N, nev=8, correction=correction)
    print(f"case device A_00 {name} {kw} N={N} {correction}", flush=True)
    r = np.cos(np.linspace(0.0, 7.0, st.n))
    zh = device.HipTwoLevelSchwarz(st, "double").apply(r)
    ctx = device.HipContext()
    ctx.upload_setup(st, "double", coarse_on_device=True)
    zd = ctx.apply(r)
    zo = orc.OracleSchwarz(st).apply(r)
    assert np.linalg.norm(zd - zh) / np.linalg.norm(zh) < 1e-12
    assert np.linalg.norm(zd - zo) / np.linalg.norm(zo) < 1e-12
