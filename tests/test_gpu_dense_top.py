"""Dense tops (s2l_set_dense_top): the skeleton above a subdomain's subtree tasks replaced by the dense
inverse of its Schur complement, S^-1 = U_SS^-1 L_SS^-1, its rows split over several workgroups.
The local solves are those of the sparse skeleton up to rounding (LocalSolver::solve,
src/linalg/local_solver.cpp:119-126): checked against the sparse skeleton, the oracle and the
oracle's local solves, on Cholesky and LU factors, in every precision and launch mode."""
import numpy as np
import pytest

from amp_amd import hostsetup as hs

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _setup(name, kw, N, **opts):
    A, _, xyz, sym, bs = hs.generate(name, **kw)
    return hs.build_setup(A, hs.random_rhs(A.shape[0], 12345), N, xyz=xyz, symmetric=sym, block_size=bs, **opts)


@pytest.mark.parametrize("precision", ["double", "single", "mixed"])
@pytest.mark.parametrize("name,kw,N", [("poisson3d", dict(nx=16), 8), ("convdiff2d", dict(nx=24), 4)])
def test_dense_top_equals_sparse_skeleton(name, kw, N, precision):
    from amp_amd import device
    st = _setup(name, kw, N, nev=8, overlap=2)
    r = np.cos(np.linspace(0.0, 7.0, st.n))
    zo = orc.OracleSchwarz(st).apply(r)
    for t in (2, 4):
        print(f"case dense top {name} {kw} N={N} {precision} tasks {t}", flush=True)
        Md = device.HipTwoLevelSchwarz(st, precision, tasks_per_subdomain=t)
        Ms = device.HipTwoLevelSchwarz(st, precision, tasks_per_subdomain=t, dense_top=0)
        zd, zs = Md.apply(r), Ms.apply(r)
        assert np.array_equal(zd, Md.apply(r))
        # FP32 storage (single, mixed) rounds S^-1 where the sparse skeleton rounds its factor blocks
        tol = 1e-13 if precision == "double" else 2e-6
        assert _rel(zd, zs) < tol, (t, _rel(zd, zs))
        if precision == "double":
            assert _rel(zd, zo) < 1e-12
        # one launch per subdomain (a workgroup runs the parts of a dense top one after another)
        Md.ctx.set_launch_mode(1)
        assert np.array_equal(Md.apply(r), zd)


def test_dense_top_local_solves_match_the_oracle():
    from amp_amd import device
    st = _setup("poisson3d", dict(nx=24), 4, nev=8, overlap=2)
    M = device.HipTwoLevelSchwarz(st, "double", tasks_per_subdomain=4)
    rng = np.random.default_rng(5)
    for i in range(4):
        f = st.factors[i]
        b = rng.standard_normal(f.n)
        assert _rel(M.ctx.local_solve(i, b), orc.OracleLocalFactor(f, "double").solve(b)) < 1e-12, i


@pytest.mark.parametrize("precision", ["double", "mixed", "single"])
@pytest.mark.parametrize("tasks", [1, 2])
def test_slots_partly_in_lds_match_the_oracle(precision, tasks):
    """Subdomains of about 7000 rows in one or two subtree tasks: a task's slots outnumber the LDS
    room (stats: lds_slot_cap < max_task_slots; in FP32 only with one task), so the first slot_cap of
    them stay in LDS and the rest in HBM. The local solves match the oracle's (LocalSolver::solve,
    local_solver.cpp:119-126), and a second apply gives the same bits."""
    from amp_amd import device
    st = _setup("poisson3d", dict(nx=24), 2, nev=4, overlap=1)
    print(f"case partial LDS slots {precision} tasks {tasks}", flush=True)
    M = device.HipTwoLevelSchwarz(st, precision, tasks_per_subdomain=tasks)
    s = M.ctx.stats()
    assert s.lds_slot_cap > 0 and s.lds_slots > 0
    if tasks == 1 or precision != "single":
        assert s.lds_slot_cap < s.max_task_slots, (s.lds_slot_cap, s.max_task_slots)
    rng = np.random.default_rng(11)
    # FP32 storage: the device rounds the inverted diagonal blocks, the oracle the factor entries
    tol = {"double": 1e-12, "mixed": 2e-6, "single": 1e-4}[precision]
    for i in range(2):
        f = st.factors[i]
        b = rng.standard_normal(f.n)
        assert _rel(M.ctx.local_solve(i, b), orc.OracleLocalFactor(f, precision).solve(b)) < tol, i
    r = np.cos(np.linspace(0.0, 7.0, st.n))
    z = M.apply(r)
    assert np.array_equal(z, M.apply(r))
    if precision == "double":
        assert _rel(z, orc.OracleSchwarz(st).apply(r)) < 1e-12


def test_dense_top_gmres_iterations_and_paired_rhs():
    """GMRES with dense tops takes the oracle's iterations (FP64); in FP32 two right-hand sides
    sharing a workgroup give the bits of each alone."""
    from amp_amd import device
    st = _setup("poisson3d", dict(nx=20), 8, nev=12, overlap=2)
    M = device.HipTwoLevelSchwarz(st, "double", tasks_per_subdomain=4)
    x, res = device.HipGmres(50, 1e-8, 500).solve(M, st.dec.b)
    O = orc.OracleSchwarz(st)
    xo, ro = orc.gmres(O.spmv, O.apply, st.dec.b, restart=50, rtol=1e-8, maxit=500)
    assert res.converged and res.iterations == ro.iterations
    assert _rel(x, xo) < 1e-8
    M32 = device.HipTwoLevelSchwarz(st, "single", tasks_per_subdomain=4)
    r1 = np.cos(np.linspace(0.0, 7.0, st.n))
    r2 = np.sin(np.linspace(-3.0, 2.0, st.n))
    z = M32.ctx.apply(np.stack([r1, r2]))
    assert np.array_equal(z[0], M32.apply(r1)) and np.array_equal(z[1], M32.apply(r2))


@pytest.mark.parametrize("precision", ["single", "mixed"])
def test_autotune_keeps_the_fastest_schedule(precision):
    """s2l_autotune_local times every candidate schedule and keeps the fastest: the kept schedule is
    the candidate with the least time, the context's later applies use it (equal to a context built
    directly with that schedule, bitwise), and a refresh of the factors keeps it."""
    from amp_amd import device
    st = _setup("poisson3d", dict(nx=20), 8, nev=8, overlap=2)
    M = device.HipTwoLevelSchwarz(st, precision)
    cands, best = M.autotune(nrhs=1, applies=5)
    print("autotune", precision, cands, best, flush=True)
    assert len(cands) >= 4 and 0 <= best < len(cands)
    assert all(c["us"] > 0 for c in cands)
    assert cands[best]["us"] == min(c["us"] for c in cands)
    assert any(c["dense_top"] == 0 for c in cands) and any(c["dense_top"] > 0 for c in cands)
    assert {c["threads"] for c in cands} <= {512, 1024}
    w = cands[best]
    r = np.cos(np.linspace(0.0, 7.0, st.n))
    Mw = device.HipTwoLevelSchwarz(st, precision, tasks_per_subdomain=w["tasks"], dense_top=w["dense_top"],
                                   local_threads=w["threads"])
    if w["launch_mode"] != 0:
        Mw.ctx.set_launch_mode(w["launch_mode"])
    z = M.apply(r)
    assert np.array_equal(z, Mw.apply(r))
    M.refresh(st)
    assert np.array_equal(M.apply(r), z) or w["launch_mode"] != 0


@pytest.mark.parametrize("precision", ["double", "single", "mixed"])
def test_512_thread_workgroups_agree(precision):
    """Two 512-thread workgroups per compute unit (s2l_set_local_threads) run the same schedule
    with half the threads: FP64 to 1e-13 of the 1024-thread workgroups and of the oracle, every
    precision bitwise repeatable and in the serial launch mode equal to the grid mode."""
    from amp_amd import device
    st = _setup("poisson3d", dict(nx=16), 8, nev=8, overlap=2)
    r = np.cos(np.linspace(0.0, 7.0, st.n))
    for t in (2, 8):
        print(f"case 512 threads {precision} tasks {t}", flush=True)
        M5 = device.HipTwoLevelSchwarz(st, precision, tasks_per_subdomain=t, local_threads=512)
        M1 = device.HipTwoLevelSchwarz(st, precision, tasks_per_subdomain=t)
        z5 = M5.apply(r)
        assert np.array_equal(z5, M5.apply(r))
        assert _rel(z5, M1.apply(r)) < (1e-13 if precision == "double" else 2e-6)
        if precision == "double":
            assert _rel(z5, orc.OracleSchwarz(st).apply(r)) < 1e-12
        M5.ctx.set_launch_mode(1)
        assert np.array_equal(M5.apply(r), z5)
