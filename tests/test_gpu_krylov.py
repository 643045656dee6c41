"""GPU parity of the device Krylov loops (BiCGStab, several right-hand sides) against the CPU oracle,
through the C-ABI."""
import numpy as np
import pytest

from oracle import oracle as orc

from cases import items, setup

pytestmark = pytest.mark.gpu

CASES = [("laplace2d", dict(nx=20), 4, "random"), ("convdiff2d", dict(nx=24), 4, "random"),
         ("bubbles3d", dict(nx=16, nbubbles=6, radius=0.12), 4, "generated")]


def _ctx(st, precision="double"):
    from amp_amd import device
    return device.HipTwoLevelSchwarz(st, precision)


def _st(name, kw, N, rhs):
    print(f"case {name} {kw} N={N}", flush=True)
    return setup(name, items(kw), N, items(dict(nev=8)), rhs=rhs)


@pytest.mark.parametrize("name,kw,N,rhs", CASES)
def test_bicgstab_fp64_matches_oracle(name, kw, N, rhs):
    st = _st(name, kw, N, rhs)
    M = _ctx(st)
    O = orc.OracleSchwarz(st)
    b = st.dec.b
    xo, ro = orc.bicgstab(O.spmv, O.apply, b, rtol=1e-8, maxit=1000)
    x, r = M.ctx.bicgstab(b, None, 1e-8, 1000)
    assert r.converged == ro.converged and r.breakdown == ro.breakdown
    assert r.iterations == ro.iterations and r.restarts == ro.restarts
    assert len(r.history) == r.iterations + 1 == len(ro.history)
    assert np.allclose(r.history, ro.history, rtol=1e-6, atol=1e-14)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
    t2 = np.linalg.norm(b - st.dec.A @ x) / np.linalg.norm(b)
    assert t2 <= 1.1e-8 and abs(t2 - r.true_rel_residual) < 1e-12


@pytest.mark.parametrize("rtol", [1e-1, 1e-2, 1e-4, 1e-8])
def test_bicgstab_openfoam_matches_oracle(rtol):
    st = _st(*CASES[2])
    M = _ctx(st)
    O = orc.OracleSchwarz(st)
    b = st.dec.b
    xo, ro = orc.bicgstab(O.spmv, O.apply, b, rtol=rtol, maxit=500, stop="openfoam")
    x, r = M.ctx.bicgstab(b, None, rtol, 500, "openfoam")
    assert r.converged and r.iterations == ro.iterations and r.restarts == ro.restarts == 0
    assert abs(r.initial_residual_of - 1.0) < 1e-12
    t1 = np.abs(b - st.dec.A @ x).sum() / np.abs(b).sum()
    assert t1 <= rtol * (1 + 1e-10) and abs(r.final_residual_of - t1) < 1e-12 * max(1.0, t1 / rtol)


def test_bicgstab_openfoam_normalisation_nonzero_guess():
    st = _st(*CASES[2])
    M = _ctx(st)
    O = orc.OracleSchwarz(st)
    b = st.dec.b
    x0 = 0.7 + 0.3 * np.sin(0.37 * np.arange(st.n))
    xo, ro = orc.bicgstab(O.spmv, O.apply, b, x0, rtol=1e-3, maxit=500, stop="openfoam")
    x, r = M.ctx.bicgstab(b, x0, 1e-3, 500, "openfoam")
    assert r.converged and r.iterations == ro.iterations
    assert abs(r.initial_residual_of - ro.initial_residual_of) < 1e-12 * ro.initial_residual_of
    assert abs(r.final_residual_of - ro.final_residual_of) < 1e-9 * ro.initial_residual_of


def test_bicgstab_budget_and_zero_rhs():
    st = _st(*CASES[0])
    M = _ctx(st)
    O = orc.OracleSchwarz(st)
    b = st.dec.b
    xo, ro = orc.bicgstab(O.spmv, O.apply, b, rtol=1e-12, maxit=3)
    x, r = M.ctx.bicgstab(b, None, 1e-12, 3)
    assert not r.converged and r.iterations == 3 == ro.iterations
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-10
    tb = np.linalg.norm(b - st.dec.A @ x) / np.linalg.norm(b)
    assert abs(tb - r.true_rel_residual) < 1e-12 and tb < 1.0
    x, r = M.ctx.bicgstab(np.zeros(st.n), np.ones(st.n), 1e-8, 100)
    assert r.converged and r.iterations == 0 and len(r.history) == 1 and not x.any()


def test_bicgstab_single_precision_factors():
    st = _st(*CASES[1])
    M32, M64 = _ctx(st, "single"), _ctx(st, "double")
    b = st.dec.b
    x32, r32 = M32.ctx.bicgstab(b, None, 1e-8, 1000)
    x64, r64 = M64.ctx.bicgstab(b, None, 1e-8, 1000)
    assert r32.converged and r32.true_rel_residual <= 1.1e-8
    assert r32.iterations <= 1.1 * r64.iterations + 2
    assert np.linalg.norm(x32 - x64) / np.linalg.norm(x64) < 1e-6


@pytest.mark.parametrize("method", ["gmres", "bicgstab"])
def test_several_rhs_equal_single_solves(method):
    """Each right-hand side of a batched solve gets the same result as when it is solved alone."""
    from amp_amd import hostsetup as hs
    st = _st(*CASES[1])
    M = _ctx(st)
    B = np.stack([hs.random_rhs(st.n, 100 + k) for k in range(3)])
    B[1] *= 1e-3
    X, R = M.ctx.krylov(B, None, method, 30, 1e-8, 1000)
    for k in range(3):
        x, r = M.ctx.krylov(B[k], None, method, 30, 1e-8, 1000)
        assert R[k].iterations == r.iterations and R[k].converged == r.converged
        assert np.array_equal(X[k], x)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_flexible_gmres_matches_oracle(precision):
    st = _st(*CASES[1])
    M = _ctx(st, precision)
    O = orc.OracleSchwarz(st, precision)
    b = st.dec.b
    xo, ro = orc.gmres(O.spmv, O.apply, b, restart=30, rtol=1e-8, maxit=500, flexible=True)
    x, r = M.ctx.gmres(b, None, 30, 1e-8, 500, flexible=True)
    assert r.converged and r.true_rel_residual <= 1.1e-8
    if precision == "double":
        assert r.iterations == ro.iterations
        assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
    else:
        assert abs(r.iterations - ro.iterations) <= 1
        assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-6


@pytest.mark.parametrize("precision", ["double", "single"])
def test_graph_replays_equal_eager_launches(precision):
    """GMRES whose apply + SpMV per iteration is replayed from captured HIP graphs gives the same
    bits as launching every kernel by itself: three rounds of solves (each replays the graphs of
    all its iterations, with the Krylov kernels and host polls in between), standard and flexible,
    one and three right-hand sides; a re-upload of the factors drops the graphs."""
    from amp_amd import hostsetup as hs
    st = _st(*CASES[1])
    M = _ctx(st, precision)
    b = st.dec.b
    B = np.stack([hs.random_rhs(st.n, 300 + k) for k in range(3)])
    cases = [(b, False), (b, True), (B, False)]
    M.ctx.set_graphs(False)
    ref = [M.ctx.krylov(rhs, None, "gmres", 30, 1e-8, 500, flexible=flex) for rhs, flex in cases]
    M.ctx.set_graphs(True)
    for rnd in range(3):
        if rnd == 2:
            M.ctx.upload_setup(st, precision)
        for (rhs, flex), (xe, re_) in zip(cases, ref):
            x, r = M.ctx.krylov(rhs, None, "gmres", 30, 1e-8, 500, flexible=flex)
            assert np.array_equal(x, xe), (rnd, flex)
            its = [q.iterations for q in (r if isinstance(r, list) else [r])]
            its_e = [q.iterations for q in (re_ if isinstance(re_, list) else [re_])]
            assert its == its_e


@pytest.mark.parametrize("precision", ["double", "single"])
def test_many_rhs_gmres_matches_single_solves(precision):
    """GMRES with 8 right-hand sides (coarse products on the matrix cores; in FP32 the local solves
    of two right-hand sides share a workgroup) converges every one of them in the iterations it
    takes alone (to one) and to the same solution within the tolerance; a right-hand side that
    converges early is masked while the others continue."""
    from amp_amd import hostsetup as hs
    st = setup("highcontrast3d", items(dict(nx=14, seed=5, contrast=1e6, channel_fraction=0.1)), 8, items(dict(nev=12)))
    print(f"case highcontrast3d 14^3 N=8 8 rhs {precision}", flush=True)
    M = _ctx(st, precision)
    B = np.stack([hs.random_rhs(st.n, 500 + k) for k in range(8)])
    B[3] *= 1e-6
    X, R = M.ctx.krylov(B, None, "gmres", 30, 1e-8, 500)
    for k in range(8):
        x, r = M.ctx.krylov(B[k], None, "gmres", 30, 1e-8, 500)
        assert R[k].converged and r.converged
        assert abs(R[k].iterations - r.iterations) <= 1
        assert np.linalg.norm(X[k] - x) / np.linalg.norm(x) < (1e-7 if precision == "double" else 1e-6)
        t = np.linalg.norm(B[k] - st.dec.A @ X[k]) / np.linalg.norm(B[k])
        assert t <= 1.1e-8


def test_refactorize_equals_fresh_setup_and_oracle():
    """Refactorize (new values of A, same pattern and coarse basis: test/test_reuse.cpp:113-140):
    uploading the refactorized host setup into the existing context gives bitwise the apply of a
    new context built from it, and matches the oracle on the new values (1e-12)."""
    from amp_amd import device, hostsetup as hs
    kw = dict(nx=16, nbubbles=6, radius=0.12)
    A0, _, xyz, sym, bs = hs.generate("bubbles3d", **kw, time=0.0)
    st0 = hs.build_setup(A0, hs.random_rhs(A0.shape[0], 1), 4, xyz=xyz, symmetric=sym, block_size=bs, nev=6)
    M = device.HipTwoLevelSchwarz(st0, "double")
    A1, _, _, _, _ = hs.generate("bubbles3d", **kw, time=3.0)
    assert (A1 != A0).nnz > 0
    st1 = hs.refactorize(st0, A1, hs.random_rhs(A1.shape[0], 2))
    print("case bubbles3d 16^3 N=4 refactorize", flush=True)
    assert M.update("refactorize", st1) == "refactorize"
    r = np.cos(np.linspace(0.0, 7.0, st1.n))
    z = M.apply(r)
    assert np.array_equal(z, device.HipTwoLevelSchwarz(st1, "double").apply(r))
    zo = orc.OracleSchwarz(st1).apply(r)
    assert np.linalg.norm(z - zo) / np.linalg.norm(zo) < 1e-12
    assert M.update("reuse") == "reuse"
    with pytest.raises(ValueError):
        M.update("rebuild")


@pytest.mark.parametrize("reuse", ["auto", "refactor", "off"])
def test_time_loop_with_reuse_policy(reuse):
    """A sequence of 4 steps x 2 correctors on rising bubbles (src/solver/time_loop.cpp:57-195): the
    first solve rebuilds; the second corrector of a step reuses (the matrix did not change) except
    with the Off policy; the first corrector of a later step refactorizes or rebuilds as the policy
    says; every solve converges with a true residual <= 1.1e-8."""
    from amp_amd.timeloop import run_time_loop
    print(f"case time loop bubbles3d 16^3 {reuse}", flush=True)
    recs = run_time_loop("bubbles3d", dict(nx=16, nbubbles=6, radius=0.12), nparts=4, steps=4, correctors=2,
                         reuse=reuse, precision="single", restart=30, nev=6)
    assert len(recs) == 8 and recs[0]["level"] == "rebuild"
    for rec in recs:
        assert rec["converged"] and rec["true_rel_residual"] <= 1.1e-8, rec
    for rec in recs[1:]:
        if reuse == "off":
            assert rec["level"] == "rebuild"
        elif rec["corrector"] == 1:
            assert rec["level"] == "reuse"
        elif reuse == "refactor":
            assert rec["level"] == "refactorize"
        else:
            assert rec["level"] in ("refactorize", "rebuild")



def test_time_loop_from_matrix_files_equals_generator(tmp_path):
    """The same sequence read from Matrix Market files through a %02d step pattern gives the same
    levels and iterations as the generator (test/test_reuse.cpp:447-476); both partition by the
    graph alone, and the host setup is reproducible, so the residuals are the same too."""
    from amp_amd import hostsetup as hs
    from amp_amd.timeloop import run_time_loop
    print("case time loop from files bubbles3d 16^3 refactor", flush=True)
    kw = dict(nx=16, nbubbles=6, radius=0.12)
    for s in range(3):
        hs.write_matrix(tmp_path / f"seq_{s:02d}.mtx", hs.generate("bubbles3d", **kw, time=float(s))[0])
    common = dict(nparts=4, steps=3, correctors=2, reuse="refactor", precision="single", restart=30, nev=6,
                  coordinates=False)
    ra = run_time_loop("bubbles3d", kw, **common)
    rf = run_time_loop(matrix=str(tmp_path / "seq_%02d.mtx"), **common)
    assert [r["level"] for r in rf] == [r["level"] for r in ra]
    assert [r["iterations"] for r in rf] == [r["iterations"] for r in ra]
    for a, f in zip(ra, rf):
        assert f["converged"] and f["true_rel_residual"] == a["true_rel_residual"]


def test_serving_mix_small():
    """A small serving mix end to end: every request converges, the families' measured device
    footprints are positive, and a budget below their sum forces evictions (and rebuilds) by the
    least-recently-used rule without breaking any request."""
    from amp_amd import serving as sv
    print("case serving mix 3 families n in [2e3, 2e4], 24 requests", flush=True)
    fams, reqs = sv.draw_mix(24, 3, (2e3, 2e4), 4, seed=11)
    build = lambda k: sv.build_family(fams[k], precision="single", rows_per_subdomain=1500, nev=6)
    big = sv.ResidentSet(1 << 40, build)
    recs = sv.serve(fams, reqs, big, restart=30)
    assert all(r["converged"] and r["true_rel_residual"] <= 1.1e-8 for r in recs)
    assert big.builds == len({q.family for q in reqs}) and big.evictions == 0
    fp = {k: big.footprint(M) for k, M in big.live.items()}
    assert all(v > 0 for v in fp.values())
    small = sv.ResidentSet(max(fp.values()) + min(fp.values()) // 2, build)
    recs2 = sv.serve(fams, reqs, small, restart=30)
    assert small.evictions > 0 and small.builds > big.builds
    assert [r["iterations"] for r in recs2] == [r["iterations"] for r in recs]
    t = sv.timeline(recs2, 0.7)
    assert t["p99_ms"] >= t["p50_ms"] > 0 and t["solves_per_s"] > 0


def test_serving_results_same_alone_and_in_the_live_mix():
    """Continuous batching on the wall clock: three families, each stepped by its own thread with a
    slot solver, requests admitted between restart cycles while others run (arrivals compressed so
    that they overlap). Every request's right-hand sides have the iterations and the solution bits
    they have when the request is served alone, and each converges to a true residual <= 1.1e-8."""
    from amp_amd import serving as sv
    print("case live serving mix, 3 families, 24 requests", flush=True)
    fams, reqs = sv.draw_mix(24, 3, (2e3, 2e4), 4, seed=5)
    rs = sv.ResidentSet(1 << 40, lambda k: sv.build_family(fams[k], precision="single", rows_per_subdomain=1500,
                                                          nev=6))
    alone = sv.serve_alone(fams, reqs, rs, slots=8, restart=30)
    mean = float(np.mean([r["service"] for r in alone]))
    live = sv.serve_live(fams, reqs, rs, scale=0.3 * mean, slots=8, restart=30)
    same = sv.same_as_alone(alone, live)
    assert same["differing"] == 0, same
    for r in live:
        assert r["converged"] and r["true_rel_residual"] <= 1.1e-8
    assert len({r["family"] for r in live}) == 3
    overlap = sum(1 for a in live for b in live if a["id"] < b["id"] and a["admitted"] < b["finish"]
                  and b["admitted"] < a["finish"])
    print(f"live mix: {overlap} overlapping request pairs, {sv.live_summary(live)}", flush=True)
    assert overlap > 0


@pytest.mark.parametrize("rtol", [1e-1, 1e-2, 1e-4, 1e-8])
def test_gmres_openfoam_matches_oracle(rtol):
    """GMRES with the OpenFOAM stopping norm (gmres.cpp:106-122: the 1-norm of the GMRES residual
    g_{j+1} V Q^T e_{j+1} against rtol sum|r0|): the oracle's iterations and solution, the
    normalised residuals, and the true 1-norm within the threshold (10 % slack of gmres.cpp:149)."""
    st = _st(*CASES[2])
    M = _ctx(st)
    O = orc.OracleSchwarz(st)
    b = st.dec.b
    xo, ro = orc.gmres(O.spmv, O.apply, b, restart=30, rtol=rtol, maxit=500, stop="openfoam")
    x, r = M.ctx.gmres(b, None, 30, rtol, 500, stop="openfoam")
    assert r.converged == ro.converged and r.iterations == ro.iterations
    assert np.allclose(r.history, ro.history, rtol=1e-6, atol=1e-14)
    assert abs(r.initial_residual_of - 1.0) < 1e-12
    t1 = np.abs(b - st.dec.A @ x).sum() / np.abs(b).sum()
    # sum|b - A x| at a residual of 1e-8 carries the cancellation of b - A x: 1e-6 relative
    assert t1 <= 1.1 * rtol and abs(r.final_residual_of - t1) < 1e-6 * t1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8


def test_gmres_openfoam_nonzero_guess_and_atol():
    """Nonzero initial guess (normalisation with the mean of x0 and A 1), and an absolute tolerance
    on the normalised residual (OpenFOAM `tolerance`): the oracle's iterations and F."""
    st = _st(*CASES[2])
    M = _ctx(st)
    O = orc.OracleSchwarz(st)
    b = st.dec.b
    x0 = 0.7 + 0.3 * np.sin(0.37 * np.arange(st.n))
    xo, ro = orc.gmres(O.spmv, O.apply, b, x0, restart=30, rtol=1e-3, maxit=500, stop="openfoam")
    x, r = M.ctx.gmres(b, x0, 30, 1e-3, 500, stop="openfoam")
    assert r.converged and r.iterations == ro.iterations
    assert abs(r.initial_residual_of - ro.initial_residual_of) < 1e-12 * ro.initial_residual_of
    assert abs(r.final_residual_of - ro.final_residual_of) < 1e-9 * ro.initial_residual_of
    xa, ra = orc.gmres(O.spmv, O.apply, b, x0, restart=30, rtol=0.0, atol=1e-4 * ro.initial_residual_of, maxit=500,
                       stop="openfoam")
    x, r = M.ctx.gmres(b, x0, 30, 0.0, 500, stop="openfoam", atol=1e-4 * ro.initial_residual_of)
    assert r.converged and r.iterations == ra.iterations and r.final_residual_of <= 1.1e-4 * ro.initial_residual_of


@pytest.mark.parametrize("precision", ["double", "single"])
def test_slot_solver_same_bits_alone_and_in_a_mix(precision):
    """Continuous batching (s2l_solver): a right-hand side solved alone in an otherwise empty solver
    and the same right-hand side admitted into a solver whose other slots are busy with solves that
    started earlier or later give the same bits, iterations and history; in FP64 the oracle's
    iterations and solution."""
    from amp_amd import hostsetup as hs
    st = _st(*CASES[1])
    M = _ctx(st, precision)
    B = [hs.random_rhs(st.n, 700 + k) for k in range(6)]
    B[2] = B[2] * 1e-4
    alone = []
    sv = M.ctx.solver(4, restart=30, rtol=1e-8, maxit=500)
    for bk in B:
        sv.admit(0, bk)
        while 0 not in sv.step():
            pass
        alone.append(sv.take(0))
    # the mix: slots filled as they free up, new requests admitted between cycles
    queue = list(range(len(B)))
    owner = {}
    mixed = {}
    steps = 0
    while queue or owner:
        for slot in range(4):
            if slot not in owner and queue and (steps % 2 == 0 or slot > 0):
                k = queue.pop(0)
                sv.admit(slot, B[k])
                owner[slot] = k
        for slot in sv.step():
            mixed[owner.pop(int(slot))] = sv.take(int(slot))
        steps += 1
    O = orc.OracleSchwarz(st)
    for k in range(len(B)):
        (xa, ra), (xm, rm) = alone[k], mixed[k]
        assert np.array_equal(xa, xm) and ra.iterations == rm.iterations and ra.history == rm.history, k
        assert rm.converged and rm.true_rel_residual <= 1.1e-8
        if precision == "double":
            xo, ro = orc.gmres(O.spmv, O.apply, B[k], restart=30, rtol=1e-8, maxit=500)
            assert rm.iterations == ro.iterations and np.linalg.norm(xm - xo) / np.linalg.norm(xo) < 1e-8
    sv.close()


def test_phase_timers_count_the_reference_timers():
    """Per-phase event timing under the reference's timer names: one solve.preconditioner interval
    per application (iterations + one per cycle), one solve.spmv and one solve.orthogonalization per
    iteration, one solve.stopping_test per iteration with the OpenFOAM norm; the coarse parts nest
    inside the preconditioner; the results are the same bits with the timers on."""
    st = _st(*CASES[1])
    M = _ctx(st)
    b = st.dec.b
    x0, r0 = M.ctx.gmres(b, None, 30, 1e-8, 500, stop="openfoam")
    M.ctx.phase_timing(True)
    x1, r1 = M.ctx.gmres(b, None, 30, 1e-8, 500, stop="openfoam")
    ph = M.ctx.phase_timing_read()
    M.ctx.phase_timing(False)
    assert np.array_equal(x0, x1) and r0.iterations == r1.iterations
    its = r1.iterations
    cycles = (its + 29) // 30
    # the device loop polls the convergence flags every 4 iterations: up to 3 launches of a cycle
    # after convergence run masked (they return at once) and are timed as well
    loop = ph["solve.spmv"][1]
    assert its <= loop <= its + 3 * cycles
    assert ph["solve.orthogonalization"][1] == ph["solve.stopping_test"][1] == loop
    assert ph["solve.preconditioner"][1] == loop + cycles
    assert ph["solve.preconditioner.coarse"][1] == ph["solve.preconditioner.local"][1] == loop + cycles
    inner = sum(ph[k][0] for k in ("solve.preconditioner.coarse", "solve.preconditioner.spmv",
                                   "solve.preconditioner.local"))
    assert 0 < inner <= ph["solve.preconditioner"][0] * 1.05
    parts = sum(ph[k][0] for k in ("solve.preconditioner.coarse.restrict", "solve.preconditioner.coarse.solve",
                                   "solve.preconditioner.coarse.prolong"))
    assert 0 < parts <= ph["solve.preconditioner.coarse"][0] * 1.05


@pytest.mark.parametrize("method", ["gmres", "bicgstab"])
def test_non_finite_right_hand_side_stops_as_a_breakdown(method):
    """A NaN in the right-hand side (and an Inf in another) ends the solve unconverged with the
    breakdown flag after at most one iteration, instead of restarting until maxit (BiCGStab:
    bicgstab.cpp:97-100; GMRES: no reference counterpart); the finite right-hand side solved in the
    same batch is untouched (the same bits as alone)."""
    st = _st(*CASES[1])
    M = _ctx(st)
    b = st.dec.b.copy()
    bn = b.copy()
    bn[len(b) // 3] = np.nan
    bi = b.copy()
    bi[len(b) // 2] = np.inf
    X, R = M.ctx.krylov(np.stack([bn, b, bi]), None, method, 30, 1e-8, 1000)
    for k in (0, 2):
        assert not R[k].converged and R[k].breakdown and R[k].iterations <= 1, (k, R[k])
    x, r = M.ctx.krylov(b, None, method, 30, 1e-8, 1000)
    assert R[1].converged and R[1].iterations == r.iterations and np.array_equal(X[1], x)


def test_non_finite_factor_is_refused():
    """A local factor with a NaN entry is refused at upload (S2L_ENUMERIC)."""
    from amp_amd import device
    st = _st(*CASES[1])
    f = st.factors[0]
    bad = type(f)(**{**f.__dict__, "l_val": f.l_val.copy()})
    bad.l_val[len(bad.l_val) // 2] = np.nan
    ctx = device.HipContext()
    with pytest.raises(RuntimeError, match="non-finite"):
        ctx.upload_setup(type(st)(st.dec, st.subs, [bad] + list(st.factors[1:]), st.coarse), "double")
