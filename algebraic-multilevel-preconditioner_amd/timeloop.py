"""A sequence of solves with a reuse policy (src/solver/time_loop.cpp:57-195, restated for one GPU):
every step generates the matrix of a moving field (the bubbles rise with the step) or reads it from
the step's Matrix Market file (a ``p_%04d.mtx`` pattern, domain_decomposer.cpp:80-97,143-150), every
corrector of a step solves it for a new right-hand side (time_loop.cpp:31-43), and the reuse
policy decides before each solve whether the preconditioner is reused, refactorized (new local
factors and A_00 with the same coarse basis, uploaded into the same device context) or rebuilt (a
new host setup and context)."""
import time

from . import device, hostsetup as hs
from .reuse import REBUILD, REFACTORIZE, REUSE, ReusePolicy


def run_time_loop(family: str = "bubbles3d", gen_kw: dict | None = None, nparts: int = 4, steps: int = 6,
                  correctors: int = 2, reuse: str = "auto", growth: float = 4.0, max_age: int = 0,
                  precision: str = "single", restart: int = 50, rtol: float = 1e-8, maxit: int = 1000,
                  seed: int = 12345, dt: float = 1.0, overlap: int = 1, pou: str = "boolean", nev: int = 8,
                  log=None, matrix: str | None = None, coordinates: bool = True, krylov: str = "gmres",
                  eig_workers: int = 0, device_refactor: bool = True, eig_backend: str = "host") -> list:
    """Runs steps x correctors solves; returns one record per solve (step, corrector, the level
    performed before it, iterations, convergence, true relative residual, update and solve times).
    With ``matrix`` (a file name or a step pattern) the matrices are read, not generated; a step
    whose size, pattern or symmetry differs from the first one raises ValueError on refactorize.
    ``coordinates=False`` partitions generated matrices by their graph only, as read ones are.
    ``krylov``: "gmres" (GMRES(restart)) or "bicgstab". ``eig_workers``: processes of the local
    eigenproblems of a rebuild (forked only while this process has no GPU context).
    ``device_refactor``: refactorize on the device (s2l_refactorize) when the factors allow it
    (Cholesky local factors and a symmetric A_00), else on the host. ``eig_backend``: "device" solves
    the local eigenproblems of a rebuild on the GPU (symmetric matrices), "host" with ARPACK."""
    gen_kw = dict(gen_kw or {})
    policy = ReusePolicy(reuse, growth, max_age)
    setup = M = None
    records = []
    for step in range(steps):
        if matrix is not None:
            A, sym = hs.read_matrix(hs.step_file_name(matrix, step))
            xyz, bs = None, 1
            sym = sym or hs.is_numerically_symmetric(A)
        else:
            A, _, xyz, sym, bs = hs.generate(family, **{**gen_kw, "time": step * dt})
            xyz = xyz if coordinates else None
        for c in range(correctors):
            b = hs.random_rhs_seq(A.shape[0], seed, step, c)
            level = REBUILD if M is None else policy.decide(c == 0)
            t0 = time.perf_counter()
            if level == REBUILD:
                setup = hs.build_setup(A, b, nparts, xyz=xyz, symmetric=sym, block_size=bs, overlap=overlap, pou=pou,
                                       nev=nev, eig_workers=eig_workers,
                                       eig_backend=eig_backend if sym else "host")
                M = device.HipTwoLevelSchwarz(setup, precision)
            elif level == REFACTORIZE:
                if device_refactor and sym and setup.factors[0].kind == 0:
                    M.refactorize_device(A)
                else:
                    setup = hs.refactorize(setup, A, b, overlap=overlap, pou=pou, symmetric=sym)
                    M.refresh(setup)
            t_update = time.perf_counter() - t0 if level != REUSE else 0.0
            t0 = time.perf_counter()
            solver = device.HipGmres(restart, rtol, maxit) if krylov == "gmres" else device.HipBiCgStab(rtol, maxit)
            x, res = solver.solve(M, b[setup.dec.new_to_old])
            t_solve = time.perf_counter() - t0
            # the solves start from 0: the reduction is the final relative residual. Every corrector is
            # the same kind of solve (one reference q): the reference gives the correctors their own
            # kinds only with a warm start or final tolerances (time_loop.cpp:154-158), neither of
            # which this loop has
            policy.record(level, res.iterations, res.true_rel_residual, bool(res.converged), t_update, t_solve, kind=0)
            rec = dict(step=step, corrector=c, level=level, iterations=res.iterations, converged=bool(res.converged),
                       true_rel_residual=res.true_rel_residual, t_update=t_update, t_solve=t_solve,
                       wasted=policy.wasted())
            records.append(rec)
            if log:
                log(rec)
    return records
