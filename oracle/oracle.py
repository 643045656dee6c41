"""CPU oracle of the two-level Schwarz apply path (FP64 restatement of the reference algorithm).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by bench.py's CPU-baseline
leg, as the checker of the HIP path. It is never the thing measured on the GPU and never shipped
as a fallback.

Each function names the reference lines it restates. Kernels with sequential sums (local
triangular solves, SpMV) are in oracle.c; the composition and the Krylov loops are here,
line by line in the reference's order of operations:

* ``OracleSchwarz.apply``    TwoLevelSchwarz::apply  src/preconditioner/two_level_schwarz.cpp:134-163
* ``OracleSchwarz.coarse``   CoarseSpace::apply      src/preconditioner/coarse_space.cpp:208-246
                             RootCoarseSolver::solve src/preconditioner/coarse_solver.cpp:166-179
* ``OracleSchwarz.one_level`` OneLevelSchwarz::apply src/preconditioner/one_level_schwarz.cpp:80-99
* ``OracleSchwarz.spmv``     DistributedMatrix::apply include/linalg/distributed_matrix.hpp:33-38
* ``reverse_add``            HaloExchange::reverseAdd src/linalg/halo_exchange.cpp:86-107
* ``gmres``                  Gmres::solve            src/solver/gmres.cpp:18-161
* ``bicgstab``               BiCgStab::solve         src/solver/bicgstab.cpp:15-154
* ``norm1/openfoam_norm_factor`` src/solver/krylov.cpp:40-57

Pins (tests/test_oracle_*.py): dense sequential reference of the apply (rel. err < 1e-9, as
test/test_preconditioner.cpp:104-184), backward errors of the local solves
(test/test_factorization.cpp:95-98,141), the GMRES minimal-residual property against an
explicit Krylov least-squares problem, the global SpMV (test/test_subdomain.cpp:77-83), the
BiCGStab accounting (test/test_bicgstab.cpp:82-109), planted-mistake checks.
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def build(force: bool = False) -> str:
    """Compiles liboracle.so next to this file (gcc, no GPU)."""
    so = os.path.join(HERE, "liboracle.so")
    src = os.path.join(HERE, "oracle.c")
    if force or not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", so, src, "-lm"])
    return so


def lib():
    global _lib
    if _lib is None:
        so = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        P = ctypes.POINTER
        i32p, i64p, dp, fp = P(ctypes.c_int32), P(ctypes.c_int64), P(ctypes.c_double), P(ctypes.c_float)
        L.orc_local_solve_d.argtypes = [ctypes.c_int32, i32p, i64p, i32p, dp, i64p, i32p, dp, dp, dp, dp]
        L.orc_local_solve_f.argtypes = [ctypes.c_int32, i32p, i64p, i32p, fp, i64p, i32p, fp, dp, dp, fp]
        L.orc_spmv_owned.argtypes = [ctypes.c_int32, i64p, i32p, dp, i64p, dp, dp]
        L.orc_dense_chol.argtypes = [ctypes.c_int32, dp]
        L.orc_dense_chol.restype = ctypes.c_int
        L.orc_dense_chol_solve.argtypes = [ctypes.c_int32, dp, dp]
        L.orc_dense_lu.argtypes = [ctypes.c_int32, dp, i32p]
        L.orc_dense_lu.restype = ctypes.c_int
        L.orc_dense_lu_solve.argtypes = [ctypes.c_int32, dp, i32p, dp]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class OracleLocalFactor:
    """LocalSolver::solve restated on one factor, FP64 or FP32 values (local_solver.cpp:77-126).

    precision "mixed" (no reference counterpart) keeps the factor entries rounded to FP32 and
    solves in FP64: the error it leaves is the rounding of the factors alone."""

    def __init__(self, f, precision: str = "double"):
        self.n = f.n
        self.single = precision == "single"
        vt = np.float32 if self.single else np.float64

        def values(v):
            if v is None:
                return None
            return np.asarray(v, np.float32).astype(np.float64) if precision == "mixed" else v

        self.perm = np.ascontiguousarray(f.perm, np.int32)
        self.lcp = np.ascontiguousarray(f.l_colptr, np.int64)
        self.lr = np.ascontiguousarray(f.l_row, np.int32)
        self.lv = np.ascontiguousarray(values(f.l_val), vt)
        self.ucp = np.ascontiguousarray(f.ut_colptr, np.int64)
        self.ur = np.ascontiguousarray(f.ut_row, np.int32)
        self.uv = np.ascontiguousarray(values(f.ut_val), vt)

    def solve(self, b: np.ndarray) -> np.ndarray:
        L = lib()
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty(self.n)
        if self.single:
            w = np.empty(self.n, np.float32)
            L.orc_local_solve_f(self.n, _p(self.perm, ctypes.c_int32), _p(self.lcp, ctypes.c_int64),
                                _p(self.lr, ctypes.c_int32), _p(self.lv, ctypes.c_float), _p(self.ucp, ctypes.c_int64),
                                _p(self.ur, ctypes.c_int32), _p(self.uv, ctypes.c_float), _p(b, ctypes.c_double),
                                _p(x, ctypes.c_double), _p(w, ctypes.c_float))
        else:
            w = np.empty(self.n)
            L.orc_local_solve_d(self.n, _p(self.perm, ctypes.c_int32), _p(self.lcp, ctypes.c_int64),
                                _p(self.lr, ctypes.c_int32), _p(self.lv, ctypes.c_double), _p(self.ucp, ctypes.c_int64),
                                _p(self.ur, ctypes.c_int32), _p(self.uv, ctypes.c_double), _p(b, ctypes.c_double),
                                _p(x, ctypes.c_double), _p(w, ctypes.c_double))
        return x


class OracleSchwarz:
    """Two-level Schwarz preconditioner restated on the CPU from a host setup.

    ``setup`` is a ``hostsetup.SchwarzSetup``. ``threads`` > 1 solves the subdomains in parallel
    (one subdomain per thread, like one MPI rank per subdomain in the reference)."""

    def __init__(self, setup, precision: str = "double", one_level: str | None = None,
                 correction: str | None = None, threads: int = 1):
        self.s = setup
        self.n = setup.n
        self.one_level_kind = one_level or setup.one_level
        self.correction = correction or setup.correction
        self.subs = setup.subs
        self.gids = [np.ascontiguousarray(sd.global_ids(), np.int64) for sd in self.subs]
        self.factors = [OracleLocalFactor(f, precision) for f in setup.factors]
        self.threads = max(1, threads)
        self.pool = cf.ThreadPoolExecutor(self.threads) if self.threads > 1 else None
        cs = setup.coarse
        self.V = cs.V
        self.offsets = cs.offsets
        self.n0 = cs.n0
        # Root coarse solver: factorization of A_00 (Cholesky for symmetric A, else LU).
        self.coarse_sym = bool(setup.dec.symmetric)
        if self.n0 > 0:
            E = np.ascontiguousarray(cs.E, np.float64).copy()
            if self.coarse_sym:
                if lib().orc_dense_chol(self.n0, _p(E, ctypes.c_double)) != 0:
                    self.coarse_sym = False
                    E = np.ascontiguousarray(cs.E, np.float64).copy()
            if not self.coarse_sym:
                self.piv = np.empty(self.n0, np.int32)
                if lib().orc_dense_lu(self.n0, _p(E, ctypes.c_double), _p(self.piv, ctypes.c_int32)) != 0:
                    raise RuntimeError("coarse matrix factorization failed")
            self.Ef = E
        # Owned-row matrices for the SpMV, in local column numbering.
        self.aown = []
        for sd in self.subs:
            Ao = sd.Aowned.tocsr()
            Ao.sort_indices()
            self.aown.append((np.ascontiguousarray(Ao.indptr, np.int64), np.ascontiguousarray(Ao.indices, np.int32),
                              np.ascontiguousarray(Ao.data, np.float64)))
        # Receivers of reverseAdd: for every subdomain, (neighbour q, ghost positions in q, owned positions here),
        # neighbours in increasing index (halo_exchange.cpp:98-106).
        self._rev = None

    # ------------------------------------------------------------------ halo
    def _reverse_plan(self):
        if self._rev is None:
            own = self.s.dec.ownership
            plan = [[] for _ in self.subs]
            for q, sd in enumerate(self.subs):
                owner = np.searchsorted(own, sd.ghosts, side="right") - 1
                for p in np.unique(owner):
                    sel = np.nonzero(owner == p)[0]
                    plan[p].append((q, sd.n_owned + sel, (sd.ghosts[sel] - own[p]).astype(np.int64)))
            for p in range(len(plan)):
                plan[p].sort(key=lambda t: t[0])
            self._rev = plan
        return self._rev

    def reverse_add(self, yloc: list) -> list:
        """owner entries += ghost copies of the neighbours, neighbours in increasing order."""
        plan = self._reverse_plan()
        out = []
        for p, sd in enumerate(self.subs):
            y = yloc[p].copy()
            for q, gpos, opos in plan[p]:
                np.add.at(y, opos, yloc[q][gpos])
            out.append(y)
        return out

    # ------------------------------------------------------------------ operator
    def spmv(self, x: np.ndarray) -> np.ndarray:
        """y = A x, each owned row summed in the local column order (distributed_matrix.hpp:33-38)."""
        L = lib()
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n)
        for sd, gid, (rp, col, val) in zip(self.subs, self.gids, self.aown):
            seg = y[sd.begin:sd.begin + sd.n_owned]
            L.orc_spmv_owned(sd.n_owned, _p(rp, ctypes.c_int64), _p(col, ctypes.c_int32), _p(val, ctypes.c_double),
                             _p(gid, ctypes.c_int64), _p(x, ctypes.c_double), _p(seg, ctypes.c_double))
        return y

    # ------------------------------------------------------------------ coarse
    def coarse(self, r: np.ndarray) -> np.ndarray:
        """x0 = R_0^T A_00^{-1} R_0 r (coarse_space.cpp:208-246, coarse_solver.cpp:166-179)."""
        c = np.empty(self.n0)
        for i, (sd, gid, V) in enumerate(zip(self.subs, self.gids, self.V)):
            o0, o1 = self.offsets[i], self.offsets[i + 1]
            if sd.boolean:
                c[o0:o1] = V[:sd.n_owned].T @ r[sd.begin:sd.begin + sd.n_owned]
            else:
                c[o0:o1] = V.T @ r[gid]
        y = c.copy()
        if self.coarse_sym:
            lib().orc_dense_chol_solve(self.n0, _p(self.Ef, ctypes.c_double), _p(y, ctypes.c_double))
        else:
            lib().orc_dense_lu_solve(self.n0, _p(self.Ef, ctypes.c_double), _p(self.piv, ctypes.c_int32),
                                     _p(y, ctypes.c_double))
        x0 = np.empty(self.n)
        if self.subs[0].boolean:
            for i, (sd, V) in enumerate(zip(self.subs, self.V)):
                x0[sd.begin:sd.begin + sd.n_owned] = V[:sd.n_owned] @ y[self.offsets[i]:self.offsets[i + 1]]
        else:
            loc = [V @ y[self.offsets[i]:self.offsets[i + 1]] for i, V in enumerate(self.V)]
            loc = self.reverse_add(loc)
            for sd, l in zip(self.subs, loc):
                x0[sd.begin:sd.begin + sd.n_owned] = l[:sd.n_owned]
        return x0

    # ------------------------------------------------------------------ one level
    def local_solves(self, r: np.ndarray) -> list:
        """y_i = A_ii^{-1} R_i r for every subdomain."""
        tasks = [(f, r[gid]) for f, gid in zip(self.factors, self.gids)]
        if self.pool is not None:
            return list(self.pool.map(lambda t: t[0].solve(t[1]), tasks))
        return [f.solve(b) for f, b in tasks]

    def one_level(self, r: np.ndarray) -> np.ndarray:
        """RAS: z = sum R_i^T D_i A_ii^{-1} R_i r; ASM: without D_i (one_level_schwarz.cpp:80-99)."""
        yl = self.local_solves(r)
        z = np.empty(self.n)
        if self.one_level_kind == "ras":
            yl = [y * sd.D for y, sd in zip(yl, self.subs)]
            if not self.subs[0].boolean:
                yl = self.reverse_add(yl)
        else:
            yl = self.reverse_add(yl)
        for sd, y in zip(self.subs, yl):
            z[sd.begin:sd.begin + sd.n_owned] = y[:sd.n_owned]
        return z

    # ------------------------------------------------------------------ two level
    def apply(self, r: np.ndarray) -> np.ndarray:
        """z = M^{-1} r (two_level_schwarz.cpp:134-163)."""
        r = np.asarray(r, np.float64)
        if self.n0 == 0:
            return self.one_level(r)
        x0 = self.coarse(r)
        if self.correction == "additive":
            z = self.one_level(r)
            return z + x0
        w = self.spmv(x0)
        w = r - w
        z1 = self.one_level(w)
        return x0 + z1


# ---------------------------------------------------------------------- Krylov
@dataclass
class KrylovResult:
    """Fields of KrylovResult (include/solver/krylov.hpp:43-54)."""

    converged: bool = False
    breakdown: bool = False
    iterations: int = 0
    restarts: int = 0
    rel_residual: float = 0.0
    true_rel_residual: float = 0.0
    initial_residual_of: float = 0.0
    final_residual_of: float = 0.0
    history: list = field(default_factory=list)


def norm1(v):
    return float(np.abs(v).sum())


def openfoam_norm_factor(A, b, x0, Ax0):
    """F = sum|A x0 - xm A 1| + sum|b - xm A 1| + 1e-20 (krylov.cpp:47-57)."""
    xm = x0.sum() / x0.size if x0.size else 0.0
    if xm == 0.0:
        return norm1(Ax0) + norm1(b) + 1e-20
    rowsum = A(np.ones(x0.size))
    return norm1(Ax0 - xm * rowsum) + norm1(b - xm * rowsum) + 1e-20


def gmres(A, M, b, x=None, restart=30, rtol=1e-8, maxit=1000, stop="l2", atol=0.0, flexible=False) -> tuple:
    """Right-preconditioned restarted GMRES(m) with CGS2 and Givens rotations (gmres.cpp:18-161).

    A, M: callables (vector -> vector). Returns (x, KrylovResult). flexible=True keeps
    Z_j = M(v_j) and forms x = x0 + Z y (flexible GMRES, Saad 1993) instead of x0 + M(V y); no
    reference counterpart."""
    n = b.size
    of = stop == "openfoam"
    res = KrylovResult()
    bnorm = math.sqrt(float(b @ b))
    x = np.zeros(n) if x is None or x.size != n else x.copy()
    if bnorm == 0.0:
        res.converged = True
        res.history.append(0.0)
        return np.zeros(n), res
    m = restart
    V = np.zeros((m + 1, n))
    Z = np.zeros((m, n)) if flexible else None
    H = np.zeros((m + 1, m))
    cs = np.zeros(m)
    sn = np.zeros(m)
    g = np.zeros(m + 1)
    r = A(x)
    normF = 1.0
    l1_0 = l1 = 0.0
    if of:
        normF = openfoam_norm_factor(A, b, x, r)
    r = b - r
    beta = math.sqrt(float(r @ r))
    res.history.append(beta / bnorm)
    res.rel_residual = beta / bnorm
    thr = 0.0
    if of:
        l1_0 = l1 = norm1(r)
        res.initial_residual_of = l1_0 / normF
        thr = max(rtol * l1_0, atol * normF)
        if l1_0 <= thr:
            res.converged = True
    elif beta <= rtol * bnorm:
        res.converged = True
    while not res.converged and res.iterations < maxit:
        V[0] = r / beta
        g[:] = 0.0
        g[0] = beta
        H[:] = 0.0
        k = 0
        breakdown = False
        for j in range(m):
            if res.iterations >= maxit:
                break
            z = M(V[j])
            if flexible:
                Z[j] = z
            w = A(z)
            h = V[:j + 1] @ w
            w = w - V[:j + 1].T @ h
            h2 = V[:j + 1] @ w
            w = w - V[:j + 1].T @ h2
            h = h + h2
            hn = math.sqrt(float(w @ w))
            H[:j + 1, j] = h
            H[j + 1, j] = hn
            if hn > 0.0:
                V[j + 1] = w / hn
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            den = math.hypot(H[j, j], H[j + 1, j])
            cs[j] = H[j, j] / den if den > 0.0 else 1.0
            sn[j] = H[j + 1, j] / den if den > 0.0 else 0.0
            H[j, j] = den
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            k += 1
            res.iterations += 1
            res.rel_residual = abs(g[j + 1]) / bnorm
            res.history.append(res.rel_residual)
            if of:
                if hn == 0.0 or g[j + 1] == 0.0:
                    l1 = 0.0
                else:
                    q = np.zeros(j + 2)
                    q[j + 1] = 1.0
                    for i in range(j, -1, -1):
                        a, c = q[i], q[i + 1]
                        q[i] = cs[i] * a - sn[i] * c
                        q[i + 1] = sn[i] * a + cs[i] * c
                    rj = V[:j + 2].T @ q
                    l1 = abs(g[j + 1]) * norm1(rj)
                done = l1 <= thr
            else:
                done = res.rel_residual <= rtol
            if done:
                res.converged = True
                break
            if hn == 0.0:
                breakdown = True
                break
        y = np.zeros(k)
        for i in range(k - 1, -1, -1):   # upper-triangular solve H_k y = g_k
            y[i] = (g[i] - H[i, i + 1:k] @ y[i + 1:k]) / H[i, i]
        if flexible:
            z = Z[:k].T @ y
        else:
            u = V[:k].T @ y
            z = M(u)
        x = x + z
        r = A(x)
        r = b - r
        beta = math.sqrt(float(r @ r))
        if of:
            l1 = norm1(r)
        within = (l1 <= thr) if of else (beta <= rtol * bnorm)
        within_slack = (l1 <= 1.1 * thr) if of else (beta <= 1.1 * rtol * bnorm)
        if breakdown:
            res.converged = within or res.converged
        if res.converged and not within_slack:
            res.converged = False
        if beta == 0.0:
            res.converged = True
    res.true_rel_residual = beta / bnorm
    if of:
        res.final_residual_of = l1 / normF
    return x, res


def bicgstab(A, M, b, x=None, rtol=1e-8, maxit=1000, stop="l2", atol=0.0) -> tuple:
    """Right-preconditioned BiCGStab, OpenFOAM's PBiCGStab variant (bicgstab.cpp:15-154)."""
    n = b.size
    of = stop == "openfoam"
    res = KrylovResult()
    bnorm = math.sqrt(float(b @ b))
    x = np.zeros(n) if x is None or x.size != n else x.copy()
    if bnorm == 0.0:
        res.converged = True
        res.history.append(0.0)
        return np.zeros(n), res
    p = np.zeros(n)
    v = np.zeros(n)
    r = A(x)
    normF = 1.0
    l1_0 = 0.0
    if of:
        normF = openfoam_norm_factor(A, b, x, r)
    r = b - r
    st = {"rn": 0.0, "l1": 0.0}

    def residual_norms(w, extra=None):
        d0 = float(w @ w)
        d1 = norm1(w) if of else 0.0
        d2 = float(extra @ w) if extra is not None else 0.0
        st["rn"] = math.sqrt(d0)
        st["l1"] = d1
        return d2

    thr = 0.0

    def satisfied(slack):
        return (st["l1"] <= slack * thr) if of else (st["rn"] <= slack * rtol * bnorm)

    residual_norms(r)
    if of:
        l1_0 = st["l1"]
        res.initial_residual_of = l1_0 / normF
        thr = max(rtol * l1_0, atol * normF)
    res.history.append(st["rn"] / bnorm)
    res.rel_residual = st["rn"] / bnorm
    while True:
        if satisfied(1.0):
            res.converged = True
            break
        if res.iterations >= maxit:
            break
        rh = r.copy()
        rhn = st["rn"]
        rho, rho_old, alpha, omega = st["rn"] * st["rn"], 1.0, 1.0, 1.0
        first, candidate = True, False
        while res.iterations < maxit:
            if first:
                p = r.copy()
            else:
                p = r + ((rho / rho_old) * (alpha / omega)) * (p - omega * v)
            first = False
            y = M(p)
            v = A(y)
            res.iterations += 1
            rhv = float(rh @ v)
            if rhv == 0.0 or not math.isfinite(rhv):
                res.breakdown = True
                break
            alpha = rho / rhv
            s = r - alpha * v
            residual_norms(s)
            res.rel_residual = st["rn"] / bnorm
            res.history.append(res.rel_residual)
            if satisfied(1.0) or res.iterations >= maxit:
                x = x + alpha * y
                r = s
                candidate = satisfied(1.0)
                break
            z = M(s)
            t = A(z)
            res.iterations += 1
            d0, d1 = float(t @ t), float(t @ s)
            omega = d1 / d0 if d0 > 0.0 else 0.0
            x = x + (alpha * y + omega * z)
            r = s - omega * t
            rho_old = rho
            rho = residual_norms(r, rh)
            res.rel_residual = st["rn"] / bnorm
            res.history.append(res.rel_residual)
            if satisfied(1.0):
                candidate = True
                break
            if omega == 0.0 or abs(rho) < 1e-14 * rhn * st["rn"]:
                break
        t = A(x)
        r = b - t
        residual_norms(r)
        if res.breakdown:
            break
        if candidate and satisfied(1.1):
            res.converged = True
            break
        if res.iterations >= maxit:
            break
        res.restarts += 1
    res.true_rel_residual = st["rn"] / bnorm
    if of:
        res.final_residual_of = st["l1"] / normF
    return x, res
