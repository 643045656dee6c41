"""Python mirror of the reference's preconditioner / Krylov interface over the C-ABI of the HIP path.

* ``HipTwoLevelSchwarz``  -- prec::Preconditioner (include/preconditioner/preconditioner.hpp:64-96)
  implemented on the device: ``apply(r) -> z``, ``name()``, ``update(level)``, ``up_to_date()``.
* ``HipGmres``            -- solv::Gmres (include/solver/gmres.hpp:39-71): ``solve(A, M, b, x)``,
  ``name() == "GMRES(m)"``; the whole loop runs on the device.
* ``make_krylov_solver``  -- solv::makeKrylovSolver (src/solver/krylov.cpp:23-36).
* ``KrylovResult``        -- include/solver/krylov.hpp:43-54.

Errors raise ``RuntimeError`` with the library's message, as the reference throws
``std::runtime_error`` (src/preconditioner/one_level_schwarz.cpp:49-53). There is no CPU
fallback: without ``libs2l_hip.so`` or a GPU every call fails.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib

P = ctypes.POINTER
_i32p, _i64p, _dp = P(ctypes.c_int32), P(ctypes.c_int64), P(ctypes.c_double)


class KrylovOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("restart", ctypes.c_int), ("rtol", ctypes.c_double),
                ("atol", ctypes.c_double), ("maxit", ctypes.c_int), ("stop_norm", ctypes.c_int),
                ("flexible", ctypes.c_int)]


class KrylovResultC(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("breakdown", ctypes.c_int), ("iterations", ctypes.c_int),
                ("restarts", ctypes.c_int), ("rel_residual", ctypes.c_double),
                ("true_rel_residual", ctypes.c_double), ("initial_residual_of", ctypes.c_double),
                ("final_residual_of", ctypes.c_double), ("history_len", ctypes.c_int)]


class TuneEntry(ctypes.Structure):
    """s2l_tune_entry of include/s2l_hip.h."""
    _fields_ = [("tasks", ctypes.c_int32), ("dense_top", ctypes.c_int32), ("launch_mode", ctypes.c_int32),
                ("threads", ctypes.c_int32), ("us", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("nsub", ctypes.c_int32), ("n0", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("nnz_a", ctypes.c_int64), ("factor_entries_fwd", ctypes.c_int64),
                ("factor_entries_bwd", ctypes.c_int64), ("local_rows", ctypes.c_int64),
                ("coarse_rows", ctypes.c_int64), ("levels_fwd_max", ctypes.c_int32),
                ("levels_bwd_max", ctypes.c_int32), ("supernodes", ctypes.c_int32),
                ("bytes_local_solve", ctypes.c_double), ("bytes_apply", ctypes.c_double),
                ("bytes_spmv", ctypes.c_double), ("refactor_ms", ctypes.c_double * 4),
                ("lds_slot_cap", ctypes.c_int32), ("max_task_slots", ctypes.c_int32), ("lds_slots", ctypes.c_int64)]


_hip = None
_METHODS = {"gmres": 0, "bicgstab": 1}
_STOPS = {"l2": 0, "openfoam": 1}

#: every symbol declared in include/s2l_hip.h
EXPORTS = ["s2l_ctx_create", "s2l_ctx_destroy", "s2l_last_error", "s2l_upload_system", "s2l_upload_subdomains",
           "s2l_upload_local_factor", "s2l_finalize_local", "s2l_upload_coarse", "s2l_apply", "s2l_spmv",
           "s2l_krylov_solve", "s2l_apply_host", "s2l_spmv_host", "s2l_krylov_solve_host", "s2l_local_solve_host",
           "s2l_device_alloc", "s2l_device_free", "s2l_memcpy_h2d", "s2l_memcpy_d2h", "s2l_copy_d2d", "s2l_synchronize",
           "s2l_get_stats", "s2l_device_footprint", "s2l_timing_enable", "s2l_timing_read", "s2l_set_local_launch_mode",
           "s2l_set_tasks_per_subdomain", "s2l_set_graphs", "s2l_solver_create", "s2l_solver_destroy",
           "s2l_solver_admit", "s2l_solver_step", "s2l_solver_take", "s2l_phase_name", "s2l_phase_timing_enable",
           "s2l_phase_timing_read", "s2l_apply_noalloc", "s2l_refactorize", "s2l_local_eigensolve", "s2l_set_dense_top",
           "s2l_autotune_local", "s2l_set_local_threads"]
#: S2L_NPHASES of include/s2l_hip.h
NPHASES = 10


def hip_lib():
    """libs2l_hip.so with the argument types of include/s2l_hip.h (loading touches no GPU)."""
    global _hip
    if _hip is not None:
        return _hip
    L = _lib._load("libs2l_hip.so")
    vp = ctypes.c_void_p
    sig = {
        "s2l_ctx_create": [ctypes.c_int, P(vp)],
        "s2l_ctx_destroy": [vp],
        "s2l_upload_system": [vp, ctypes.c_int32, _i32p, _i32p, _dp, ctypes.c_int32, _i32p],
        "s2l_upload_subdomains": [vp, ctypes.c_int32, _i32p, _i32p, _dp, ctypes.c_int],
        "s2l_upload_local_factor": [vp, ctypes.c_int32, ctypes.c_int32, _i32p, _i64p, _i32p, _dp, _i64p, _i32p, _dp],
        "s2l_finalize_local": [vp, ctypes.c_int],
        "s2l_upload_coarse": [vp, ctypes.c_int32, _i64p, _dp, _dp, ctypes.c_int, ctypes.c_int],
        "s2l_apply": [vp, vp, vp, ctypes.c_int32, vp],
        "s2l_spmv": [vp, vp, vp, ctypes.c_int32, vp],
        "s2l_apply_noalloc": [vp, vp, vp, ctypes.c_int32, vp],
        "s2l_refactorize": [vp, _dp, ctypes.c_int],
        "s2l_local_eigensolve": [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                 _i32p, _dp, _dp, _i32p, _i32p],
        "s2l_krylov_solve": [vp, P(KrylovOpts), vp, vp, ctypes.c_int32, P(KrylovResultC), _dp, vp],
        "s2l_apply_host": [vp, _dp, _dp, ctypes.c_int32],
        "s2l_spmv_host": [vp, _dp, _dp, ctypes.c_int32],
        "s2l_krylov_solve_host": [vp, P(KrylovOpts), _dp, _dp, ctypes.c_int32, P(KrylovResultC), _dp],
        "s2l_local_solve_host": [vp, ctypes.c_int32, _dp, _dp],
        "s2l_device_alloc": [vp, ctypes.c_int64, P(vp)],
        "s2l_device_free": [vp, vp],
        "s2l_memcpy_h2d": [vp, vp, vp, ctypes.c_int64],
        "s2l_memcpy_d2h": [vp, vp, vp, ctypes.c_int64],
        "s2l_copy_d2d": [vp, vp, vp, ctypes.c_int64, vp],
        "s2l_synchronize": [vp],
        "s2l_get_stats": [vp, P(Stats)],
        "s2l_device_footprint": [vp, P(ctypes.c_int64)],
        "s2l_timing_enable": [vp, ctypes.c_int],
        "s2l_timing_read": [vp, P(ctypes.c_int64), P(ctypes.c_double), ctypes.c_int],
        "s2l_set_local_launch_mode": [vp, ctypes.c_int],
        "s2l_set_tasks_per_subdomain": [vp, ctypes.c_int],
        "s2l_set_dense_top": [vp, ctypes.c_int32],
        "s2l_set_local_threads": [vp, ctypes.c_int32],
        "s2l_autotune_local": [vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(TuneEntry), ctypes.c_int32,
                               ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)],
        "s2l_set_graphs": [vp, ctypes.c_int],
        "s2l_solver_create": [vp, P(KrylovOpts), ctypes.c_int32, P(vp)],
        "s2l_solver_destroy": [vp],
        "s2l_solver_admit": [vp, ctypes.c_int32, vp, vp, ctypes.c_int],
        "s2l_solver_step": [vp, _i32p],
        "s2l_solver_take": [vp, ctypes.c_int32, P(KrylovResultC), _dp, vp, ctypes.c_int],
    }
    sig["s2l_phase_timing_enable"] = [vp, ctypes.c_int]
    sig["s2l_phase_timing_read"] = [vp, _dp, _i64p, ctypes.c_int]
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.s2l_last_error.argtypes = [vp]
    L.s2l_last_error.restype = ctypes.c_char_p
    L.s2l_phase_name.argtypes = [ctypes.c_int]
    L.s2l_phase_name.restype = ctypes.c_char_p
    _hip = L
    return L


def _p(a, t):
    return a.ctypes.data_as(P(t))


# local factor precision: "double" (FP64), "single" (FP32, the reference's solveIn<float>), "mixed"
# (FP32 factor entries applied in FP64 arithmetic, no reference counterpart)
_PRECISIONS = {"double": 0, "single": 1, "mixed": 2}


def setup_arrays(setup, precision: str = "double", one_level: str | None = None,
                 correction: str | None = None) -> dict:
    """The arrays of the C-ABI upload (include/s2l_hip.h) from a hostsetup.SchwarzSetup: the
    operator (CSR, renumbered order), ownership, Omega_i maps, D_i (absent for the Boolean
    partition of unity), factors f{i}_* (U^T only for LU), coarse offsets, V blocks and A_00."""
    dec = setup.dec
    A = dec.A.tocsr()
    d = {"n": np.int64(A.shape[0]),
         "rowptr": np.ascontiguousarray(A.indptr, np.int32),
         "col": np.ascontiguousarray(A.indices, np.int32),
         "val": np.ascontiguousarray(A.data, np.float64),
         "ownership": np.ascontiguousarray(dec.ownership, np.int32)}
    gids = [sd.global_ids() for sd in setup.subs]
    lptr = np.zeros(len(gids) + 1, np.int32)
    np.cumsum([len(g) for g in gids], out=lptr[1:])
    d["local_ptr"] = lptr
    d["gid"] = np.ascontiguousarray(np.concatenate(gids), np.int32)
    if not setup.subs[0].boolean:
        d["pou"] = np.ascontiguousarray(np.concatenate([sd.D for sd in setup.subs]), np.float64)
    for i, f in enumerate(setup.factors):
        d[f"f{i}_perm"] = np.ascontiguousarray(f.perm, np.int32)
        d[f"f{i}_lcp"] = np.ascontiguousarray(f.l_colptr, np.int64)
        d[f"f{i}_lrow"] = np.ascontiguousarray(f.l_row, np.int32)
        d[f"f{i}_lval"] = np.ascontiguousarray(f.l_val, np.float64)
        if f.kind != 0:
            d[f"f{i}_ucp"] = np.ascontiguousarray(f.ut_colptr, np.int64)
            d[f"f{i}_urow"] = np.ascontiguousarray(f.ut_row, np.int32)
            d[f"f{i}_uval"] = np.ascontiguousarray(f.ut_val, np.float64)
    cs = setup.coarse
    d["n0"] = np.int64(cs.n0 if cs is not None else 0)
    d["coarse_offsets"] = np.ascontiguousarray(cs.offsets if cs is not None else np.zeros(len(setup.subs) + 1), np.int64)
    if cs is not None and cs.n0 > 0:
        blocks = []
        for sd, V in zip(setup.subs, cs.V):
            rows = V[:sd.n_owned] if sd.boolean else V
            blocks.append(np.asfortranarray(rows).reshape(-1, order="F"))
        d["V"] = np.ascontiguousarray(np.concatenate(blocks), np.float64)
        d["E"] = np.ascontiguousarray(cs.E, np.float64)
    d["symmetric"] = np.int64(1 if dec.symmetric else 0)
    d["one_level"] = one_level or setup.one_level
    d["correction"] = correction or setup.correction
    d["precision"] = precision
    return d


class DeviceArray:
    """A device buffer of float64 owned by a context (for inputs resident in HBM)."""

    def __init__(self, ctx: "HipContext", n: int):
        self.ctx = ctx
        self.n = n
        self.ptr = ctypes.c_void_p()
        ctx._check(hip_lib().s2l_device_alloc(ctx.h, 8 * n, ctypes.byref(self.ptr)))

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, np.float64)
        assert a.size == self.n
        self.ctx._check(hip_lib().s2l_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(ctypes.c_void_p), 8 * self.n))

    def download(self) -> np.ndarray:
        out = np.empty(self.n)
        self.ctx._check(hip_lib().s2l_memcpy_d2h(self.ctx.h, out.ctypes.data_as(ctypes.c_void_p), self.ptr, 8 * self.n))
        return out

    def free(self):
        if self.ptr:
            hip_lib().s2l_device_free(self.ctx.h, self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HipContext:
    """Owner of one s2l_ctx (all subdomains of one problem on one GPU)."""

    def __init__(self, device: int = 0):
        self.L = hip_lib()
        self.h = ctypes.c_void_p()
        _lib.gpu_initialised = True
        rc = self.L.s2l_ctx_create(device, ctypes.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"s2l_ctx_create: {self.L.s2l_last_error(None).decode()}")

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"s2l error {rc}: {self.L.s2l_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.L.s2l_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ uploads
    def upload_setup(self, setup, precision: str = "double", one_level: str | None = None,
                     correction: str | None = None, tasks_per_subdomain: int | None = None, coarse: bool = True,
                     dense_top: int | None = None, coarse_on_device: bool = False, local_threads: int | None = None):
        """Uploads everything a hostsetup.SchwarzSetup holds (the reference's setup products);
        ``coarse=False`` stops after the local factors (a setup whose coarse space is not built yet).
        ``dense_top``: widest skeleton replaced by a dense top (s2l_set_dense_top; None = the
        library's default, 0 = none). ``coarse_on_device``: A_00 = W^T A W formed by the library on
        the device from the uploaded operator and basis (E = NULL) instead of the setup's."""
        if tasks_per_subdomain is not None:
            self._check(self.L.s2l_set_tasks_per_subdomain(self.h, tasks_per_subdomain))
        if dense_top is not None:
            self._check(self.L.s2l_set_dense_top(self.h, dense_top))
        if local_threads is not None:
            self._check(self.L.s2l_set_local_threads(self.h, local_threads))
        d = setup_arrays(setup, precision, one_level, correction)
        rp, col, val, own = d["rowptr"], d["col"], d["val"], d["ownership"]
        N = len(own) - 1
        self._check(self.L.s2l_upload_system(self.h, int(d["n"]), _p(rp, ctypes.c_int32), _p(col, ctypes.c_int32),
                                             _p(val, ctypes.c_double), N, _p(own, ctypes.c_int32)))
        lptr, gid, pou = d["local_ptr"], d["gid"], d.get("pou")
        self._check(self.L.s2l_upload_subdomains(self.h, N, _p(lptr, ctypes.c_int32), _p(gid, ctypes.c_int32),
                                                 None if pou is None else _p(pou, ctypes.c_double),
                                                 0 if d["one_level"] == "ras" else 1))
        for i, f in enumerate(setup.factors):
            self.upload_factor(i, f)
        self.finalize(precision)
        if coarse:
            self.upload_coarse(setup, d["correction"], coarse_on_device)
        self._keep = d

    def upload_factor(self, i, f):
        perm = np.ascontiguousarray(f.perm, np.int32)
        lcp = np.ascontiguousarray(f.l_colptr, np.int64)
        lr = np.ascontiguousarray(f.l_row, np.int32)
        lv = np.ascontiguousarray(f.l_val, np.float64)
        if f.kind == 0:
            args = (None, None, None)
        else:
            ucp = np.ascontiguousarray(f.ut_colptr, np.int64)
            ur = np.ascontiguousarray(f.ut_row, np.int32)
            uv = np.ascontiguousarray(f.ut_val, np.float64)
            args = (_p(ucp, ctypes.c_int64), _p(ur, ctypes.c_int32), _p(uv, ctypes.c_double))
        self._check(self.L.s2l_upload_local_factor(self.h, i, f.n, _p(perm, ctypes.c_int32), _p(lcp, ctypes.c_int64),
                                                   _p(lr, ctypes.c_int32), _p(lv, ctypes.c_double), *args))

    def finalize(self, precision: str = "double"):
        self._check(self.L.s2l_finalize_local(self.h, _PRECISIONS[precision]))

    def upload_coarse(self, setup, correction: str = "deflated", coarse_on_device: bool = False):
        cs = setup.coarse
        n0 = cs.n0
        off = np.ascontiguousarray(cs.offsets, np.int64)
        if n0 > 0:
            blocks = []
            for sd, V in zip(setup.subs, cs.V):
                rows = V[:sd.n_owned] if sd.boolean else V
                blocks.append(np.asfortranarray(rows).reshape(-1, order="F"))
            Vc = np.ascontiguousarray(np.concatenate(blocks), np.float64)
            E = None if coarse_on_device else np.ascontiguousarray(cs.E, np.float64)
            self._check(self.L.s2l_upload_coarse(self.h, n0, _p(off, ctypes.c_int64), _p(Vc, ctypes.c_double),
                                                 None if E is None else _p(E, ctypes.c_double),
                                                 1 if setup.dec.symmetric else 0, 1 if correction == "additive" else 0))
            self._keep_coarse = (Vc, E, off)
        else:
            self._check(self.L.s2l_upload_coarse(self.h, 0, _p(off, ctypes.c_int64), None, None, 1,
                                                 1 if correction == "additive" else 0))

    def refactorize(self, val: np.ndarray):
        """s2l_refactorize: new values of A in the order of the uploaded arrays (host memory)."""
        val = np.ascontiguousarray(val, np.float64)
        self._check(self.L.s2l_refactorize(self.h, _p(val, ctypes.c_double), 1))

    def local_eigensolve(self, nev: int, ncv: int, shift_rel: float = 1e-10, tol: float = 1e-8,
                         max_restarts: int = 100, skip=None):
        """s2l_local_eigensolve: the nev largest eigenpairs of B_i u = lambda (A~_ii + sigma_i I) u of
        every subdomain not in ``skip`` (a boolean per subdomain). Returns (lam nsub x nev,
        Z nev x local_total, converged per subdomain, info [restarts, steps, attempts, converged])."""
        lptr = self._keep["local_ptr"]
        nsub, nl = len(lptr) - 1, int(lptr[-1])
        lam = np.zeros((nsub, nev))
        Z = np.zeros((nev, nl))
        conv = np.zeros(nsub, np.int32)
        info = np.zeros(4, np.int32)
        sk = None if skip is None else np.ascontiguousarray(skip, np.int32)
        self._check(self.L.s2l_local_eigensolve(self.h, nev, ncv, shift_rel, tol, max_restarts,
                                                None if sk is None else _p(sk, ctypes.c_int32), _p(lam, ctypes.c_double),
                                                _p(Z, ctypes.c_double), _p(conv, ctypes.c_int32),
                                                _p(info, ctypes.c_int32)))
        return lam, Z, conv, info

    # ------------------------------------------------------------------ operations (host memory)
    def apply(self, r: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, np.float64)
        nrhs = 1 if r.ndim == 1 else r.shape[0]
        z = np.empty_like(r)
        self._check(self.L.s2l_apply_host(self.h, _p(r, ctypes.c_double), _p(z, ctypes.c_double), nrhs))
        return z

    def spmv(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        nrhs = 1 if x.ndim == 1 else x.shape[0]
        y = np.empty_like(x)
        self._check(self.L.s2l_spmv_host(self.h, _p(x, ctypes.c_double), _p(y, ctypes.c_double), nrhs))
        return y

    def local_solve(self, i: int, b: np.ndarray) -> np.ndarray:
        b = np.ascontiguousarray(b, np.float64)
        y = np.empty_like(b)
        self._check(self.L.s2l_local_solve_host(self.h, i, _p(b, ctypes.c_double), _p(y, ctypes.c_double)))
        return y

    def krylov(self, b: np.ndarray, x: np.ndarray | None = None, method: str = "gmres", restart=30, rtol=1e-8,
               maxit=1000, stop: str = "l2", atol: float = 0.0, flexible: bool = False):
        """KrylovSolver::solve on the device through s2l_krylov_solve_host (host arrays in and out)."""
        b = np.ascontiguousarray(b, np.float64)
        nrhs = 1 if b.ndim == 1 else b.shape[0]
        x = np.zeros_like(b) if x is None else np.ascontiguousarray(x, np.float64).copy()
        o = KrylovOpts(_METHODS[method], restart, rtol, atol, maxit, _STOPS[stop], 1 if flexible else 0)
        res = (KrylovResultC * nrhs)()
        hist = np.zeros(nrhs * (maxit + 1))
        self._check(self.L.s2l_krylov_solve_host(self.h, ctypes.byref(o), _p(b, ctypes.c_double),
                                                 _p(x, ctypes.c_double), nrhs, res, _p(hist, ctypes.c_double)))
        out = []
        for r in range(nrhs):
            R = res[r]
            out.append(KrylovResult(bool(R.converged), bool(R.breakdown), R.iterations, R.restarts, R.rel_residual,
                                    R.true_rel_residual, R.initial_residual_of, R.final_residual_of,
                                    list(hist[r * (maxit + 1): r * (maxit + 1) + R.history_len])))
        return x, (out[0] if nrhs == 1 else out)

    def gmres(self, b: np.ndarray, x: np.ndarray | None = None, restart=30, rtol=1e-8, maxit=1000,
              flexible: bool = False, stop: str = "l2", atol: float = 0.0):
        return self.krylov(b, x, "gmres", restart, rtol, maxit, stop, atol, flexible=flexible)

    def bicgstab(self, b: np.ndarray, x: np.ndarray | None = None, rtol=1e-8, maxit=1000, stop="l2", atol=0.0):
        return self.krylov(b, x, "bicgstab", 0, rtol, maxit, stop, atol)

    # ------------------------------------------------------------------ operations (device memory)
    def alloc(self, n: int) -> DeviceArray:
        return DeviceArray(self, n)

    def apply_device(self, r: DeviceArray, z: DeviceArray, nrhs: int = 1):
        self._check(self.L.s2l_apply(self.h, r.ptr, z.ptr, nrhs, None))

    def spmv_device(self, x: DeviceArray, y: DeviceArray, nrhs: int = 1):
        self._check(self.L.s2l_spmv(self.h, x.ptr, y.ptr, nrhs, None))

    def gmres_device(self, b: DeviceArray, x: DeviceArray, nrhs=1, restart=30, rtol=1e-8, maxit=1000, method="gmres",
                     flexible: bool = False):
        o = KrylovOpts(_METHODS[method], restart, rtol, 0.0, maxit, 0, 1 if flexible else 0)
        res = (KrylovResultC * nrhs)()
        self._check(self.L.s2l_krylov_solve(self.h, ctypes.byref(o), b.ptr, x.ptr, nrhs, res, None, None))
        return [res[r] for r in range(nrhs)]

    def solver(self, slots: int, restart=30, rtol=1e-8, maxit=1000, stop: str = "l2", atol: float = 0.0,
               flexible: bool = False) -> "SlotSolver":
        """A GMRES run over ``slots`` right-hand-side slots (s2l_solver_*, continuous batching)."""
        return SlotSolver(self, slots, KrylovOpts(0, restart, rtol, atol, maxit, _STOPS[stop], 1 if flexible else 0))

    def copy_device(self, src: DeviceArray, dst: DeviceArray, count: int | None = None):
        count = min(src.n, dst.n) if count is None else count
        self._check(self.L.s2l_copy_d2d(self.h, dst.ptr, src.ptr, count, None))

    def synchronize(self):
        self._check(self.L.s2l_synchronize(self.h))

    def stats(self) -> Stats:
        s = Stats()
        self._check(self.L.s2l_get_stats(self.h, ctypes.byref(s)))
        return s

    def footprint(self) -> int:
        """Bytes of device memory the context holds (s2l_device_footprint)."""
        v = ctypes.c_int64(0)
        self._check(self.L.s2l_device_footprint(self.h, ctypes.byref(v)))
        return int(v.value)

    def timing(self, on: bool):
        self._check(self.L.s2l_timing_enable(self.h, 1 if on else 0))

    def timing_read(self, reset=True):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        self._check(self.L.s2l_timing_read(self.h, ctypes.byref(n), ctypes.byref(ms), 1 if reset else 0))
        return n.value, ms.value

    def phase_timing(self, on: bool):
        """Per-phase event timing under the reference's timer names (s2l_phase_timing_enable)."""
        self._check(self.L.s2l_phase_timing_enable(self.h, 1 if on else 0))

    def phase_timing_read(self, reset=True) -> dict:
        """{timer name: (milliseconds summed, intervals)} (s2l_phase_timing_read)."""
        ms = np.zeros(NPHASES)
        cnt = np.zeros(NPHASES, np.int64)
        self._check(self.L.s2l_phase_timing_read(self.h, _p(ms, ctypes.c_double), _p(cnt, ctypes.c_int64),
                                                 1 if reset else 0))
        return {self.L.s2l_phase_name(q).decode(): (float(ms[q]), int(cnt[q])) for q in range(NPHASES)}

    def set_graphs(self, on: bool):
        self._check(self.L.s2l_set_graphs(self.h, 1 if on else 0))

    def set_launch_mode(self, mode: int):
        self._check(self.L.s2l_set_local_launch_mode(self.h, mode))

    def autotune(self, nrhs: int = 1, applies: int = 20):
        """s2l_autotune_local: times the candidate local-solve schedules and keeps the fastest.
        Returns (candidates as dicts with tasks, dense_top, launch_mode, us; index of the kept one)."""
        ent = (TuneEntry * 32)()
        cnt = ctypes.c_int32()
        best = ctypes.c_int32()
        self._check(self.L.s2l_autotune_local(self.h, nrhs, applies, ent, 32, ctypes.byref(cnt), ctypes.byref(best)))
        out = [dict(tasks=e.tasks, dense_top=e.dense_top, launch_mode=e.launch_mode, threads=e.threads, us=e.us)
               for e in ent[:cnt.value]]
        return out, best.value


class SlotSolver:
    """s2l_solver: GMRES over a fixed number of right-hand-side slots sharing every launch. A right-hand
    side enters a free slot (admit) and starts at the next step, which runs one restart cycle of
    every busy slot; a finished slot's result is taken (take), which frees it. A slot's results do
    not depend on the other slots. ctypes releases the GIL during the calls, so solvers of different
    contexts stepped from different threads overlap on the device."""

    def __init__(self, ctx: HipContext, slots: int, opts: KrylovOpts):
        self.ctx, self.slots, self.maxit, self.n = ctx, slots, opts.maxit, None
        self.h = ctypes.c_void_p()
        self.L = ctx.L
        ctx._check(self.L.s2l_solver_create(ctx.h, ctypes.byref(opts), slots, ctypes.byref(self.h)))
        self.n = ctx.stats().n
        self._fin = np.zeros(slots, np.int32)

    def admit(self, slot: int, b: np.ndarray, x0: np.ndarray | None = None):
        b = np.ascontiguousarray(b, np.float64)
        x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        self.ctx._check(self.L.s2l_solver_admit(self.h, slot, b.ctypes.data_as(ctypes.c_void_p),
                                                None if x0 is None else x0.ctypes.data_as(ctypes.c_void_p), 1))

    def step(self) -> np.ndarray:
        """One restart cycle; returns the slots that hold a finished solve."""
        self.ctx._check(self.L.s2l_solver_step(self.h, _p(self._fin, ctypes.c_int32)))
        return np.flatnonzero(self._fin)

    def take(self, slot: int):
        R = KrylovResultC()
        hist = np.zeros(self.maxit + 1)
        x = np.empty(self.n)
        self.ctx._check(self.L.s2l_solver_take(self.h, slot, ctypes.byref(R), _p(hist, ctypes.c_double),
                                               x.ctypes.data_as(ctypes.c_void_p), 1))
        return x, KrylovResult(bool(R.converged), bool(R.breakdown), R.iterations, R.restarts, R.rel_residual,
                               R.true_rel_residual, R.initial_residual_of, R.final_residual_of,
                               list(hist[:R.history_len]))

    def close(self):
        if self.h:
            self.L.s2l_solver_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class KrylovResult:
    """include/solver/krylov.hpp:43-54."""

    converged: bool = False
    breakdown: bool = False
    iterations: int = 0
    restarts: int = 0
    rel_residual: float = 0.0
    true_rel_residual: float = 0.0
    initial_residual_of: float = 0.0
    final_residual_of: float = 0.0
    history: list = field(default_factory=list)


class HipTwoLevelSchwarz:
    """prec::Preconditioner of the reference, applied on the GPU (two_level_schwarz.cpp:134-163)."""

    def __init__(self, setup, precision: str = "double", device: int = 0, one_level: str | None = None,
                 correction: str | None = None, tasks_per_subdomain: int | None = None, dense_top: int | None = None,
                 local_threads: int | None = None):
        self.setup = setup
        self.precision = precision
        self.one_level = one_level or setup.one_level
        self.correction = correction or setup.correction
        self.tasks_per_subdomain = tasks_per_subdomain
        self.dense_top = dense_top
        self.local_threads = local_threads
        self.ctx = HipContext(device)
        self.ctx.upload_setup(setup, precision, self.one_level, self.correction, tasks_per_subdomain, dense_top=dense_top,
                              local_threads=local_threads)

    def apply(self, r):
        return self.ctx.apply(r)

    def autotune(self, nrhs: int = 1, applies: int = 20):
        """Picks the local-solve schedule by timing (HipContext.autotune) and keeps it for later
        refreshes. Call before refactorize_device (it repacks the uploaded host factors)."""
        cands, best = self.ctx.autotune(nrhs, applies)
        self.tasks_per_subdomain = cands[best]["tasks"]
        self.dense_top = cands[best]["dense_top"]
        self.local_threads = cands[best]["threads"]
        return cands, best

    def name(self) -> str:
        """OneLevelSchwarz::name + ",deflated"/",additive" (two_level_schwarz.cpp:129-132)."""
        return self.one_level.upper() + ("," + self.correction)

    def refactorize_device(self, A_new) -> str:
        """Refactorize on the device (s2l_refactorize): ``A_new`` in the original ordering with the
        sparsity pattern the setup was built from (ValueError otherwise: a rebuild); the local
        factors and A_00 = W^T A W are recomputed on the GPU with the same subdomains and coarse
        basis. The host setup keeps its old factors (``self.setup.factors`` no longer match)."""
        import scipy.sparse as sp
        A = sp.csr_matrix(A_new)
        A.sort_indices()
        if getattr(self, "_ren", None) is None:
            dec = self.setup.dec
            idx = sp.csr_matrix((np.arange(A.nnz, dtype=np.float64), A.indices, A.indptr), shape=A.shape)
            ren = idx[dec.new_to_old][:, dec.new_to_old].tocsr()
            ren.sort_indices()
            Ar = dec.A.tocsr()
            if not (np.array_equal(ren.indptr, Ar.indptr) and np.array_equal(ren.indices, Ar.indices)):
                raise ValueError("refactorize: the sparsity pattern changed; rebuild instead")
            self._ren = ren.data.astype(np.int64)
            self._pattern = (A.indptr.copy(), A.indices.copy())
        elif not (np.array_equal(A.indptr, self._pattern[0]) and np.array_equal(A.indices, self._pattern[1])):
            raise ValueError("refactorize: the sparsity pattern changed; rebuild instead")
        self.ctx.refactorize(A.data[self._ren])
        return "refactorize"

    def refresh(self, setup) -> str:
        """Refactorize: uploads the operator, the local factors and A_00 of ``setup`` (new values of
        A with the same pattern and coarse basis, hostsetup.refactorize) into this context."""
        self.setup = setup
        self.ctx.upload_setup(setup, self.precision, self.one_level, self.correction, self.tasks_per_subdomain,
                              dense_top=self.dense_top, local_threads=self.local_threads)
        return "refactorize"

    def update(self, level: str = "reuse", setup=None) -> str:
        """UpdateLevel semantics (preconditioner.hpp:30): reuse keeps everything; refactorize
        uploads the refactorized host setup given; rebuild needs a new object (the host setup of a
        new coarse basis is not part of this class): ValueError."""
        if level not in ("reuse", "refactorize", "rebuild"):
            raise ValueError(f"unknown update level '{level}'")
        if level == "reuse":
            return "reuse"
        if level == "refactorize" and setup is not None:
            return self.refresh(setup)
        raise ValueError(f"update level '{level}' needs " + ("the refactorized setup" if level == "refactorize"
                                                              else "a new preconditioner object"))

    def up_to_date(self) -> bool:
        return True


class HipGmres:
    """solv::Gmres (gmres.hpp:39-71) with the loop on the device; flexible=True stores M^{-1} v_j
    (flexible GMRES, an extension without reference counterpart)."""

    def __init__(self, restart=30, rtol=1e-8, maxit=1000, stop="l2", atol=0.0, flexible=False):
        if stop not in _STOPS:
            raise ValueError(f"unknown stopping norm '{stop}' (l2|openfoam)")
        if atol > 0.0 and stop != "openfoam":
            raise ValueError("an absolute tolerance needs the OpenFOAM stopping norm")
        self.m, self.rtol, self.maxit, self.flexible = restart, rtol, maxit, flexible
        self.stop, self.atol = stop, atol

    def name(self) -> str:
        return f"GMRES({self.m})"

    def solve(self, M: HipTwoLevelSchwarz, b, x=None):
        return M.ctx.gmres(b, x, self.m, self.rtol, self.maxit, flexible=self.flexible, stop=self.stop, atol=self.atol)


class HipBiCgStab:
    """solv::BiCgStab (bicgstab.hpp:14-40, bicgstab.cpp:15-154) with the vector work on the device."""

    def __init__(self, rtol=1e-8, maxit=1000, stop="l2", atol=0.0):
        if stop not in _STOPS:
            raise ValueError(f"unknown stopping norm '{stop}' (l2|openfoam)")
        self.rtol, self.maxit, self.stop, self.atol = rtol, maxit, stop, atol

    def name(self) -> str:
        return "BiCGStab"

    def solve(self, M: HipTwoLevelSchwarz, b, x=None):
        return M.ctx.bicgstab(b, x, self.rtol, self.maxit, self.stop, self.atol)


def make_krylov_solver(krylov="gmres", restart=30, rtol=1e-8, maxit=1000, stop="l2", atol=0.0):
    """makeKrylovSolver (src/solver/krylov.cpp:23-36)."""
    if atol > 0.0 and stop != "openfoam":
        raise ValueError("an absolute tolerance needs the OpenFOAM stopping norm")
    if krylov == "gmres":
        return HipGmres(restart, rtol, maxit, stop, atol)
    if krylov == "bicgstab":
        return HipBiCgStab(rtol, maxit, stop, atol)
    raise ValueError(f"unknown Krylov solver '{krylov}' (gmres|bicgstab)")
