"""ctypes bindings of the native libraries of the package.

* ``libs2l_setup.so``: host setup stand-in (generators, partition, local factorizations),
  declared in ``include/s2l_setup.h``.
* ``libs2l_hip.so``: the MI355X apply path (device context, uploads, HIP kernels, device
  Krylov loops), declared in ``include/s2l_hip.h``.
* ``libs2l_backend.so``: the C++ host classes; its C-ABI of the reuse policy is declared in
  ``include/s2l_reuse.h``.

All are built in-tree by ``__graft_entry__.build()``. A missing library raises
``RuntimeError`` at first use: there is no fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))

_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_dp = ctypes.POINTER(ctypes.c_double)


class GenOpts(ctypes.Structure):
    """Mirror of ``s2l_gen_opts`` (include/s2l_setup.h)."""

    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("nu", ctypes.c_double), ("ncomp", ctypes.c_int), ("time", ctypes.c_double),
        ("pulse", ctypes.c_double), ("pulse_period", ctypes.c_int),
        ("rho_ratio", ctypes.c_double), ("nbubbles", ctypes.c_int), ("radius", ctypes.c_double),
        ("cfl", ctypes.c_double), ("buoyancy", ctypes.c_int), ("seed", ctypes.c_uint),
        ("rmin_cells", ctypes.c_double), ("rmax_cells", ctypes.c_double),
        ("contrast", ctypes.c_double), ("channel_fraction", ctypes.c_double),
        ("aniso", ctypes.c_double),
    ]


_setup = None
_hip = None
_backend = None

#: set once this process has created a device context (the HIP runtime is initialised)
gpu_initialised = False


def fork_is_safe() -> bool:
    """Whether the host setup may fork worker processes: not after this process initialised the HIP
    runtime, and not under rocprofv3, whose preloaded library initialises the GPU before the program
    starts and aborts in forked children at their exit."""
    if gpu_initialised:
        return False
    if "rocprof" in os.environ.get("LD_PRELOAD", "") or any(k.startswith("ROCPROF") for k in os.environ):
        return False
    return True


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{name} is not built (expected at {path}); run __graft_entry__.build()")
    return ctypes.CDLL(path)


def setup_lib():
    """Handle of libs2l_setup.so with argument types declared."""
    global _setup
    if _setup is not None:
        return _setup
    L = _load("libs2l_setup.so")
    L.s2l_setup_last_error.restype = ctypes.c_char_p
    L.s2l_gen_defaults.argtypes = [ctypes.POINTER(GenOpts)]
    L.s2l_generate.restype = ctypes.c_void_p
    L.s2l_generate.argtypes = [ctypes.c_char_p, ctypes.POINTER(GenOpts)]
    L.s2l_problem_free.argtypes = [ctypes.c_void_p]
    for f in ("s2l_problem_n", "s2l_problem_nnz"):
        getattr(L, f).restype = ctypes.c_int64
        getattr(L, f).argtypes = [ctypes.c_void_p]
    for f in ("s2l_problem_symmetric", "s2l_problem_block_size", "s2l_problem_dim"):
        getattr(L, f).restype = ctypes.c_int
        getattr(L, f).argtypes = [ctypes.c_void_p]
    L.s2l_problem_copy.argtypes = [ctypes.c_void_p, _c_i64p, _c_i64p, _c_dp, _c_dp, _c_dp]
    L.s2l_random_rhs.argtypes = [ctypes.c_int64, ctypes.c_uint64, _c_dp]
    L.s2l_random_rhs_seq.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _c_dp]
    L.s2l_read_matrix.restype = ctypes.c_void_p
    L.s2l_read_matrix.argtypes = [ctypes.c_char_p]
    L.s2l_read_vector.restype = ctypes.c_int
    L.s2l_read_vector.argtypes = [ctypes.c_char_p, _c_dp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    L.s2l_write_matrix.restype = ctypes.c_int
    L.s2l_write_matrix.argtypes = [ctypes.c_char_p, ctypes.c_int64, _c_i64p, _c_i64p, _c_dp]
    L.s2l_write_vector.restype = ctypes.c_int
    L.s2l_write_vector.argtypes = [ctypes.c_char_p, ctypes.c_int64, _c_dp]
    L.s2l_step_file_name.restype = ctypes.c_int64
    L.s2l_step_file_name.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int64]
    L.s2l_hash_pattern.restype = ctypes.c_uint64
    L.s2l_hash_pattern.argtypes = [ctypes.c_int64, _c_i64p, _c_i64p]
    L.s2l_from_ldu.restype = ctypes.c_void_p
    L.s2l_from_ldu.argtypes = [ctypes.c_int64, ctypes.c_int64, _c_i64p, _c_i64p, _c_dp, _c_dp, _c_dp, _c_i64p, _c_dp]
    L.s2l_problem_rhs_size.restype = ctypes.c_int64
    L.s2l_problem_rhs_size.argtypes = [ctypes.c_void_p]
    L.s2l_partition.restype = ctypes.c_int
    L.s2l_partition.argtypes = [ctypes.c_int64, _c_i64p, _c_i64p, ctypes.c_int, _c_dp, ctypes.c_int, _c_i32p]
    L.s2l_factorize.restype = ctypes.c_void_p
    L.s2l_factorize.argtypes = [ctypes.c_int32, _c_i32p, _c_i32p, _c_dp, ctypes.c_int, ctypes.c_int]
    L.s2l_factor_free.argtypes = [ctypes.c_void_p]
    L.s2l_factor_info.argtypes = [ctypes.c_void_p, _c_i32p, _c_i64p, _c_i64p, _c_i32p,
                                  ctypes.POINTER(ctypes.c_int), _c_dp]
    L.s2l_factor_copy.argtypes = [ctypes.c_void_p, _c_i32p, _c_i64p, _c_i32p, _c_dp, _c_i64p, _c_i32p, _c_dp]
    L.s2l_factor_supernodes.argtypes = [ctypes.c_void_p, _c_i32p]
    _setup = L
    return L


def backend_lib():
    """Handle of libs2l_backend.so with the argument types of its reuse-policy C-ABI declared."""
    global _backend
    if _backend is not None:
        return _backend
    L = _load("libs2l_backend.so")
    P = ctypes.c_void_p
    L.s2l_reuse_create.restype = P
    L.s2l_reuse_create.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_int]
    L.s2l_reuse_destroy.argtypes = [P]
    L.s2l_reuse_last_error.restype = ctypes.c_char_p
    L.s2l_reuse_decide.restype = ctypes.c_int
    L.s2l_reuse_decide.argtypes = [P, ctypes.c_int]
    L.s2l_reuse_record.restype = ctypes.c_int
    L.s2l_reuse_record.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int]
    for f in ("s2l_reuse_last", "s2l_reuse_wasted", "s2l_reuse_rebuild_premium"):
        getattr(L, f).restype = ctypes.c_double
        getattr(L, f).argtypes = [P]
    L.s2l_reuse_reference.restype = ctypes.c_double
    L.s2l_reuse_reference.argtypes = [P, ctypes.c_int]
    L.s2l_reuse_age.restype = ctypes.c_int
    L.s2l_reuse_age.argtypes = [P]
    L.s2l_reuse_per_decade.restype = ctypes.c_double
    L.s2l_reuse_per_decade.argtypes = [ctypes.c_int, ctypes.c_double]
    L.s2l_reuse_lower_median.restype = ctypes.c_double
    L.s2l_reuse_lower_median.argtypes = [_c_dp, ctypes.c_int]
    _backend = L
    return L


def ptr(a, ctype):
    """ctypes pointer to the data of a contiguous numpy array (None stays None)."""
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))
