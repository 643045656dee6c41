"""The C-ABI libraries: they load, export every symbol their headers declare, report errors through
status codes, and the product path refuses to run without the HIP library (CPU only)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from amp_amd import _lib, device

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "algebraic-multilevel-preconditioner_amd")


def declared(header):
    """Function names declared in a C header (comments stripped)."""
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(s2l_\w+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("s2l_hip.h", "libs2l_hip.so"), ("s2l_setup.h", "libs2l_setup.so"),
                                        ("s2l_reuse.h", "libs2l_backend.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = declared(header)
    assert len(names) > 10
    L = ctypes.CDLL(os.path.join(PKG, lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_cpp_backend_links_and_runs_without_a_gpu_call():
    """libs2l_backend.so (the C++ host classes) loads; the C++ test program prints its usage."""
    ctypes.CDLL(os.path.join(PKG, "libs2l_backend.so"))
    exe = os.path.join(REPO, "tests", "cpp", "test_backend")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "usage" in p.stderr


def test_cpp_setup_checks_without_a_gpu(tmp_path):
    """The pattern condition of a refactorization and the SolverConfig factory of the C++ backend."""
    import shutil
    if shutil.which("g++") is None:
        pytest.skip("g++ is missing")
    exe = str(tmp_path / "test_setup_checks")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(REPO, "include"), "-o", exe,
                           os.path.join(REPO, "tests", "cpp", "test_setup_checks.cpp"), "-L", PKG, "-ls2l_backend",
                           "-ls2l_hip", f"-Wl,-rpath,{PKG}"])
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "0 failure(s)" in p.stdout, p.stdout + p.stderr


def test_setup_file_round_trip(tmp_path):
    """The setup-file writer produces the layout the C++ reader expects (header, records)."""
    import struct
    from setup_io import write_arrays
    path = tmp_path / "s.bin"
    write_arrays(str(path), {"a": np.arange(3, dtype=np.int32), "b": np.float64(2.5), "t": "ras"})
    raw = path.read_bytes()
    assert raw[:8] == b"S2LSETUP" and struct.unpack("<I", raw[8:12])[0] == 3
    assert raw[12:16] == struct.pack("<I", 1) and raw[16:17] == b"a" and raw[17] == 0
    assert struct.unpack("<Q", raw[18:26])[0] == 3


def test_python_binding_covers_the_header():
    assert sorted(device.EXPORTS) == declared("s2l_hip.h")
    device.hip_lib()   # sets every signature; loading touches no GPU


def test_ctx_create_reports_errors_by_status():
    """Without a usable device the context creation returns S2L_EHIP and a message, no exception
    crosses the ABI. Invalid arguments give S2L_EINVAL."""
    L = device.hip_lib()
    assert L.s2l_ctx_create(0, None) == -1
    assert L.s2l_last_error(None)
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and not _gpu_present():
        h = ctypes.c_void_p()
        rc = L.s2l_ctx_create(0, ctypes.byref(h))
        assert rc == -3 and not h.value
        assert b"hip" in L.s2l_last_error(None).lower()


def _gpu_present():
    return os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK)


def test_null_context_is_rejected():
    L = device.hip_lib()
    for rc in (L.s2l_apply(None, None, None, 1, None), L.s2l_spmv(None, None, None, 1, None),
               L.s2l_finalize_local(None, 0), L.s2l_synchronize(None),
               L.s2l_solver_create(None, None, 1, None), L.s2l_solver_step(None, None),
               L.s2l_solver_admit(None, 0, None, None, 1),
               L.s2l_autotune_local(None, 1, 1, None, 0, None, None), L.s2l_set_dense_top(None, 0)):
        assert rc < 0


def test_missing_library_fails_loudly():
    """The product path has no CPU fallback: a missing native library is an error."""
    with pytest.raises(RuntimeError):
        _lib._load("libs2l_does_not_exist.so")


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 2 and "one GPU" in p.stderr and p.stdout == ""


def test_setup_errors_are_reported():
    from amp_amd import hostsetup as hs
    with pytest.raises(ValueError):
        hs.generate("laplace4d", nx=4)
    with pytest.raises(ValueError):
        hs.generate("laplace2d", nx=4, not_an_option=1)
    A, b, xyz, sym, bs = hs.generate("laplace2d", nx=6)
    dec = hs.decompose(A, b, hs.partition(A, 2, xyz, bs), sym, bs)
    with pytest.raises(ValueError):
        hs.build_subdomains(dec, 1, "unknown")
    with pytest.raises(ValueError):
        hs.build_subdomains(dec, 0, "boolean")


def test_factorization_matches_matrix():
    """The host factor reproduces A_ii: P^T L U P = A_ii with perm[k] the local index of position k."""
    from amp_amd import hostsetup as hs
    import scipy.sparse as sp
    A, b, xyz, sym, bs = hs.generate("convdiff2d", nx=9)
    for kind in (0, 1):
        M = A if kind == 1 else (A + A.T) * 0.5 + sp.identity(A.shape[0]) * 0.0
        f = hs.factorize(sp.csr_matrix(M), kind)
        n = f.n
        L = sp.csc_matrix((f.l_val, f.l_row, f.l_colptr), shape=(n, n)).toarray()
        Ut = sp.csc_matrix((f.ut_val, f.ut_row, f.ut_colptr), shape=(n, n)).toarray() if kind == 1 else L
        Ap = M.toarray()[np.ix_(f.perm, f.perm)]
        assert np.abs(L @ Ut.T - Ap).max() < 1e-12 * np.abs(Ap).max()
