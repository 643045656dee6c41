"""In-tree build of the native libraries (no cmake): hipcc for the gfx950 device library, g++ for
the host setup stand-in. Run ``python build.py`` or call :func:`build_all`.

Outputs (next to this file): ``libs2l_hip.so``, ``libs2l_setup.so`` and ``libs2l_backend.so``; the
C++ test program ``tests/cpp/test_backend``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
INCLUDE = os.path.join(REPO, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled on every HIP translation unit.
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_setup():
    src = os.path.join(HERE, "csrc", "setup")
    files = [os.path.join(src, f) for f in ("generators.cpp", "partition.cpp", "factor.cpp", "mmio.cpp", "ldu.cpp", "setup_api.cpp")]
    out = os.path.join(HERE, "libs2l_setup.so")
    _run(["g++", "-std=c++17", "-O3", "-march=x86-64-v3", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-I", INCLUDE, "-o", out] + files)
    return out


def build_device():
    src = os.path.join(HERE, "csrc", "device")
    objdir = os.path.join(HERE, "csrc", "device", "_obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for f in ("kernels.hip", "lsolve.hip", "refactor.hip", "eigen.hip", "api.hip", "pack.cpp", "lsched.cpp", "refactor.cpp",
              "eigen.cpp"):
        o = os.path.join(objdir, f + ".o")
        lang = ["-x", "hip"] if f.endswith(".hip") else []
        _run([HIPCC] + HIP_FLAGS + ["-I", INCLUDE, "-I", src] + lang + ["-c", os.path.join(src, f), "-o", o])
        objs.append(o)
    out = os.path.join(HERE, "libs2l_hip.so")
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", out] + objs +
         ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    return out


def build_backend():
    """libs2l_backend.so: the C++17 host classes over the C-ABI (include/s2l_backend.hpp)."""
    src = os.path.join(HERE, "csrc", "host", "backend.cpp")
    out = os.path.join(HERE, "libs2l_backend.so")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", "-I", INCLUDE, "-o", out, src,
          "-L", HERE, "-ls2l_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_backend_test():
    """tests/cpp/test_backend: the C++ test program of the host classes (run by a GPU test)."""
    tdir = os.path.join(REPO, "tests", "cpp")
    out = os.path.join(tdir, "test_backend")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", INCLUDE, "-I", tdir, "-o", out,
          os.path.join(tdir, "test_backend.cpp"), "-L", HERE, "-ls2l_backend", "-ls2l_hip",
          "-Wl,-rpath,$ORIGIN/../../algebraic-multilevel-preconditioner_amd"])
    return out


def build_all():
    build_setup()
    build_device()
    build_backend()
    build_backend_test()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what or "setup" in what:
        build_setup()
    if "all" in what or "device" in what:
        build_device()
    if "all" in what or "backend" in what:
        build_backend()
        build_backend_test()
