"""The C++ host classes (include/s2l_backend.hpp) on the GPU: the program tests/cpp/test_backend
reads a setup written here together with the oracle's results and checks apply, GMRES,
BiCGStab, configuration and update semantics the way the reference's tests do."""
import os
import subprocess

import numpy as np
import pytest

from amp_amd import device
from oracle import oracle as orc

from cases import items, setup
from setup_io import write_arrays

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_backend")

CASES = [("laplace2d", dict(nx=20), 4, dict(nev=8), "double"),
         ("convdiff2d", dict(nx=24), 4, dict(nev=8, pou="multiplicity", one_level="asm", correction="additive"), "double"),
         ("poisson3d", dict(nx=16), 8, dict(nev=8, overlap=2), "single")]


@pytest.mark.parametrize("name,kw,N,opts,precision", CASES)
def test_cpp_backend(tmp_path, name, kw, N, opts, precision):
    print(f"case {name} {kw} N={N} {opts} {precision}", flush=True)
    st = setup(name, items(kw), N, items(opts))
    d = device.setup_arrays(st, precision)
    O = orc.OracleSchwarz(st, precision)
    r = np.cos(np.linspace(0.0, 7.0, st.n))
    d["r"] = r
    d["z_expected"] = O.apply(r)
    d["b"] = st.dec.b
    x, res = orc.gmres(O.spmv, O.apply, st.dec.b, restart=30, rtol=1e-8, maxit=500)
    d["gmres_iterations"] = np.int64(res.iterations)
    d["x_expected"] = x
    _, bres = orc.bicgstab(O.spmv, O.apply, st.dec.b, rtol=1e-8, maxit=1000)
    d["bicgstab_iterations"] = np.int64(bres.iterations)
    xf, rf = orc.gmres(O.spmv, O.apply, st.dec.b, restart=30, rtol=1e-6, maxit=500, stop="openfoam")
    d["gmres_of_iterations"] = np.int64(rf.iterations)
    d["name"] = d["one_level"].upper() + "," + d["correction"]
    path = str(tmp_path / "setup.bin")
    write_arrays(path, d)
    args = [BIN, path]
    if name == "laplace2d":
        # update levels through a setup source: the values doubled (Refactorize: same pattern and
        # coarse basis) and a rebuild of the tripled matrix with another coarse space size
        from amp_amd import hostsetup as hs
        A, _, xyz, sym, bs = hs.generate(name, **kw)
        b = hs.random_rhs(A.shape[0], 12345)
        st1 = hs.refactorize(st, 2.0 * A, b, overlap=opts.get("overlap", 1))
        st2 = hs.build_setup(3.0 * A, b, N, xyz=xyz, symmetric=sym, block_size=bs, nev=5)
        for k, stk in ((1, st1), (2, st2)):
            dk = device.setup_arrays(stk, precision)
            dk["z_expected"] = orc.OracleSchwarz(stk, precision).apply(r)
            write_arrays(str(tmp_path / f"setup{k}.bin"), dk)
        args += [str(tmp_path / "setup1.bin"), str(tmp_path / "setup2.bin")]
    p = subprocess.run(args, capture_output=True, text=True, timeout=300)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0 and "ALL PASSED" in p.stdout
