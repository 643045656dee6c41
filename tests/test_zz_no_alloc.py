"""s2l_apply_noalloc (the apply whose scratch is a static device array, no allocation) against
s2l_apply through the C-ABI: two calls in a row, the second with fewer right-hand sides, on the
task-queue and level-synchronous schedules, FP64 and FP32 local factors, the Boolean RAS (direct
output) and the multiplicity ASM (local buffer and reverse add) variants; a request beyond the
scratch size is refused with S2L_ELIMIT before any launch."""
import ctypes

import numpy as np
import pytest

from cases import items, setup

pytestmark = pytest.mark.gpu

S2L_ELIMIT = -4


def _apply(M, fn, R):
    nrhs = R.shape[0]
    dr = M.ctx.alloc(R.size)
    dz = M.ctx.alloc(R.size)
    dr.upload(R)
    M.ctx._check(fn(M.ctx.h, dr.ptr, dz.ptr, nrhs, None))
    M.ctx.synchronize()
    return dz.download().reshape(R.shape)


@pytest.mark.parametrize("precision,mode,opts", [("double", 0, dict(nev=8)), ("single", 0, dict(nev=8)),
                                                 ("double", 2, dict(nev=8)),
                                                 ("double", 0, dict(nev=8, pou="multiplicity", one_level="asm",
                                                                    correction="additive"))])
def test_noalloc_apply_equals_apply(precision, mode, opts):
    from amp_amd import device
    print(f"case no-alloc apply convdiff2d 24^2 N=4 {precision} mode {mode} {opts}", flush=True)
    st = setup("convdiff2d", items(dict(nx=24)), 4, items(opts))
    M = device.HipTwoLevelSchwarz(st, precision)
    M.ctx.set_launch_mode(mode)
    L = device.hip_lib()
    rng = np.random.default_rng(17)
    for nrhs in (3, 1):
        R = rng.standard_normal((nrhs, st.n))
        z_ref = _apply(M, L.s2l_apply, R)
        z = _apply(M, L.s2l_apply_noalloc, R)
        assert np.array_equal(z, z_ref), nrhs
    # beyond the scratch size: refused before any launch, the context keeps working
    big = 1 << 20
    rc = L.s2l_apply_noalloc(M.ctx.h, ctypes.c_void_p(1), ctypes.c_void_p(1), big, None)
    assert rc == S2L_ELIMIT and b"scratch" in L.s2l_last_error(M.ctx.h)
    R = rng.standard_normal((2, st.n))
    assert np.array_equal(_apply(M, L.s2l_apply_noalloc, R), _apply(M, L.s2l_apply, R))
