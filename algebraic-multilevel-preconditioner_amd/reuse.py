"""prec::ReusePolicy (include/preconditioner/reuse_policy.hpp:20-136,
src/preconditioner/reuse_policy.cpp:14-84): the update level before each solve of a sequence of
matrices with one sparsity pattern, for the Python time loop. There is one implementation, the C++
class of the host classes (include/s2l_backend.hpp); this is its binding through the C-ABI of
include/s2l_reuse.h (libs2l_backend.so).

Off rebuilds before every solve; Refactor never rebuilds after the first setup; Auto reuses an
unchanged matrix, refactorizes a changed one and rebuilds once the time lost to the aging coarse
basis reaches the measured extra cost of a rebuild (lower medians of the rebuild and
refactorization times), or when a safety net fires (no convergence, q above growth x its
reference, more than max_age refactorizations). q is iterations per decade of residual reduction;
each kind of solve (e.g. a corrector index) has its own reference q."""
import numpy as np

from . import _lib

REUSE, REFACTORIZE, REBUILD = "reuse", "refactorize", "rebuild"
_LEVELS = (REUSE, REFACTORIZE, REBUILD)   # UpdateLevel 0, 1, 2


class ReusePolicy:
    def __init__(self, mode: str = "auto", growth: float = 4.0, max_age: int = 0):
        L = _lib.backend_lib()
        h = L.s2l_reuse_create(str(mode).encode(), float(growth), int(max_age))
        if not h:
            raise ValueError(L.s2l_reuse_last_error().decode())
        self._L, self._h = L, h
        self.mode, self.growth, self.max_age = mode, growth, max_age

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.s2l_reuse_destroy(h)
            self._h = None

    @staticmethod
    def per_decade(iterations: int, reduction: float) -> float:
        return _lib.backend_lib().s2l_reuse_per_decade(int(iterations), float(reduction))

    @staticmethod
    def lower_median(v) -> float:
        a = np.ascontiguousarray(v, dtype=np.float64)
        return _lib.backend_lib().s2l_reuse_lower_median(_lib.ptr(a, _lib.ctypes.c_double), int(a.size))

    def rebuild_premium(self) -> float:
        return self._L.s2l_reuse_rebuild_premium(self._h)

    def reference(self, kind: int = 0) -> float:
        return self._L.s2l_reuse_reference(self._h, int(kind))

    def last(self) -> float:
        return self._L.s2l_reuse_last(self._h)

    def age(self) -> int:
        return self._L.s2l_reuse_age(self._h)

    def wasted(self) -> float:
        return self._L.s2l_reuse_wasted(self._h)

    def decide(self, matrix_changed: bool) -> str:
        return _LEVELS[self._L.s2l_reuse_decide(self._h, int(bool(matrix_changed)))]

    def record(self, done: str, iterations: int, reduction: float, converged: bool, t_update: float, t_solve: float,
               kind: int = 0):
        if done not in _LEVELS:
            raise ValueError(f"unknown update level '{done}' ({'|'.join(_LEVELS)})")
        self._L.s2l_reuse_record(self._h, _LEVELS.index(done), int(iterations), float(reduction), int(bool(converged)),
                                 float(t_update), float(t_solve), int(kind))
