"""prec::ReusePolicy (include/preconditioner/reuse_policy.hpp:20-136,
src/preconditioner/reuse_policy.cpp:14-84): the update level before each solve of a sequence of
matrices with one sparsity pattern. The C++ class of the same name in include/s2l_backend.hpp is
the drop-in; this is its Python counterpart for the Python driver.

Off rebuilds before every solve; Refactor never rebuilds after the first setup; Auto reuses an
unchanged matrix, refactorizes a changed one and rebuilds once the time lost to the aging coarse
basis reaches the measured extra cost of a rebuild (lower medians of the rebuild and
refactorization times), or when a safety net fires (no convergence, q above growth x its
reference, more than max_age refactorizations). q is iterations per decade of residual reduction;
each kind of solve (e.g. a corrector index) has its own reference q."""
import math

REUSE, REFACTORIZE, REBUILD = "reuse", "refactorize", "rebuild"


class ReusePolicy:
    def __init__(self, mode: str = "auto", growth: float = 4.0, max_age: int = 0):
        if mode not in ("off", "refactor", "auto"):
            raise ValueError(f"unknown reuse mode '{mode}' (off|refactor|auto)")
        self.mode, self.growth, self.max_age = mode, growth, max_age
        self._ref = []            # per kind: q of the first solve after the last rebuild (0 = none yet)
        self._last = 0.0
        self._last_ref = 0.0
        self._age = 0
        self._converged = True
        self._wasted = 0.0        # seconds lost to the old basis since the last rebuild
        self._t_rebuild = []
        self._t_refactor = []

    @staticmethod
    def per_decade(iterations: int, reduction: float) -> float:
        if iterations > 0 and 0.0 < reduction < 1.0:
            return iterations / math.log10(1.0 / reduction)
        return 0.0

    @staticmethod
    def lower_median(v) -> float:
        v = sorted(v)
        return v[(len(v) - 1) // 2] if v else 0.0

    def rebuild_premium(self) -> float:
        rebuild = self.lower_median(self._t_rebuild)
        return max(0.0, rebuild - self.lower_median(self._t_refactor)) if rebuild > 0.0 else 0.0

    def reference(self, kind: int = 0) -> float:
        return self._ref[kind] if 0 <= kind < len(self._ref) else 0.0

    def last(self) -> float:
        return self._last

    def age(self) -> int:
        return self._age

    def wasted(self) -> float:
        return self._wasted

    def decide(self, matrix_changed: bool) -> str:
        if self.mode == "off":
            return REBUILD
        if not matrix_changed:
            return REUSE
        if self.mode == "refactor":
            return REFACTORIZE
        premium = self.rebuild_premium()
        if (not self._converged or (self.max_age > 0 and self._age >= self.max_age)
                or (self.growth > 0.0 and self._last_ref > 0.0 and self._last > self.growth * self._last_ref)
                or (premium > 0.0 and self._wasted >= premium)):
            return REBUILD
        return REFACTORIZE

    def record(self, done: str, iterations: int, reduction: float, converged: bool, t_update: float, t_solve: float,
               kind: int = 0):
        if done == REBUILD:
            self._age = 0
            self._wasted = 0.0
            self._last_ref = 0.0
            self._ref = [0.0] * len(self._ref)
            if t_update > 0.0:
                self._t_rebuild.append(t_update)
        elif done == REFACTORIZE:
            self._age += 1
            if t_update > 0.0:
                self._t_refactor.append(t_update)
        self._converged = converged
        q = self.per_decade(iterations, reduction)
        if not q > 0.0:
            return
        self._last = q
        k = max(kind, 0)
        if len(self._ref) <= k:
            self._ref += [0.0] * (k + 1 - len(self._ref))
        if not self._ref[k] > 0.0:
            self._ref[k] = q
            self._last_ref = q
            return
        self._last_ref = self._ref[k]
        extra = iterations - self._ref[k] * math.log10(1.0 / reduction)
        if extra > 0.0:
            self._wasted += extra * (t_solve / iterations)
