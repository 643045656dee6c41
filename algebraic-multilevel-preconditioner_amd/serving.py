"""A request loop over resident matrix families on one GPU: a seeded stream of solve requests (Poisson
arrivals; each request names one of a few resident families and brings 1-8 right-hand sides), served
first in, first out, with the families' preconditioners kept resident while their device footprint
fits a budget (least recently used evicted first).

The reference runs one solve, or one time loop of solves (src/solver/time_loop.cpp:57-195), per
process; the request loop is this library's: it reuses the reference's per-solve path unchanged
(GMRES(m) or BiCGStab with the two-level preconditioner; several right-hand sides of a request share
every launch). Service times are measured on the device request by request, back to back; the
queue is then replayed on a virtual clock for the offered load (start = max(arrival, previous
finish)), which is the timeline of a single FIFO server with those service times, without idle
waits on the GPU. serve_batched is the continuous-batching server: requests waiting for the same
family join one solve that shares every launch, on the same virtual clock with measured times."""
from __future__ import annotations

import math
import time
from collections import OrderedDict
from dataclasses import dataclass

import numpy as np

from . import device, hostsetup as hs

KINDS = ("poisson2d", "poisson3d", "aniso2d")


@dataclass(frozen=True)
class Family:
    """A resident matrix family: generator kind and grid size."""

    kind: str
    nx: int

    @property
    def dim(self) -> int:
        return 3 if self.kind.endswith("3d") else 2

    @property
    def n(self) -> int:
        return self.nx ** self.dim


@dataclass
class Request:
    id: int
    family: int
    nrhs: int
    arrival: float      # in units of the mean inter-arrival time until scaled by the offered load
    seed: int


def draw_mix(n_requests: int = 2000, n_families: int = 6, n_range=(1e4, 2e6), max_rhs: int = 8,
             seed: int = 2024) -> tuple[list, list]:
    """The families (kind uniform over KINDS, n log-uniform in n_range, rounded to a grid) and the
    requests (family uniform, 1..max_rhs right-hand sides, exponential inter-arrival times of mean 1)."""
    rng = np.random.default_rng(seed)
    fams = []
    while len(fams) < n_families:
        kind = KINDS[int(rng.integers(len(KINDS)))]
        n = math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))
        d = 3 if kind.endswith("3d") else 2
        f = Family(kind, max(8, int(round(n ** (1.0 / d)))))
        if f not in fams:
            fams.append(f)
    gaps = rng.exponential(1.0, n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = [Request(i, int(rng.integers(n_families)), int(rng.integers(1, max_rhs + 1)), float(arrivals[i]),
                    int(rng.integers(1 << 31))) for i in range(n_requests)]
    return fams, reqs


class ResidentSet:
    """Preconditioners of the families kept on the device while their measured footprint
    (s2l_device_footprint, Krylov workspace included once grown) fits ``budget`` bytes; the least
    recently used family is evicted first. ``build(family) -> preconditioner`` builds a missing one."""

    def __init__(self, budget: int, build, footprint=lambda M: M.ctx.footprint()):
        self.budget = int(budget)
        self.build = build
        self.footprint = footprint
        self.live: OrderedDict = OrderedDict()
        self.builds = 0
        self.evictions = 0
        self.build_s = 0.0

    def used(self) -> int:
        return sum(self.footprint(M) for M in self.live.values())

    def get(self, key):
        if key in self.live:
            self.live.move_to_end(key)
            return self.live[key]
        t0 = time.perf_counter()
        M = self.build(key)
        self.build_s += time.perf_counter() - t0
        self.builds += 1
        self.live[key] = M
        self.fit(key)
        return M

    def fit(self, keep):
        """Evicts least recently used families other than ``keep`` until the set fits the budget;
        RuntimeError when ``keep`` alone does not."""
        while self.used() > self.budget:
            victim = next((k for k in self.live if k != keep), None)
            if victim is None:
                del self.live[keep]
                raise RuntimeError(f"family {keep} alone needs more than the device budget of {self.budget} bytes")
            del self.live[victim]
            self.evictions += 1


def build_family(f: Family, *, precision: str = "single", rows_per_subdomain: int = 4096, overlap: int = 1,
                 nev: int = 16):
    """Host setup and device context of a family (subdomains of about rows_per_subdomain rows)."""
    A, rhs, xyz, sym, bs = hs.generate(f.kind, nx=f.nx)
    nparts = max(2, int(round(f.n / rows_per_subdomain)))
    setup = hs.build_setup(A, rhs, nparts, xyz=xyz, symmetric=sym, block_size=bs, overlap=overlap, nev=nev)
    return device.HipTwoLevelSchwarz(setup, precision)


def serve(fams: list, reqs: list, resident: ResidentSet, *, krylov: str = "gmres", restart: int = 50,
          rtol: float = 1e-8, maxit: int = 1000, log=None) -> list:
    """Serves the requests in arrival order, back to back; returns per request the service time
    (preconditioner lookup, right-hand sides and solve), iterations (max over its right-hand sides),
    convergence (all) and whether the family had to be built."""
    out = []
    for q in reqs:
        t0 = time.perf_counter()
        built = resident.builds
        M = resident.get(q.family)
        built = resident.builds != built
        t1 = time.perf_counter()
        n = M.setup.n
        B = np.stack([hs.random_rhs(n, q.seed + r) for r in range(q.nrhs)])
        B = B[:, M.setup.dec.new_to_old]
        solver = device.HipGmres(restart, rtol, maxit) if krylov == "gmres" else device.HipBiCgStab(rtol, maxit)
        _, res = solver.solve(M, B if q.nrhs > 1 else B[0])
        res = res if isinstance(res, list) else [res]
        resident.fit(q.family)  # the Krylov workspace of this request counts from now on
        t2 = time.perf_counter()
        rec = dict(id=q.id, family=q.family, nrhs=q.nrhs, arrival=q.arrival, service=t2 - t0, build=t1 - t0 if built
                   else 0.0, built=built, iterations=max(r.iterations for r in res),
                   converged=all(r.converged for r in res),
                   true_rel_residual=max(r.true_rel_residual for r in res))
        out.append(rec)
        if log:
            log(rec)
    return out


def timeline(records: list, load: float, *, exclude_builds: bool = True) -> dict:
    """FIFO queue replayed at offered load ``load`` (mean inter-arrival = mean service / load).
    Returns the arrival rate, p50 / p99 / max latency, right-hand sides solved per second over the
    run (first arrival to last finish) and requests per second. Builds of families admitted before
    the stream are excluded from the service times when ``exclude_builds``."""
    svc = np.array([r["service"] - (r["build"] if exclude_builds else 0.0) for r in records])
    mean = float(svc.mean())
    scale = mean / load
    arr = np.array([r["arrival"] for r in records]) * scale
    start = np.empty_like(arr)
    fin = np.empty_like(arr)
    free = -np.inf
    for i in range(arr.size):
        start[i] = max(arr[i], free)
        fin[i] = start[i] + svc[i]
        free = fin[i]
    lat = fin - arr
    span = float(fin[-1] - arr[0])
    nrhs = sum(r["nrhs"] for r in records)
    return dict(load=load, rate_per_s=1.0 / scale, mean_service_ms=1e3 * mean, p50_ms=1e3 * float(np.percentile(lat, 50)),
                p99_ms=1e3 * float(np.percentile(lat, 99)), max_ms=1e3 * float(lat.max()), solves_per_s=nrhs / span,
                requests_per_s=len(records) / span, capacity_solves_per_s=nrhs / float(svc.sum()))


def serve_batched(fams: list, reqs: list, resident: ResidentSet, *, scale: float, max_rhs: int = 16,
                  krylov: str = "gmres", restart: int = 50, rtol: float = 1e-8, maxit: int = 1000,
                  solve=None) -> list:
    """Continuous batching on a virtual clock: arrivals at ``arrival * scale`` seconds; whenever the
    server is free it takes the oldest waiting request and every other waiting request of the same
    family (up to ``max_rhs`` right-hand sides), solves them together (each right-hand side's
    result is that of its solve alone: the launches are shared, the arithmetic per right-hand side
    is not) and advances the clock by the measured time. Returns per request its arrival, finish,
    latency, batch size, iterations, convergence and true residual. ``solve(M, B) -> list`` replaces
    the device solve in tests."""
    def device_solve(M, B):
        solver = device.HipGmres(restart, rtol, maxit) if krylov == "gmres" else device.HipBiCgStab(rtol, maxit)
        _, res = solver.solve(M, B if B.shape[0] > 1 else B[0])
        return res if isinstance(res, list) else [res]
    solve = solve or device_solve
    pending = sorted(reqs, key=lambda q: q.arrival)
    waiting: list = []
    out = []
    t = 0.0
    i = 0
    while i < len(pending) or waiting:
        while i < len(pending) and pending[i].arrival * scale <= t:
            waiting.append(pending[i])
            i += 1
        if not waiting:
            t = pending[i].arrival * scale
            continue
        fam = waiting[0].family
        batch, rhs = [], 0
        for q in waiting:
            if q.family == fam and rhs + q.nrhs <= max(max_rhs, waiting[0].nrhs):
                batch.append(q)
                rhs += q.nrhs
        waiting = [q for q in waiting if q not in batch]
        t0 = time.perf_counter()
        M = resident.get(fam)
        n = M.setup.n
        B = np.stack([hs.random_rhs(n, q.seed + r) for q in batch for r in range(q.nrhs)])
        if hasattr(M, "setup") and hasattr(M.setup, "dec"):
            B = B[:, M.setup.dec.new_to_old]
        res = solve(M, B)
        resident.fit(fam)
        t += time.perf_counter() - t0
        k = 0
        for q in batch:
            rr = res[k:k + q.nrhs]
            k += q.nrhs
            out.append(dict(id=q.id, family=fam, nrhs=q.nrhs, arrival=q.arrival * scale, finish=t,
                            latency=t - q.arrival * scale, batch=rhs, iterations=max(r.iterations for r in rr),
                            converged=all(r.converged for r in rr),
                            true_rel_residual=max(r.true_rel_residual for r in rr)))
    return sorted(out, key=lambda r: r["id"])


def batched_summary(records: list) -> dict:
    lat = np.array([r["latency"] for r in records])
    span = max(r["finish"] for r in records) - min(r["arrival"] for r in records)
    return dict(p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
                max_ms=1e3 * float(lat.max()), solves_per_s=sum(r["nrhs"] for r in records) / span,
                mean_batch_rhs=float(np.mean([r["batch"] for r in records])))
