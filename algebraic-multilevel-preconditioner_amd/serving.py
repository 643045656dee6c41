"""A request loop over resident matrix families on one GPU: a seeded stream of solve requests (Poisson
arrivals; each request names one of a few resident families and brings 1-8 right-hand sides), served
first in, first out, with the families' preconditioners kept resident while their device footprint
fits a budget (least recently used evicted first).

The reference runs one solve, or one time loop of solves (src/solver/time_loop.cpp:57-195), per
process; the request loop is this library's: it reuses the reference's per-solve path unchanged
(GMRES(m) or BiCGStab with the two-level preconditioner; several right-hand sides of a request share
every launch). Service times are measured on the device request by request, back to back; the
queue is then replayed on a virtual clock for the offered load (start = max(arrival, previous
finish)), which is the timeline of a single FIFO server with those service times.

serve_live is the server on the wall clock: requests arrive in real time at the offered load. In
"batched" mode (continuous batching) every family has a slot solver (s2l_solver: GMRES over a
fixed number of right-hand-side slots sharing every launch) stepped by a thread of its own; between
two restart cycles the thread admits the requests of its family that have arrived into free slots,
and the families' threads run concurrently, so the solves of different families overlap on the
device (one context and stream per family). A slot's results do not depend on the other slots, so
every request gets the bits it gets alone (serve_alone). In "fifo" mode one thread serves the
requests one at a time in arrival order."""
from __future__ import annotations

import hashlib
import math
import threading
import time
from collections import OrderedDict, deque
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


def _rhs(M, q):
    """The right-hand sides of a request in the family's renumbered order."""
    n = M.setup.n
    B = np.stack([hs.random_rhs(n, q.seed + r) for r in range(q.nrhs)])
    return B[:, M.setup.dec.new_to_old]


def _digest(x: np.ndarray) -> str:
    return hashlib.blake2b(np.ascontiguousarray(x).tobytes(), digest_size=8).hexdigest()


def slot_solver(M, slots: int, restart: int, rtol: float, maxit: int):
    """The family's slot solver (created once, kept with its preconditioner; its device memory counts
    in the family's footprint)."""
    sv = getattr(M, "slot_solver", None)
    key = (slots, restart, rtol, maxit)
    if sv is None or M.slot_solver_key != key:
        M.slot_solver = sv = M.ctx.solver(slots, restart=restart, rtol=rtol, maxit=maxit)
        M.slot_solver_key = key
    return sv


def _finish(rec, results):
    rec["iterations"] = [r.iterations for _, r in results]
    rec["converged"] = all(r.converged for _, r in results)
    rec["true_rel_residual"] = max(r.true_rel_residual for _, r in results)
    rec["digest"] = [_digest(x) for x, _ in results]


def serve_alone(fams: list, reqs: list, resident: ResidentSet, *, slots: int = 16, restart: int = 50,
                rtol: float = 1e-8, maxit: int = 1000, log=None) -> list:
    """Every request by itself through its family's slot solver, one after the other (no other slot
    busy): the reference results of the continuous-batching server (iterations and solution digests
    per right-hand side) and the service time of each request."""
    out = []
    for q in reqs:
        t0 = time.perf_counter()
        M = resident.get(q.family)
        sv = slot_solver(M, slots, restart, rtol, maxit)
        B = _rhs(M, q)
        for r in range(q.nrhs):
            sv.admit(r, B[r])
        done = set()
        results = [None] * q.nrhs
        while len(done) < q.nrhs:
            for slot in sv.step():
                slot = int(slot)
                if slot not in done:
                    results[slot] = sv.take(slot)
                    done.add(slot)
        resident.fit(q.family)
        rec = dict(id=q.id, family=q.family, nrhs=q.nrhs, service=time.perf_counter() - t0)
        _finish(rec, results)
        out.append(rec)
        if log:
            log(rec)
    return out


def serve_live(fams: list, reqs: list, resident: ResidentSet, *, scale: float, mode: str = "batched",
               slots: int = 16, restart: int = 50, rtol: float = 1e-8, maxit: int = 1000, log=None) -> list:
    """Serves the requests on the wall clock: request q arrives at q.arrival * scale seconds after the
    start. "batched": one thread per family steps its slot solver one restart cycle at a time and
    admits, between cycles, the arrived requests of its family (first come, first served) whose
    right-hand sides fit the free slots; "fifo": one thread solves the requests one at a time in
    arrival order (s2l_krylov_solve with the request's right-hand sides). Every family must be
    resident (built before the call). Returns per request its arrival, admission and finish times
    [s], latency, iterations per right-hand side, convergence, the largest true residual and the
    digests of the solutions."""
    if mode not in ("batched", "fifo"):
        raise ValueError(f"unknown serving mode '{mode}' (batched|fifo)")
    members = {k: resident.get(k) for k in sorted({q.family for q in reqs})}
    solvers = {k: slot_solver(M, slots, restart, rtol, maxit) for k, M in members.items()} if mode == "batched" else {}
    recs = {q.id: dict(id=q.id, family=q.family, nrhs=q.nrhs, arrival=q.arrival * scale) for q in reqs}
    done_count = [0]
    lock = threading.Lock()
    errors = []
    t_start = time.perf_counter()

    def clock():
        return time.perf_counter() - t_start

    def completed(q, results):
        rec = recs[q.id]
        rec["finish"] = clock()
        rec["latency"] = rec["finish"] - rec["arrival"]
        _finish(rec, results)
        with lock:
            done_count[0] += 1
            if log and done_count[0] % 100 == 0:
                log(f"{mode}: {done_count[0]} requests done at {rec['finish']:.1f} s")

    def family_worker(k):
        M, sv = members[k], solvers[k]
        pending = deque(sorted((q for q in reqs if q.family == k), key=lambda q: q.arrival))
        free = list(range(slots))
        owner = {}      # slot -> (request, index of its right-hand side)
        partial = {}    # request id -> results so far
        while pending or owner:
            now = clock()
            while pending and pending[0].arrival * scale <= now and len(free) >= pending[0].nrhs:
                q = pending.popleft()
                recs[q.id]["admitted"] = now
                B = _rhs(M, q)
                partial[q.id] = [None] * q.nrhs
                for r in range(q.nrhs):
                    slot = free.pop(0)
                    sv.admit(slot, B[r])
                    owner[slot] = (q, r)
            if not owner:
                time.sleep(max(0.0, pending[0].arrival * scale - clock()))
                continue
            for slot in sv.step():
                slot = int(slot)
                q, r = owner.pop(slot)
                partial[q.id][r] = sv.take(slot)
                free.append(slot)
                if all(v is not None for v in partial[q.id]):
                    completed(q, partial.pop(q.id))

    def fifo_worker():
        for q in sorted(reqs, key=lambda q: q.arrival):
            wait = q.arrival * scale - clock()
            if wait > 0:
                time.sleep(wait)
            recs[q.id]["admitted"] = clock()
            M = members[q.family]
            B = _rhs(M, q)
            X, res = M.ctx.gmres(B if q.nrhs > 1 else B[0], None, restart, rtol, maxit)
            res = res if isinstance(res, list) else [res]
            X = X.reshape(q.nrhs, -1)
            completed(q, list(zip(X, res)))

    def guard(fn, *a):
        try:
            fn(*a)
        except BaseException as e:  # reported after the join
            errors.append(e)

    threads = ([threading.Thread(target=guard, args=(family_worker, k)) for k in members] if mode == "batched"
               else [threading.Thread(target=guard, args=(fifo_worker,))])
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return [recs[q.id] for q in sorted(reqs, key=lambda q: q.id)]


def live_summary(records: list) -> dict:
    """Wall-clock latency percentiles, right-hand sides solved per second over the run (first
    arrival to last finish) and requests per second."""
    lat = np.array([r["latency"] for r in records])
    span = max(r["finish"] for r in records) - min(r["arrival"] for r in records)
    return dict(p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
                max_ms=1e3 * float(lat.max()), solves_per_s=sum(r["nrhs"] for r in records) / span,
                requests_per_s=len(records) / span, span_s=span)


def same_as_alone(alone: list, live: list) -> dict:
    """Per request: the same iterations and the same solution bits (digests) for every right-hand
    side, alone and in the mix. Returns the count of requests that differ and their ids."""
    diff = [b["id"] for a, b in zip(alone, live)
            if a["id"] != b["id"] or a["iterations"] != b["iterations"] or a["digest"] != b["digest"]]
    return dict(requests=len(live), differing=len(diff), differing_ids=diff[:20])
