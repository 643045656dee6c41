"""The request loop's host logic (mix drawing, admission by footprint with LRU eviction, the FIFO
timeline), CPU only, with stand-in preconditioners where a device would be needed."""
import numpy as np
import pytest

from amp_amd import serving as sv


def test_mix_is_seeded_and_in_range():
    fams, reqs = sv.draw_mix(500, 6, (1e4, 2e6), 8, seed=7)
    fams2, reqs2 = sv.draw_mix(500, 6, (1e4, 2e6), 8, seed=7)
    assert fams == fams2 and [(q.family, q.nrhs, q.arrival, q.seed) for q in reqs] == \
        [(q.family, q.nrhs, q.arrival, q.seed) for q in reqs2]
    assert len(set(fams)) == 6 and all(f.kind in sv.KINDS for f in fams)
    assert all(0.5e4 <= f.n <= 2.2e6 for f in fams)
    assert {q.nrhs for q in reqs} == set(range(1, 9))
    gaps = np.diff([q.arrival for q in reqs])
    assert np.all(gaps >= 0) and abs(gaps.mean() - 1.0) < 0.15  # exponential of mean 1


class _Fake:
    def __init__(self, size):
        self.size = size


def test_resident_set_evicts_least_recently_used():
    sizes = {0: 40, 1: 30, 2: 50, 3: 120}
    rs = sv.ResidentSet(100, build=lambda k: _Fake(sizes[k]), footprint=lambda M: M.size)
    rs.get(0), rs.get(1)
    assert list(rs.live) == [0, 1] and rs.used() == 70
    rs.get(0)                       # 0 becomes the most recent
    rs.get(2)                       # 120 > 100: evicts 1 (least recent), keeps 0
    assert list(rs.live) == [0, 2] and rs.evictions == 1 and rs.builds == 3
    with pytest.raises(RuntimeError, match="budget"):
        rs.get(3)                   # alone above the budget
    assert 3 not in rs.live and rs.used() <= 100


def test_timeline_is_a_fifo_queue():
    """Hand-checked FIFO: service 1, 3, 1 at arrivals 0, 1, 2 (load such that the scale is 1)."""
    recs = [dict(service=s, build=0.0, arrival=a, nrhs=k) for s, a, k in ((1.0, 0.0, 1), (3.0, 0.6, 2), (1.0, 1.2, 1))]
    t = sv.timeline(recs, load=5.0 / 3.0 / 0.6 * 0.6)   # mean service 5/3; scale = mean / load = 1
    # starts 0, 1, 4; finishes 1, 4, 5; latencies 1, 3.4, 3.8
    assert t["max_ms"] == pytest.approx(3800.0)
    assert t["p50_ms"] == pytest.approx(3400.0)
    assert t["solves_per_s"] == pytest.approx(4 / 5.0)
    assert t["capacity_solves_per_s"] == pytest.approx(4 / 5.0)
    # at a light load nobody waits: latency = service
    light = sv.timeline(recs, load=1e-6)
    assert light["max_ms"] == pytest.approx(3000.0)


class _FakeSolver:
    """Slot solver stand-in: a slot finishes two steps after its admission; x = b."""

    def __init__(self, slots):
        self.slots, self.busy, self.log = slots, {}, []

    def admit(self, slot, b, x0=None):
        assert 0 <= slot < self.slots and slot not in self.busy
        self.busy[slot] = [2, np.array(b)]

    def step(self):
        self.log.append(sorted(self.busy))
        for v in self.busy.values():
            v[0] -= 1
        return np.array([s for s, v in self.busy.items() if v[0] <= 0], dtype=np.int64)

    def take(self, slot):
        from types import SimpleNamespace as NS
        v = self.busy.pop(slot)
        return v[1].copy(), NS(iterations=7, converged=True, true_rel_residual=1e-9)


class _FakeM:
    def __init__(self, k):
        from types import SimpleNamespace as NS
        self.k = k
        self.setup = NS(n=4, dec=NS(new_to_old=np.arange(4)))
        self.ctx = NS(solver=lambda slots, **kw: _FakeSolver(slots))


def test_live_server_admits_between_cycles_and_matches_alone():
    """The wall-clock server: a family's requests enter free slots between steps, first come first
    served, a request that does not fit waits for slots; every request is completed with its own
    results, and the digests and iterations equal those of the requests served alone."""
    reqs = [sv.Request(0, 0, 2, 0.0, 1), sv.Request(1, 1, 1, 0.0, 2), sv.Request(2, 0, 2, 0.0, 3),
            sv.Request(3, 0, 1, 0.0, 4), sv.Request(4, 1, 3, 0.0, 5)]
    rs = sv.ResidentSet(1 << 30, build=_FakeM, footprint=lambda M: 1)
    alone = sv.serve_alone([None, None], reqs, rs, slots=3)
    rs.live[0].slot_solver.log.clear()
    live = sv.serve_live([None, None], reqs, rs, scale=0.0, slots=3)
    assert [r["id"] for r in live] == [0, 1, 2, 3, 4]
    assert sv.same_as_alone(alone, live)["differing"] == 0
    # family 0 with 3 slots: requests 0 (2 rhs) and 3 cannot both... request 0 takes slots 0, 1;
    # request 2 (2 rhs) waits; request 3 (1 rhs) waits behind it (first come first served)
    log0 = rs.live[0].slot_solver.log
    assert log0[0] == [0, 1] and log0[2] == [0, 1, 2]
    assert live[2]["admitted"] >= live[0]["finish"] - 1e-9 and live[3]["admitted"] >= live[2]["admitted"]
    s = sv.live_summary(live)
    assert s["p99_ms"] >= s["p50_ms"] >= 0 and s["solves_per_s"] > 0
    # arrivals spread over the wall clock: a request is admitted no earlier than it arrives
    late = [sv.Request(0, 0, 1, 0.0, 1), sv.Request(1, 0, 1, 2.0, 2)]
    rec = sv.serve_live([None], late, sv.ResidentSet(1 << 30, build=_FakeM, footprint=lambda M: 1), scale=0.05, slots=2)
    assert rec[1]["admitted"] >= 0.1 - 1e-3 and rec[1]["latency"] >= 0.0
