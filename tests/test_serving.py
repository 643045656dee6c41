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


def test_batched_server_groups_waiting_requests_of_a_family():
    """Requests waiting for the same family share one solve, oldest first, within the right-hand
    side cap; every request gets its own results back; an idle server waits for the next arrival."""
    from types import SimpleNamespace as NS
    reqs = [sv.Request(0, 0, 2, 0.0, 1), sv.Request(1, 1, 1, 0.1, 2), sv.Request(2, 0, 3, 0.2, 3),
            sv.Request(3, 0, 8, 0.3, 4), sv.Request(4, 1, 1, 50.0, 5)]
    calls = []

    class FakeM:
        def __init__(self, k):
            self.setup = NS(n=4)
            self.k = k

    def solve(M, B):
        calls.append((M.k, B.shape[0]))
        return [NS(iterations=10 + M.k, converged=True, true_rel_residual=1e-9) for _ in range(B.shape[0])]

    rs = sv.ResidentSet(1 << 30, build=FakeM, footprint=lambda M: 1)
    # scale 1, each solve costs almost nothing: request 0 is served alone at t=0, then at t~0.1+
    # requests 1 (family 1) ... ; with a huge scale nothing waits
    recs = sv.serve_batched([None, None], reqs, rs, scale=1.0, max_rhs=8, solve=solve)
    assert [r["id"] for r in recs] == [0, 1, 2, 3, 4]
    assert all(r["iterations"] == 10 + r["family"] for r in recs)
    # make every request wait: arrivals compressed to 0 -> one batch per family (within the cap)
    calls.clear()
    recs = sv.serve_batched([None, None], reqs, rs, scale=0.0, max_rhs=8, solve=solve)
    assert calls == [(0, 5), (1, 2), (0, 8)]
    assert [r["batch"] for r in recs] == [5, 2, 5, 8, 2]
    s = sv.batched_summary(recs)
    assert s["p99_ms"] >= s["p50_ms"] >= 0
