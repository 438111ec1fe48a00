"""Host logic of the multi-process server (``tenet.network.serve.MultiProcessServer``): the
dispatcher releases requests at their arrival times, every request is solved exactly once,
a worker never runs more solves than its slots or beyond its memory budget, and a worker
that fails to start is reported.  The workers here run a stand-in solver (a fixed number of
"bond steps" per request, no device) through the same worker loop as the GPU server."""

import time
from functools import partial

import pytest

from tenet.network.serve import MultiProcessServer, SolveRequest, SolveResult, _Active, footprint_bytes


class FakeServer:
    """The worker interface of ``DMRGServer`` with a host-only solve: ``n_sites // 4`` steps
    of 2 ms each, energy ``-delta * n_sites``.  Each step records how many solves the worker
    holds and their summed footprint."""

    def __init__(self, max_active, memory_budget):
        self.streams = list(range(max_active))
        self.memory_budget = memory_budget
        self.live = {}

    def _start(self, req, stream, now):
        def gen():
            for _ in range(req.n_sites // 4):
                time.sleep(0.002)
                yield len(self.live), sum(self.live.values())
            return req

        self.live[req.rid] = footprint_bytes(req)
        return _Active(req, gen(), stream, now, footprint_bytes(req))

    def _step(self, a, clock):
        try:
            n, used = next(a.gen)
            assert n <= len(self.streams) and used <= max(self.memory_budget, a.bytes)
            return None
        except StopIteration:
            del self.live[a.req.rid]
            return SolveResult(a.req.rid, -a.req.delta * a.req.n_sites, a.req.n_sites // 4, [],
                               a.req.arrival, a.admitted, clock())


def failing_factory():
    raise RuntimeError("no device")


def _reqs(n=14):
    return [SolveRequest(i, 8 + 4 * (i % 5), 0.1 * i, 16 + 16 * (i % 3), arrival=0.01 * i) for i in range(n)]


@pytest.mark.parametrize("workers,slots", [(1, 3), (3, 2)])
def test_every_request_solved_once_after_its_arrival(workers, slots):
    reqs = _reqs()
    budget = 2 * max(footprint_bytes(r) for r in reqs)
    srv = MultiProcessServer(workers=workers, max_active=workers * slots,
                             factory=partial(FakeServer, slots, budget))
    res = srv.serve(reqs, boot_timeout=120)
    assert [r.rid for r in res] == [r.rid for r in reqs]
    for q, r in zip(reqs, res):
        assert r.energy == -q.delta * q.n_sites and r.sweeps == q.n_sites // 4
        assert r.arrival == q.arrival
        # released no earlier than its arrival (small slack for the two processes' clocks)
        assert r.admitted >= q.arrival - 1e-3 and r.finished >= r.admitted
    assert srv.boot_s is not None and srv.boot_s > 0


def test_one_slot_one_worker_runs_solves_back_to_back():
    reqs = _reqs(6)
    srv = MultiProcessServer(workers=1, max_active=1, factory=partial(FakeServer, 1, 1 << 40))
    res = srv.serve(reqs, boot_timeout=120)
    spans = sorted((r.admitted, r.finished) for r in res)
    for (_, f0), (a1, _) in zip(spans, spans[1:]):
        assert a1 >= f0 - 1e-6


def test_worker_start_failure_is_reported():
    srv = MultiProcessServer(workers=2, max_active=2, factory=failing_factory)
    with pytest.raises(RuntimeError, match="failed to start"):
        srv.serve(_reqs(2), boot_timeout=120)


def test_bad_worker_counts_are_refused():
    with pytest.raises(ValueError):
        MultiProcessServer(workers=4, max_active=2)
    with pytest.raises(ValueError):
        MultiProcessServer(workers=0)


class BrokenServer(FakeServer):
    def _step(self, a, clock):
        raise ValueError("bond step failed")


def test_worker_failure_mid_stream_is_reported():
    srv = MultiProcessServer(workers=2, max_active=2, factory=partial(BrokenServer, 1, 1 << 40))
    with pytest.raises(RuntimeError, match="bond step failed"):
        srv.serve(_reqs(3), boot_timeout=120)


def test_empty_stream_returns_at_once():
    srv = MultiProcessServer(workers=2, max_active=2, factory=partial(FakeServer, 1, 1 << 40))
    t0 = time.time()
    assert srv.serve([], boot_timeout=120) == []
    assert time.time() - t0 < 30  # the workers stop on their message, not on a join timeout
