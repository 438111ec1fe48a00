"""Serving: continuous batching of DMRG solves with memory-based admission.  Every solve's
result is bit-identical alone and in a mix (independent streams, deterministic kernels),
and energies match exact diagonalization."""

import pytest

from oracle import ed

pytestmark = pytest.mark.gpu


def _reqs():
    from tenet.network.serve import SolveRequest

    return [
        SolveRequest(0, 10, 1.0, 24, seed=11, arrival=0.0),
        SolveRequest(1, 12, 0.4, 32, seed=12, arrival=0.0),
        SolveRequest(2, 8, 1.7, 16, seed=13, arrival=0.05),
    ]


def test_results_identical_alone_and_in_the_mix():
    from tenet.network.serve import DMRGServer

    mix = DMRGServer(max_active=3).serve(_reqs())
    for r in _reqs():
        alone = DMRGServer(max_active=1).serve([r])[0]
        got = mix[r.rid]
        assert got.energy == alone.energy
        assert got.history == alone.history
        assert abs(got.energy - ed.xxz_sz0_energy(r.n_sites, r.delta)) < 1e-8 * abs(got.energy)


def test_admission_respects_the_memory_budget():
    from tenet.network.serve import DMRGServer, footprint_bytes

    reqs = _reqs()
    budget = max(footprint_bytes(r) for r in reqs)  # room for one solve at a time
    res = DMRGServer(max_active=3, memory_budget=budget).serve(reqs)
    spans = sorted((r.admitted, r.finished) for r in res)
    for (a0, f0), (a1, f1) in zip(spans, spans[1:]):
        assert a1 >= f0 - 1e-9  # never two solves resident together


def test_multiprocess_server_matches_the_in_process_server():
    from tenet.network.serve import DMRGServer, MultiProcessServer

    srv = MultiProcessServer(workers=2, max_active=4)
    got = srv.serve(_reqs(), boot_timeout=300)
    ref = DMRGServer(max_active=3).serve(_reqs())
    assert [r.rid for r in got] == [r.rid for r in ref]
    for g, r in zip(got, ref):
        assert g.sweeps == r.sweeps
        assert abs(g.energy - r.energy) <= 1e-12 * abs(r.energy)
        assert g.admitted >= g.arrival - 1e-3 and g.finished > g.admitted


def test_shared_batches_agree_alone_and_in_the_mix():
    """The shared mode runs the pending Lanczos steps of every active solve as one batched
    Krylov solve per tick; each solve's energy agrees alone and in the mix (the batched
    GEMMs may slice K differently, so agreement is to rounding, not bitwise) and matches
    exact diagonalization."""
    from tenet.network.serve import DMRGServer

    srv = DMRGServer(max_active=3, mode="shared")
    mix = srv.serve(_reqs())
    alone_ticks = 0
    for r in _reqs():
        solo = DMRGServer(max_active=1, mode="shared")
        alone = solo.serve([r])[0]
        alone_ticks += solo.ticks
        got = mix[r.rid]
        assert abs(got.energy - alone.energy) <= 1e-10 * abs(alone.energy)
        assert abs(got.energy - ed.xxz_sz0_energy(r.n_sites, r.delta)) < 1e-8 * abs(got.energy)
    assert srv.ticks < alone_ticks  # bonds of different solves did share batches
