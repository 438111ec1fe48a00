"""Continuous-batching service for DMRG ground-state solves on one GPU.

Requests (chain length, anisotropy, bond cap, tolerances) arrive over time.  The server
keeps every admitted solve's MPS and environments resident in HBM and admits queued
requests as soon as (a) they have arrived, (b) a slot is free and (c) their estimated HBM
footprint fits the memory budget beside the solves already running.  Each active solve is
a resumable DMRG (``dmrg_steps``: one yield per bond) bound to its own HIP stream; the
scheduler round-robins bond steps across the active solves, so one solve's host-side
planning and synchronisation overlap other solves' kernels, and kernels of different
solves run concurrently on the device.  A solve's arithmetic never depends on the others
(no shared reductions, deterministic kernels), so its result is bit-identical alone and in
any mix.

``poisson_requests`` draws the synthetic request stream; ``serve`` runs it and returns the
per-request results; ``latency_report`` gives p50/p99 time-to-converged and solves/hour.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from tenet.network.dmrg import dmrg_steps
from tenet.network.models import heisenberg_mpo, random_chain

__all__ = ["SolveRequest", "SolveResult", "DMRGServer", "poisson_requests", "latency_report",
           "footprint_bytes"]


@dataclass
class SolveRequest:
    rid: int
    n_sites: int
    delta: float
    chi: int
    cutoff: float = 1e-10
    energy_tol: float = 1e-12
    schmidt_tol: float = 1e-8
    max_sweeps: int = 30
    seed: int = 0
    arrival: float = 0.0  # seconds after the stream starts


@dataclass
class SolveResult:
    rid: int
    energy: float
    sweeps: int
    history: list
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def footprint_bytes(req: SolveRequest, itemsize: int = 8) -> int:
    """Upper estimate of a solve's resident HBM: MPS (L x 2 x chi^2), both environment
    chains (L x 5 x chi^2), FP32 operator copies and two-site work buffers (~40 chi^2)."""
    chi2 = req.chi * req.chi
    return int(itemsize * (req.n_sites * (2 + 5) * chi2 + 40 * chi2))


@dataclass
class _Active:
    req: SolveRequest
    gen: object
    stream: torch.cuda.Stream
    admitted: float
    bytes: int


class DMRGServer:
    """Admission by memory footprint, one stream per active solve, bond-level interleaving."""

    def __init__(self, *, memory_budget: int | None = None, max_active: int = 8,
                 precision: str = "fp32", device="cuda") -> None:
        if memory_budget is None:
            free, _ = torch.cuda.mem_get_info()
            memory_budget = int(0.8 * free)
        self.memory_budget = memory_budget
        self.max_active = max_active
        self.precision = precision
        self.device = device
        self.streams = [torch.cuda.Stream(device=device) for _ in range(max_active)]

    def _start(self, req: SolveRequest, stream, now: float) -> _Active:
        with torch.cuda.stream(stream):
            psi = random_chain(req.n_sites, max(8, req.chi // 4), 3.0, seed=req.seed, device=self.device)
            h = heisenberg_mpo(req.n_sites, req.delta, device=self.device)
            gen = dmrg_steps(psi, h, chi=req.chi, cutoff=req.cutoff, energy_tol=req.energy_tol,
                             schmidt_tol=req.schmidt_tol, max_sweeps=req.max_sweeps,
                             precision=self.precision)
        return _Active(req, gen, stream, now, footprint_bytes(req))

    def serve(self, requests, *, clock=time.perf_counter) -> list:
        """Run a request stream to completion; arrival times are relative to the start."""
        queue = sorted(requests, key=lambda r: (r.arrival, r.rid))
        t0 = clock()
        active: list = []
        done: list = []
        free_streams = list(self.streams)
        used = 0
        qi = 0
        while qi < len(queue) or active:
            now = clock() - t0
            # admission: arrived, a slot free, footprint fits
            while (qi < len(queue) and queue[qi].arrival <= now and free_streams
                   and (not active or used + footprint_bytes(queue[qi]) <= self.memory_budget)):
                req = queue[qi]
                qi += 1
                a = self._start(req, free_streams.pop(), now)
                used += a.bytes
                active.append(a)
            if not active:
                time.sleep(max(0.0, min(0.01, queue[qi].arrival - now)))
                continue
            # one bond step of every active solve, round robin
            for a in list(active):
                with torch.cuda.stream(a.stream):
                    try:
                        next(a.gen)
                    except StopIteration as stop:
                        out = stop.value
                        a.stream.synchronize()
                        done.append(SolveResult(a.req.rid, out.energy, out.sweeps, out.history,
                                                a.req.arrival, a.admitted, clock() - t0))
                        active.remove(a)
                        free_streams.append(a.stream)
                        used -= a.bytes
        return sorted(done, key=lambda r: r.rid)


def poisson_requests(count: int, rate: float, *, sizes=(32, 64, 128), chis=(128, 256, 512, 1024),
                     seed: int = 0, **kw) -> list:
    """``count`` requests with exponential inter-arrival times of mean ``1/rate`` seconds,
    L drawn from ``sizes``, Delta ~ U[0, 2], bond cap drawn from ``chis``."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        out.append(SolveRequest(i, int(rng.choice(sizes)), float(rng.uniform(0.0, 2.0)),
                                int(rng.choice(chis)), seed=seed + 1000 + i, arrival=t, **kw))
    return out


def latency_report(results, wall: float | None = None) -> dict:
    lat = np.array([r.latency for r in results])
    span = wall if wall is not None else max(r.finished for r in results) - min(r.arrival for r in results)
    return {
        "solves": len(results),
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
        "mean_s": float(lat.mean()),
        "solves_per_hour": 3600.0 * len(results) / span if span > 0 else math.inf,
        "span_s": span,
    }
