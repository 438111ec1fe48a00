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

A solve's per-bond host work (planning, Python orchestration, the keep-rule readback) is
larger than its kernels at these sizes, so one process leaves the GPU idle most of the
time.  ``MultiProcessServer`` runs ``workers`` such servers as separate processes on the
same GPU (each with its own streams and HBM budget share): a dispatcher releases every
request into a shared queue at its arrival time and the first worker with a free slot takes
it.  Each solve runs the same kernels on the same inputs as in the single-process server,
so its energy agrees with that server's to rounding.
"""

from __future__ import annotations

import math
import multiprocessing as mp
import queue as queue_mod
import time
from dataclasses import dataclass, field, replace
from functools import partial

import numpy as np
import torch

from tenet.network.dmrg import dmrg_steps
from tenet.network.models import heisenberg_mpo, random_chain

__all__ = ["SolveRequest", "SolveResult", "DMRGServer", "poisson_requests", "latency_report",
           "footprint_bytes", "MultiProcessServer"]


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
    stream: torch.cuda.Stream | None
    admitted: float
    bytes: int
    pending: object = None  # shared mode: the solve's next LanczosRequest


class DMRGServer:
    """Admission by memory footprint, bond-level interleaving of the active solves.

    ``mode="streams"``: every active solve runs its bond steps (its own Lanczos included)
    on its own HIP stream, round robin.  ``mode="shared"``: on every scheduler tick the
    pending Lanczos steps of ALL active solves run as one batch -- one ``BatchedHeff2``
    matvec per Krylov iteration for every solve together, segmented Krylov scalars, one
    batched tridiagonal solve and the batched FP64 refinement (``BatchedLanczos``) -- and
    then each solve finishes its bond (split, environments) and hands in its next Lanczos
    step.  A solve's Krylov arithmetic per problem is the same in any mix except the K
    slicing of the batched GEMMs, so its result agrees alone and in the mix to rounding."""

    def __init__(self, *, memory_budget: int | None = None, max_active: int = 8,
                 precision: str = "fp32", device="cuda", mode: str = "streams") -> None:
        if mode not in ("streams", "shared"):
            raise ValueError(f"unknown serving mode {mode!r}")
        if memory_budget is None:
            free, _ = torch.cuda.mem_get_info()
            memory_budget = int(0.8 * free)
        self.memory_budget = memory_budget
        self.max_active = max_active
        self.precision = precision
        self.device = device
        self.mode = mode
        self.streams = [torch.cuda.Stream(device=device) for _ in range(max_active)]
        self.ticks = 0  # batched Lanczos rounds run by the shared mode

    def _start(self, req: SolveRequest, stream, now: float) -> _Active:
        with torch.cuda.stream(stream):
            psi = random_chain(req.n_sites, max(8, req.chi // 4), 3.0, seed=req.seed, device=self.device)
            h = heisenberg_mpo(req.n_sites, req.delta, device=self.device)
            gen = dmrg_steps(psi, h, chi=req.chi, cutoff=req.cutoff, energy_tol=req.energy_tol,
                             schmidt_tol=req.schmidt_tol, max_sweeps=req.max_sweeps,
                             precision=self.precision)
        return _Active(req, gen, stream, now, footprint_bytes(req))

    def _step(self, a: _Active, clock) -> SolveResult | None:
        """One bond step of ``a``; its result once the solve has converged."""
        with torch.cuda.stream(a.stream):
            try:
                next(a.gen)
                return None
            except StopIteration as stop:
                out = stop.value
                a.stream.synchronize()
                return SolveResult(a.req.rid, out.energy, out.sweeps, list(out.history),
                                   a.req.arrival, a.admitted, clock())

    def _solve_batch(self, reqs) -> list:
        """One ``BatchedLanczos`` per (operator precision, refinement) group of the pending
        Lanczos requests; returns ``(energy, vector)`` per request, in order."""
        from tenet.network.batched import BatchedHeff2
        from tenet.network.batched_lanczos import BatchedLanczos

        groups: dict = {}
        for k, r in enumerate(reqs):
            groups.setdefault((r.op.precision, r.refine is not None, r.ncv), []).append(k)
        answers = [None] * len(reqs)
        for (prec, has_ref, ncv), ks in groups.items():
            batch = BatchedHeff2(None, prec, operators=[(reqs[k].op, reqs[k].aa.structure) for k in ks])
            ref = (BatchedHeff2(None, "fp64", operators=[(reqs[k].refine, reqs[k].aa.structure) for k in ks])
                   if has_ref else None)
            vals, vecs = BatchedLanczos(batch, ncv=ncv, refine=ref, graphs=False).solve([reqs[k].aa for k in ks])
            for k, e, v in zip(ks, vals, vecs):
                answers[k] = (e, v)
        self.ticks += 1
        return answers

    def _start_shared(self, req: SolveRequest, now: float) -> _Active:
        psi = random_chain(req.n_sites, max(8, req.chi // 4), 3.0, seed=req.seed, device=self.device)
        h = heisenberg_mpo(req.n_sites, req.delta, device=self.device)
        gen = dmrg_steps(psi, h, chi=req.chi, cutoff=req.cutoff, energy_tol=req.energy_tol,
                         schmidt_tol=req.schmidt_tol, max_sweeps=req.max_sweeps,
                         precision=self.precision, external=True)
        a = _Active(req, gen, None, now, footprint_bytes(req))
        a.pending = next(gen)
        return a

    def _serve_shared(self, requests, clock, on_result) -> list:
        queue = sorted(requests, key=lambda r: (r.arrival, r.rid))
        t0 = clock()
        active, done, used, qi = [], [], 0, 0
        while qi < len(queue) or active:
            now = clock() - t0
            while (qi < len(queue) and queue[qi].arrival <= now and len(active) < self.max_active
                   and (not active or used + footprint_bytes(queue[qi]) <= self.memory_budget)):
                a = self._start_shared(queue[qi], now)
                qi += 1
                used += a.bytes
                active.append(a)
            if not active:
                time.sleep(max(0.0, min(0.01, queue[qi].arrival - now)))
                continue
            answers = self._solve_batch([a.pending for a in active])
            for a, ans in zip(list(active), answers):
                try:
                    a.pending = a.gen.send(ans)
                except StopIteration as stop:
                    out = stop.value
                    torch.cuda.synchronize()
                    res = SolveResult(a.req.rid, out.energy, out.sweeps, list(out.history),
                                      a.req.arrival, a.admitted, clock() - t0)
                    done.append(res)
                    if on_result is not None:
                        on_result(res)
                    active.remove(a)
                    used -= a.bytes
        return sorted(done, key=lambda r: r.rid)

    def serve(self, requests, *, clock=time.perf_counter, on_result=None) -> list:
        """Run a request stream to completion; arrival times are relative to the start.
        ``on_result`` is called with each ``SolveResult`` as its solve converges."""
        if self.mode == "shared":
            return self._serve_shared(requests, clock, on_result)
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
                res = self._step(a, lambda: clock() - t0)
                if res is not None:
                    done.append(res)
                    if on_result is not None:
                        on_result(res)
                    active.remove(a)
                    free_streams.append(a.stream)
                    used -= a.bytes
        return sorted(done, key=lambda r: r.rid)


def _gpu_server(max_active: int, precision: str, share: float, mode: str = "streams") -> DMRGServer:
    """A worker's server: ``share`` of the HBM free when it starts."""
    free, _ = torch.cuda.mem_get_info()
    return DMRGServer(memory_budget=int(0.8 * free * share), max_active=max_active, precision=precision,
                      mode=mode)


def _worker_main(wid: int, req_q, res_q, factory) -> None:
    """One serving process: continuous batching over the requests it takes from ``req_q``
    (``None``: no more requests).  A taken request that does not fit the memory budget
    beside the running solves is held until it does.  Time stamps are wall-clock
    (``time.time``); the dispatcher makes them relative."""
    try:
        server = factory()
    except BaseException as exc:
        res_q.put(("error", wid, f"failed to start: {exc!r}"))
        raise
    res_q.put(("ready", wid, None))
    try:
        _worker_loop(wid, server, req_q, res_q)
    except BaseException as exc:
        res_q.put(("error", wid, repr(exc)))
        raise
    res_q.put(("done", wid, None))


def _worker_loop(wid: int, server, req_q, res_q) -> None:
    active, free_streams, used = [], list(server.streams), 0
    held, closing = None, False
    while True:
        while free_streams and not closing:
            if held is None:
                try:
                    item = req_q.get_nowait() if active else req_q.get(timeout=0.05)
                except queue_mod.Empty:
                    break
                if item is None:
                    closing = True
                    break
                held = item
            if active and used + footprint_bytes(held) > server.memory_budget:
                break
            a = server._start(held, free_streams.pop(), time.time())
            held = None
            used += a.bytes
            active.append(a)
        if not active:
            if closing and held is None:
                break
            continue
        for a in list(active):
            res = server._step(a, time.time)
            if res is not None:
                res_q.put(("result", wid, res))
                active.remove(a)
                free_streams.append(a.stream)
                used -= a.bytes


class MultiProcessServer:
    """``workers`` serving processes on the one GPU with ``max_active`` solves in flight in
    total (``max_active // workers`` each) and an equal share of the free HBM each.

    Workers are spawned (a fresh interpreter each, never a fork of this process) and boot
    before the request stream's clock starts; ``boot_s`` is how long that took.  ``serve``
    releases each request into the shared queue at its arrival time; the first worker with a
    free slot takes it.  ``factory`` (a picklable callable returning an object with the
    ``DMRGServer`` worker interface) replaces the GPU server in host-only tests."""

    def __init__(self, *, workers: int = 4, max_active: int = 8, precision: str = "fp32",
                 factory=None, mode: str = "streams") -> None:
        """``mode``: each worker's ``DMRGServer`` mode (``"shared"``: its active solves'
        Lanczos steps run as one batched Krylov solve per tick)."""
        if workers < 1 or max_active < workers:
            raise ValueError(f"need 1 <= workers <= max_active, got {workers}, {max_active}")
        self.workers = workers
        self.factory = factory or partial(_gpu_server, max_active // workers, precision, 1.0 / workers, mode)
        self.boot_s = None

    def serve(self, requests, *, boot_timeout: float = 600.0, on_result=None) -> list:
        ctx = mp.get_context("spawn")
        req_q, res_q = ctx.Queue(), ctx.Queue()
        procs = [ctx.Process(target=_worker_main, args=(w, req_q, res_q, self.factory), daemon=True)
                 for w in range(self.workers)]
        tb = time.time()
        for p in procs:
            p.start()
        try:
            ready = 0
            while ready < len(procs):
                try:
                    tag, wid, val = res_q.get(timeout=1.0)
                except queue_mod.Empty:
                    dead = [w for w, p in enumerate(procs) if p.exitcode is not None]
                    if dead:
                        raise RuntimeError(f"serving worker(s) {dead} exited while starting "
                                           f"(exit codes {[procs[w].exitcode for w in dead]})")
                    if time.time() - tb > boot_timeout:
                        raise RuntimeError(f"serving workers not ready after {boot_timeout:.0f} s")
                    continue
                if tag != "ready":
                    raise RuntimeError(f"serving worker {wid}: {val}")
                ready += 1
            t0 = time.time()
            self.boot_s = t0 - tb
            pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
            results, qi, closed = [], 0, False
            while len(results) < len(pending):
                now = time.time() - t0
                while qi < len(pending) and pending[qi].arrival <= now:
                    req_q.put(pending[qi])
                    qi += 1
                if qi == len(pending) and not closed:
                    for _ in procs:
                        req_q.put(None)  # after every request: a worker that takes it has no more to do
                    closed = True
                wait = pending[qi].arrival - now if qi < len(pending) else 1.0
                try:
                    tag, wid, val = res_q.get(timeout=max(0.001, min(wait, 1.0)))
                except queue_mod.Empty:
                    if not any(p.is_alive() for p in procs):
                        raise RuntimeError("every serving worker exited before the stream finished")
                    continue
                if tag == "error":
                    raise RuntimeError(f"serving worker {wid} failed: {val}")
                if tag == "result":
                    results.append(replace(val, admitted=val.admitted - t0, finished=val.finished - t0))
                    if on_result is not None:
                        on_result(results[-1])
            if not closed:  # an empty stream: the workers still need their stop message
                for _ in procs:
                    req_q.put(None)
            for p in procs:
                p.join(timeout=60)
            return sorted(results, key=lambda r: r.rid)
        finally:
            for p in procs:
                if p.is_alive():
                    p.terminate()
                    p.join(timeout=10)


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
