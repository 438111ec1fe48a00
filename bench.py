"""Benchmark: two-site H_eff matvecs/s of the U(1) Heisenberg chain (L=64, D=2048 at the
centre bond, Gaussian sector profile sigma=6, MPO bond 5, physical dim 2) on one MI355X.

One step = one H_eff matvec on the centre bond (all four contractions), inputs resident
in HBM.  Arms timed on the same operands:

* ``fp32``       -- the engine with FP32 MFMA GEMMs (the headline ``value``)
* ``bf16x3``     -- the engine with split-BF16x3 MFMA GEMMs (FP32 accumulation)
* ``fp64``       -- the engine with FP64 MFMA GEMMs
* ``fp64_blocks``-- the plain FP64 per-block GPU path (the reference's lowering run with
  one torch call per block/sector)
* ``cpu_baseline`` -- the oracle's NumPy FP64 restatement on the host cores

Energy check: a restarted FP64-Krylov Lanczos on this H_eff with FP32 (and BF16x3)
matvecs plus a closing FP64 Rayleigh quotient, against the same solve with FP64
matvecs; ``dE`` is the relative difference.

Usage: ``python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]``; prints ONE
JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "tenet-py_amd"))

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6, "bf16x3": 2500.0}

# HBM traffic per launch of the H_eff GEMM kernels (keys: the events' step numbers), read
# from `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` (KiB per dispatch) of
# `bench.py --steps 20 --warmup 3 --no-cpu --no-blocks --no-energy --arms fp32`, and
# scaled by what the same counters read for `bench.py --workload copy` (a 1 GiB device
# copy: FETCH_SIZE 524302 KiB for 1048576 KiB read, WRITE_SIZE 1048576 KiB written).
PMC_TRAFFIC = {
    (64, 2048, "fp32"): {"fetch_kib": {0: 32795.0, 3: 37416.6}, "write_kib": {0: 30349.4, 3: 24009.3},
                         "fetch_scale": 0.5, "write_scale": 1.0,
                         "source": "profiles/heff2_L64_D2048_fp32_traffic.md"},
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_arm(fn, steps: int, warmup: int):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def restarted_lanczos(op, aa, refine=None, max_restarts: int = 40, ncv: int = 8, tol: float = 1e-14):
    """Thick-free restarted Lanczos to convergence of the Ritz value, each restart replaying
    the captured hipGraphs of its iterations; returns (E, matvecs)."""
    from tenet.network.graph import KrylovGraph

    kg = KrylovGraph(op, aa.structure, ncv=ncv)
    v, e, used = aa, None, 0
    for _ in range(max_restarts):
        e_new, v = kg.solve(v, refine=refine)
        used += ncv + (2 if refine else 0)
        if e is not None and abs(e_new - e) <= tol * max(1.0, abs(e_new)):
            e = e_new
            break
        e = e_new
    return e, used


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _median_reps(fn, reps: int, budget_s: float):
    """One warm-up call, then ``reps`` timed calls (fewer, but at least 5, when the warm-up
    says ``reps`` would exceed ``budget_s``); returns (median seconds, reps, warm-up out)."""
    import statistics

    t0 = time.perf_counter()
    out = fn()
    first = time.perf_counter() - t0
    n = reps if first * reps <= budget_s else max(5, int(budget_s / max(first, 1e-6)))
    times = []
    for _ in range(n):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times), n, out


def cpu_baseline(fl, w1, w2, fr, aa, reps: int = 20, budget_s: float = 40.0):
    """The oracle's NumPy FP64 H_eff on the host cores, two arms: every core of this
    process's affinity set (BLAS threads = that count) and one core (BLAS threads = 1),
    each a warm-up and the median of ``reps`` whole matvecs."""
    from threadpoolctl import threadpool_limits

    sys.path.insert(0, str(ROOT))
    from oracle import u1_network as N
    from tests.helpers import to_oracle

    cores = len(os.sched_getaffinity(0))
    ofl, ow1, ow2, ofr, oaa = (to_oracle(t) for t in (fl, w1, w2, fr, aa))
    fn = lambda: N.heff2(ofl, ow1, ow2, ofr, oaa)  # noqa: E731
    with threadpool_limits(limits=cores):
        t_all, n_all, out = _median_reps(fn, reps, budget_s)
    with threadpool_limits(limits=1):
        t_one, n_one, _ = _median_reps(fn, reps, budget_s)
    model = _cpu_model()
    return {"value": 1.0 / t_all, "unit": "matvecs/s", "cores": cores, "kind": "port",
            "cpu_model": model,
            "sample": f"median of {n_all} whole H_eff matvecs after 1 warm-up, NumPy float64, "
                      f"BLAS threads = {cores} (threadpoolctl)",
            "single_core": {"value": 1.0 / t_one, "unit": "matvecs/s", "cores": 1,
                            "sample": f"median of {n_one} matvecs after 1 warm-up, BLAS threads = 1"}}, out.flat()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--sites", type=int, default=64)
    ap.add_argument("--bond", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-blocks", action="store_true")
    ap.add_argument("--no-energy", action="store_true")
    ap.add_argument("--arms", default="fp32,bf16x3,fp64")
    ap.add_argument("--workload", default="heff2",
                    choices=["heff2", "einsum", "batched", "dmrg", "serve", "su2", "copy"])
    ap.add_argument("--count", type=int, default=None)
    ap.add_argument("--chi", type=int, default=128)
    ap.add_argument("--rate", type=float, default=2.0)
    ap.add_argument("--max-sweeps", type=int, default=30)
    ap.add_argument("--serve-arms", default="mp,single,serial",
                    help="serve workload: mp (worker processes), mp_shared (worker processes, each batching "
                         "its solves' Krylov steps), shared (one process, batched Krylov steps), single "
                         "(one process, a stream per solve), serial (one-at-a-time FP64)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the H_eff of one bond runs on one device)",
              file=sys.stderr)
        return 2

    if args.workload == "einsum":
        return bench_einsum(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "su2":
        return bench_su2(args)
    if args.count is None:
        args.count = 12 if args.workload == "serve" else 256
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "dmrg":
        return bench_dmrg(args)
    if args.workload == "serve":
        return bench_serve(args)
    return bench_heff2(args)


def bench_heff2(args) -> int:
    import numpy as np
    import torch

    import tenet
    from tenet.baseline_blocks import BlockTensor, heff2 as blocks_heff2
    from tenet.network import Heff2
    from tenet.network.graph import GraphedMatvec
    from tenet.network.workloads import heff2_problem

    dev = torch.device("cuda")
    log(f"building the L={args.sites} D={args.bond} problem")
    t0 = time.perf_counter()
    fl, w1, w2, fr, aa = heff2_problem(args.sites, args.bond, 6.0, seed=0, device=dev)
    torch.cuda.synchronize()
    log(f"built in {time.perf_counter() - t0:.1f}s: aa {aa.structure.size} elements, "
        f"{aa.structure.num_blocks} blocks; FL {fl.structure.size}; FR {fr.structure.size}")

    arms = [a for a in args.arms.split(",") if a]
    ops_ = {p: Heff2(fl, w1, w2, fr, p) for p in set(arms) | {"fp64"}}
    flops = ops_["fp64"].flops(aa.structure)
    results = {}
    outs = {}
    for p in arms:
        op = ops_[p]
        op.apply_data(aa.structure, aa.data)  # compile
        torch.cuda.synchronize()
        dt = time_arm(lambda: op.apply_data(aa.structure, aa.data), args.steps, args.warmup)
        outs[p] = op.apply_data(aa.structure, aa.data)
        results[p] = {"matvecs_per_s": 1.0 / dt, "ms": dt * 1e3,
                      "effective_tflops": flops / dt / 1e12}
        log(f"arm {p}: {dt * 1e3:.3f} ms/matvec, {1 / dt:.1f} matvecs/s")

    # accuracy of each arm's single matvec against the FP64 engine
    ref = outs.get("fp64", ops_["fp64"].apply_data(aa.structure, aa.data))
    for p in arms:
        d = (outs[p] - ref).norm() / ref.norm()
        results[p]["matvec_rel_err_vs_fp64"] = float(d)

    if not args.no_blocks:
        bfl, bw1, bw2, bfr, baa = (BlockTensor.from_symmetric(t) for t in (fl, w1, w2, fr, aa))
        fn = lambda: blocks_heff2(bfl, bw1, bw2, bfr, baa)
        dt = time_arm(fn, max(3, args.steps // 10), 1)
        o = fn()
        got = torch.cat([b.reshape(-1) for b in o.blocks])
        results["fp64_blocks"] = {"matvecs_per_s": 1.0 / dt, "ms": dt * 1e3,
                                  "matvec_rel_err_vs_fp64": float((got - ref).norm() / ref.norm())}
        log(f"arm fp64_blocks: {dt * 1e3:.3f} ms/matvec")

    # GEMM kernel timing over a timed pass of an arm: HIP events recorded on the stream
    # the kernels are launched on, around each launch
    aux = ("recouple", "reduce", "entry", "exit")

    def kernel_times(p):
        ev: dict = {}
        for _ in range(args.steps):
            ops_[p].apply_data(aa.structure, aa.data, events=ev)
        torch.cuda.synchronize()
        per_step = {k: (sum(e0.elapsed_time(e1) for e0, e1, _ in v) / len(v), v[0][2]) for k, v in ev.items()
                    if v and not str(k).startswith(aux)}
        reduce_ms = {str(k): sum(e0.elapsed_time(e1) for e0, e1, _ in v) / len(v) for k, v in ev.items()
                     if v and str(k).startswith(aux)}
        big = [k for k in per_step if per_step[k][1] >= 0.05 * flops]
        return per_step, reduce_ms, big, sum(per_step[k][0] for k in big), sum(per_step[k][1] for k in big)

    head = "fp32" if "fp32" in arms else arms[0]
    per_step, reduce_ms, big, gemm_ms, gemm_flops = kernel_times(head)
    achieved = gemm_flops / (gemm_ms * 1e-3) / 1e12
    pmc = PMC_TRAFFIC.get((args.sites, args.bond, head))
    traffic = None
    if pmc is not None:
        traffic = sum(pmc["fetch_kib"][k] / pmc["fetch_scale"] + pmc["write_kib"][k] / pmc["write_scale"]
                      for k in big) * 1024.0
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s",
                "frac": achieved / PEAK_TFLOPS[head], "traffic": traffic,
                "kernel": f"gemm_pipe_kernel ({head}; step 1 through tn_gemm_grouped_rows, step 3 through "
                          "tn_gemm_grouped): the two H_eff GEMM launches of one "
                          f"matvec, steps {sorted(big)}; achieved = their flops / their summed average "
                          "kernel durations, traffic = their summed HBM bytes",
                "avg_launch_ms": {str(k): per_step[k][0] for k in per_step},
                "flops_per_launch": {str(k): per_step[k][1] for k in per_step},
                # the passes around the GEMMs: exit (split-K slice sum + scatter into the
                # vector's blocks + cast; the FP32 step 1 reads the vector's blocks itself)
                "aux_passes_avg_ms": reduce_ms,
                "traffic_source": pmc["source"] if pmc else None}

    # the split-BF16x3 arm against its own peak: each FP32 product is three BF16 MFMA
    # products (hi*hi + hi*lo + lo*hi), so the BF16 rate is 3x the FP32-equivalent one
    roofline_bf16 = None
    if "bf16x3" in arms:
        b_steps, _, b_big, b_ms, b_flops = kernel_times("bf16x3")
        eq = b_flops / (b_ms * 1e-3) / 1e12
        roofline_bf16 = {"bound": "mfma", "achieved": 3 * eq, "peak": PEAK_TFLOPS["bf16x3"], "unit": "TFLOP/s",
                         "frac": 3 * eq / PEAK_TFLOPS["bf16x3"],
                         "fp32_equivalent_tflops": eq, "fp32_equivalent_frac_of_fp32_peak": eq / PEAK_TFLOPS["fp32"],
                         "avg_launch_ms": {str(k): b_steps[k][0] for k in b_big},
                         "kernel": "gemm_pipe_kernel (bf16x3): the two H_eff GEMM launches; "
                                   "achieved counts 3 BF16 products per FP32 flop"}

    # the headline arm's matvec captured as a hipGraph and replayed on resident buffers
    gm = GraphedMatvec(ops_[head], aa.structure)
    gm.x.copy_(aa.data)
    dt = time_arm(gm.graph.replay, args.steps, args.warmup)
    gerr = float((gm.y - ref).norm() / ref.norm())
    results[f"{head}_graph"] = {"matvecs_per_s": 1.0 / dt, "ms": dt * 1e3,
                                "effective_tflops": flops / dt / 1e12, "matvec_rel_err_vs_fp64": gerr}
    log(f"arm {head}_graph: {dt * 1e3:.3f} ms/matvec, {1 / dt:.1f} matvecs/s")

    # energy agreement: FP32/BF16x3 Krylov + FP64 Rayleigh quotient vs all-FP64, each a
    # restarted Lanczos whose iterations replay captured hipGraphs
    energy = {}
    if not args.no_energy:
        e64, n64 = restarted_lanczos(ops_["fp64"], aa)
        hv = ops_["fp64"](aa)
        scale = float(tenet.norm(hv)) / float(tenet.norm(aa))  # a lower bound of |H|
        energy = {"E_fp64": e64, "matvecs_fp64": n64, "operator_scale": scale}
        for p in arms:
            if p == "fp64":
                continue
            ep, npv = restarted_lanczos(ops_[p], aa, refine=ops_["fp64"])
            energy[f"E_{p}"] = ep
            energy[f"matvecs_{p}"] = npv
            energy[f"dE_rel_{p}"] = abs(ep - e64) / abs(e64)
            energy[f"dE_rel_scale_{p}"] = abs(ep - e64) / scale
    log(f"energies {energy}")

    cpu = None
    if not args.no_cpu:
        log("timing the CPU baseline (oracle, NumPy FP64)")
        cpu, ref_cpu = cpu_baseline(fl, w1, w2, fr, aa)
        got64 = ref.to("cpu").numpy()
        results["fp64"]["rel_err_vs_oracle"] = float(np.linalg.norm(got64 - ref_cpu) / np.linalg.norm(ref_cpu))
        log(f"cpu baseline {cpu['value']:.3f} matvecs/s ({cpu['cores']} cores), "
            f"{cpu['single_core']['value']:.3f} matvecs/s (1 core), {cpu['cpu_model']}")

    # the headline is the faster of the two launch modes of the same kernels
    best = f"{head}_graph" if results[f"{head}_graph"]["matvecs_per_s"] > results[head]["matvecs_per_s"] else head
    v = results[best]["matvecs_per_s"]
    line = {
        "metric": "H_eff matvecs/s FP32-MFMA vs FP64 at |dE|<1e-8",
        "value": v,
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": results[best]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if head != "fp64" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"two-site H_eff, U(1) Heisenberg L={args.sites}, D={args.bond} centre bond "
                        f"(22 odd-charge sectors, Gaussian sigma=6), MPO bond 5, phys dim 2",
            "sites": "random N(0,1) MPS sites, each normalised (not canonized: the operator's block "
                     "structure and flops are those of a canonized chain, only the values differ), "
                     "environments folded from them and normalised",
            "aa_elements": aa.structure.size,
            "flops_per_matvec": flops,
            "launch": "hipGraph replay of the captured matvec" if best != head else "eager launches",
        },
        "arms": results,
        "speedup_fp32_vs_fp64_blocks": (v / results["fp64_blocks"]["matvecs_per_s"]) if "fp64_blocks" in results else None,
        "speedup_fp32_vs_cpu": (v / cpu["value"]) if cpu else None,
        "energy": energy,
        "roofline": roofline,
        "roofline_bf16x3": roofline_bf16,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_su2(args) -> int:
    """SU(2) Heisenberg chain, L=64, 512 multiplets per bond (2j <= 12 even / 13 odd,
    Gaussian profile sigma=6 in 2j): the generic H_eff (four contractions with their
    multi-term recoupling launches) at the centre bond per arithmetic, the graph-captured
    FP32 Lanczos energy against FP64, and full two-site DMRG sweeps FP32 vs FP64."""
    import torch

    from tenet import ops
    from tenet.network import Heff2, dmrg_
    from tenet.network.graph import KrylovGraph
    from tenet.network.models import su2_chain_bonds, su2_heisenberg_mpo, SU2_PHYS
    from tenet.network.mps import MPS
    from tenet.network.workloads import su2_heff2_problem

    t0 = time.perf_counter()
    fl, w1, w2, fr, aa = su2_heff2_problem(args.sites, 512, 6.0, seed=0, device="cuda")
    torch.cuda.synchronize()
    dense_dim = sum(m * (a.two_j + 1) for a, m in aa.legs[0].space.sectors)
    log(f"SU(2) problem built in {time.perf_counter() - t0:.1f}s: aa {aa.structure.size} elements, "
        f"{aa.structure.num_blocks} blocks; left bond {dense_dim} dense-equivalent")
    arms = [a for a in args.arms.split(",") if a in ("fp64", "fp32", "bf16x3")]
    ops_ = {p: Heff2(fl, w1, w2, fr, p) for p in set(arms) | {"fp64"}}
    flops = ops_["fp64"].flops(aa.structure)
    res, outs = {}, {}
    for p in arms:
        op = ops_[p]
        t0 = time.perf_counter()
        op.apply_data(aa.structure, aa.data)
        torch.cuda.synchronize()
        compile_s = time.perf_counter() - t0
        dt = time_arm(lambda: op.apply_data(aa.structure, aa.data), args.steps, args.warmup)
        outs[p] = op.apply_data(aa.structure, aa.data)
        res[p] = {"matvecs_per_s": 1 / dt, "ms": dt * 1e3, "effective_tflops": flops / dt / 1e12,
                  "first_call_s": compile_s}
        log(f"su2 {p}: {dt * 1e3:.3f} ms/matvec (first call {compile_s:.2f}s)")
    ref = outs.get("fp64", ops_["fp64"].apply_data(aa.structure, aa.data))
    for p in arms:
        res[p]["matvec_rel_err_vs_fp64"] = float((outs[p] - ref).norm() / ref.norm())
    # recouple kernel (multi-term SU(2) bends) and GEMM launch timing inside one pass
    head = "fp32" if "fp32" in arms else arms[0]
    # the same matvec replayed from one captured hipGraph: the generic chain is a sequence of
    # short launches, so the host's launch cost is a visible part of an eager matvec
    from tenet.network.graph import GraphedMatvec

    gm = GraphedMatvec(ops_[head], aa.structure)
    gm.x.copy_(aa.data)
    dt = time_arm(gm.graph.replay, args.steps, args.warmup)
    res[f"{head}_graph"] = {"matvecs_per_s": 1 / dt, "ms": dt * 1e3, "effective_tflops": flops / dt / 1e12,
                            "matvec_rel_err_vs_fp64": float((gm.y - ref).norm() / ref.norm())}
    log(f"su2 {head}_graph: {dt * 1e3:.3f} ms/matvec")
    launch = "graph replay" if res[f"{head}_graph"]["ms"] < res[head]["ms"] else "eager launches"
    ev: dict = {}
    for _ in range(args.steps):
        ops_[head].apply_data(aa.structure, aa.data, events=ev)
    torch.cuda.synchronize()
    kern = {}
    for k, v in ev.items():
        if not v:
            continue
        ms = sum(e0.elapsed_time(e1) for e0, e1, _ in v) / len(v)
        kern[str(k)] = {"avg_ms": ms, "work": v[0][2]}
    rec = {k: x for k, x in kern.items() if k.startswith("recouple")}
    gem = {k: x for k, x in kern.items() if not k.startswith(("recouple", "reduce"))}
    rec_bytes = sum(x["work"] for x in rec.values())
    rec_ms = sum(x["avg_ms"] for x in rec.values())
    gemm_flops = sum(x["work"] for x in gem.values())
    gemm_ms = sum(x["avg_ms"] for x in gem.values())
    # energies: graph-captured Lanczos restarted to convergence, FP32 + FP64 Rayleigh quotient
    energy = {}
    for p in ("fp64", head):
        kg = KrylovGraph(ops_[p], aa.structure, ncv=8)
        v, e = aa, None
        for it in range(60):
            e_new, v = kg.solve(v, refine=ops_["fp64"] if p != "fp64" else None)
            if e is not None and abs(e_new - e) <= 1e-14 * abs(e_new):
                e = e_new
                break
            e = e_new
        energy[f"E_{p}"] = e
        energy[f"restarts_{p}"] = it + 1
    energy[f"dE_rel_{head}"] = abs(energy[f"E_{head}"] - energy["E_fp64"]) / abs(energy["E_fp64"])
    # full two-site sweeps from one random start, FP64 and the headline arithmetic
    sweeps = {}
    chi = 2858
    for p in (("fp64", head) if args.max_sweeps > 0 else ()):
        psi = MPS.random(SU2_PHYS, su2_chain_bonds(args.sites, 512, 6.0, max_two_j=12), seed=5, device="cuda")
        times, last = [], [time.perf_counter()]

        def cb(out):
            torch.cuda.synchronize()
            now = time.perf_counter()
            times.append(now - last[0])
            last[0] = now
            log(f"su2 {p} sweep {out.sweeps}: E={out.energy:.12f} ({times[-1]:.1f}s)")

        out = dmrg_(psi, su2_heisenberg_mpo(args.sites, device="cuda"), chi=chi,
                    max_sweeps=args.max_sweeps, precision=p, callback=cb, polish_tol=0.0)
        ref = MPSKIT.get(args.sites, {}).get(256)
        sweeps[p] = {"energy": out.energy, "sweeps": out.sweeps, "s_per_sweep": times,
                     "max_discarded_weight": out.max_discarded_weight,
                     "mpskit_chi256_energy": ref, "dE_vs_mpskit": abs(out.energy - ref) if ref else None}
    if sweeps:
        sweeps["dE_rel"] = abs(sweeps[head]["energy"] - sweeps["fp64"]["energy"]) / abs(sweeps["fp64"]["energy"])
    best = f"{head}_graph" if launch == "graph replay" else head
    line = {"metric": "SU(2) H_eff matvecs/s FP32-MFMA vs FP64", "value": res[best]["matvecs_per_s"],
            "unit": "matvecs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": res[best]["ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32" if head != "fp64" else "f64", "data": "synthetic",
            "config": {"workload": f"SU(2) Heisenberg L={args.sites}, 512 multiplets per bond "
                                   f"(2j<=12 even, <=13 odd, sigma=6), centre-bond two-site H_eff",
                       "aa_elements": aa.structure.size, "flops_per_matvec": flops,
                       "dense_equivalent_bond": dense_dim, "launch": launch},
            "arms": res, "kernels": kern, "energy": energy, "sweeps": sweeps,
            "roofline": {"bound": "mfma", "achieved": gemm_flops / (gemm_ms * 1e-3) / 1e12,
                         "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s",
                         "frac": gemm_flops / (gemm_ms * 1e-3) / 1e12 / PEAK_TFLOPS[head], "traffic": None},
            "recouple_roofline": {"bound": "hbm", "achieved": (rec_bytes / (rec_ms * 1e-3) / 1e9) if rec_ms else None,
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": (rec_bytes / (rec_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if rec_ms else None}}
    print(json.dumps(line), flush=True)
    return 0


def bench_copy(args) -> int:
    """A plain device copy of 1 GiB (FP64), the calibration kernel of the HBM-traffic
    counters: its bytes read and written are known exactly."""
    import torch

    n = (1 << 30) // 8
    src = torch.randn(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dt = time_arm(lambda: dst.copy_(src), args.steps, args.warmup)
    gbs = 2 * n * 8 / dt / 1e9
    log(f"copy 1 GiB: {dt * 1e3:.3f} ms, {gbs:.0f} GB/s")
    line = {"metric": "device copy GB/s (counter calibration)", "value": gbs, "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": {"workload": "copy of 1 GiB (134217728 float64)",
                                            "bytes_read": n * 8, "bytes_written": n * 8}}
    print(json.dumps(line), flush=True)
    return 0


def bench_einsum(args) -> int:
    """abc,cde->abde with q in {-2..2}, D_q = {4,19,32,19,4}: HBM-bound (mostly C writes)."""
    import numpy as np
    import torch

    import tenet
    from tenet import IN, OUT, GradedSpace, Leg
    from tenet.engine import contraction_exec
    from tenet.symmetry import U1, U1Sector

    V = GradedSpace.new(U1, {U1Sector(q): int(round(32 * np.exp(-q * q / 2))) for q in range(-2, 3)})
    a = tenet.SymmetricTensor.random((Leg(V, OUT), Leg(V, OUT), Leg(V, IN)), seed=0, device="cuda")
    b = tenet.SymmetricTensor.random((Leg(V, OUT), Leg(V, IN), Leg(V, IN)), seed=1, device="cuda")
    results = {}
    for p in args.arms.split(","):
        x, y = (a, b) if p == "fp64" else (a.to(dtype=torch.float32), b.to(dtype=torch.float32))
        if p != "fp64":
            tenet.set_gemm_precision(p)
        fn = lambda: tenet.einsum("abc,cde->abde", x, y)
        dt = time_arm(fn, args.steps, args.warmup)
        ex = contraction_exec(x.structure, y.structure, ((2,), (0,)), (0, 1, 2, 3))
        isz = 8 if p == "fp64" else 4
        nbytes = (x.structure.size + y.structure.size + ex.out_structure.size) * isz
        results[p] = {"einsums_per_s": 1 / dt, "ms": dt * 1e3, "GB_per_s": nbytes / dt / 1e9,
                      "hbm_frac": nbytes / dt / 1e9 / HBM_PEAK_GBS, "flops": ex.flops}
        log(f"einsum {p}: {dt * 1e3:.3f} ms, {nbytes / dt / 1e9:.0f} GB/s")
    tenet.set_gemm_precision("fp32")
    head = "fp64" if "fp64" in results else next(iter(results))
    line = {"metric": "einsum abc,cde->abde per second", "value": results[head]["einsums_per_s"],
            "unit": "einsums/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": results[head]["ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64" if head == "fp64" else "fp32", "data": "synthetic",
            "config": {"workload": "U(1) einsum abc,cde->abde, q in {-2..2}, D_q={4,19,32,19,4}"},
            "arms": results,
            "roofline": {"bound": "hbm", "achieved": results[head]["GB_per_s"], "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": results[head]["hbm_frac"], "traffic": None}}
    print(json.dumps(line), flush=True)
    return 0


def bench_batched(args) -> int:
    """256 independent U(1) XXZ two-site H_eff problems (Delta ~ U[0,2], D log-uniform in
    [64, 1024]) in one launch per step; autotuned tiles against a fixed 64x64 tile."""
    import numpy as np
    import torch

    from tenet import autotune
    from tenet.engine import GemmProgram
    from tenet.network.batched import BatchedHeff2, batch_workload

    t0 = time.perf_counter()
    insts, dims, deltas = batch_workload(args.count, device="cuda")
    log(f"built {len(insts)} problems in {time.perf_counter() - t0:.1f}s (D median {int(np.median(dims))})")
    results = {}
    for p in [a for a in args.arms.split(",") if a in ("fp32", "bf16x3", "fp64")]:
        t0 = time.perf_counter()
        bh = BatchedHeff2(insts, p)
        log(f"{p}: planned in {time.perf_counter() - t0:.1f}s, {bh.flops / 1e9:.2f} GF per batched matvec")
        x = bh.pack([i[4] for i in insts])
        fixed = {"gemm1": (64, 64), "gemm4": (64, 64)}
        dt_fixed = time_arm(lambda: bh.apply(x, fixed), args.steps, args.warmup)
        t1 = torch.empty(bh.t1_size, dtype=x.dtype, device=x.device)
        # step 1 of the FP32 / BF16x3 sector layout reads the arena in place (row tables)
        g1 = bh.gemm1_rows if (p != "fp64" and getattr(bh, "gemm1_rows", None) is not None) else bh.gemm1
        best1 = autotune.tune(g1, x, bh.fr, t1, p,
                              candidates=GemmProgram.ROWS_TILES if g1 is not bh.gemm1 else autotune._TILES)
        t3 = torch.rand(bh.t3_size, dtype=x.dtype, device=x.device)
        # the sector layout's step-3 GEMM writes its K slices to a workspace
        o = torch.empty(max(bh.gemm4.ws_size, x.numel()), dtype=x.dtype, device=x.device)
        best4 = autotune.tune(bh.gemm4, bh.fl, t3, o, p)
        tuned = {"gemm1": best1, "gemm4": best4}
        dt_tuned = time_arm(lambda: bh.apply(x, tuned), args.steps, args.warmup)
        results[p] = {"batched_matvecs_per_s": 1 / dt_tuned, "ms": dt_tuned * 1e3,
                      "problem_matvecs_per_s": len(insts) / dt_tuned,
                      "effective_tflops": bh.flops / dt_tuned / 1e12,
                      "fixed_tile_ms": dt_fixed * 1e3, "tuned_tiles": {k: list(v) for k, v in tuned.items()},
                      "speedup_tuned_vs_fixed": dt_fixed / dt_tuned, "flops": bh.flops}
        log(f"{p}: fixed {dt_fixed * 1e3:.3f} ms, tuned {dt_tuned * 1e3:.3f} ms ({tuned})")
        # the Lanczos step of every problem at once: ncv graph-replayed iterations (one
        # batched matvec each, segmented FP64 scalars) and the batched tridiagonal finish;
        # the reduced-precision arms close with the batched FP64 residual refinement
        from tenet.network.batched_lanczos import BatchedLanczos

        start = torch.cat([i[4].data for i in insts])
        bl = BatchedLanczos(bh, ncv=3)
        bl.start_(start)
        dt_l = time_arm(bl.run, max(2, args.steps // 4), 2)
        results[p]["lanczos"] = {"ncv": 3, "ms_per_batched_step": dt_l * 1e3,
                                 "problem_lanczos_steps_per_s": len(insts) / dt_l,
                                 "launch": "one hipGraph per iteration and one for the finish"}
        log(f"{p}: batched Lanczos step (ncv 3) {dt_l * 1e3:.3f} ms for {len(insts)} problems")
        if p != "fp64":
            blr = BatchedLanczos(bh, ncv=3, refine=BatchedHeff2(insts, "fp64"))
            blr.start_(start)
            dt_r = time_arm(blr.run, max(2, args.steps // 4), 2)
            results[p]["lanczos"]["ms_per_batched_step_with_fp64_refine"] = dt_r * 1e3
            log(f"{p}: with the batched FP64 refinement {dt_r * 1e3:.3f} ms")
        if p == "fp32":
            # one problem at a time: the single-problem graph-captured Krylov step, a sample
            from tenet.network import Heff2
            from tenet.network.graph import KrylovGraph

            sample = insts[:: max(1, len(insts) // 16)]
            kgs = [(KrylovGraph(Heff2(*i[:4], "fp32"), i[4].structure, ncv=3), i[4]) for i in sample]
            for kg, v in kgs:
                kg.start_(v.data)

            def serial():
                for kg, _ in kgs:
                    kg.replay()

            dt_s = time_arm(serial, max(2, args.steps // 4), 2) * len(insts) / len(sample)
            results[p]["lanczos"]["one_at_a_time_ms"] = dt_s * 1e3
            results[p]["lanczos"]["speedup_batched_vs_one_at_a_time"] = dt_s / dt_l
            log(f"fp32: one problem at a time {dt_s * 1e3:.3f} ms for {len(insts)} problems")
        if p == "fp32":
            # the same batch with one GEMM problem per block pair (no sector stacking)
            bb = BatchedHeff2(insts, p, layout="blocks")
            xb = bb.pack([i[4] for i in insts])
            dt_b = time_arm(lambda: bb.apply(xb, fixed), args.steps, args.warmup)
            d = (bb.apply(xb, fixed) - bh.apply(x, tuned)).norm() / bb.apply(xb, fixed).norm()
            results["fp32_blocks"] = {"batched_matvecs_per_s": 1 / dt_b, "ms": dt_b * 1e3,
                                      "rel_diff_vs_sector": float(d)}
            log(f"fp32 block layout: {dt_b * 1e3:.3f} ms")
    head = "fp32" if "fp32" in results else next(iter(results))
    line = {"metric": "batched H_eff matvecs/s (256 problems per batched matvec)",
            "value": results[head]["batched_matvecs_per_s"], "unit": "batched matvecs/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": results[head]["ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32" if head != "fp64" else "f64", "data": "synthetic",
            "config": {"workload": f"{len(insts)} U(1) XXZ two-site H_eff, L=32 centre bond, D log-uniform "
                                   "[64,1024], Delta~U[0,2], MPO bond 5", "D_median": int(np.median(dims)),
                       "layout": "sector-stacked"},
            "arms": results,
            "roofline": {"bound": "mfma", "achieved": results[head]["effective_tflops"],
                         "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s",
                         "frac": results[head]["effective_tflops"] / PEAK_TFLOPS[head], "traffic": None}}
    print(json.dumps(line), flush=True)
    return 0


MPSKIT = {32: {64: -13.99731561797256, 128: -13.997315618224851, 256: -13.997315618224897},
          64: {64: -28.175424792530414, 128: -28.17542485963937, 256: -28.175424859745323}}


def bench_dmrg(args) -> int:
    """One ground-state solve of the U(1) Heisenberg chain; s/sweep and energy against the
    MPSKit fixture energies the reference's tests hold."""
    import torch

    from tenet.network import dmrg_
    from tenet.network.models import heisenberg_mpo, random_chain

    L = args.sites if args.sites != 64 or args.chi in (64, 128, 256) else args.sites
    res = {}
    for p in [a for a in args.arms.split(",") if a in ("fp64", "fp32", "bf16x3")]:
        psi = random_chain(L, 16, 3.0, seed=1, device="cuda")
        h = heisenberg_mpo(L, device="cuda")
        times = []
        last = [time.perf_counter()]

        def cb(out):
            torch.cuda.synchronize()
            now = time.perf_counter()
            times.append(now - last[0])
            last[0] = now
            log(f"{p} sweep {out.sweeps}: E={out.energy:.15f} dE={out.denergy:.2e} dS={out.max_dSchmidt:.2e} ({times[-1]:.2f}s)")

        t0 = time.perf_counter()
        out = dmrg_(psi, h, chi=args.chi, max_sweeps=args.max_sweeps, precision=p, callback=cb)
        torch.cuda.synchronize()
        total = time.perf_counter() - t0
        ref = MPSKIT.get(L, {}).get(args.chi)
        res[p] = {"energy": out.energy, "sweeps": out.sweeps, "total_s": total,
                  "s_per_sweep_steady": sorted(times)[len(times) // 2] if times else None,
                  "first_sweep_s": times[0] if times else None,
                  "mpskit_energy": ref, "dE_vs_mpskit": (abs(out.energy - ref) if ref else None)}
    if "fp64" in res:
        for p in res:
            res[p]["dE_rel_vs_fp64"] = abs(res[p]["energy"] - res["fp64"]["energy"]) / abs(res["fp64"]["energy"])
    head = "fp32" if "fp32" in res else next(iter(res))
    line = {"metric": "DMRG seconds per steady sweep", "value": res[head]["s_per_sweep_steady"],
            "unit": "s/sweep", "n_gpus": 1, "steps": res[head]["sweeps"], "warmup": 0,
            "ms_per_step": 1e3 * res[head]["s_per_sweep_steady"], "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32" if head != "fp64" else "f64",
            "data": "synthetic", "config": {"workload": f"U(1) Heisenberg OBC L={L} chi={args.chi} DMRG"},
            "arms": res}
    print(json.dumps(line), flush=True)
    return 0


def bench_serve(args) -> int:
    """Poisson stream of XXZ ground-state solves through the continuous-batching server,
    against the same stream solved one at a time in FP64."""
    from tenet.network.serve import DMRGServer, MultiProcessServer, latency_report, poisson_requests

    sizes = tuple(int(x) for x in os.environ.get("TENET_SERVE_SIZES", "32,64,128").split(","))
    chis = tuple(int(x) for x in os.environ.get("TENET_SERVE_CHIS", "128,256,512,1024").split(","))
    workers = int(os.environ.get("TENET_SERVE_WORKERS", "4"))
    reqs = poisson_requests(args.count, args.rate, sizes=sizes, chis=chis, seed=0,
                            max_sweeps=args.max_sweeps)
    out = {}
    # the multi-process arm first: its workers are spawned before this process uses the GPU
    arms = [("mp", f"continuous_fp32_{workers}proc", lambda: MultiProcessServer(workers=workers, max_active=8,
                                                                                precision="fp32")),
            ("mp_shared", f"shared_batches_fp32_{workers}proc",
             lambda: MultiProcessServer(workers=workers, max_active=8, precision="fp32", mode="shared")),
            ("shared", "shared_batches_fp32", lambda: DMRGServer(max_active=8, precision="fp32", mode="shared")),
            ("single", "continuous_fp32", lambda: DMRGServer(max_active=8, precision="fp32")),
            ("serial", "one_at_a_time_fp64", lambda: DMRGServer(max_active=1, precision="fp64"))]
    chosen = set(args.serve_arms.split(","))
    arms = [(name, make) for key, name, make in arms if key in chosen]
    if not arms:
        print(f"--serve-arms {args.serve_arms!r} names no arm (mp, single, serial)", file=sys.stderr)
        return 2
    for name, make in arms:
        srv = make()
        t0 = time.perf_counter()
        res = srv.serve(reqs, on_result=lambda r, n=name: log(
            f"{n}: request {r.rid} done at {r.finished:.1f}s ({r.sweeps} sweeps, E={r.energy:.15f})"))
        wall = time.perf_counter() - t0
        boot = getattr(srv, "boot_s", None)
        rep = latency_report(res, wall - (boot or 0.0))
        if boot is not None:
            rep["worker_boot_s"] = boot
        rep["energies"] = [r.energy for r in res]
        out[name] = rep
        log(f"{name}: p50 {rep['p50_s']:.2f}s p99 {rep['p99_s']:.2f}s, {rep['solves_per_hour']:.0f} solves/h")
    de = None
    if "one_at_a_time_fp64" in out and len(out) > 1:
        e64 = out["one_at_a_time_fp64"]["energies"]
        de = max(max(abs(a - b) / abs(b) for a, b in zip(out[k]["energies"], e64))
                 for k in out if k != "one_at_a_time_fp64")
    head = out[arms[0][0]]
    line = {"metric": "serve p50/p99 time-to-solve", "value": head["solves_per_hour"],
            "unit": "solves/hour", "n_gpus": 1, "steps": args.count, "warmup": 0,
            "ms_per_step": 1e3 * head["mean_s"], "higher_is_better": True, "scaling": "strong",
            "vs_baseline": None, "dtype": "f64" if arms[0][0] == "one_at_a_time_fp64" else "fp32",
            "data": "synthetic",
            "config": {"workload": f"Poisson stream of {args.count} XXZ DMRG solves, rate {args.rate}/s, "
                                   f"L in {sizes}, chi in {chis}, Delta~U[0,2], cutoff 1e-10",
                       "server": (f"{workers} worker processes x {8 // workers} slots"
                                  if arms[0][0].endswith("proc") else arms[0][0])},
            "p50_s": head["p50_s"], "p99_s": head["p99_s"], "max_dE_rel_vs_fp64_serial": de,
            "arms": out}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
