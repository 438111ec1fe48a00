"""Tile-shape autotuner for the block-sparse grouped GEMM.

The key of a GEMM program is the flop-weighted histogram of its problem sizes over
log2 bins of (M, N, K), quantized to 5% steps -- two launches with the same sector-size
profile share one decision.  ``tune`` times every candidate tile on the live buffers with
device events (warm-up, then the median of ``reps`` launches), keeps the fastest, and
records it on the program and in the process-wide cache.
"""

from __future__ import annotations

import math

import torch

from tenet.engine import _TILES, GemmProgram

__all__ = ["histogram_key", "tune", "CACHE"]

CACHE: dict = {}


def histogram_key(g: GemmProgram) -> tuple:
    hist: dict = {}
    total = 0
    for M, N, K in zip(g.Ms, g.Ns, g.Ks):
        w = M * N * K
        b = (int(math.log2(max(M, 1))), int(math.log2(max(N, 1))), int(math.log2(max(K, 1))))
        hist[b] = hist.get(b, 0) + w
        total += w
    total = total or 1
    return tuple(sorted((b, round(20 * w / total)) for b, w in hist.items() if round(20 * w / total) > 0))


def tune(g: GemmProgram, A, B, C, precision: str, candidates=_TILES, reps: int = 5) -> tuple:
    key = (precision, g.trans_a, g.trans_b, histogram_key(g))
    hit = CACHE.get(key)
    if hit is not None:
        g.tile = hit[0]
        return hit[0]
    timings = {}
    for tile in candidates:
        g.run(A, B, C, precision, tile=tile)
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            g.run(A, B, C, precision, tile=tile)
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        timings[tile] = sorted(ts)[len(ts) // 2]
    best = min(timings, key=timings.get)
    CACHE[key] = (best, timings)
    g.tile = best
    return best
