"""Autotuner of the batched sweep: staging length per chain, then tile size and
workgroup size of the LDS tile kernel.

Staging length.  ReshapeLinear adapts its slice count to keep the staging
acceptance in [0.6, 0.8] (NumbOfSlices, src/updates/helper.jl:42-52).  Here
the acceptance of every chain is measured once for each candidate length,
which bounds each chain's length from above: the longest candidate whose
staging acceptance is at least `min_acc`.

Throughput.  For each cap on that bound and each (tile_beads, threads) pair,
the same measuring sweeps run on a copy of the workload, and the combination
with the most proposed beads per second of kernel time wins.  Trial sweeps
use their own seed and start from the workload's initial paths, so tuning
never changes the production chains.
"""
from __future__ import annotations

import copy
import time
from dataclasses import dataclass, field

import numpy as np

from ._capi import Context, Options


@dataclass
class TuneResult:
    ell: list
    tile_beads: int
    block_threads: int
    timings: dict = field(default_factory=dict)
    acceptance: dict = field(default_factory=dict)
    seconds: float = 0.0


def _options(wl, spl, tile_beads=0, block_threads=0, seed=0x5EED):
    opt = Options()
    opt.precision = wl.precision
    opt.sweeps_per_launch = spl
    opt.radial_bins = wl.radial_bins
    opt.radial_rmax = wl.radial_rmax
    opt.cart_bins = wl.cart_bins
    opt.derivatives = wl.derivatives
    opt.tile_beads = tile_beads
    opt.block_threads = block_threads
    opt.seed = seed
    return opt


def _context(wl, descs, opt):
    ctx = Context(descs, opt)
    for c, r in enumerate(wl.init):
        ctx.set_paths(c, r)
    return ctx


def autotune(wl, trial_sweeps: int = 100, ells=(4, 8, 16, 32), tiles=(2048, 4096, 8192),
             threads=(128, 256), min_acc: float = 0.6, tune_ell: bool = True) -> TuneResult:
    t0 = time.perf_counter()
    descs = [copy.copy(d) for d in wl.descs]
    n = len(descs)
    half = max(1, trial_sweeps // 2)
    acc_table = {}
    if tune_ell:
        best = [min(2 if d.M < 4 else 4, d.M) for d in descs]
        for ell in ells:
            trial = [copy.copy(d) for d in descs]
            for d in trial:
                d.ell = min(ell, d.M)
            ctx = _context(wl, trial, _options(wl, half))
            ctx.run(half, measure=False)
            before = ctx.all_counters()
            ctx.run(half, measure=False)
            after = ctx.all_counters()
            ctx.close()
            acc = (after[:, 0] - before[:, 0]) / np.maximum(1, after[:, 1] - before[:, 1])
            acc_table[ell] = float(acc.mean())
            for c in range(n):
                if acc[c] >= min_acc and ell <= descs[c].M:
                    best[c] = max(best[c], min(ell, descs[c].M))
        for c in range(n):
            descs[c].ell = best[c]
    timings = {}
    best_key, best_rate, best_ell = None, -1.0, None
    caps = [e for e in ells if e >= 8] if tune_ell else [None]
    for cap in caps:
        trial = [copy.copy(d) for d in descs]
        if cap is not None:
            for d in trial:
                d.ell = max(2, min(d.ell, cap, d.M))
        for tb in tiles:
            for th in threads:
                try:
                    ctx = _context(wl, trial, _options(wl, half, tb, th))
                except Exception:  # a tile that does not fit this workgroup size
                    timings[(cap, tb, th)] = None
                    continue
                ctx.run(half, measure=True)
                cnt = ctx.run(half, measure=True)
                ctx.close()
                rate = cnt.bead_moves / max(1e-9, cnt.kernel_ms)
                timings[(cap, tb, th)] = rate
                if rate > best_rate:
                    best_rate, best_key, best_ell = rate, (tb, th), [d.ell for d in trial]
    return TuneResult(best_ell, best_key[0], best_key[1],
                      {f"ell<={k[0]} tile={k[1]} threads={k[2]}": v for k, v in timings.items()}, acc_table,
                      time.perf_counter() - t0)
