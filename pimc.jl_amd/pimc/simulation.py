"""run! and apply! (src/simulation.jl:12-42)."""
from __future__ import annotations

import numpy as np

from .measurement import measurement_Z_sector
from .updates import UpdateC, acceptance, adjust, queue


def apply(s, f) -> None:
    """apply! (src/simulation.jl:12-27): measurements are called in order;
    an update is called, its try queued, and every `adj` tries its variable
    adjusted from the sliding-window acceptance."""
    if isinstance(f, (list, tuple)):
        for m in f:
            m(s)
        return
    acc = f(s)
    queue(f.counter, acc)
    if f.var is not None and f.counter_var.tries % f.counter_var.adj == 0:
        adjust(f.var, acceptance(f.counter_var.queue))


def run(s, n: int, updates, Zmeasurements=()) -> None:
    """run!(s, n, updates; Zmeasurements) (src/simulation.jl:29-42): n
    iterations, each applying one update drawn with weight 1/every."""
    every = np.array([e for e, _ in updates], dtype=float)
    fs = [f for _, f in updates]
    w = (1.0 / every)
    w = w / w.sum()
    cum = np.cumsum(w)
    for _ in range(n):
        k = int(np.searchsorted(cum, s.rng.random(), side="right"))
        k = min(k, len(fs) - 1)
        apply(s, fs[k])
        measurement_Z_sector(s, list(Zmeasurements))
