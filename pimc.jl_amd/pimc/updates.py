"""Updates as callable structs (src/updates/*.jl), executed on the GPU.

Each functor draws its random choices on the host (the reference draws them
from Julia's global RNG) and hands the move to the device: the bridge, the
action differences, the Metropolis test and the commit run in
libpimc_hip.so.
"""
from __future__ import annotations

import math
from collections import deque

import numpy as np


class Update:
    pass


class UpdateC(Update):
    pass


class Counter:
    """src/updates/helper.jl:6-11: sliding window of accept flags."""

    def __init__(self, adj: int = 10, range: int = 100_000):
        self.queue: deque = deque(maxlen=range)
        self.tries = 0
        self.range = range
        self.adj = adj


def queue(c: Counter, acc: bool) -> None:
    """queue! (src/simulation.jl:3-10)."""
    c.tries += 1
    c.queue.append(1 if acc else 0)


def acceptance(q) -> float:
    """src/simulation.jl:1."""
    return sum(q) / len(q) if len(q) else 0.0


class Step:
    """src/updates/helper.jl:14-20."""

    def __init__(self, size, minstep, maxstep, minacc, maxacc):
        self.size, self.minstep, self.maxstep, self.minacc, self.maxacc = size, minstep, maxstep, minacc, maxacc


class NumbOfSlices:
    """src/updates/helper.jl:34-40."""

    def __init__(self, m, minslices, maxslices, minacc, maxacc):
        self.m, self.minslices, self.maxslices, self.minacc, self.maxacc = m, minslices, maxslices, minacc, maxacc


def adjust(var, acc: float) -> None:
    """adjust! for Step (src/updates/helper.jl:22-32) and NumbOfSlices (:42-52)."""
    if isinstance(var, Step):
        if acc < var.minacc:
            var.size *= 0.9
        elif acc > var.maxacc:
            var.size *= 1.1
        var.size = min(var.maxstep, max(var.minstep, var.size))
    elif isinstance(var, NumbOfSlices):
        if acc < var.minacc:
            var.m -= 1
        elif acc > var.maxacc:
            var.m += 1
        var.m = min(var.maxslices, max(var.minslices, var.m))


class ReshapeLinear(UpdateC):
    """ReshapeLinear(s, slices) (src/updates/reshape.jl:7-91): Levy staging of
    m = min(maxslices, rand(2:var.m)) links from a random bead."""

    def __init__(self, s, slices: int, minslices: int = 2, maxslices: int | None = None, minacc: float = 0.6,
                 maxacc: float = 0.8, adj: int = 10, range: int = 10_000):
        maxslices = s.M - 2 if maxslices is None else maxslices
        self.counter = Counter()
        self.var = NumbOfSlices(min(slices, maxslices), minslices, maxslices, minacc, maxacc)
        self.counter_var = Counter(adj=adj, range=range)

    def __call__(self, s) -> bool:
        n = int(s.rng.integers(s.N))
        j0 = int(s.rng.integers(s.M))
        m = min(self.var.maxslices, int(s.rng.integers(2, self.var.m + 1)))
        normals = s.rng.standard_normal((m - 1, s.dim))
        u = float(s.rng.random())
        acc = s.ctx.reshape_linear(0, n, j0, m, normals, u)
        queue(self.counter_var, acc)
        return acc


class SingleCenterOfMass(UpdateC):
    """SingleCenterOfMass(s, step) (src/updates/com.jl:112-224): rigid shift of
    one worldline by d uniform in [-step, step]^dim."""

    def __init__(self, s, step: float, minstep: float = 1e-1, maxstep: float | None = None, minacc: float = 0.4,
                 maxacc: float = 0.6, adj: int = 10, range: int = 10_000):
        maxstep = s.L / 2 if maxstep is None else maxstep
        self.counter = Counter()
        self.var = Step(step, minstep, maxstep, minacc, maxacc)
        self.counter_var = Counter(adj=adj, range=range)

    def __call__(self, s) -> bool:
        n = int(s.rng.integers(s.N))
        d = self.var.size * 2 * (s.rng.random(s.dim) - 0.5)
        u = float(s.rng.random())
        acc = s.ctx.center_of_mass(0, n, d, u)
        queue(self.counter_var, acc)
        return acc


class HIPSweep(UpdateC):
    """One batched sweep on the device: every particle's ring re-grown by
    staging segments of length `ell` at a random offset, then one COM try per
    particle.  Pushes one flag (any move accepted) per call."""

    def __init__(self, s, adj: int = 10, range: int = 10_000):
        self.counter = Counter()
        self.var = None
        self.counter_var = Counter(adj=adj, range=range)
        self.last = None

    def __call__(self, s) -> bool:
        cnt = s.ctx.run(1, measure=False)
        self.last = cnt
        acc = (cnt.levy_acc + cnt.com_acc) > 0
        queue(self.counter_var, acc)
        return acc


_SWEEP_PASS = {"ReshapeLinear": "staging_every", "SingleCenterOfMass": "com_every",
               "ReshapeSwapLinear": "swap_every"}


def sweep_weights(updates) -> dict:
    """The move weights of run!(s, n, updates) (src/simulation.jl:33-36) as the
    pass periods of the batched sweep.

    run! draws one update per iteration with weight 1/every.  The batched
    sweep runs each pass (staging, centre of mass, swap tries) on the sweeps s
    with s % every == 0, so the passes keep the relative frequencies
    1/every.  The weights are relative, so the periods are divided by their
    greatest common divisor: [(2, SingleCenterOfMass), (1, ReshapeLinear),
    (1, ReshapeSwapLinear)], the reference's own test mix
    (test/testsystem.jl:39-43), gives com_every=2, staging_every=1,
    swap_every=1.  Entries name their update by class, instance or class
    name.  A sweep always holds a staging and a centre-of-mass pass, so both
    kinds must be listed; a list without ReshapeSwapLinear leaves
    swap_every at 1 (give the System no swap tries).
    """
    found: dict[str, int] = {}
    for every, f in updates:
        name = f if isinstance(f, str) else (f.__name__ if isinstance(f, type) else type(f).__name__)
        if name not in _SWEEP_PASS:
            raise ValueError(f"{name} is not a pass of the batched sweep")
        if int(every) != every or every < 1:
            raise ValueError("every must be a positive integer")
        key = _SWEEP_PASS[name]
        if key in found:
            raise ValueError(f"{name} listed twice")
        found[key] = int(every)
    for key in ("staging_every", "com_every"):
        if key not in found:
            raise ValueError("the batched sweep needs both ReshapeLinear and SingleCenterOfMass")
    g = 0
    for v in found.values():
        g = math.gcd(g, v)
    out = {k: v // g for k, v in found.items()}
    out.setdefault("swap_every", 1)
    if max(out.values()) > 65536:
        raise ValueError("every / gcd must be <= 65536")
    return out
