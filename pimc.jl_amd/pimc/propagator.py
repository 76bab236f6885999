"""The pair-product propagator of a 2D contact interaction as a device table.

build_prop_int(L, g0, tau) (src/propagator.jl:79-89) tabulates the integral
terms of the relative propagator on a 600 x 600 grid of (|r1|, |r2|) in
[1e-20, L] (prop_rel_interpolate_terms, :36-70) and returns

    prop_int(r1_rel, r2_rel, tau) = 1 + terms(|r1|, |r2|) / prop_rel0(r1_rel, r2_rel, tau),
    prop_rel0 = exp(-|r1_rel - r2_rel|^2 / (4 tau)) / (4 pi tau),

which System turns into lnU = log(prop_int) (-mu where prop_int < 0,
src/system.jl:26-28).  Here the table is computed on the GPU
(pimc_prop_int_terms) and handed to a PIMC_PAIR_TABLE chain, whose kernels
interpolate it bilinearly for every neighbour pair and link.
"""
from __future__ import annotations

import math

import numpy as np

from ._capi import PAIR_TABLE, prop_int_terms

R0 = 1e-20  # first grid radius (the integrals diverge at r = 0, src/propagator.jl:39-41)


class PropInt:
    """A tabulated pair-product propagator; pass it as System(pair=...) or
    System(interactions=True, propint=...)."""
    kind = PAIR_TABLE
    sigma = 1.0

    def __init__(self, terms: np.ndarray, r0: float, r1: float, g0: float, tau: float):
        self.terms = np.ascontiguousarray(terms, dtype=np.float64)
        self.r0, self.r1, self.g0, self.tau = float(r0), float(r1), float(g0), float(tau)
        self.g = self.g0
        self.n = self.terms.shape[0]

    def _diag(self, r: float) -> float:
        """terms(r, r) by the bilinear rule of the device tables."""
        n, h = self.n, (self.r1 - self.r0) / (self.n - 1)
        p = min(max((r - self.r0) / h, 0.0), n - 1.0)
        i = min(int(math.floor(p)), n - 2)
        f = p - i
        t = self.terms
        a0 = (1.0 - f) * t[i, i] + f * t[i, i + 1]
        a1 = (1.0 - f) * t[i + 1, i] + f * t[i + 1, i + 1]
        return (1.0 - f) * a0 + f * a1

    def nn_range(self, level: float = 0.999) -> float:
        """determine_nnrange (src/system.jl:10-15): the radius where
        prop_int([r], [r], tau) = 1 + 4 pi tau terms(r, r) crosses `level`,
        by bisection between the minimum of the tabulated diagonal and r1
        (the reference minimises with Optim first, then brackets with Roots)."""
        grid = self.r0 + np.arange(self.n) * ((self.r1 - self.r0) / (self.n - 1))
        c = 4.0 * math.pi * self.tau
        f = lambda r: 1.0 + c * self._diag(r) - level
        lo = float(grid[int(np.argmin(np.diag(self.terms)))])
        hi = self.r1
        if f(lo) > 0 or f(hi) < 0:
            raise ValueError("prop_int(r, r) does not cross the level inside the table")
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if f(mid) < 0:
                lo = mid
            else:
                hi = mid
            if hi - lo <= 1e-15 * max(1.0, hi):
                break
        return 0.5 * (lo + hi)


def build_prop_int(L: float, g0: float, tau: float, n: int = 600) -> PropInt:
    """build_prop_int(L, g0, tau) (src/propagator.jl:79-89) with the terms
    computed on the GPU; n is the reference's grid size Delta."""
    return PropInt(prop_int_terms(n, R0, L, g0, tau), R0, L, g0, tau)
