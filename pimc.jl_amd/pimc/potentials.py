"""Potentials as callables that also describe themselves to the device.

In the reference `System(potential::Function; dV::Function)` takes Julia
closures (src/system.jl:129-131).  A GPU kernel cannot call a host closure,
so each potential here is a callable object (usable on the host exactly like
the closure) that also carries the enum and parameters of its device form.
"""
from __future__ import annotations

import numpy as np

from ._capi import PAIR_GAUSSIAN, POT_HARMONIC, POT_LATTICE, POT_NONE


class NoPotential:
    kind = POT_NONE

    def __call__(self, r):
        return 0.0

    def dV(self, r):
        return np.zeros_like(np.asarray(r, dtype=float))

    def params(self):
        return {}


class HarmonicTrap:
    """V(r) = omega^2 |r|^2 / 2 (mass 1), e.g. examples/energy_2d_harmonically_trapped_bose_gas.jl:5."""
    kind = POT_HARMONIC

    def __init__(self, omega: float = 1.0):
        self.omega = float(omega)

    def __call__(self, r):
        r = np.asarray(r, dtype=float)
        return 0.5 * self.omega ** 2 * float(np.dot(r, r))

    def dV(self, r):
        return self.omega ** 2 * np.asarray(r, dtype=float)

    def params(self):
        return {"omega": self.omega}


class SquareLattice:
    """V(x, y) = V0 (cos^2(pi x / a) + cos^2(pi y / a))."""
    kind = POT_LATTICE

    def __init__(self, V0: float = 4.0, a: float = 1.0):
        self.V0, self.a = float(V0), float(a)

    def __call__(self, r):
        r = np.asarray(r, dtype=float)
        return self.V0 * float(np.sum(np.cos(np.pi * r / self.a) ** 2))

    def dV(self, r):
        r = np.asarray(r, dtype=float)
        return -self.V0 * (np.pi / self.a) * np.sin(2 * np.pi * r / self.a)

    def params(self):
        return {"V0": self.V0, "a": self.a}


class GaussianPair:
    """v(r) = g exp(-r^2 / (2 sigma^2)), entering the reference's lnU hook as a
    primitive link action lnU = -tau/2 (v(|r1_rel|) + v(|r2_rel|))."""
    kind = PAIR_GAUSSIAN

    def __init__(self, g: float = 1.0, sigma: float = 0.5, cutoff: float | None = None):
        self.g, self.sigma = float(g), float(sigma)
        self.cutoff = float(cutoff) if cutoff is not None else 4.0 * self.sigma

    def __call__(self, r):
        return self.g * np.exp(-float(r) ** 2 / (2 * self.sigma ** 2))
