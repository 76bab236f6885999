"""Potentials as callables that also describe themselves to the device.

In the reference `System(potential::Function; dV::Function)` takes Julia
closures (src/system.jl:129-131).  A GPU kernel cannot call a host closure,
so each potential here is a callable object (usable on the host exactly like
the closure) that also carries the enum and parameters of its device form.
"""
from __future__ import annotations

import numpy as np

import math

from ._capi import MAX_BEAMS, PAIR_GAUSSIAN, POT_HARMONIC, POT_LATTICE, POT_NONE, POT_OPTICAL


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


class OpticalLattice:
    """V(r) = depth * normalized_intensity(r, angles, scale; helical)
    (examples/tools/potentialtools.jl:1-16, generate_V :18-23): the squared
    mean of the beams' plane waves exp(i 2 pi scale (x sin a + y cos a)),
    2D, at most 16 beams.  A negative depth is generate_V's attractive form."""
    kind = POT_OPTICAL

    def __init__(self, angles, scale: float = 1.0, depth: float = 1.0, helical: bool = False):
        self.angles = [float(a) for a in angles]
        if not 1 <= len(self.angles) <= MAX_BEAMS:
            raise ValueError(f"1 to {MAX_BEAMS} beams")
        self.scale, self.depth, self.helical = float(scale), float(depth), bool(helical)
        self.V0 = self.depth

    def _sums(self, r):
        x, y = float(r[0]), float(r[1])
        s = c = 0.0
        ds, dc = np.zeros(2), np.zeros(2)
        for a in self.angles:
            k = 2 * math.pi * self.scale * np.array([math.sin(a), math.cos(a)])
            t = k[0] * x + k[1] * y + (a if self.helical else 0.0)
            s += math.sin(t)
            c += math.cos(t)
            ds += math.cos(t) * k
            dc -= math.sin(t) * k
        return s, c, ds, dc

    def __call__(self, r):
        s, c, _, _ = self._sums(r)
        n = len(self.angles)
        return self.depth * ((s / n) ** 2 + (c / n) ** 2)

    def dV(self, r):
        s, c, ds, dc = self._sums(r)
        n = len(self.angles)
        return 2 * self.depth / n ** 2 * (s * ds + c * dc)

    def params(self):
        return {"angles": self.angles, "scale": self.scale, "depth": self.depth, "helical": self.helical}


# beam directions of generate_V(scale, depth, sym) (examples/tools/potentialtools.jl:25-39)
BEAMS = {
    "cubic": [2 * math.pi * k / 4 for k in range(4)],
    "l25": [2.214297435588181, 0.9272952180016122, -0.6435011087932844, 0.6435011087932844, -2.498091544796509,
            3.141592653589793, 2.498091544796509, 0.0, 1.5707963267948966, -2.2142974355881813,
            -1.5707963267948968, -0.9272952180016123],
    "l65": [0.5191461142465229, 1.695151321341658, -0.12435499454676144, -1.4464413322481353, 2.0899424410414196,
            2.62244653934327, 0.12435499454676144, -1.0516502125483738, -1.695151321341658, 1.446441332248135,
            -2.0899424410414196, -3.017237659043032, -0.5191461142465229, 3.017237659043032, -2.62244653934327,
            1.0516502125483738],
}


def generate_V(scale: float, depth: float, sym: str, attractive: bool = True):
    """generate_V(scale, depth, sym; attractive) (examples/tools/potentialtools.jl:25-39):
    the square (cubic) and the 12- and 16-beam superlattices (l25, l65), or the harmonic trap."""
    if sym == "harmonic":
        return HarmonicTrap(1.0)
    return OpticalLattice(BEAMS[sym], scale, -depth if attractive else depth)


class GaussianPair:
    """v(r) = g exp(-r^2 / (2 sigma^2)), entering the reference's lnU hook as a
    primitive link action lnU = -tau/2 (v(|r1_rel|) + v(|r2_rel|))."""
    kind = PAIR_GAUSSIAN

    def __init__(self, g: float = 1.0, sigma: float = 0.5, cutoff: float | None = None):
        self.g, self.sigma = float(g), float(sigma)
        self.cutoff = float(cutoff) if cutoff is not None else 4.0 * self.sigma

    def __call__(self, r):
        return self.g * np.exp(-float(r) ** 2 / (2 * self.sigma ** 2))
