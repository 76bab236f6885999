"""Measurements as callable structs (src/measurement.jl), evaluated on the GPU."""
from __future__ import annotations

import numpy as np


class Measurement:
    pass


class ZMeasurement(Measurement):
    pass


def measurement_Z_sector(s, measurements) -> None:
    """src/measurement.jl:1-17: every Ncycle attempts, apply the measurements."""
    if not measurements:
        return
    if s.N not in s.Nctr:
        s.Nctr[s.N] = 0
        s.N_MC[s.N] = 0
    s.Nctr[s.N] += 1
    if s.Nctr[s.N] == s.Ncycle:
        s.N_MC[s.N] += 1
        for m in measurements:
            m(s)
        s.Nctr[s.N] = 0


class Density(ZMeasurement):
    """Density(s; nbins=500) (src/measurement.jl:31-55), the Cartesian
    histogram with the reference's bin convention (the first bin is dropped
    and the histogram is shifted by one bin)."""

    def __init__(self, s, nbins: int = 500):
        self.nbins = nbins
        self.bin = (2 * s.L) / nbins
        self.dens = np.zeros((nbins,) * s.dim, order="F")
        self.ndata = 0

    def __call__(self, s):
        self.dens, self.ndata = s.ctx.density(0, self.nbins, self.dens, self.ndata)


class Energy(ZMeasurement):
    """Energy(s, n=20000) (src/measurement.jl:78-122): thermodynamic and
    virial energy per configuration, stored per particle number.  With a pair
    potential the pair energy and pair virial are added (the reference has no
    pair term)."""

    def __init__(self, s, n: int = 20_000):
        self.energy: dict[int, list] = {0: []}
        self.energy_virial: dict[int, list] = {0: []}
        self.capacity = n

    def __call__(self, s):
        E, Ev, Ep, Evp = s.ctx.energy(0)
        self.energy.setdefault(s.N, []).append(E + Ep)
        self.energy_virial.setdefault(s.N, []).append(Ev + Evp)
