"""Closed forms for the trapped ideal gas used to pin the oracle and the GPU path.

Primitive action of one free ring of M beads in a harmonic trap, one
coordinate, lambda = hbar^2/2m:

    S = sum_j (x_j - x_j+1)^2 / (4 lambda tau) + tau omega^2 x_j^2 / 2

is the quadratic form x^T A x / 2 with eigenvalues
a_k = (2 - 2 cos(2 pi k / M)) / (2 lambda tau) + tau omega^2.  With the
kinetic normalisation (4 pi lambda tau)^(-M/2),
ln Z_1 = -(M/2) ln(4 pi lambda tau) - (1/2) sum_k ln a_k + const, so the
thermodynamic energy per coordinate is

    E_1 = -d ln Z_1 / d beta = (1/M) [ M/(2 tau) + (1/2) sum_k a_k'(tau) / a_k ]

with a_k' = -(2 - 2 cos(2 pi k/M)) / (2 lambda tau^2) + omega^2, and
<x^2> = (1/M) sum_k 1 / a_k.  Distinguishable particles (no permutation
moves) in d dimensions: E = d N E_1.  Continuum limit: E_1 = (omega/2)
coth(beta omega / 2).  Bosons: the canonical recursion of
examples/tools/energytools.jl:3-21 (Z_N = (1/N) sum_k z(k beta) Z_{N-k}).
"""
from __future__ import annotations

import math

import numpy as np


def ring_modes(M: int, beta: float, omega: float = 1.0, lam: float = 0.5):
    tau = beta / M
    k = np.arange(M)
    c = 2.0 - 2.0 * np.cos(2.0 * np.pi * k / M)
    a = c / (2.0 * lam * tau) + tau * omega ** 2
    da = -c / (2.0 * lam * tau ** 2) + omega ** 2
    return tau, a, da


def energy_1d_primitive(M: int, beta: float, omega: float = 1.0, lam: float = 0.5) -> float:
    tau, a, da = ring_modes(M, beta, omega, lam)
    return (M / (2.0 * tau) + 0.5 * np.sum(da / a)) / M


def x2_primitive(M: int, beta: float, omega: float = 1.0, lam: float = 0.5) -> float:
    tau, a, _ = ring_modes(M, beta, omega, lam)
    return float(np.sum(1.0 / a) / M)


def energy_boltzmann_primitive(N: int, M: int, beta: float, dim: int = 2, omega: float = 1.0,
                               lam: float = 0.5) -> float:
    return dim * N * energy_1d_primitive(M, beta, omega, lam)


def energy_boltzmann_continuum(N: int, beta: float, dim: int = 2, omega: float = 1.0) -> float:
    return dim * N * 0.5 * omega / math.tanh(beta * omega / 2.0)


def denergy_domega_continuum(N: int, beta: float, dim: int = 2, omega: float = 1.0) -> float:
    """d/d omega of d N (omega/2) coth(beta omega / 2)."""
    x = beta * omega / 2.0
    return dim * N * (0.5 / math.tanh(x) - 0.5 * x / math.sinh(x) ** 2)


def denergy_domega_primitive(N: int, M: int, beta: float, dim: int = 2, omega: float = 1.0,
                             lam: float = 0.5, h: float = 1e-5) -> float:
    return (energy_boltzmann_primitive(N, M, beta, dim, omega + h, lam)
            - energy_boltzmann_primitive(N, M, beta, dim, omega - h, lam)) / (2 * h)


def energy_boson_continuum(N: int, beta: float, dim: int = 2) -> float:
    """examples/tools/energytools.jl:15-21 with omega = 1, derivative by central difference."""
    def z(b):
        return (math.exp(-0.5 * b) / (1.0 - math.exp(-b))) ** dim

    def Z(n, b):
        if n == 0:
            return 1.0
        return sum(z(k * b) * Z(n - k, b) for k in range(1, n + 1)) / n

    h = 1e-6
    return -(math.log(Z(N, beta + h)) - math.log(Z(N, beta - h))) / (2 * h)


def _ln_ring(K: int, tau: float, omega: float, lam: float) -> float:
    """ln of one coordinate's primitive-action partition function of a ring of
    K beads at time step tau (kinetic normalisation per bead included, a
    constant per bead dropped): -1/2 sum_k ln(2 - 2 cos(2 pi k/K) + 2 lam tau^2 omega^2)."""
    k = np.arange(K)
    return float(-0.5 * np.sum(np.log(2.0 - 2.0 * np.cos(2.0 * np.pi * k / K) + 2.0 * lam * tau * tau * omega ** 2)))


def ln_z_bosons_primitive(N: int, M: int, beta: float, dim: int = 2, omega: float = 1.0, lam: float = 0.5) -> float:
    """ln Z_N of N bosons in the trap at M slices: the canonical recursion
    Z_n = (1/n) sum_k z_k Z_{n-k}, with z_k the d-dimensional partition
    function of one k-cycle, a ring of k M beads at the same tau = beta/M."""
    tau = beta / M
    lnz = [0.0] + [dim * _ln_ring(k * M, tau, omega, lam) for k in range(1, N + 1)]
    lnZ = [0.0]
    for n in range(1, N + 1):
        terms = np.array([lnz[k] + lnZ[n - k] for k in range(1, n + 1)])
        mx = terms.max()
        lnZ.append(float(mx + math.log(np.exp(terms - mx).sum()) - math.log(n)))
    return lnZ[N]


def energy_boson_primitive(N: int, M: int, beta: float, dim: int = 2, omega: float = 1.0, lam: float = 0.5,
                           h: float = 1e-5) -> float:
    """-d ln Z_N / d beta at fixed M (the thermodynamic estimator's mean)."""
    return -(ln_z_bosons_primitive(N, M, beta + h, dim, omega, lam)
             - ln_z_bosons_primitive(N, M, beta - h, dim, omega, lam)) / (2 * h)


def radial_expected(N: int, M: int, nb: int, rmax: float, x2: float) -> np.ndarray:
    """Expected counts per radial bin and per measurement of N*M beads whose
    coordinates are i.i.d. Normal(0, x2) in 2D (Rayleigh distribution)."""
    edges = np.linspace(0.0, rmax, nb + 1)
    cdf = 1.0 - np.exp(-edges ** 2 / (2.0 * x2))
    return N * M * np.diff(cdf)
