"""Batched chains in one launch, and the synthetic workloads of the sampling path.

Workloads (lambda = hbar^2/2m = 0.5, 2D, hbar = m = k_B = 1; L is the
reference's half-width, box [-L, L)^2):

* trapped_plain   N=10, M=64, beta=1, omega=1, one chain, FP64.  Beads
                  i.i.d. Normal(0, 1/sqrt(omega)).
* trapped_batch   4096 chains, N ~ U[4,64], M ~ U{32,64,128,256},
                  beta ~ U[0.5,4], omega=1, FP32 paths, FP64 estimators.
* gauss_box       N=256 in a periodic box of side 16 (L=8), Gaussian pair
                  g=1, sigma=0.5, cut-off 4 sigma through the cell list,
                  M=128, beta=2, 256 chains, uniform start replicated over slices.
* trapped_gauss   N=64, M=256, beta=2, omega=1, Gaussian g=0.5, sigma=0.5,
                  512 chains (the differentiation workload).
* lattice         N=400, M=512, beta=4, V0=4, a=1, box 20x20 (L=10), Gaussian
                  g=1, sigma=0.3, 64 chains, beads at lattice minima plus
                  Normal(0, 0.1), 256x256 density histogram.
* contact_box     gauss_box's geometry with the reference's own interaction:
                  the pair-product propagator of a contact coupling g0 = 2
                  (build_prop_int on a 600 x 600 grid up to ceil(sqrt(2) L)),
                  hard-sphere radius exp(-2 pi / g0), r_a from the table's
                  0.999 crossing; 256 chains.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._capi import FP32, FP64, Context, Options
from .potentials import GaussianPair, HarmonicTrap, NoPotential, SquareLattice
from .system import chain_desc


@dataclass
class Workload:
    name: str
    descs: list
    init: list            # per chain N x dim x M initial paths (built lazily)
    precision: int
    radial_bins: int = 0
    radial_rmax: float = 0.0
    cart_bins: int = 0
    derivatives: int = 0
    meta: dict = field(default_factory=dict)
    pair_table: tuple | None = None  # (terms, r0, r1, mu) for every PIMC_PAIR_TABLE chain


def _normal_start(rng, N, M, omega, L):
    r = rng.standard_normal((N, 2, M)) / math.sqrt(omega)
    return np.clip(r, -L * 0.999, L * 0.999)


def trapped_plain(seed: int = 1, precision: int = FP64, ell: int = 8) -> Workload:
    rng = np.random.default_rng(seed)
    d = chain_desc(HarmonicTrap(1.0), dim=2, M=64, N=10, L=100.0, beta=1.0, lam=0.5, ell=ell, com_step=1.0)
    return Workload("trapped_plain", [d], [_normal_start(rng, 10, 64, 1.0, 100.0)], precision,
                    radial_bins=64, radial_rmax=4.0)


def trapped_batch(n_chains: int = 4096, seed: int = 2, precision: int = FP32, ell: int = 8) -> Workload:
    rng = np.random.default_rng(seed)
    Ns = rng.integers(4, 65, size=n_chains)
    Ms = rng.choice([32, 64, 128, 256], size=n_chains)
    betas = rng.uniform(0.5, 4.0, size=n_chains)
    descs, init = [], []
    for N, M, b in zip(Ns, Ms, betas):
        descs.append(chain_desc(HarmonicTrap(1.0), dim=2, M=int(M), N=int(N), L=100.0, beta=float(b), lam=0.5,
                                ell=ell, com_step=1.0))
        init.append(_normal_start(rng, int(N), int(M), 1.0, 100.0))
    return Workload("trapped_batch", descs, init, precision, radial_bins=64, radial_rmax=4.0,
                    meta={"N": Ns, "M": Ms, "beta": betas})


def gauss_box(n_chains: int = 256, seed: int = 3, precision: int = FP32, N: int = 256, M: int = 128,
              ell: int = 8) -> Workload:
    rng = np.random.default_rng(seed)
    pair = GaussianPair(1.0, 0.5)
    L = 8.0
    descs, init = [], []
    for _ in range(n_chains):
        descs.append(chain_desc(NoPotential(), dim=2, M=M, N=N, L=L, beta=2.0, lam=0.5, pair=pair, ell=ell,
                                com_step=1.0))
        x = rng.uniform(-L, L, size=(N, 2))
        init.append(np.repeat(x[:, :, None], M, axis=2))
    return Workload("gauss_box", descs, init, precision)


def trapped_gauss(n_chains: int = 512, seed: int = 4, precision: int = FP32, N: int = 64, M: int = 256,
                  ell: int = 8) -> Workload:
    rng = np.random.default_rng(seed)
    pair = GaussianPair(0.5, 0.5)
    descs, init = [], []
    for _ in range(n_chains):
        descs.append(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=100.0, beta=2.0, lam=0.5, pair=pair,
                                ell=ell, com_step=1.0))
        init.append(_normal_start(rng, N, M, 1.0, 100.0))
    return Workload("trapped_gauss", descs, init, precision, derivatives=1)


def lattice(n_chains: int = 64, seed: int = 5, precision: int = FP32, M: int = 512, ell: int = 8) -> Workload:
    rng = np.random.default_rng(seed)
    pair = GaussianPair(1.0, 0.3)
    L = 10.0
    sites = np.array([(x + 0.5, y + 0.5) for y in range(-10, 10) for x in range(-10, 10)])  # 400 minima
    descs, init = [], []
    for _ in range(n_chains):
        descs.append(chain_desc(SquareLattice(4.0, 1.0), dim=2, M=M, N=400, L=L, beta=4.0, lam=0.5, pair=pair,
                                ell=ell, com_step=0.5))
        r = sites[:, :, None] + rng.normal(0.0, 0.1, size=(400, 2, M))
        r = (r + L) % (2 * L) - L
        init.append(r)
    return Workload("lattice", descs, init, precision, cart_bins=256)


def contact_box(n_chains: int = 256, seed: int = 6, precision: int = FP32, N: int = 256, M: int = 128,
                ell: int = 8, g0: float = 2.0) -> Workload:
    from .propagator import build_prop_int
    rng = np.random.default_rng(seed)
    L, beta = 8.0, 2.0
    p = build_prop_int(float(math.ceil(math.sqrt(2.0) * L)), g0, beta / M)
    ra, a = p.nn_range(), math.exp(-2.0 * math.pi / g0)
    descs, init = [], []
    for _ in range(n_chains):
        descs.append(chain_desc(NoPotential(), dim=2, M=M, N=N, L=L, beta=beta, lam=0.5, pair=p, r_a=ra, ell=ell,
                                com_step=1.0, hs_a=a))
        x = rng.uniform(-L, L, size=(N, 2))
        init.append(np.repeat(x[:, :, None], M, axis=2))
    return Workload("contact_box", descs, init, precision, pair_table=(p.terms, p.r0, p.r1, 0.0),
                    meta={"r_a": ra, "hs_a": a})


class Batch:
    """Many independent chains in one context (the batched HIPSweep form).

    autotune=True picks the staging length per chain and the tile and
    workgroup sizes first (pimc.autotune); use_graph=True replays the launch
    sequence from a captured hipGraph; measure_every=k takes the estimators
    after every k-th sweep (the reference's measurement cycle, counted in
    sweeps); measure_kernel chooses where whole chains take their estimators
    (pimc_options.measure_kernel: 0 automatic, 1 in the sweep kernel, 2 a
    wide kernel of their own); pair_kernel chooses the kernels of Gaussian-pair
    chains (pimc_options.pair_kernel: 0 automatic, 1 whole-chain, 2 windows);
    adapt_ell=True lets every chain adapt its staging length between
    thermalising launches (adjust!(::NumbOfSlices), pimc_options.adapt_ell)."""

    def __init__(self, wl: Workload, sweeps_per_launch: int = 50, seed: int = 12345, tile_beads: int = 0,
                 block_threads: int = 0, autotune: bool = False, use_graph: bool = False, tune_sweeps: int = 100,
                 measure_every: int = 1, measure_kernel: int = 0, pair_kernel: int = 0, adapt_ell: bool = False):
        import copy
        descs = [copy.copy(d) for d in wl.descs]
        self.tune = None
        if autotune:
            from .autotune import autotune as _tune
            self.tune = _tune(wl, trial_sweeps=tune_sweeps)
            for d, ell in zip(descs, self.tune.ell):
                d.ell = ell
            tile_beads, block_threads = self.tune.tile_beads, self.tune.block_threads
        opt = Options()
        opt.precision = wl.precision
        opt.sweeps_per_launch = sweeps_per_launch
        opt.radial_bins = wl.radial_bins
        opt.radial_rmax = wl.radial_rmax
        opt.cart_bins = wl.cart_bins
        opt.derivatives = wl.derivatives
        opt.tile_beads = tile_beads
        opt.block_threads = block_threads
        opt.use_graph = 1 if use_graph else 0
        opt.seed = seed
        opt.measure_every = measure_every
        opt.measure_kernel = measure_kernel
        opt.pair_kernel = pair_kernel
        opt.adapt_ell = 1 if adapt_ell else 0
        self.wl = wl
        self.descs = descs
        self.ctx = Context(descs, opt)
        for c, r in enumerate(wl.init):
            self.ctx.set_paths(c, r)
        if wl.pair_table is not None:
            from ._capi import PAIR_TABLE
            for c, d in enumerate(descs):
                if d.pair == PAIR_TABLE:
                    self.ctx.set_pair_table(c, *wl.pair_table)

    def run(self, n_sweeps: int, measure: bool = False):
        return self.ctx.run(n_sweeps, measure)

    def estimates(self, chain: int):
        return self.ctx.estimates(chain)

    def close(self):
        self.ctx.close()
