"""Estimators of the batched sweep kernels against the oracle's (FP64, same
Philox draws, so the kernels and the oracle walk the same trajectory).

* Tile kernel (chains without a pair action): the radial histogram in LDS and
  the reference's Cartesian Density (src/measurement.jl:45-55, first bin
  dropped) bin for bin against o_radial / o_density summed over the oracle's
  measured sweeps; the same after one FP32 sweep against the oracle's
  histograms of the paths the kernel left.
* Whole-chain kernel (pair chains: cell lists, all-pairs with the path in HBM
  and in LDS; permutation cycles): E, Ev and Epair blocks (the reference's
  Energy, src/measurement.jl:92-122, plus the pair energy and pair virial)
  against o_energy + o_pair_energy, and both histograms, with the estimators
  inside the sweep kernel (measure_kernel 1) and in the wide estimator kernel
  (measure_kernel 2).
"""
import numpy as np
import pytest

import oracle as O
from pimc import _capi as A
from pimc.potentials import GaussianPair, HarmonicTrap
from pimc.system import chain_desc
from test_gpu_parity import assert_blocks, oracle_meas, oracle_sys, rand_paths

pytestmark = pytest.mark.gpu


def make(descs, precision, seed, spl, radial_bins, radial_rmax, cart_bins, measure_kernel=0, pair_kernel=0):
    opt = A.Options()
    opt.precision = precision
    opt.sweeps_per_launch = spl
    opt.seed = seed
    opt.radial_bins = radial_bins
    opt.radial_rmax = radial_rmax
    opt.cart_bins = cart_bins
    opt.measure_kernel = measure_kernel
    opt.pair_kernel = pair_kernel
    return A.Context(descs, opt)


def oracle_run(d, r, c, seed, n_sweeps, rb, rmax, cb, swaps=False):
    """The oracle's sweeps of one chain with every sweep measured: the energy
    records and the radial and Cartesian histograms summed over the sweeps."""
    s = oracle_sys(d, r)
    meas, rad, nd_r = [], np.zeros(rb), 0
    cart, nd_c = np.zeros((cb,) * d.dim), 0
    for sw in range(n_sweeps):
        if swaps:
            s.sweep_perm(c, sw, seed, d.ell, d.com_step, d.swaps, d.swap_m, d.swap_rule)
        else:
            s.sweep(c, sw, seed, d.ell, d.com_step)
        meas.append(oracle_meas(s))
        rad, nd_r = s.radial(rb, rmax, rad, nd_r)
        cart, nd_c = s.density(cb, cart, nd_c)
    return s, meas, rad, nd_r, cart, nd_c


def test_tile_histograms_match_oracle_fp64():
    """Tile kernel, FP64, uneven chains in a trap inside a small box (the
    Cartesian bins cover the cloud): radial and Cartesian histograms over two
    launches of three measured sweeps equal the oracle's bin for bin."""
    rng = np.random.default_rng(101)
    descs, inits = [], []
    for N, M, beta, ell in [(5, 32, 1.0, 8), (7, 24, 2.0, 4), (3, 40, 0.7, 6)]:
        descs.append(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=4.0, beta=beta, lam=0.5, ell=ell,
                                com_step=1.0))
        inits.append(np.clip(rng.normal(0, 1.0, size=(N, 2, M)), -3.9, 3.9))
    seed, spl, n_sweeps, rb, rmax, cb = 17, 3, 6, 16, 3.0, 12
    ctx = make(descs, A.FP64, seed, spl, rb, rmax, cb)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    ctx.run(n_sweeps, measure=True)
    for c, (d, r) in enumerate(zip(descs, inits)):
        s, meas, rad, nd_r, cart, nd_c = oracle_run(d, r, c, seed, n_sweeps, rb, rmax, cb)
        assert np.max(np.abs(ctx.get_paths(c) - s.r)) < 1e-9, c
        h, nd = ctx.radial(c)
        assert nd == nd_r == n_sweeps * d.M
        assert np.array_equal(h, rad), (c, h, rad)
        hc, ndc = ctx.cartesian(c)
        assert ndc == nd_c
        assert np.array_equal(hc, cart), c
        assert 0 < cart.sum() <= n_sweeps * d.N * d.M
        assert_blocks(ctx, c, meas, spl)


@pytest.mark.parametrize("cart_bins", [0, 12])
def test_tile_histograms_fp32_one_sweep(cart_bins):
    """FP32 tile kernel, one measured sweep per call: the histograms added by
    that sweep equal the oracle's histograms of the paths the kernel left.
    With cart_bins = 0 the segments that are one skew block of the LDS rows
    take the kernel's unrolled estimator path; the Cartesian binning is FP64
    arithmetic as in the oracle, the radial one FP32, so a bead within
    rounding of a radial bin edge may fall on the other side (at most two)."""
    rng = np.random.default_rng(103)
    descs = [chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=4.0, beta=1.0, lam=0.5, ell=ell, com_step=1.0)
             for N, M, ell in [(6, 32, 8), (4, 64, 16), (5, 24, 6)]]
    rb, rmax = 16, 3.0
    ctx = make(descs, A.FP32, 5, 1, rb, rmax, cart_bins)
    for c, d in enumerate(descs):
        ctx.set_paths(c, np.clip(rng.normal(0, 1.0, size=(d.N, 2, d.M)), -3.9, 3.9))
    prev = [(np.zeros(rb), np.zeros((cart_bins,) * 2) if cart_bins else None) for _ in descs]
    for it in range(3):
        ctx.run(1, measure=True)
        for c, d in enumerate(descs):
            s = oracle_sys(d, ctx.get_paths(c))
            h, _ = ctx.radial(c)
            want, _ = s.radial(rb, rmax)
            got = h - prev[c][0]
            assert got.sum() == want.sum(), (it, c)
            assert np.abs(got - want).sum() <= 4, (it, c, got, want)
            hc = None
            if cart_bins:
                hc, _ = ctx.cartesian(c)
                wc, _ = s.density(cart_bins)
                assert np.array_equal(hc - prev[c][1], wc), (it, c)
            prev[c] = (h, hc)


def pair_descs():
    pair = GaussianPair(1.0, 0.5)
    return [
        # cell lists (N > 96)
        chain_desc(HarmonicTrap(0.3), dim=2, M=16, N=110, L=5.0, beta=2.0, lam=0.5, pair=pair, ell=8, r_a=2.0,
                   com_step=0.5),
        # all-pairs scan, path in LDS
        chain_desc(HarmonicTrap(0.3), dim=2, M=16, N=30, L=3.0, beta=2.0, lam=0.5, pair=pair, ell=8, r_a=1.5,
                   com_step=0.5),
        # all-pairs scan, path in HBM (too large for LDS next to the kernel's buffers)
        chain_desc(HarmonicTrap(0.3), dim=2, M=96, N=96, L=5.0, beta=2.0, lam=0.5, pair=pair, ell=8, r_a=2.0,
                   com_step=0.5),
        # permutation cycles with swap tries, no pair action
        chain_desc(HarmonicTrap(1.0), dim=2, M=12, N=4, L=5.0, beta=2.0, lam=0.5, ell=4, com_step=0.8, swaps=3,
                   swap_m=4),
    ]


@pytest.mark.parametrize("measure_kernel,pair_kernel", [(1, 1), (2, 1), (1, 0)])
def test_whole_chain_estimators_and_histograms_match_oracle_fp64(measure_kernel, pair_kernel):
    """Pair and permutation chains, FP64: the E, Ev and Epair blocks and both
    histograms equal the oracle's, with the whole-chain kernel's estimators
    inside the sweep kernel and in the wide estimator kernel, and (pair_kernel
    0) with the cell-list chain on the window kernels and their estimator
    kernel (k_meas_win)."""
    rng = np.random.default_rng(107)
    descs = pair_descs()
    inits = [rng.uniform(-5.0, 5.0, size=(110, 2, 16)), rng.uniform(-3.0, 3.0, size=(30, 2, 16)),
             rng.uniform(-5.0, 5.0, size=(96, 2, 96)), rand_paths(rng, 4, 12, 5.0, scale=0.6)]
    seed, spl, n_sweeps, rb, rmax, cb = 23, 2, 4, 16, 5.0, 10
    ctx = make(descs, A.FP64, seed, spl, rb, rmax, cb, measure_kernel=measure_kernel, pair_kernel=pair_kernel)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    ctx.run(n_sweeps, measure=True)
    for c, (d, r) in enumerate(zip(descs, inits)):
        s, meas, rad, nd_r, cart, nd_c = oracle_run(d, r, c, seed, n_sweeps, rb, rmax, cb, swaps=d.swaps > 0)
        assert np.max(np.abs(ctx.get_paths(c) - s.r)) < 1e-9, c
        assert_blocks(ctx, c, meas, spl)
        h, nd = ctx.radial(c)
        assert nd == nd_r and np.array_equal(h, rad), (c, h, rad)
        hc, ndc = ctx.cartesian(c)
        assert ndc == nd_c and np.array_equal(hc, cart), c
        if d.pair:
            assert ctx.blocks(c)[2][0] > 0.0
