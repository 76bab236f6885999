"""Move weights of run! (src/simulation.jl:33-36) on the batched sweep: the
staging pass, the centre-of-mass pass and the swap tries run on the sweeps s
with s % every == 0.  The tile kernel, the whole-chain kernel (pair and swap
chains) and the window kernels against the oracle's o_sweep_every in FP64:
paths to 1e-9, every counter, the adapted COM steps and the estimator blocks;
launches whose sweeps start off a multiple of every, and launches in which a
pass never runs (no COM tries: the step stays)."""
import numpy as np
import pytest

import oracle as O
from pimc import _capi as A
from pimc.potentials import GaussianPair, HarmonicTrap
from pimc.system import chain_desc
from test_gpu_parity import assert_blocks, make_ctx, oracle_meas, oracle_sys, rand_paths

pytestmark = pytest.mark.gpu


def every_of(d):
    return (max(1, d.staging_every), max(1, d.com_every), max(1, d.swap_every))


def oracle_weighted(descs, inits, seed, n_sweeps, spl, measure, sweep0=0):
    """The oracle's sweeps with the chains' move weights, launches of spl
    sweeps, the COM-step rule between thermalising launches on the tries that
    ran; per chain (paths, next, totals, step, measurements)."""
    out = []
    for c, (d, r) in enumerate(zip(descs, inits)):
        s = oracle_sys(d, r)
        step, sweep, left = d.com_step, sweep0, n_sweeps
        tot = dict(levy_acc=0, levy_tries=0, com_acc=0, com_tries=0, swap_acc=0, swap_tries=0, bead_moves=0)
        meas = []
        while left > 0:
            k = min(spl, left)
            ca = ct = 0
            for _ in range(k):
                st = s.sweep_every(c, sweep, seed, d.ell, step, every_of(d), d.swaps, d.swap_m or d.ell,
                                   d.swap_rule)
                for key in tot:
                    tot[key] += getattr(st, key)
                ca += st.com_acc
                ct += st.com_tries
                sweep += 1
                if measure:
                    meas.append(oracle_meas(s))
            if not measure:
                step = O.adapt_step(step, ca, ct, d.L)
            left -= k
        out.append((s.r, s.next, tot, step, meas))
    return out


def check(ctx, cnt, descs, ref, measure, spl):
    for c, d in enumerate(descs):
        r, nxt, tot, step, meas = ref[c]
        assert np.max(np.abs(ctx.get_paths(c) - r)) < 1e-9, c
        if d.swaps:
            assert np.array_equal(ctx.get_next(c), nxt), c
        assert list(ctx.counters(c)) == [tot["levy_acc"], tot["levy_tries"], tot["com_acc"], tot["com_tries"]], c
        assert abs(ctx.com_step(c) - step) < 1e-12, (c, ctx.com_step(c), step)
        if measure:
            assert_blocks(ctx, c, meas, spl)
    for key in ("levy_acc", "levy_tries", "com_acc", "com_tries", "swap_acc", "swap_tries"):
        assert getattr(cnt, key) == sum(x[2][key] for x in ref), key


@pytest.mark.parametrize("measure", [False, True])
def test_weights_tile_chains_fp64(measure):
    """Tile chains of one (M, ell) with different weights (separate tiles)
    and one of another shape: launches of 5 sweeps from sweep 3 (14 sweeps
    thermalising, the last launch shorter; 15 measuring, whole blocks)."""
    rng = np.random.default_rng(61)
    specs = [(5, 32, 8, 1, 1), (9, 32, 8, 2, 3), (3, 40, 6, 3, 1), (7, 32, 8, 1, 2), (4, 32, 8, 2, 3)]
    descs, inits = [], []
    for N, M, ell, se, ce in specs:
        descs.append(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=100.0, beta=1.5, lam=0.5, ell=ell,
                                com_step=1.0, staging_every=se, com_every=ce))
        inits.append(rand_paths(rng, N, M, 100.0))
    seed, n_sweeps, spl, s0 = 17, 15 if measure else 14, 5, 3
    ctx = make_ctx(descs, A.FP64, seed=seed, spl=spl)
    ctx.set_sweep(s0)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    cnt = ctx.run(n_sweeps, measure=measure)
    ref = oracle_weighted(descs, inits, seed, n_sweeps, spl, measure, sweep0=s0)
    check(ctx, cnt, descs, ref, measure, spl)
    assert cnt.bead_moves == sum(x[2]["bead_moves"] for x in ref)


def test_weights_tile_chains_fp32_counts_and_rigid_rings():
    """FP32 tiles (the lane-group COM path, ell = 16 dividing M): with the
    staging pass off every ring moves rigidly; the tries are those of the
    sweeps that ran each pass."""
    rng = np.random.default_rng(62)
    descs = [chain_desc(HarmonicTrap(1.0), dim=2, M=64, N=12, L=100.0, beta=2.0, lam=0.5, ell=16, com_step=1.0,
                        staging_every=1000),
             chain_desc(HarmonicTrap(1.0), dim=2, M=64, N=12, L=100.0, beta=2.0, lam=0.5, ell=16, com_step=1.0,
                        com_every=3)]
    inits = [rand_paths(rng, 12, 64, 100.0, scale=0.5) for _ in descs]
    ctx = make_ctx(descs, A.FP32, seed=5, spl=10)
    ctx.set_sweep(1)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    ctx.run(10, measure=True)
    d = ctx.get_paths(0) - inits[0]
    assert np.allclose(d, d[:, :, :1], atol=1e-5)
    assert np.abs(d).max() > 0.01
    c0, c1 = ctx.counters(0), ctx.counters(1)
    assert c0[1] == 0 and c0[3] == 10 * 12
    assert c1[1] == 10 * 12 * 4 and c1[3] == 3 * 12  # sweeps 3, 6, 9 of 1..10


@pytest.mark.parametrize("pair_kernel", [1, 2])
@pytest.mark.parametrize("measure", [False, True])
def test_weights_pair_chains_fp64(pair_kernel, measure):
    """Gaussian-pair chains with weights (2, 3) and (3, 2) in one launch, on
    the whole-chain kernel and on the window kernels (N = 110, cell lists)."""
    rng = np.random.default_rng(63)
    pair = GaussianPair(1.0, 0.5)
    descs, inits = [], []
    for se, ce in [(2, 3), (3, 2)]:
        descs.append(chain_desc(HarmonicTrap(0.3), dim=2, M=16, N=110, L=5.0, beta=2.0, lam=0.5, pair=pair, ell=8,
                                r_a=2.0, com_step=0.5, staging_every=se, com_every=ce))
        inits.append(rng.uniform(-5.0, 5.0, size=(110, 2, 16)))
    seed, n_sweeps, spl = 23, 6, 3
    ctx = make_ctx(descs, A.FP64, seed=seed, spl=spl, pair_kernel=pair_kernel)
    ctx.set_sweep(1)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    cnt = ctx.run(n_sweeps, measure=measure)
    ref = oracle_weighted(descs, inits, seed, n_sweeps, spl, measure, sweep0=1)
    check(ctx, cnt, descs, ref, measure, spl)


def test_weights_swap_chains_fp64():
    """Permutation cycles and swap tries with weights: paths, next[], every
    counter against the oracle."""
    rng = np.random.default_rng(64)
    specs = [(4, 12, 4, 3, 4, 0, 2.0, (1, 2, 3)), (5, 8, 4, 4, 3, 1, 3.0, (2, 1, 2))]
    descs, inits = [], []
    for N, M, ell, nsw, m, rule, beta, (se, ce, we) in specs:
        descs.append(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=100.0, beta=beta, lam=0.5, ell=ell,
                                com_step=0.8, swaps=nsw, swap_m=m, swap_rule=rule, staging_every=se, com_every=ce,
                                swap_every=we))
        inits.append(rand_paths(rng, N, M, 100.0, scale=0.6))
    seed, n_sweeps, spl = 31, 12, 12
    ctx = make_ctx(descs, A.FP64, seed=seed, spl=spl)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    cnt = ctx.run(n_sweeps, measure=True)
    ref = oracle_weighted(descs, inits, seed, n_sweeps, spl, True)
    check(ctx, cnt, descs, ref, True, spl)
    assert cnt.swap_tries == 3 * 4 + 4 * 6 and cnt.swap_acc > 0
