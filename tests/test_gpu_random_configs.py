"""Seeded random small configurations against the oracle, FP64: a potential
(harmonic trap or square lattice), a pair action (none, Gaussian, tabulated
propagator with or without hard spheres), swap tries (0-4, random window
length, both acceptance rules), pair_in_accept 0 or 1, random move weights
and staging lengths.  Two thermalising launches and a measuring one: paths,
next[], every counter and the first energy block equal the oracle's."""
import math

import numpy as np
import pytest

import oracle as O
from pimc import _capi as A
from pimc.potentials import GaussianPair, HarmonicTrap, SquareLattice
from pimc.propagator import PropInt
from pimc.system import chain_desc
from test_gpu_parity import oracle_meas, oracle_sys

pytestmark = pytest.mark.gpu


def case(trial):
    rng = np.random.default_rng(1000 + trial)
    M = int(rng.choice([8, 12, 16, 24]))
    N = int(rng.integers(2, 40))
    L = float(rng.uniform(1.5, 4.0))
    beta = float(rng.uniform(0.5, 3.0))
    ell = int(rng.integers(2, min(9, M - 1)))
    kind = str(rng.choice(["gauss", "table", "none"]))
    swaps, sm, rule, pia = int(rng.integers(0, 5)), int(rng.integers(2, M)), int(rng.integers(0, 2)), int(rng.integers(0, 2))
    every = tuple(int(x) for x in rng.integers(1, 4, size=3))
    hs = float(rng.uniform(0.05, 0.2)) if (kind == "table" and rng.random() < 0.5) else 0.0
    ra = 2 * L / int(rng.integers(2, 5))
    pot = HarmonicTrap(float(rng.uniform(0.3, 1.5))) if rng.random() < 0.6 else SquareLattice(2.0, 1.0)
    if kind == "gauss":
        pair = GaussianPair(float(rng.uniform(0.5, 3.0)), float(rng.uniform(0.2, 0.6)))
    elif kind == "table":
        pair = PropInt(np.zeros((2, 2)), 0.0, 1.0, 4.0, beta / M)  # kind marker; the table is set separately
    else:
        pair = None
    d = chain_desc(pot, dim=2, M=M, N=N, L=L, beta=beta, lam=0.5, pair=pair, ell=ell, r_a=ra if pair else 0.0,
                   com_step=0.5, swaps=swaps, swap_m=sm, swap_rule=rule, hs_a=hs, pair_in_accept=bool(pia),
                   staging_every=every[0], com_every=every[1], swap_every=every[2])
    r = rng.uniform(-L, L, size=(N, 1, 1)).repeat(2, axis=1).repeat(M, axis=2) + rng.normal(0, 0.1, size=(N, 2, M))
    return d, (r + L) % (2 * L) - L, kind, every


def oracle_of(d, r, kind, tab, r1):
    if kind == "table":
        return O.System(r, d.lam, d.L, d.beta, O.pot_from_desc(d),
                        O.make_pair(O.PAIR_TABLE, table=tab, r0=1e-20, r1=r1, mu=0.4), d.r_a, bool(d.pair_in_accept),
                        hs_a=d.hs_a)
    return oracle_sys(d, r)


def table_of(d):
    r1 = float(math.ceil(math.sqrt(2.0) * d.L))
    return O.prop_int_terms(48, 1e-20, r1, 4.0, d.beta / d.M), r1


def run_and_check(chains, seed, spl, pair_kernel=0):
    """chains: (desc, paths, kind, every).  Two thermalising launches and a
    measuring one; every chain against the oracle under its own chain index."""
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed, opt.pair_kernel = A.FP64, spl, seed, pair_kernel
    ctx = A.Context([d for d, _, _, _ in chains], opt)
    refs = []
    for c, (d, r, kind, every) in enumerate(chains):
        ctx.set_paths(c, r)
        tab, r1 = table_of(d) if kind == "table" else (None, None)
        if kind == "table":
            ctx.set_pair_table(c, tab, 1e-20, r1, 0.4)
        refs.append(oracle_of(d, r, kind, tab, r1))
    c1 = ctx.run(2 * spl, measure=False)
    c2 = ctx.run(spl, measure=True)
    keys = ("levy_acc", "levy_tries", "com_acc", "com_tries", "swap_acc", "swap_tries")
    grand = dict.fromkeys(keys, 0)
    for c, ((d, _, _, every), s) in enumerate(zip(chains, refs)):
        step, meas = d.com_step, []
        tot = dict.fromkeys(keys, 0)
        for launch in range(3):
            s.hs_ctr = 10000 if launch < 2 else 1000
            ca = ct = 0
            for k in range(spl):
                st = s.sweep_every(c, launch * spl + k, seed, d.ell, step, every, d.swaps, d.swap_m or d.ell,
                                   d.swap_rule)
                for key in tot:
                    tot[key] += getattr(st, key)
                ca += st.com_acc
                ct += st.com_tries
                if launch == 2:
                    meas.append(oracle_meas(s))
            if launch < 2:
                step = O.adapt_step(step, ca, ct, d.L)
        assert np.max(np.abs(ctx.get_paths(c) - s.r)) < 1e-9, c
        if d.swaps:
            assert np.array_equal(ctx.get_next(c), s.next), c
        assert list(ctx.counters(c)) == [tot["levy_acc"], tot["levy_tries"], tot["com_acc"], tot["com_tries"]], c
        Eo = np.mean([m[0] for m in meas])
        assert abs(ctx.blocks(c)[0][0] - Eo) < 1e-8 * max(1.0, abs(Eo)), c
        for key in keys:
            grand[key] += tot[key]
    for key in keys:
        assert getattr(c1, key) + getattr(c2, key) == grand[key], key


@pytest.mark.parametrize("trial", range(24))
def test_random_configuration_matches_oracle_fp64(trial):
    d, r, kind, every = case(trial)
    run_and_check([(d, r, kind, every)], 77 + trial, 3)


def big_gauss(trial, pair_in_accept=True):
    """A Gaussian-pair chain with cell lists (N = 110, 2L / r_a = 5 cells a
    side): the window kernels take it by default when its pair action enters
    the acceptance."""
    rng = np.random.default_rng(5000 + trial)
    every = tuple(int(x) for x in rng.integers(1, 4, size=3))
    d = chain_desc(HarmonicTrap(0.3), dim=2, M=16, N=110, L=5.0, beta=2.0, lam=0.5,
                   pair=GaussianPair(1.0, 0.5), ell=int(rng.integers(2, 9)), r_a=2.0, com_step=0.5,
                   staging_every=every[0], com_every=every[1], swap_every=every[2], pair_in_accept=pair_in_accept)
    return d, rng.uniform(-5.0, 5.0, size=(110, 2, 16)), "gauss", every


@pytest.mark.parametrize("trial", range(8))
def test_random_mixed_contexts_match_oracle_fp64(trial):
    """Four random chains and a window-kernel chain in one context (tile,
    whole-chain, table, swap and window launches side by side), pair_kernel
    0 or 2 (small Gaussian chains without swaps then run on the window
    kernels too): every chain equals the oracle under its own chain index."""
    chains = [case(100 + 4 * trial + k) for k in range(4)] + [big_gauss(trial)]
    run_and_check(chains, 300 + trial, 3, pair_kernel=0 if trial % 2 else 2)


@pytest.mark.parametrize("pair_kernel", [0, 2])
def test_reference_pair_semantics_match_oracle_fp64(pair_kernel):
    """A Gaussian-pair chain with cell lists and pair_in_accept = 0 (the
    reference's interaction_action! sums computed and dropped from the
    acceptance) beside one with the pair action in the acceptance: the first
    samples without its pair action whichever pair kernel is asked for, and
    both equal the oracle."""
    run_and_check([big_gauss(11, pair_in_accept=False), big_gauss(12)], 401, 3, pair_kernel=pair_kernel)
