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


@pytest.mark.parametrize("trial", range(24))
def test_random_configuration_matches_oracle_fp64(trial):
    d, r, kind, every = case(trial)
    seed, spl = 77 + trial, 3
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed = A.FP64, spl, seed
    ctx = A.Context([d], opt)
    ctx.set_paths(0, r)
    if kind == "table":
        r1 = float(math.ceil(math.sqrt(2.0) * d.L))
        tab = O.prop_int_terms(48, 1e-20, r1, 4.0, d.beta / d.M)
        ctx.set_pair_table(0, tab, 1e-20, r1, 0.4)
        s = O.System(r, d.lam, d.L, d.beta, O.pot_from_desc(d),
                     O.make_pair(O.PAIR_TABLE, table=tab, r0=1e-20, r1=r1, mu=0.4), d.r_a, bool(d.pair_in_accept),
                     hs_a=d.hs_a)
    else:
        s = oracle_sys(d, r)
    c1 = ctx.run(2 * spl, measure=False)
    c2 = ctx.run(spl, measure=True)
    step, meas = d.com_step, []
    tot = dict(levy_acc=0, levy_tries=0, com_acc=0, com_tries=0, swap_acc=0, swap_tries=0)
    for launch in range(3):
        s.hs_ctr = 10000 if launch < 2 else 1000
        ca = ct = 0
        for k in range(spl):
            st = s.sweep_every(0, launch * spl + k, seed, d.ell, step, every, d.swaps, d.swap_m or d.ell, d.swap_rule)
            for key in tot:
                tot[key] += getattr(st, key)
            ca += st.com_acc
            ct += st.com_tries
            if launch == 2:
                meas.append(oracle_meas(s))
        if launch < 2:
            step = O.adapt_step(step, ca, ct, d.L)
    assert np.max(np.abs(ctx.get_paths(0) - s.r)) < 1e-9
    if d.swaps:
        assert np.array_equal(ctx.get_next(0), s.next)
    for key in tot:
        assert getattr(c1, key) + getattr(c2, key) == tot[key], key
    Eo = np.mean([m[0] for m in meas])
    assert abs(ctx.blocks(0)[0][0] - Eo) < 1e-8 * max(1.0, abs(Eo))
