"""Seeded random contexts against the oracle, FP64: one dimension (1 or 2)
per context, 2-5 chains of every class (no pair, Gaussian pair, tabulated
propagator with or without hard spheres; swap tries under either rule and
pair mode; move weights; traps, square and optical lattices, no potential;
cut-off radii down to one cell a side), a Gaussian chain the window kernels
take beside them in half the contexts, and random run options:
sweeps_per_launch, measure_every, adapt_ell, graph replay, pair_kernel,
measure_kernel, reweighting derivatives and the starting sweep word.
Thermalising launches then measuring ones: paths, next[], every counter, the
adapted COM steps and staging lengths and the estimator blocks equal the
oracle's schedule."""
import math

import numpy as np
import pytest

import oracle as O
from pimc import _capi as A
from pimc.potentials import GaussianPair, HarmonicTrap, NoPotential, OpticalLattice, SquareLattice
from pimc.propagator import PropInt
from pimc.system import chain_desc
from test_gpu_parity import oracle_meas, oracle_sys

pytestmark = pytest.mark.gpu

KEYS = ("levy_acc", "levy_tries", "com_acc", "com_tries", "swap_acc", "swap_tries")


def any_chain(rng, dim):
    M = int(rng.choice([6, 8, 12, 16, 24]))
    N = int(rng.integers(1, 30))
    L = float(rng.uniform(1.5, 4.0))
    beta = float(rng.uniform(0.5, 3.0))
    ell = int(rng.integers(2, M))
    kind = str(rng.choice(["gauss", "table", "none"]))
    swaps = int(rng.integers(0, 5)) if N >= 2 else 0
    every = tuple(int(x) for x in rng.integers(1, 4, size=3))
    hs = float(rng.uniform(0.05, 0.2)) if (kind == "table" and rng.random() < 0.4) else 0.0
    ra = 2 * L / int(rng.integers(1, 5))
    u = rng.random()
    if u < 0.4:
        pot = HarmonicTrap(float(rng.uniform(0.3, 1.5)))
    elif u < 0.6:
        pot = SquareLattice(float(rng.uniform(0.5, 3.0)), float(rng.uniform(0.5, 1.5)))
    elif u < 0.8 or dim == 1:
        pot = NoPotential()
    else:
        pot = OpticalLattice([2 * math.pi * k / 4 for k in range(4)], 0.5, float(rng.uniform(1.0, 4.0)))
    if kind == "gauss":
        pair = GaussianPair(float(rng.uniform(0.5, 3.0)), float(rng.uniform(0.2, 0.6)))
    elif kind == "table":
        pair = PropInt(np.zeros((2, 2)), 0.0, 1.0, 4.0, beta / M)  # kind marker; the table is set separately
    else:
        pair = None
    d = chain_desc(pot, dim=dim, M=M, N=N, L=L, beta=beta, lam=0.5, pair=pair, ell=ell, r_a=ra if pair else 0.0,
                   com_step=float(rng.uniform(0.2, 1.0)), swaps=swaps, swap_m=int(rng.integers(2, M)),
                   swap_rule=int(rng.integers(0, 2)), hs_a=hs, pair_in_accept=bool(rng.integers(0, 2)),
                   staging_every=every[0], com_every=every[1], swap_every=every[2])
    r = rng.uniform(-L, L, size=(N, dim, 1)).repeat(M, axis=2) + rng.normal(0, 0.1, size=(N, dim, M))
    return d, (r + L) % (2 * L) - L, kind, every


def window_chain(rng, dim):
    """A Gaussian chain with cell lists (N = 110, five cells a side), the
    pair action in the acceptance: the window kernels' by default."""
    every = tuple(int(x) for x in rng.integers(1, 4, size=3))
    d = chain_desc(HarmonicTrap(0.3), dim=dim, M=16, N=110, L=5.0, beta=2.0, lam=0.5, pair=GaussianPair(1.0, 0.5),
                   ell=int(rng.integers(2, 9)), r_a=2.0, com_step=0.5, staging_every=every[0],
                   com_every=every[1], swap_every=every[2])
    return d, rng.uniform(-5.0, 5.0, size=(110, dim, 16)), "gauss", every


def table_of(d):
    r1 = float(math.ceil(math.sqrt(2.0) * d.L))
    return O.prop_int_terms(48, 1e-20, r1, 4.0, d.beta / d.M), r1


def context(trial):
    rng = np.random.default_rng(9000 + trial)
    dim = int(rng.integers(1, 3))
    chains = [any_chain(rng, dim) for _ in range(int(rng.integers(2, 6)))]
    if rng.random() < 0.5:
        chains.append(window_chain(rng, dim))
    spl = int(rng.choice([2, 3, 4]))
    opt = dict(spl=spl, measure_every=int(rng.choice([k for k in (1, 2, 3, 4) if spl % k == 0])),
               adapt_ell=int(rng.random() < 0.4), use_graph=int(rng.random() < 0.5),
               pair_kernel=int(rng.integers(0, 3)), measure_kernel=int(rng.integers(0, 3)),
               sweep0=int(rng.integers(0, 7)), n_therm=int(rng.integers(1, 4)), n_meas=int(rng.integers(1, 3)))
    # reweighting derivatives (every chain then runs whole; refused with a tabulated propagator)
    opt["derivatives"] = int(rng.random() < 0.3 and all(k != "table" for _, _, k, _ in chains))
    return chains, opt


@pytest.mark.parametrize("trial", range(24))
def test_random_context_with_options_matches_oracle_fp64(trial):
    chains, o = context(trial)
    seed, spl = 600 + trial, o["spl"]
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed = A.FP64, spl, seed
    opt.measure_every, opt.adapt_ell, opt.use_graph = o["measure_every"], o["adapt_ell"], o["use_graph"]
    opt.pair_kernel, opt.measure_kernel, opt.derivatives = o["pair_kernel"], o["measure_kernel"], o["derivatives"]
    print(f"trial {trial}: options {o}; chains {[(k, d.N, d.M, d.dim, d.swaps, d.hs_a > 0) for d, _, k, _ in chains]}",
          flush=True)
    ctx = A.Context([d for d, _, _, _ in chains], opt)
    ctx.set_sweep(o["sweep0"])
    refs = []
    for c, (d, r, kind, _) in enumerate(chains):
        ctx.set_paths(c, r)
        if kind == "table":
            tab, r1 = table_of(d)
            ctx.set_pair_table(c, tab, 1e-20, r1, 0.4)
            refs.append(O.System(r, d.lam, d.L, d.beta, O.pot_from_desc(d),
                                 O.make_pair(O.PAIR_TABLE, table=tab, r0=1e-20, r1=r1, mu=0.4), d.r_a,
                                 bool(d.pair_in_accept), hs_a=d.hs_a))
        else:
            refs.append(oracle_sys(d, r))
    n_therm, n_meas, me = o["n_therm"], o["n_meas"], o["measure_every"]
    c1 = ctx.run(n_therm * spl, measure=False)
    ells = [ctx.ell(c) for c in range(len(chains))]
    c2 = ctx.run(n_meas * spl, measure=True)
    grand = dict.fromkeys(KEYS, 0)
    for c, ((d, _, _, every), s) in enumerate(zip(chains, refs)):
        step, ell, sweep = d.com_step, d.ell, o["sweep0"]
        tot = dict.fromkeys(KEYS, 0)
        blocks = []
        for launch in range(n_therm + n_meas):
            measure = launch >= n_therm
            s.hs_ctr = 1000 if measure else 10000
            la = lt = ca = ct = 0
            meas = []
            for _ in range(spl):
                st = s.sweep_every(c, sweep, seed, ell, step, every, d.swaps, d.swap_m, d.swap_rule)
                for key in tot:
                    tot[key] += getattr(st, key)
                la += st.levy_acc
                lt += st.levy_tries
                ca += st.com_acc
                ct += st.com_tries
                if measure and (sweep + 1) % me == 0:
                    meas.append(oracle_meas(s))
                sweep += 1
            if measure:
                blocks.append(np.mean([m[0] for m in meas]))
            else:
                step = O.adapt_step(step, ca, ct, d.L)
                if o["adapt_ell"]:
                    ell = O.adapt_ell(ell, la, lt, d.M)
        assert np.max(np.abs(ctx.get_paths(c) - s.r)) < 1e-9, c
        if d.swaps:
            assert np.array_equal(ctx.get_next(c), s.next), c
        assert list(ctx.counters(c)) == [tot["levy_acc"], tot["levy_tries"], tot["com_acc"], tot["com_tries"]], c
        assert ctx.com_step(c) == pytest.approx(step, rel=1e-12), c
        assert ells[c] == ell and ctx.ell(c) == ell, (c, ells[c], ell)
        E = ctx.blocks(c)[0]
        assert len(E) == n_meas, c
        for b in range(n_meas):
            assert abs(E[b] - blocks[b]) < 1e-8 * max(1.0, abs(blocks[b])), (c, b, E[b], blocks[b])
        for key in KEYS:
            grand[key] += tot[key]
    for key in KEYS:
        assert getattr(c1, key) + getattr(c2, key) == grand[key], key
