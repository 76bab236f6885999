"""Physics checks of the HIP path on the GPU.

* the potential action and its adjoint (dS/d beads, dS/d omega, dS/d g)
  against the oracle's action and central finite differences of it
  (h = 1e-4, FP64);
* sampled energies against the closed-form primitive-action energies of the
  trapped ideal gas (FP32 batched and FP64 plain forms);
* d<E>/d omega by reweighting against the closed form;
* thermodynamic and virial energy estimators agree for the interacting
  trapped gas (the virial uses the pair gradient of the adjoint).
"""
import math

import numpy as np
import pytest

import analytic as AN
import oracle as O
from pimc import _capi as A
from pimc.batch import Batch, Workload
from pimc.potentials import GaussianPair, HarmonicTrap, NoPotential, OpticalLattice, SquareLattice
from pimc.system import chain_desc

pytestmark = pytest.mark.gpu


def ctx_for(descs, precision=A.FP64, seed=3, spl=10, derivatives=0):
    opt = A.Options()
    opt.precision = precision
    opt.sweeps_per_launch = spl
    opt.seed = seed
    opt.derivatives = derivatives
    return A.Context(descs, opt)


def oracle_action(d, r):
    s = O.System(r, d.lam, d.L, d.beta, O.pot_from_desc(d),
                 O.make_pair(d.pair, d.g, d.sigma), d.r_a, True)
    return -(s.total_lnV() + s.total_lnU()), s


CASES = {
    "trapped_pair_allpairs": lambda: chain_desc(HarmonicTrap(1.3), dim=2, M=12, N=20, L=100.0, beta=2.0, lam=0.5,
                                                pair=GaussianPair(0.5, 0.5)),
    "box_pair_cells": lambda: chain_desc(NoPotential(), dim=2, M=8, N=120, L=5.0, beta=2.0, lam=0.5,
                                         pair=GaussianPair(1.0, 0.5)),
    "lattice_pair_cells": lambda: chain_desc(SquareLattice(4.0, 1.0), dim=2, M=8, N=110, L=5.0, beta=4.0,
                                             lam=0.5, pair=GaussianPair(1.0, 0.3)),
    "optical_helical": lambda: chain_desc(OpticalLattice([0.0, 1.2, 2.5, 4.0], 0.8, -3.0, helical=True), dim=2,
                                          M=8, N=12, L=5.0, beta=2.0, lam=0.5),
}


def init_for(d, rng):
    if d.potential == A.POT_HARMONIC:
        return np.clip(rng.normal(0, 1.5, size=(d.N, 2, d.M)), -d.L * 0.99, d.L * 0.99)
    return rng.uniform(-d.L, d.L, size=(d.N, 2, d.M))


@pytest.mark.parametrize("case", sorted(CASES))
def test_action_and_adjoint_against_oracle_fd(case):
    d = CASES[case]()
    rng = np.random.default_rng(21)
    r = init_for(d, rng)
    ctx = ctx_for([d])
    ctx.set_paths(0, r)
    S_gpu, grad, par = ctx.adjoint(0)
    S_or, osys = oracle_action(d, r)
    assert S_gpu == pytest.approx(S_or, rel=1e-12, abs=1e-12)
    # beads whose pair distances stay clear of the cut-off under a 1e-4 move
    h = 1e-4
    hc = osys.h
    checked = 0
    for _ in range(200):
        n, e, j = int(rng.integers(d.N)), int(rng.integers(2)), int(rng.integers(d.M))
        rj = r[n, :, j]
        near = False
        for jj in (j,):
            for k in range(d.N):
                if k == n:
                    continue
                dd = math.sqrt(sum(O.distance(r[k, x, jj], rj[x], d.L) ** 2 for x in range(2)))
                if abs(dd - hc) < 10 * h:
                    near = True
        if near:
            continue
        rp, rm = r.copy(), r.copy()
        rp[n, e, j] += h
        rm[n, e, j] -= h
        fd = (oracle_action(d, rp)[0] - oracle_action(d, rm)[0]) / (2 * h)
        assert grad[n, e, j] == pytest.approx(fd, rel=2e-6, abs=2e-8), (n, e, j)
        checked += 1
        if checked >= 25:
            break
    assert checked >= 10
    # parameter derivatives
    if d.potential == A.POT_HARMONIC:
        dp, dm = CASES[case](), CASES[case]()
        dp.omega += h
        dm.omega -= h
        fd = (oracle_action(dp, r)[0] - oracle_action(dm, r)[0]) / (2 * h)
        assert par[0] == pytest.approx(fd, rel=1e-7)
    dp, dm = CASES[case](), CASES[case]()
    dp.g += h
    dm.g -= h
    fd = (oracle_action(dp, r)[0] - oracle_action(dm, r)[0]) / (2 * h)
    assert par[1] == pytest.approx(fd, rel=1e-7)


def test_batched_fp32_energies_within_errors():
    """Trapped ideal-gas chains of the batched form (FP32 paths, FP64
    estimators) against the closed-form primitive-action energy per chain."""
    from pimc.batch import trapped_batch
    wl = trapped_batch(n_chains=256, seed=9, precision=A.FP32)
    b = Batch(wl, sweeps_per_launch=100, seed=5)
    b.run(2000, measure=False)
    b.run(8000, measure=True)
    z = []
    for c, d in enumerate(wl.descs):
        e = b.estimates(c)
        z.append((e.E - AN.energy_boltzmann_primitive(d.N, d.M, d.beta)) / e.E_err)
    z = np.array(z)
    print("pull mean", z.mean(), "chi2/chain", np.mean(z ** 2), "within 2 sigma", np.mean(np.abs(z) < 2))
    assert abs(z.mean()) < 4.0 / math.sqrt(len(z)) * 1.5
    assert np.mean(z ** 2) < 1.6
    assert np.mean(np.abs(z) < 2) > 0.88


def test_plain_fp64_energy_and_virial():
    """One chain N=10, M=64, beta=1 (the plain form), thermodynamic and virial
    estimators against 21.63906... (and the continuum 21.63953...)."""
    d = chain_desc(HarmonicTrap(1.0), dim=2, M=64, N=10, L=100.0, beta=1.0, lam=0.5, ell=8, com_step=1.0)
    ctx = ctx_for([d] * 32, seed=11, spl=100)
    rng = np.random.default_rng(4)
    for c in range(32):
        ctx.set_paths(c, rng.normal(size=(10, 2, 64)))
    ctx.run(2000, measure=False)
    ctx.run(10000, measure=True)
    E = np.array([ctx.estimates(c).E for c in range(32)])
    Ev = np.array([ctx.estimates(c).Ev for c in range(32)])
    exact = AN.energy_boltzmann_primitive(10, 64, 1.0)
    for x in (E, Ev):
        m, e = x.mean(), x.std(ddof=1) / math.sqrt(len(x))
        print(m, e, exact)
        assert abs(m - exact) < 3.5 * e


def test_reweighting_domega_ideal():
    """d<E>/d omega = <dE/d omega> - cov(E, dS/d omega) against the closed form."""
    N, M, beta = 4, 32, 1.0
    d = chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=100.0, beta=beta, lam=0.5, ell=8, com_step=1.0)
    nch = 128
    ctx = ctx_for([d] * nch, seed=12, spl=100, derivatives=1)
    rng = np.random.default_rng(5)
    for c in range(nch):
        ctx.set_paths(c, rng.normal(size=(N, 2, M)))
    ctx.run(1000, measure=False)
    ctx.run(6000, measure=True)
    vals = np.array([ctx.estimates(c).dE_domega for c in range(nch)])
    m, e = vals.mean(), vals.std(ddof=1) / math.sqrt(nch)
    exact = AN.denergy_domega_primitive(N, M, beta)
    print(m, e, exact)
    assert abs(m - exact) < 3.5 * e


def test_virial_matches_thermodynamic_with_pairs():
    """Trapped bosons with the Gaussian repulsion (the differentiation
    workload, scaled down): the virial estimator, built from the gradient of
    the pair action, agrees with the thermodynamic estimator."""
    d = chain_desc(HarmonicTrap(1.0), dim=2, M=64, N=16, L=100.0, beta=2.0, lam=0.5, pair=GaussianPair(0.5, 0.5),
                   ell=8, com_step=1.0)
    nch = 96
    ctx = ctx_for([d] * nch, precision=A.FP32, seed=13, spl=100)
    rng = np.random.default_rng(6)
    for c in range(nch):
        ctx.set_paths(c, np.clip(rng.normal(size=(16, 2, 64)), -99, 99))
    ctx.run(1000, measure=False)
    ctx.run(4000, measure=True)
    est = [ctx.estimates(c) for c in range(nch)]
    E = np.array([e.E for e in est])
    Ev = np.array([e.Ev for e in est])
    Ep = np.array([e.Epair for e in est])
    diff = E - Ev
    m, err = diff.mean(), diff.std(ddof=1) / math.sqrt(nch)
    print("E", E.mean(), "Ev", Ev.mean(), "Epair", Ep.mean(), "diff", m, err)
    assert Ep.mean() > 0.5
    assert abs(m) < 3.5 * err


def _boson_run(rule, nch, sweeps=3000):
    descs = [chain_desc(HarmonicTrap(1.0), dim=2, M=64, N=10, L=100.0, beta=1.0, lam=0.5, ell=8, com_step=1.0,
                        swaps=10, swap_m=8, swap_rule=rule) for _ in range(nch)]
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed, opt.use_graph = A.FP32, 100, 2024 + rule, 1
    ctx = A.Context(descs, opt)
    ctx.init_paths(-1, 7)
    ctx.run(1000, measure=False)
    cnt = ctx.run(sweeps, measure=True)
    E = np.array([ctx.estimates(c).E for c in range(nch)])
    return E.mean(), E.std(ddof=1) / math.sqrt(nch), cnt


def test_reference_swap_rule_is_biased_fp32():
    """The reference's swap acceptance exp(w_updated - w_initial)
    (src/updates/reshape.jl:266) leaves out the ratio of the partner-choice
    normalisations; at N = 10, M = 64, beta = 1 its boson energy is low by
    about 0.5 (kept behind swap_rule = 1, off by default)."""
    m, err, _ = _boson_run(1, 128)
    assert AN.energy_boson_primitive(10, 64, 1.0) - m > 8 * err, (m, err)


def test_boson_energy_with_swaps_fp32():
    """N = 10 trapped bosons, M = 64, beta = 1 (the headline single-chain
    configuration): with swap moves the thermodynamic energy is the
    primitive-action boson value (canonical recursion, 13.50), not the
    distinguishable-particle 21.64."""
    nch = 64
    descs = [chain_desc(HarmonicTrap(1.0), dim=2, M=64, N=10, L=100.0, beta=1.0, lam=0.5, ell=8, com_step=1.0,
                        swaps=10, swap_m=8) for _ in range(nch)]
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed, opt.use_graph = A.FP32, 100, 2024, 1
    ctx = A.Context(descs, opt)
    ctx.init_paths(-1, 7)
    ctx.run(1000, measure=False)
    cnt = ctx.run(3000, measure=True)
    assert cnt.swap_acc > 0.05 * cnt.swap_tries
    E = np.array([ctx.estimates(c).E for c in range(nch)])
    m, err = E.mean(), E.std(ddof=1) / math.sqrt(nch)
    exact = AN.energy_boson_primitive(10, 64, 1.0)
    assert abs(m - exact) < 3.0 * err, (m, err, exact)
    assert AN.energy_boltzmann_primitive(10, 64, 1.0) - m > 20 * err


def _sampled(descs, precision, seed, init, table=None, therm=600, meas=2400):
    ctx = ctx_for(descs, precision=precision, seed=seed, spl=100)
    for c, r in enumerate(init):
        ctx.set_paths(c, r)
        if table is not None:
            ctx.set_pair_table(c, *table)
    ctx.run(therm, measure=False)
    ctx.run(meas, measure=True)
    est = [ctx.estimates(c) for c in range(len(descs))]
    ctx.close()
    E = np.array([e.E for e in est])
    Ep = np.array([e.Epair for e in est])
    return E, Ep


@pytest.mark.parametrize("kind", ["gauss_allpairs", "gauss_cells", "table_hardsphere"])
def test_fp32_pair_sampling_matches_fp64(kind):
    """The FP32 form of the interacting sweeps samples the same distribution
    as FP64: chain-averaged E (and E_pair for the Gaussian) of independent FP32
    and FP64 batches agree within 3.5 combined standard errors, for all-pairs
    and cell-list scans and for the tabulated propagator with hard spheres."""
    from pimc.propagator import PropInt
    rng = np.random.default_rng(71)
    table = None
    if kind == "gauss_allpairs":
        d = chain_desc(HarmonicTrap(1.0), dim=2, M=32, N=16, L=100.0, beta=2.0, lam=0.5,
                       pair=GaussianPair(0.5, 0.5), ell=8, com_step=1.0)
        mk = lambda: np.clip(rng.normal(size=(16, 2, 32)), -99, 99)
    elif kind == "gauss_cells":
        d = chain_desc(NoPotential(), dim=2, M=16, N=110, L=5.0, beta=1.0, lam=0.5, pair=GaussianPair(1.0, 0.5),
                       ell=8, r_a=2.0, com_step=0.5)
        mk = lambda: np.repeat(rng.uniform(-5.0, 5.0, size=(110, 2))[:, :, None], 16, axis=2)
    else:
        tau = 2.0 / 32
        terms = O.prop_int_terms(200, 1e-20, 5.0, 2.0, tau)
        p = PropInt(terms, 1e-20, 5.0, 2.0, tau)
        d = chain_desc(HarmonicTrap(1.0), dim=2, M=32, N=12, L=3.5, beta=2.0, lam=0.5, pair=p, r_a=p.nn_range(),
                       ell=8, com_step=0.5, hs_a=math.exp(-2 * math.pi / 2.0))
        mk = lambda: np.clip(rng.normal(size=(12, 2, 32)), -3.4, 3.4)
        table = (terms, 1e-20, 5.0, 0.0)
    nch = 64
    init = [mk() for _ in range(nch)]
    E32, Ep32 = _sampled([d] * nch, A.FP32, 5, init, table)
    E64, Ep64 = _sampled([d] * nch, A.FP64, 6, init, table)
    se = lambda x: x.std(ddof=1) / math.sqrt(len(x))
    print(kind, E32.mean(), se(E32), E64.mean(), se(E64), Ep32.mean(), Ep64.mean())
    assert abs(E32.mean() - E64.mean()) < 3.5 * math.hypot(se(E32), se(E64))
    if kind != "table_hardsphere":
        assert Ep64.mean() > 0.1
        assert abs(Ep32.mean() - Ep64.mean()) < 3.5 * math.hypot(se(Ep32), se(Ep64))


def _pair_boson_run(ell, swaps, swap_m, seed, nch=128, therm=1000, meas=4000):
    descs = [chain_desc(HarmonicTrap(1.0), dim=2, M=32, N=4, L=100.0, beta=2.0, lam=0.5, pair=GaussianPair(2.0, 0.7),
                        ell=ell, com_step=1.0, swaps=swaps, swap_m=swap_m) for _ in range(nch)]
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed, opt.use_graph = A.FP32, 100, seed, 1
    ctx = A.Context(descs, opt)
    ctx.init_paths(-1, seed + 1)
    ctx.run(therm, measure=False)
    cnt = ctx.run(meas, measure=True)
    E = np.array([ctx.estimates(c).E for c in range(nch)])
    ex = np.array([np.any(ctx.get_next(c) != np.arange(4)) for c in range(nch)], dtype=float)
    return E.mean(), E.std(ddof=1) / math.sqrt(nch), ex.mean(), cnt


def test_pair_swaps_agree_across_move_sets_fp32():
    """Detailed balance of the swap move with a pair action (pair_in_accept =
    1: the pair action's difference): four trapped bosons with a Gaussian
    repulsion, sampled with swap windows of 4 links (staging length 4) and of
    12 links (staging length 8), give the same energy and the same fraction of
    chains in an exchanged sector; the repulsion moves the energy well away
    from the ideal bosons'."""
    mA, eA, xA, cA = _pair_boson_run(4, 4, 4, 301)
    mB, eB, xB, cB = _pair_boson_run(8, 2, 12, 302)
    assert cA.swap_acc > 0.02 * cA.swap_tries and cB.swap_acc > 0.02 * cB.swap_tries
    assert abs(mA - mB) < 4.0 * math.hypot(eA, eB), (mA, eA, mB, eB)
    sx = math.sqrt(max(xA * (1 - xA), 0.05) / 128 + max(xB * (1 - xB), 0.05) / 128)
    assert abs(xA - xB) < 4.0 * sx, (xA, xB)
    ideal = AN.energy_boson_primitive(4, 32, 2.0)
    assert mA - ideal > 8 * eA, (mA, eA, ideal)
