"""The tabulated pair-product propagator on the GPU against the oracle.

pimc_prop_int_terms (device quadrature of src/propagator.jl:36-70) against
the oracle's table; PIMC_PAIR_TABLE chains (lnU = log prop_int, src/system.jl:
26-28) through the single-move hook and the whole-chain sweep kernel, FP64,
against the oracle's link sums and sweep schedule.

Tolerances: tables to 1e-10 of max(1, |terms|) (the device's Bessel functions
and summation order differ from the host's); link sums to 1e-12 of
sum |lnU|; sweep paths to 1e-9.
"""
import math

import numpy as np
import pytest

import oracle as O
from pimc import _capi as A
from pimc.propagator import PropInt, build_prop_int
from pimc.potentials import GaussianPair, HarmonicTrap
from pimc.system import System, chain_desc
from test_gpu_parity import assert_blocks, oracle_meas, oracle_sys

pytestmark = pytest.mark.gpu


def table_for(L, g0, tau, n):
    r1 = float(math.ceil(math.sqrt(2.0) * L))
    return O.prop_int_terms(n, 1e-20, r1, g0, tau), 1e-20, r1


def table_desc(N, L, M, ra, beta=2.0, pot=None, ell=4, com_step=0.5, hs_a=0.0, dim=2):
    pi = PropInt(np.zeros((2, 2)), 0.0, 1.0, 4.0, beta / M)  # kind marker; the table is set separately
    return chain_desc(pot if pot is not None else HarmonicTrap(0.3), dim=dim, M=M, N=N, L=L, beta=beta, lam=0.5,
                      pair=pi, ell=ell, r_a=ra, com_step=com_step, hs_a=hs_a)


def oracle_table_sys(d, r, tab, r0, r1, mu):
    pot = O.pot_from_desc(d)
    pair = O.make_pair(O.PAIR_TABLE, table=tab, r0=r0, r1=r1, mu=mu)
    return O.System(r, d.lam, d.L, d.beta, pot, pair, d.r_a, bool(d.pair_in_accept), hs_a=d.hs_a)


def ctx_for(descs, precision, seed=7, spl=4):
    opt = A.Options()
    opt.precision = precision
    opt.sweeps_per_launch = spl
    opt.seed = seed
    return A.Context(descs, opt)


@pytest.mark.parametrize("n,g0,tau", [(40, 4.0, 0.05), (70, 20.0, 0.01), (130, 1.0, 0.25)])
def test_prop_int_terms_match_oracle(n, g0, tau):
    gpu = A.prop_int_terms(n, 1e-20, 4.0, g0, tau)
    ref = O.prop_int_terms(n, 1e-20, 4.0, g0, tau)
    assert np.all(np.isfinite(gpu))
    assert np.max(np.abs(gpu - ref) / np.maximum(1.0, np.abs(ref))) < 1e-10
    assert np.array_equal(gpu, gpu.T) or np.max(np.abs(gpu - gpu.T)) < 1e-12 * np.max(np.abs(gpu))


def test_reference_size_table():
    """The reference's grid (600 x 600 on [1e-20, L]) for tau = 0.01: entries
    at sampled grid points equal the oracle's point values; prop_int(r, r) < 1
    at short range (repulsive), -> 1 at long range; nn_range where it crosses
    0.999 (determine_nnrange, src/system.jl:10-15)."""
    L, g0, tau = 6.0, 4.0, 0.01
    p = build_prop_int(L, g0, tau)
    assert p.terms.shape == (600, 600)
    grid = 1e-20 + np.arange(600) * ((L - 1e-20) / 599)
    rng = np.random.default_rng(1)
    for i, k in rng.integers(0, 600, size=(12, 2)):
        ref = O.prop_int_term(grid[i], grid[k], g0, tau, L)
        assert abs(p.terms[i, k] - ref) < 1e-10 * max(1.0, abs(ref)), (i, k)
    c = 4 * math.pi * tau
    assert 1 + c * p.terms[5, 5] < 0.999 and abs(c * p.terms[-1, -1]) < 1e-6
    ra = p.nn_range()
    assert 0.0 < ra < L
    assert 1 + c * p._diag(ra) == pytest.approx(0.999, abs=1e-9)


@pytest.mark.parametrize("N,L,ra", [(30, 3.0, 1.5), (110, 5.0, 2.0)])
def test_table_interaction_hook_matches_oracle(N, L, ra):
    rng = np.random.default_rng(21)
    M, beta, mu = 8, 2.0, 0.4
    d = table_desc(N, L, M, ra, beta=beta)
    tab, r0, r1 = table_for(L, 4.0, beta / M, 48)
    r = rng.uniform(-L, L, size=(N, 2, M))
    ctx = ctx_for([d], A.FP64)
    ctx.set_paths(0, r)
    with pytest.raises(A.PimcError):
        ctx.interaction(0, 0, 0)  # no table yet
    ctx.set_pair_table(0, tab, r0, r1, mu)
    s = oracle_table_sys(d, r, tab, r0, r1, mu)
    for _ in range(25):
        n, j = int(rng.integers(N)), int(rng.integers(M))
        w = ctx.interaction(0, n, j)
        wo = s.interaction_old(n, j)
        assert abs(w - wo) <= 1e-12 * max(1.0, abs(wo) * 10), (n, j, w, wo)


@pytest.mark.parametrize("N,L,ra", [(30, 3.0, 1.5), (110, 5.0, 2.0)])
def test_table_sweeps_match_oracle_fp64(N, L, ra):
    """Whole-chain kernel with the tabulated propagator in the acceptance
    (all-pairs with LDS-resident paths for N = 30, cell lists for N = 110),
    next to a Gaussian pair chain in the same context: paths and counters
    equal the oracle's sweep schedule."""
    rng = np.random.default_rng(23)
    M, beta, mu, seed, n_sweeps = 16, 2.0, 0.4, 5, 3
    dt = table_desc(N, L, M, ra, beta=beta, ell=8)
    dg = chain_desc(HarmonicTrap(0.3), dim=2, M=M, N=N, L=L, beta=beta, lam=0.5, pair=GaussianPair(1.0, 0.5), ell=8,
                    r_a=ra, com_step=0.5)
    tab, r0, r1 = table_for(L, 4.0, beta / M, 64)
    inits = [rng.uniform(-L, L, size=(N, 2, M)) for _ in range(2)]
    ctx = ctx_for([dg, dt], A.FP64, seed=seed, spl=n_sweeps)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    with pytest.raises(A.PimcError):
        ctx.run(n_sweeps, measure=False)  # the table chain has no table yet
    ctx.set_pair_table(1, tab, r0, r1, mu)
    cnt = ctx.run(n_sweeps, measure=True)
    la = ca = 0
    for c, d in enumerate([dg, dt]):
        if c == 0:
            pot = O.pot_from_desc(d)
            s = O.System(inits[c], d.lam, d.L, d.beta, pot, O.make_pair(d.pair, d.g, d.sigma), d.r_a, True)
        else:
            s = oracle_table_sys(d, inits[c], tab, r0, r1, mu)
        meas = []
        for sw in range(n_sweeps):
            st = s.sweep(c, sw, seed, d.ell, d.com_step)
            la += st.levy_acc
            ca += st.com_acc
            meas.append(oracle_meas(s))
        assert np.max(np.abs(ctx.get_paths(c) - s.r)) < 1e-9, c
        # one block of n_sweeps measurements; the table chain has no pair energy, so its E is
        # the reference's Energy
        assert_blocks(ctx, c, meas, n_sweeps)
    assert cnt.levy_acc == la and cnt.com_acc == ca
    assert 0 < cnt.levy_acc < cnt.levy_tries
    E, Ev, Ep = ctx.blocks(1)
    assert Ep[0] == 0.0 and ctx.blocks(0)[2][0] > 0.0


@pytest.mark.parametrize("N,L,ra,hs_a", [(30, 3.0, 1.5, 0.3), (110, 5.0, 2.0, 0.25), (240, 5.0, 2.0, 0.12)])
def test_hard_sphere_sweeps_match_oracle_fp64(N, L, ra, hs_a):
    """hardspherelevy! and move_polymer! in the whole-chain kernel (first draws
    tested in parallel, a segment walked again by one wave from its first bead
    too close, keyed redraws) against the oracle's sweep schedule: thermalising
    launches (10000 draws allowed), then measuring ones (1000), FP64: paths and
    counters.  The last case has about 9.6 beads per cell, so most stencils
    hold cells with more than 8 members."""
    hard_sphere_case(N, L, ra, hs_a, 2)


@pytest.mark.parametrize("N,L,ra,hs_a", [(40, 10.0, 2.0, 0.15), (120, 20.0, 2.0, 0.1)])
def test_hard_sphere_sweeps_1d_match_oracle_fp64(N, L, ra, hs_a):
    """The same in 1D (3-cell stencils): all-pairs for N = 40, cell lists with
    about 6 beads per cell for N = 120."""
    hard_sphere_case(N, L, ra, hs_a, 1)


def hard_sphere_case(N, L, ra, hs_a, dim):
    rng = np.random.default_rng(31)
    M, beta, mu, seed, spl = 16, 2.0, 0.4, 11, 2
    d = table_desc(N, L, M, ra, beta=beta, ell=8, hs_a=hs_a, dim=dim)
    tab, r0, r1 = table_for(L, 4.0, beta / M, 64)
    r = rng.uniform(-L, L, size=(N, dim, M))
    ctx = ctx_for([d], A.FP64, seed=seed, spl=spl)
    ctx.set_paths(0, r)
    ctx.set_pair_table(0, tab, r0, r1, mu)
    with pytest.raises(A.PimcError):
        ctx.reshape_linear(0, 0, 0, 3, np.zeros((2, dim)), 0.5)  # single moves refuse hard-sphere chains
    c1 = ctx.run(2 * spl, measure=False)
    c2 = ctx.run(spl, measure=True)
    s = oracle_table_sys(d, r, tab, r0, r1, mu)
    tot = dict(levy_acc=0, levy_tries=0, com_acc=0, com_tries=0)
    step = d.com_step
    for launch in range(3):
        s.hs_ctr = 10000 if launch < 2 else 1000
        ca = ct = 0
        for k in range(spl):
            st = s.sweep(0, launch * spl + k, seed, d.ell, step)
            for key in tot:
                tot[key] += getattr(st, key)
            ca += st.com_acc
            ct += st.com_tries
        if launch < 2:
            step = O.adapt_step(step, ca, ct, d.L)
    assert np.max(np.abs(ctx.get_paths(0) - s.r)) < 1e-9
    for key in tot:
        assert getattr(c1, key) + getattr(c2, key) == tot[key], key
    assert tot["levy_acc"] > 0 and tot["com_acc"] > 0


@pytest.mark.parametrize("precision", [A.FP64, A.FP32])
def test_init_paths_with_hard_spheres_matches_oracle(precision):
    """pimc_init_paths on chains with hard spheres == the oracle's init_world
    with a > 0 (src/system.jl:44-64: particles placed in order, each redrawn
    while a bead comes closer than a to an earlier particle's at that slice):
    the same paths, no two beads of a slice closer than a; a chain without
    hard spheres keeps the plain init_world; the cell lists are rebuilt, so a
    sweep with the table runs afterwards."""
    M, beta, seed = 16, 2.0, 21
    descs = [table_desc(12, 1.5, M, 1.5, beta=beta, ell=4, hs_a=0.3),
             table_desc(30, 3.0, M, 1.5, beta=beta, ell=4),
             table_desc(110, 5.0, M, 2.0, beta=beta, ell=8, hs_a=0.25)]
    ctx = ctx_for(descs, precision, spl=2)
    ctx.init_paths(-1, seed)
    tol = 1e-12 if precision == A.FP64 else 1e-5
    redrawn = 0
    for c, d in enumerate(descs):
        ref, tries, rc = O.init_world_hs(seed, c, d.N, d.M, d.dim, d.L, d.beta / d.M, d.lam, d.hs_a)
        assert rc == 0
        redrawn += int((tries > 1).sum())
        got = ctx.get_paths(c)
        diff = np.abs(got - ref)
        diff = np.minimum(diff, 2 * d.L - diff)  # a bead at the wall may round to the other side in FP32
        assert np.max(diff) <= tol * max(1.0, d.L), (c, np.max(diff))
        if d.hs_a > 0:
            best = np.inf
            for i in range(d.N):
                dd = np.abs(got[:i] - got[i][None])
                dd = np.minimum(dd, 2 * d.L - dd)
                if i:
                    best = min(best, float(np.sqrt((dd ** 2).sum(axis=1)).min()))
            assert best >= d.hs_a - tol * d.L, (c, best)
    assert redrawn > 0  # the redraw loop ran
    for c, d in enumerate(descs):
        tab, r0, r1 = table_for(d.L, 2.0, beta / M, 64)
        ctx.set_pair_table(c, tab, r0, r1, 0.0)
    cnt = ctx.run(2, measure=False)
    assert cnt.levy_acc > 0


def test_table_chain_fp32_runs():
    rng = np.random.default_rng(29)
    N, L, M, ra = 40, 3.0, 16, 1.5
    d = table_desc(N, L, M, ra, beta=2.0, ell=8)
    tab, r0, r1 = table_for(L, 4.0, 2.0 / M, 64)
    ctx = ctx_for([d], A.FP32, spl=10)
    ctx.set_paths(0, rng.uniform(-L, L, size=(N, 2, M)))
    ctx.set_pair_table(0, tab, r0, r1, 0.4)
    ctx.run(40, measure=False)
    cnt = ctx.run(40, measure=True)
    assert 0 < cnt.levy_acc < cnt.levy_tries and 0 < cnt.com_acc < cnt.com_tries
    assert np.all(np.isfinite(ctx.get_paths(0)))


def test_table_refusals():
    d = table_desc(10, 3.0, 8, 1.5)
    ctx = ctx_for([d], A.FP64)
    with pytest.raises(A.PimcError):
        ctx.adjoint(0)
    with pytest.raises(A.PimcError):
        ctx.set_pair_table(0, np.zeros((3, 4)), 0.0, 1.0)
    with pytest.raises(A.PimcError):
        ctx.set_pair_table(0, np.zeros((1, 1)), 0.0, 1.0)
    ctx.set_pair_table(0, np.zeros((4, 4)), 0.0, 1.0)
    dg = chain_desc(HarmonicTrap(0.3), dim=2, M=8, N=10, L=3.0, beta=2.0, lam=0.5, pair=GaussianPair(1.0, 0.5),
                    r_a=1.5)
    ctx2 = ctx_for([dg], A.FP64)
    with pytest.raises(A.PimcError):
        ctx2.set_pair_table(0, np.zeros((4, 4)), 0.0, 1.0)
    d0 = table_desc(10, 3.0, 8, 1.5)
    d0.r_a = 0.0
    with pytest.raises(A.PimcError):
        ctx_for([d0], A.FP64)
    opt = A.Options()
    opt.precision = A.FP64
    opt.derivatives = 1
    with pytest.raises(A.PimcError):
        A.Context([d], opt)
    with pytest.raises(A.PimcError):
        A.prop_int_terms(1, 0.0, 1.0, 4.0, 0.1)


def test_system_hard_sphere_radius_from_g():
    """System(...; interactions=true, g) sets a = exp(-2 pi / g) (src/system.jl:148);
    without g the reference's default g = 0 gives a = 0."""
    from pimc import HIPSweep, run
    T, M, L = 1.0, 16, 4.0
    p = build_prop_int(6.0, 4.0, 1.0 / (T * M), n=200)
    s = System(HarmonicTrap(1.0), lam=0.5, M=M, N=6, L=L, T=T, interactions=True, propint=p, g=4.0, seed=5)
    assert s.a == pytest.approx(math.exp(-2 * math.pi / 4.0))
    run(s, 20, [(1, HIPSweep(s))])
    assert np.all(np.isfinite(s.paths))
    s0 = System(HarmonicTrap(1.0), lam=0.5, M=M, N=6, L=L, T=T, interactions=True, propint=p, seed=5)
    assert s0.a == 0.0


def test_system_with_interactions_and_propint():
    """System(potential; interactions=true, propint=build_prop_int(...)) as in
    examples/density_SRL_lattice.jl:16-20: r_a from determine_nnrange, the
    sweep samples with the table on the device."""
    from pimc import HIPSweep, run
    T, M, L = 1.0, 16, 4.0
    p = build_prop_int(round(math.sqrt(2 * L * L) + 0.5), 4.0, 1.0 / (T * M), n=200)
    s = System(HarmonicTrap(1.0), lam=0.5, M=M, N=6, L=L, T=T, interactions=True, propint=p, seed=3)
    assert s.r_a == pytest.approx(p.nn_range())
    sweep = HIPSweep(s)
    run(s, 20, [(1, sweep)])
    assert np.all(np.isfinite(s.paths))


@pytest.mark.parametrize("N,L,ra,hs_a,pia", [(30, 3.0, 1.5, 0.0, 1), (30, 3.0, 1.5, 0.3, 0), (110, 5.0, 2.0, 0.25, 1),
                                             (110, 5.0, 2.0, 0.25, 0)])
def test_table_swap_sweeps_match_oracle_fp64(N, L, ra, hs_a, pia):
    """Swap tries on chains with the tabulated propagator, with and without
    hard spheres (the reference's interacting bosons, System(...;
    interactions=true) with ReshapeSwapLinear): the bridges of a swap keep
    clear of every particle but the partner (hardspherelevy!(r1, s, j0,
    fpcycle2)), the pair sums as the reference writes them (pair_in_accept
    0) or as the pair action's difference (1), both swap rules, all-pairs and
    cell lists; thermalising launches (10000 draws), then a measuring one
    (1000), FP64: paths, next[], every counter and the energy blocks equal
    the oracle's."""
    rng = np.random.default_rng(71 + N)
    M, beta, mu, seed, spl = 16, 2.0, 0.4, 13, 2
    descs = []
    for rule, m in ((0, 6), (1, 9)):
        d = table_desc(N, L, M, ra, beta=beta, ell=8, hs_a=hs_a)
        d.swaps, d.swap_m, d.swap_rule, d.pair_in_accept = 6, m, rule, pia
        descs.append(d)
    tab, r0, r1 = table_for(L, 4.0, beta / M, 64)
    inits = [rng.uniform(-L, L, size=(N, 1, 1)).repeat(2, axis=1).repeat(M, axis=2) +
             rng.normal(0, 0.1, size=(N, 2, M)) for _ in descs]
    inits = [(r + L) % (2 * L) - L for r in inits]
    ctx = ctx_for(descs, A.FP64, seed=seed, spl=spl)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
        ctx.set_pair_table(c, tab, r0, r1, mu)
    c1 = ctx.run(2 * spl, measure=False)
    c2 = ctx.run(spl, measure=True)
    tot = dict(levy_acc=0, levy_tries=0, com_acc=0, com_tries=0, swap_acc=0, swap_tries=0)
    for c, (d, r) in enumerate(zip(descs, inits)):
        s = oracle_table_sys(d, r, tab, r0, r1, mu)
        step, meas = d.com_step, []
        for launch in range(3):
            s.hs_ctr = 10000 if launch < 2 else 1000
            ca = ct = 0
            for k in range(spl):
                st = s.sweep_perm(c, launch * spl + k, seed, d.ell, step, d.swaps, d.swap_m, d.swap_rule)
                for key in tot:
                    tot[key] += getattr(st, key)
                ca += st.com_acc
                ct += st.com_tries
                if launch == 2:
                    meas.append(oracle_meas(s))
            if launch < 2:
                step = O.adapt_step(step, ca, ct, d.L)
        assert np.array_equal(ctx.get_next(c), s.next), c
        assert np.max(np.abs(ctx.get_paths(c) - s.r)) < 1e-9, c
        assert_blocks(ctx, c, meas, spl)
    for key in tot:
        assert getattr(c1, key) + getattr(c2, key) == tot[key], (key, getattr(c1, key) + getattr(c2, key), tot[key])
    assert tot["swap_acc"] > 0


def test_every_chain_class_in_one_context_fp64():
    """One context holding every kind of chain: tile chains with and without
    move weights, Gaussian-pair chains on the window kernels (cell lists) and
    on the whole-chain kernel, swap chains without and with the Gaussian pair,
    tabulated-propagator chains without swaps and with swaps and hard spheres
    (the TPERM instantiation): the launch lists, tiles and instantiations
    keep every chain equal to the oracle (paths, next[], counters, COM steps,
    energy blocks) over two thermalising launches and a measuring one."""
    rng = np.random.default_rng(91)
    M, beta, mu, seed, spl = 16, 2.0, 0.4, 17, 2
    gp = GaussianPair(1.0, 0.5)
    tab, r0, r1 = table_for(3.0, 4.0, beta / M, 64)
    descs, inits, is_tab = [], [], []

    def add(d, r, t=False):
        descs.append(d)
        inits.append(r)
        is_tab.append(t)

    add(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=5, L=100.0, beta=beta, lam=0.5, ell=8, com_step=1.0),
        rng.normal(0, 0.7, size=(5, 2, M)))
    add(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=6, L=100.0, beta=beta, lam=0.5, ell=8, com_step=1.0,
                   staging_every=2, com_every=3), rng.normal(0, 0.7, size=(6, 2, M)))
    add(chain_desc(HarmonicTrap(0.3), dim=2, M=M, N=110, L=5.0, beta=beta, lam=0.5, pair=gp, ell=8, r_a=2.0,
                   com_step=0.5), rng.uniform(-5.0, 5.0, size=(110, 2, M)))
    add(chain_desc(HarmonicTrap(0.3), dim=2, M=M, N=30, L=3.0, beta=beta, lam=0.5, pair=gp, ell=8, r_a=1.5,
                   com_step=0.5), rng.uniform(-3.0, 3.0, size=(30, 2, M)))
    add(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=4, L=100.0, beta=beta, lam=0.5, ell=4, com_step=0.8, swaps=3,
                   swap_m=5), rng.normal(0, 0.6, size=(4, 2, M)))
    add(chain_desc(HarmonicTrap(0.5), dim=2, M=M, N=12, L=1.5, beta=beta, lam=0.5, pair=gp, ell=4, r_a=1.5,
                   com_step=0.4, swaps=4, swap_m=6, swap_every=2),
        (rng.uniform(-1.5, 1.5, size=(12, 1, 1)).repeat(2, axis=1).repeat(M, axis=2)
         + rng.normal(0, 0.1, size=(12, 2, M)) + 1.5) % 3.0 - 1.5)
    add(table_desc(30, 3.0, M, 1.5, beta=beta, ell=8), rng.uniform(-3.0, 3.0, size=(30, 2, M)), True)
    dt = table_desc(30, 3.0, M, 1.5, beta=beta, ell=8, hs_a=0.3)
    dt.swaps, dt.swap_m = 5, 6
    add(dt, (rng.uniform(-3.0, 3.0, size=(30, 1, 1)).repeat(2, axis=1).repeat(M, axis=2)
             + rng.normal(0, 0.1, size=(30, 2, M)) + 3.0) % 6.0 - 3.0, True)
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed = A.FP64, spl, seed
    ctx = A.Context(descs, opt)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
        if is_tab[c]:
            ctx.set_pair_table(c, tab, r0, r1, mu)
    c1 = ctx.run(2 * spl, measure=False)
    c2 = ctx.run(spl, measure=True)
    tot = dict(levy_acc=0, levy_tries=0, com_acc=0, com_tries=0, swap_acc=0, swap_tries=0)
    for c, (d, r) in enumerate(zip(descs, inits)):
        s = oracle_table_sys(d, r, tab, r0, r1, mu) if is_tab[c] else oracle_sys(d, r)
        step, meas = d.com_step, []
        every = (max(1, d.staging_every), max(1, d.com_every), max(1, d.swap_every))
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
        if d.swaps:
            assert np.array_equal(ctx.get_next(c), s.next), c
        assert np.max(np.abs(ctx.get_paths(c) - s.r)) < 1e-9, c
        assert abs(ctx.com_step(c) - step) < 1e-12, c
        assert_blocks(ctx, c, meas, spl)
    for key in tot:
        assert getattr(c1, key) + getattr(c2, key) == tot[key], (key, getattr(c1, key) + getattr(c2, key), tot[key])


def test_table_swaps_with_hard_spheres_fp32_run():
    """FP32 swap tries with the tabulated propagator and hard spheres (cell
    lists): the permutation stays a permutation, every try is counted, some
    are accepted, and the paths end finite and inside the box."""
    rng = np.random.default_rng(93)
    M, beta, L = 16, 2.0, 5.0
    d = table_desc(120, L, M, 2.0, beta=beta, ell=8, hs_a=0.2)
    d.swaps, d.swap_m = 12, 6
    tab, r0, r1 = table_for(L, 4.0, beta / M, 64)
    ctx = ctx_for([d], A.FP32, seed=5, spl=10)
    x = rng.uniform(-L, L, size=(120, 2, 1)).repeat(M, axis=2)
    ctx.set_paths(0, (x + rng.normal(0, 0.05, size=x.shape) + L) % (2 * L) - L)
    ctx.set_pair_table(0, tab, r0, r1, 0.4)
    ctx.run(40, measure=False)
    cnt = ctx.run(40, measure=True)
    assert sorted(ctx.get_next(0)) == list(range(120))
    assert cnt.swap_tries == 40 * 12 and cnt.swap_acc > 0
    r = ctx.get_paths(0)
    assert np.all(np.isfinite(r)) and np.all(r >= -L) and np.all(r < L)
