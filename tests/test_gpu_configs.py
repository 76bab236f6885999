"""The five benchmark configurations at their workload sizes on the GPU.

* C1 (N=10, M=64, beta=1, FP64; radial density of 64 bins): the radial
  histogram against the Rayleigh profile of the primitive-action <x^2>
  (the energies are in test_gpu_physics.py).
* C2 (4096 chains, N ~ U[4,64], M ~ U{32..256}, beta ~ U[0.5,4], FP32,
  autotuned, hipGraph): 2k + 10k sweeps, per-chain pull of the energy
  against the closed-form primitive-action value.
* C3 (N=256 bosons, M=128, box side 16, Gaussian g=1 sigma=0.5 through the
  cell list): two chains, three FP64 sweeps against the oracle (paths,
  counters, E/Ev/Epair blocks from the wide estimator kernel); one FP32
  sweep against the oracle's schedule on the FP32 start.
* C4 (N=64, M=256, beta=2, trapped, Gaussian g=0.5 sigma=0.5): the adjoint
  of a full-size chain against central differences of the oracle's action;
  d<E>/dg and d<E>/d omega by reweighting against finite differences of <E>
  between batches at g +- h and omega +- h; virial = thermodynamic energy.
* C5 (N=400, M=512, beta=4, square lattice V0=4 a=1, box 20x20, Gaussian
  g=1 sigma=0.3, 256x256 density): one chain, two FP64 sweeps against the
  oracle, histogram bin for bin.
"""
import copy
import math

import numpy as np
import pytest

import analytic as AN
import oracle as O
from pimc import _capi as A
from pimc import batch as B
from pimc.potentials import HarmonicTrap
from pimc.system import chain_desc
from test_gpu_parity import assert_blocks, oracle_meas, oracle_sys

pytestmark = pytest.mark.gpu


def hotelling_pvalue(per, want, ncoarse=8):
    """p-value of Hotelling's T^2 test that the chains' mean histogram (rows of
    per: one independent chain each) equals want, over ncoarse groups of
    adjacent bins with about equal expected counts (the bins of one chain are
    correlated, so a per-bin chi^2 is not calibrated)."""
    from scipy.stats import f
    cw = np.cumsum(want) / want.sum()
    groups = np.split(np.arange(len(want)), np.searchsorted(cw, np.linspace(0, 1, ncoarse + 1)[1:-1]))
    X = np.stack([per[:, g].sum(1) for g in groups], 1)
    mu = np.array([want[g].sum() for g in groups])
    n, p = X.shape
    d = X.mean(0) - mu
    T2 = n * d @ np.linalg.solve(np.cov(X, rowvar=False), d)
    return float(1.0 - f.cdf(T2 * (n - p) / (p * (n - 1)), p, n - p))


def test_c1_radial_density_against_closed_form():
    """C1 in FP64, 32 chains x (2k + 10k sweeps): the 64-bin radial histogram
    per measurement against N M times the Rayleigh bin probabilities with
    the primitive-action <x^2> = (1/M) sum_k 1/a_k per coordinate, by
    Hotelling's T^2 over the independent chains; a profile of 3% wider <x^2>
    is rejected."""
    d = chain_desc(HarmonicTrap(1.0), dim=2, M=64, N=10, L=100.0, beta=1.0, lam=0.5, ell=8, com_step=1.0)
    nch, nb, rmax = 32, 64, 4.0
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed = A.FP64, 100, 31
    opt.radial_bins, opt.radial_rmax = nb, rmax
    ctx = A.Context([d] * nch, opt)
    rng = np.random.default_rng(8)
    for c in range(nch):
        ctx.set_paths(c, rng.normal(size=(10, 2, 64)))
    ctx.run(2000, measure=False)
    ctx.run(10000, measure=True)
    per = []
    for c in range(nch):
        h, nd = ctx.radial(c)
        assert nd == 10000 * 64
        per.append(h / (nd // 64))  # counts per measurement of N M beads
    per = np.array(per)
    x2 = AN.x2_primitive(64, 1.0)
    want = AN.radial_expected(10, 64, nb, rmax, x2)
    pv = hotelling_pvalue(per, want)
    pw = hotelling_pvalue(per, AN.radial_expected(10, 64, nb, rmax, 1.03 * x2))
    print("radial Hotelling p-value", pv, "with <x^2> 3% wider", pw)
    assert pv > 1e-3
    assert pw < 1e-6


def test_c2_batched_autotuned_energies():
    """C2 at its size: 4096 chains, autotuned staging length, tile and
    workgroup size, hipGraph replay, 2k thermalising + 10k measuring sweeps;
    each chain's energy against its closed form."""
    wl = B.trapped_batch(n_chains=4096, seed=2, precision=A.FP32)
    b = B.Batch(wl, sweeps_per_launch=100, seed=20240601, autotune=True, use_graph=True)
    b.run(2000, measure=False)
    cnt = b.run(10000, measure=True)
    assert cnt.kernel_launches == 100
    z = []
    for c, d in enumerate(wl.descs):
        e = b.estimates(c)
        assert e.n_blocks == 100 and e.E_err > 0
        z.append((e.E - AN.energy_boltzmann_primitive(d.N, d.M, d.beta)) / e.E_err)
    z = np.array(z)
    print("ell", sorted(set(b.tune.ell)), "tile", b.tune.tile_beads, b.tune.block_threads,
          "pull mean", z.mean(), "chi2/chain", np.mean(z ** 2), "within 2 sigma", np.mean(np.abs(z) < 2))
    b.close()
    assert abs(z.mean()) < 0.1
    assert np.mean(z ** 2) < 1.25
    assert np.mean(np.abs(z) < 2) > 0.92


def test_c3_gauss_box_matches_oracle():
    """C3 at its size, FP64: two chains, three measured sweeps (the wide
    estimator kernel, automatic for 32768 beads) against the oracle's
    schedule; then one FP32 sweep from the FP32-rounded start: identical
    accept counts, paths to FP32 rounding of the box."""
    wl = B.gauss_box(n_chains=2, seed=3, precision=A.FP64)
    seed, n = 41, 3
    b = B.Batch(wl, sweeps_per_launch=n, seed=seed)
    cnt = b.run(n, measure=True)
    la = ca = 0
    for c, (d, r) in enumerate(zip(wl.descs, wl.init)):
        s = oracle_sys(d, r)
        meas = []
        for sw in range(n):
            st = s.sweep(c, sw, seed, d.ell, d.com_step)
            la += st.levy_acc
            ca += st.com_acc
            meas.append(oracle_meas(s))
        assert np.max(np.abs(b.ctx.get_paths(c) - s.r)) < 1e-9, c
        assert_blocks(b.ctx, c, meas, n)
    assert (cnt.levy_acc, cnt.com_acc) == (la, ca)
    assert 0 < cnt.com_acc < cnt.com_tries
    b.close()
    # FP32, one sweep
    wl32 = B.gauss_box(n_chains=2, seed=3, precision=A.FP32)
    init = [r.astype(np.float32).astype(np.float64) for r in wl32.init]
    wl32.init = init
    b = B.Batch(wl32, sweeps_per_launch=1, seed=seed)
    cnt = b.run(1, measure=False)
    la = ca = 0
    worst = 0.0
    for c, (d, r) in enumerate(zip(wl32.descs, init)):
        s = oracle_sys(d, r)
        st = s.sweep(c, 0, seed, d.ell, d.com_step)
        la += st.levy_acc
        ca += st.com_acc
        diff = b.ctx.get_paths(c) - s.r
        diff -= 2 * d.L * np.round(diff / (2 * d.L))
        worst = max(worst, float(np.max(np.abs(diff))))
    print("C3 FP32 one sweep: max |r32 - r64| =", worst)
    assert (cnt.levy_acc, cnt.com_acc) == (la, ca)
    assert worst < 1e-5 * 8.0
    b.close()


def _c4_run(g=0.5, omega=1.0, nch=256, therm=500, meas=2000, seed=77):
    wl = B.trapped_gauss(n_chains=nch, seed=4, precision=A.FP32)
    for d in wl.descs:
        d.g = g
        d.omega = omega
    b = B.Batch(wl, sweeps_per_launch=100, seed=seed, use_graph=True)
    b.run(therm, measure=False)
    b.run(meas, measure=True)
    est = [b.estimates(c) for c in range(nch)]
    b.close()
    return est


def test_c4_adjoint_of_a_full_size_chain():
    """C4 chain (N=64, M=256): S_pot, dS/d(beads), dS/d omega and dS/dg of
    the adjoint kernel against the oracle's action and its central
    differences (h = 1e-4, FP64)."""
    wl = B.trapped_gauss(n_chains=1, seed=4, precision=A.FP64)
    d, r = wl.descs[0], wl.init[0]
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed, opt.derivatives = A.FP64, 10, 3, 1
    ctx = A.Context([d], opt)
    ctx.set_paths(0, r)
    ctx.run(20, measure=False)  # a sampled configuration, not the i.i.d. start
    r = ctx.get_paths(0)
    S_gpu, grad, par = ctx.adjoint(0)

    def action(dd, rr):
        s = O.System(rr, dd.lam, dd.L, dd.beta, O.pot_from_desc(dd), O.make_pair(dd.pair, dd.g, dd.sigma),
                     dd.r_a, True)
        return -(s.total_lnV() + s.total_lnU())

    assert S_gpu == pytest.approx(action(d, r), rel=1e-12)
    h = 1e-4
    rng = np.random.default_rng(5)
    for _ in range(12):  # the cut-off (25) exceeds every distance: no neighbour set changes
        n, e, j = int(rng.integers(64)), int(rng.integers(2)), int(rng.integers(256))
        rp, rm = r.copy(), r.copy()
        rp[n, e, j] += h
        rm[n, e, j] -= h
        fd = (action(d, rp) - action(d, rm)) / (2 * h)
        assert grad[n, e, j] == pytest.approx(fd, rel=1e-6, abs=1e-8), (n, e, j)
    for k, name in enumerate(("omega", "g")):
        dp, dm = copy.copy(d), copy.copy(d)
        setattr(dp, name, getattr(d, name) + h)
        setattr(dm, name, getattr(d, name) - h)
        fd = (action(dp, r) - action(dm, r)) / (2 * h)
        assert par[k] == pytest.approx(fd, rel=1e-7), name


def test_c4_reweighting_derivatives_and_virial():
    """C4 batches (256 chains, FP32 paths, FP64 sums): d<E>/dg and
    d<E>/d omega by reweighting (<dE/dx> - cov(E, dS/dx)) against central
    differences of <E_virial> between batches at g +- 0.1 and omega +- 0.05
    (the virial estimator has the same mean as the thermodynamic one and
    less noise); at the central point the two energy estimators agree."""
    est0 = _c4_run()
    E = np.array([e.E for e in est0])
    Ev = np.array([e.Ev for e in est0])
    sem = lambda x: float(np.std(x, ddof=1) / math.sqrt(len(x)))
    diff = E - Ev
    print("C4 E", E.mean(), sem(E), "Ev", Ev.mean(), sem(Ev), "Epair", np.mean([e.Epair for e in est0]))
    assert abs(diff.mean()) < 3.5 * sem(diff)
    rw_g = np.array([e.dE_dg for e in est0])
    rw_w = np.array([e.dE_domega for e in est0])
    hg, hw = 0.1, 0.05
    ep = np.array([e.Ev for e in _c4_run(g=0.5 + hg, seed=78)])
    em = np.array([e.Ev for e in _c4_run(g=0.5 - hg, seed=79)])
    fd_g, fd_g_err = (ep.mean() - em.mean()) / (2 * hg), math.hypot(sem(ep), sem(em)) / (2 * hg)
    wp = np.array([e.Ev for e in _c4_run(omega=1.0 + hw, seed=80)])
    wm = np.array([e.Ev for e in _c4_run(omega=1.0 - hw, seed=81)])
    fd_w, fd_w_err = (wp.mean() - wm.mean()) / (2 * hw), math.hypot(sem(wp), sem(wm)) / (2 * hw)
    print("dE/dg reweighting", rw_g.mean(), sem(rw_g), "fd", fd_g, fd_g_err)
    print("dE/domega reweighting", rw_w.mean(), sem(rw_w), "fd", fd_w, fd_w_err)
    assert abs(rw_g.mean() - fd_g) < 3.5 * math.hypot(sem(rw_g), fd_g_err)
    assert abs(rw_w.mean() - fd_w) < 3.5 * math.hypot(sem(rw_w), fd_w_err)
    assert abs(fd_g) > 4 * fd_g_err and abs(fd_w) > 4 * fd_w_err  # both derivatives are resolved


def test_c5_lattice_matches_oracle_fp64():
    """C5 at its size (one chain, FP64): two measured sweeps against the
    oracle's schedule, with the 256 x 256 Cartesian density (the reference's
    Density) bin for bin and the E, Ev, Epair blocks."""
    wl = B.lattice(n_chains=1, seed=5, precision=A.FP64)
    d, r = wl.descs[0], wl.init[0]
    seed, n = 43, 2
    b = B.Batch(wl, sweeps_per_launch=n, seed=seed)
    cnt = b.run(n, measure=True)
    s = oracle_sys(d, r)
    meas, cart, ndc = [], np.zeros((256, 256)), 0
    la = ca = 0
    for sw in range(n):
        st = s.sweep(0, sw, seed, d.ell, d.com_step)
        la += st.levy_acc
        ca += st.com_acc
        meas.append(oracle_meas(s))
        cart, ndc = s.density(256, cart, ndc)
    assert (cnt.levy_acc, cnt.com_acc) == (la, ca)
    assert np.max(np.abs(b.ctx.get_paths(0) - s.r)) < 1e-9
    assert_blocks(b.ctx, 0, meas, n)
    hc, nd = b.ctx.cartesian(0)
    assert nd == ndc == n * 512
    assert np.array_equal(hc, cart)
    assert cart.sum() > 0.99 * n * 400 * 512
    b.close()


def test_adapted_staging_lengths_keep_the_energies_fp32():
    """The batched trapped workload (FP32 tiles) with adapt_ell: the chains
    leave the shared length 8 for lengths of their own (tiles rebuilt between
    thermalising launches, graph replay on), within [2, min(32, M - 2)], and the
    measured energies stay on the primitive-action closed form (per-chain
    pulls of 16 blocks)."""
    wl = B.trapped_batch(n_chains=192, seed=9)
    b = B.Batch(wl, sweeps_per_launch=20, seed=5, use_graph=True, adapt_ell=True)
    b.run(400)
    ells = np.array([b.ctx.ell(c) for c in range(len(wl.descs))])
    assert np.mean(ells != 8) > 0.5
    assert np.all(ells >= 2) and np.all(ells <= np.minimum(32, wl.meta["M"] - 2))
    b.run(16 * 20, measure=True)
    assert np.array_equal(ells, [b.ctx.ell(c) for c in range(len(wl.descs))])
    pulls = []
    for c, d in enumerate(wl.descs):
        e = b.estimates(c)
        exact = AN.energy_boltzmann_primitive(d.N, d.M, d.beta)
        pulls.append((e.E - exact) / e.E_err)
    pulls = np.array(pulls)
    assert np.mean(np.abs(pulls) < 2.5) > 0.9, np.sort(np.abs(pulls))[-10:]
    assert abs(np.mean(pulls)) < 0.35
