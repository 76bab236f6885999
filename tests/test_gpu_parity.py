"""Parity of the HIP path with the FP64 oracle, through the C ABI.

Tolerances: FP64 paths agree with the oracle to 1e-12 relative (bridge
beads) and 1e-12 of sum|terms| (action differences); FP32 paths to 1e-5.
"""
import numpy as np
import pytest

import oracle as O
from pimc import _capi as A
from pimc.system import chain_desc
from pimc.potentials import GaussianPair, HarmonicTrap, NoPotential, SquareLattice

pytestmark = pytest.mark.gpu

RTOL = {A.FP64: 1e-12, A.FP32: 1e-5}


def make_ctx(descs, precision, seed=7, **kw):
    opt = A.Options()
    opt.precision = precision
    opt.sweeps_per_launch = kw.get("spl", 4)
    opt.seed = seed
    opt.radial_bins = kw.get("radial_bins", 0)
    opt.radial_rmax = kw.get("radial_rmax", 0.0)
    opt.cart_bins = kw.get("cart_bins", 0)
    opt.derivatives = kw.get("derivatives", 0)
    opt.adapt_ell = kw.get("adapt_ell", 0)
    opt.pair_kernel = kw.get("pair_kernel", 0)
    opt.use_graph = kw.get("use_graph", 0)
    return A.Context(descs, opt)


def oracle_sys(desc, r, pair_in_accept=None):
    pot = O.pot_from_desc(desc)
    pair = O.make_pair(desc.pair, desc.g, desc.sigma)
    pia = desc.pair_in_accept if pair_in_accept is None else pair_in_accept
    return O.System(r, desc.lam, desc.L, desc.beta, pot, pair, desc.r_a, bool(pia))


def harmonic_desc(N=6, M=32, beta=1.0, ell=8, L=100.0):
    return chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=L, beta=beta, lam=0.5, ell=ell, com_step=1.0)


def rand_paths(rng, N, M, L, scale=1.0):
    r = rng.normal(0, scale, size=(N, 2, M))
    return np.clip(r, -L * 0.999, L * 0.999)


def close(a, b, rtol, scale=None):
    a, b = np.asarray(a, float), np.asarray(b, float)
    s = np.maximum(np.abs(b), 1.0) if scale is None else scale
    return np.all(np.abs(a - b) <= rtol * s)


def test_device_info():
    name, cu = A.device_info()
    print(name, cu, flush=True)
    assert "gfx950" in name
    assert cu > 0


@pytest.mark.parametrize("precision", [A.FP64, A.FP32])
def test_levy_propose_harmonic(precision):
    rng = np.random.default_rng(11)
    d = harmonic_desc(N=5, M=40)
    r = rand_paths(rng, 5, 40, 100.0)
    ctx = make_ctx([d], precision)
    ctx.set_paths(0, r)
    rr = ctx.get_paths(0)  # as stored (FP32-rounded for FP32)
    osys = oracle_sys(d, rr)
    for trial in range(25):
        n, j0, m = int(rng.integers(5)), int(rng.integers(40)), int(rng.integers(2, 12))
        nrm = rng.standard_normal((m - 1, 2))
        r_gpu, w_gpu = ctx.levy_propose(0, n, j0, m, nrm)
        r_or, w_or = osys.reshape_propose(n, j0, m, nrm)
        assert close(r_gpu, r_or, RTOL[precision]), (trial, np.max(np.abs(r_gpu - r_or)))
        dS_gpu = w_gpu[1] - w_gpu[0]
        dS_or = w_or.wV_updated - w_or.wV_initial
        scale = abs(w_or.wV_updated) + abs(w_or.wV_initial) + 1e-300
        assert abs(dS_gpu - dS_or) <= RTOL[precision] * scale, (trial, dS_gpu, dS_or)


@pytest.mark.parametrize("precision", [A.FP64, A.FP32])
def test_levy_propose_periodic_lattice(precision):
    """Small periodic box: the end-bead shift of levy! and the teleport act."""
    rng = np.random.default_rng(12)
    d = chain_desc(SquareLattice(4.0, 1.0), dim=2, M=24, N=4, L=2.0, beta=4.0, lam=0.5, ell=8)
    r = rng.uniform(-2.0, 2.0, size=(4, 2, 24))
    ctx = make_ctx([d], precision)
    ctx.set_paths(0, r)
    rr = ctx.get_paths(0)
    osys = oracle_sys(d, rr)
    tol = 1e-11 if precision == A.FP64 else 2e-5
    for trial in range(25):
        n, j0, m = int(rng.integers(4)), int(rng.integers(24)), int(rng.integers(2, 12))
        nrm = rng.standard_normal((m - 1, 2)) * 3.0
        r_gpu, w_gpu = ctx.levy_propose(0, n, j0, m, nrm)
        r_or, w_or = osys.reshape_propose(n, j0, m, nrm)
        diff = np.abs(r_gpu - r_or)
        diff = np.minimum(diff, 4.0 - diff)  # a bead exactly at the wrap may land on either side
        assert np.all(diff <= tol * 4.0), (trial, diff.max())
        scale = abs(w_or.wV_updated) + abs(w_or.wV_initial) + 1e-300
        assert abs((w_gpu[1] - w_gpu[0]) - (w_or.wV_updated - w_or.wV_initial)) <= 10 * tol * scale


@pytest.mark.parametrize("precision", [A.FP64, A.FP32])
def test_com_propose(precision):
    rng = np.random.default_rng(13)
    d = harmonic_desc(N=3, M=50)
    r = rand_paths(rng, 3, 50, 100.0)
    ctx = make_ctx([d], precision)
    ctx.set_paths(0, r)
    osys = oracle_sys(d, ctx.get_paths(0))
    for trial in range(10):
        n = int(rng.integers(3))
        disp = rng.uniform(-1, 1, size=2)
        r_gpu, w_gpu = ctx.com_propose(0, n, disp)
        r_or, w_or = osys.com_propose(n, disp)
        assert close(r_gpu, r_or, RTOL[precision])
        scale = abs(w_or.wV_updated) + abs(w_or.wV_initial)
        assert abs((w_gpu[1] - w_gpu[0]) - (w_or.wV_updated - w_or.wV_initial)) <= RTOL[precision] * scale


def test_reshape_and_com_apply_fp64():
    """The accept decisions and the committed paths follow the oracle move by move."""
    rng = np.random.default_rng(14)
    d = harmonic_desc(N=4, M=32, beta=2.0)
    r = rand_paths(rng, 4, 32, 100.0)
    ctx = make_ctx([d], A.FP64)
    ctx.set_paths(0, r)
    osys = oracle_sys(d, r)
    n_acc = 0
    for trial in range(200):
        u = float(rng.random())
        if trial % 3 == 0:
            n = int(rng.integers(4))
            disp = rng.uniform(-0.5, 0.5, size=2)
            a = ctx.center_of_mass(0, n, disp, u)
            b, _ = osys.com_apply(n, disp, u)
        else:
            n, j0, m = int(rng.integers(4)), int(rng.integers(32)), int(rng.integers(2, 9))
            nrm = rng.standard_normal((m - 1, 2))
            a = ctx.reshape_linear(0, n, j0, m, nrm, u)
            b, _ = osys.reshape_apply(n, j0, m, nrm, u)
        assert a == b, trial
        n_acc += a
    assert 20 < n_acc < 200
    assert np.max(np.abs(ctx.get_paths(0) - osys.r)) < 1e-11


@pytest.mark.parametrize("pot", ["harmonic", "lattice"])
@pytest.mark.parametrize("ell,M", [(16, 64), (16, 256), (32, 128), (8, 40)])
def test_fp32_fast_path_moves_with_injected_draws(pot, ell, M):
    """FP32 tile chains whose staging length ell (a power of two) divides M run
    the batched kernel's fast paths: levy_full_propose (compile-time bridge
    coefficients, no wrap inside the box, harmonic Delta V from sums of
    |new|^2 - |old|^2) and com_chunk_full (harmonic Delta V from the chunk sums,
    lane-group butterfly over the chunks when M / ell is a power of two).  The
    single-move hooks run exactly that code with injected normals and shifts:
    bridge beads to 1e-5 relative (of max(1, |r|)) and Delta S to 1e-5 of
    sum |terms| against the oracle, in the C2 box (L = 100) and a periodic
    lattice box; the committed move writes the proposed beads."""
    rng = np.random.default_rng(211)
    N = 3
    if pot == "harmonic":
        d = chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=100.0, beta=2.0, lam=0.5, ell=ell, com_step=1.0)
        r = rand_paths(rng, N, M, 100.0)
    else:
        d = chain_desc(SquareLattice(4.0, 1.0), dim=2, M=M, N=N, L=3.0, beta=4.0, lam=0.5, ell=ell, com_step=0.5)
        r = rng.uniform(-3.0, 3.0, size=(N, 2, M))
    ctx = make_ctx([d], A.FP32)
    ctx.set_paths(0, r)
    rr = ctx.get_paths(0)
    osys = oracle_sys(d, rr)
    L = d.L
    worst_r = worst_s = 0.0
    for trial in range(30):
        n, j0 = int(rng.integers(N)), int(rng.integers(M))
        nrm = rng.standard_normal((ell - 1, 2))
        r_gpu, w_gpu = ctx.levy_propose(0, n, j0, ell, nrm)
        r_or, w_or = osys.reshape_propose(n, j0, ell, nrm)
        diff = np.abs(r_gpu - r_or)
        diff = np.minimum(diff, 2 * L - diff)  # a bead at the wall may sit at -L in one, +L in the other
        worst_r = max(worst_r, float(np.max(diff / np.maximum(1.0, np.abs(r_or)))))
        scale = abs(w_or.wV_updated) + abs(w_or.wV_initial) + 1e-30
        worst_s = max(worst_s, abs((w_gpu[1] - w_gpu[0]) - (w_or.wV_updated - w_or.wV_initial)) / scale)
    for trial in range(20):
        n = int(rng.integers(N))
        disp = rng.uniform(-1, 1, size=2)
        r_gpu, w_gpu = ctx.com_propose(0, n, disp)
        r_or, w_or = osys.com_propose(n, disp)
        diff = np.abs(r_gpu - r_or)
        diff = np.minimum(diff, 2 * L - diff)
        worst_r = max(worst_r, float(np.max(diff / np.maximum(1.0, np.abs(r_or)))))
        scale = abs(w_or.wV_updated) + abs(w_or.wV_initial) + 1e-30
        worst_s = max(worst_s, abs((w_gpu[1] - w_gpu[0]) - (w_or.wV_updated - w_or.wV_initial)) / scale)
    print(pot, ell, M, "bridge/COM beads", worst_r, "Delta S", worst_s)
    assert worst_r <= 1e-5 and worst_s <= 1e-5, (worst_r, worst_s)
    # committing: an accepted staging move (u = 0) writes the proposed interior beads
    n, j0 = 1, 5
    nrm = rng.standard_normal((ell - 1, 2)) * 0.1
    r_prop, _ = ctx.levy_propose(0, n, j0, ell, nrm)
    assert ctx.reshape_linear(0, n, j0, ell, nrm, 0.0)
    after = ctx.get_paths(0)
    for i in range(1, ell):
        assert np.array_equal(after[n, :, (j0 + i) % M], r_prop[i].astype(np.float32).astype(np.float64))


def test_density_after_every_sweep_matches_oracle():
    """pimc_density (and pimc_density_noalloc) called right after each of 400
    one-sweep runs: every call's histogram equals the oracle's of the same
    paths bin for bin (the histogram is cleared on the library stream ahead
    of the counting kernel; a clear that landed after the counts would show
    as a deficit)."""
    rng = np.random.default_rng(16)
    d = harmonic_desc(N=8, M=64, beta=1.5, L=6.0)
    ctx = make_ctx([d], A.FP64, spl=1)
    ctx.set_paths(0, rand_paths(rng, 8, 64, 6.0, scale=1.5))
    for it in range(400):
        ctx.run(1, measure=False)
        dg, ndg = (ctx.density if it % 2 == 0 else ctx.density_noalloc)(0, 40)
        do, ndo = oracle_sys(d, ctx.get_paths(0)).density(40)
        assert ndg == ndo == 64 and np.array_equal(dg, do), it


@pytest.mark.parametrize("precision", [A.FP64, A.FP32])
def test_energy_density_hooks(precision):
    rng = np.random.default_rng(15)
    d = harmonic_desc(N=7, M=48, beta=1.5, L=6.0)
    r = rand_paths(rng, 7, 48, 6.0, scale=1.5)
    ctx = make_ctx([d], precision)
    ctx.set_paths(0, r)
    rr = ctx.get_paths(0)
    osys = oracle_sys(d, rr)
    E, Ev, Ep, Evp = ctx.energy(0)
    Eo, Evo = osys.energy()
    # E is a difference of O(dN/tau) terms: compare relative to them
    big = 2 * 7 / (2 * d.beta / d.M)
    assert abs(E - Eo) <= 10 * RTOL[precision] * big, (E, Eo)
    assert abs(Ev - Evo) <= 10 * RTOL[precision] * max(1.0, abs(Evo)), (Ev, Evo)
    assert Ep == 0.0 and Evp == 0.0
    for nb in (7, 16):
        dg, ndg = ctx.density(0, nb)
        do, ndo = osys.density(nb)
        assert ndg == ndo == 48
        assert np.array_equal(dg, do)


def pair_case(N, L, M, cells_N_threshold_hint, precision, seed=16, nbins_ra=None, g=1.0, sigma=0.5):
    rng = np.random.default_rng(seed)
    pair = GaussianPair(g, sigma)
    d = chain_desc(NoPotential(), dim=2, M=M, N=N, L=L, beta=2.0, lam=0.5, pair=pair, ell=8,
                   r_a=nbins_ra if nbins_ra else 0.0)
    r = rng.uniform(-L, L, size=(N, 2, M))
    ctx = make_ctx([d], precision)
    ctx.set_paths(0, r)
    osys = oracle_sys(d, ctx.get_paths(0))
    return d, ctx, osys, rng


@pytest.mark.parametrize("N,L,ra", [(120, 6.0, 2.0), (40, 4.0, 2.0), (30, 1.5, 1.5), (25, 2.0, 3.0)])
@pytest.mark.parametrize("precision", [A.FP64, A.FP32])
def test_interaction_sums(N, L, ra, precision):
    """interaction_action! values: cell lists (N > 96) and the all-pairs scan
    (N <= 96), including nbins = 2 and 1 where the stencil repeats cells."""
    d, ctx, osys, rng = pair_case(N, L, 16, 0, precision, nbins_ra=ra)
    tol = 1e-12 if precision == A.FP64 else 3e-5
    for trial in range(40):
        n, j = int(rng.integers(N)), int(rng.integers(16))
        g = ctx.interaction(0, n, j)
        o = osys.interaction_old(n, j)
        assert abs(g - o) <= tol * max(1e-3, abs(o)), (trial, g, o)


@pytest.mark.parametrize("precision", [A.FP64, A.FP32])
def test_pair_moves_and_energy(precision):
    d, ctx, osys, rng = pair_case(120, 6.0, 16, 0, precision, seed=17)
    tol = 1e-11 if precision == A.FP64 else 5e-5
    for trial in range(20):
        n, j0, m = int(rng.integers(120)), int(rng.integers(16)), int(rng.integers(2, 8))
        nrm = rng.standard_normal((m - 1, 2))
        _, w_gpu = ctx.levy_propose(0, n, j0, m, nrm)
        _, w_or = osys.reshape_propose(n, j0, m, nrm)
        dU_g = w_gpu[3] - w_gpu[2]
        dU_o = w_or.wU_updated - w_or.wU_initial
        assert abs(dU_g - dU_o) <= tol * (abs(w_or.wU_updated) + abs(w_or.wU_initial) + 1e-3), (trial, dU_g, dU_o)
    E, Ev, Ep, Evp = ctx.energy(0)
    Epo, Evpo = osys.pair_energy()
    assert abs(Ep - Epo) <= tol * abs(Epo) * 10
    assert abs(Evp - Evpo) <= tol * max(1.0, abs(Evpo)) * 10


def oracle_meas(s):
    """The kernels' energy record of one measurement on the oracle state:
    (E, Ev, Epair) with the pair energy and pair virial added to the
    reference's Energy (zero without a Gaussian pair)."""
    e, ev = s.energy()
    p, pv = s.pair_energy()
    return e + p, ev + pv, p


def block_means(vals, spl_meas):
    """Means of consecutive groups of spl_meas measurements (one block each)."""
    vals = np.asarray(vals, float)
    return np.array([vals[i:i + spl_meas].mean() for i in range(0, len(vals), spl_meas)])


def sweep_oracle(descs, inits, seed, n_sweeps, spl, measure=False, measure_every=1, sweep0=0):
    """Run the oracle's sweep schedule for every chain, with the COM-step
    adaptation between launches of `spl` sweeps when not measuring; when
    measuring, the (E, Ev, Epair) of oracle_meas after the global sweeps s
    with (s + 1) % measure_every == 0."""
    out = []
    for c, (d, r) in enumerate(zip(descs, inits)):
        s = oracle_sys(d, r)
        step = d.com_step
        sweep = sweep0
        acc_l = acc_c = 0
        left = n_sweeps
        Es = []
        while left > 0:
            k = min(spl, left)
            ca = 0
            for _ in range(k):
                st = s.sweep(c, sweep, seed, d.ell, step)
                acc_l += st.levy_acc
                acc_c += st.com_acc
                ca += st.com_acc
                sweep += 1
                if measure and sweep % measure_every == 0:
                    Es.append(oracle_meas(s))
            if not measure:
                step = O.adapt_step(step, ca, d.N * k, d.L)
            left -= k
        out.append((s.r, acc_l, acc_c, step, Es))
    return out


@pytest.mark.parametrize("measure", [False, True])
def test_sweep_trajectory_fp64(measure):
    """Batched tile kernel == oracle sweep schedule, FP64, several uneven chains."""
    rng = np.random.default_rng(18)
    descs, inits = [], []
    for N, M, beta, ell in [(5, 32, 1.0, 8), (9, 64, 2.0, 8), (3, 40, 0.7, 6), (12, 24, 3.0, 4)]:
        descs.append(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=100.0, beta=beta, lam=0.5, ell=ell,
                                com_step=1.0))
        inits.append(rand_paths(rng, N, M, 100.0))
    seed, n_sweeps, spl = 99, 12, 5
    ctx = make_ctx(descs, A.FP64, seed=seed, spl=spl)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    cnt = ctx.run(n_sweeps, measure=measure)
    ref = sweep_oracle(descs, inits, seed, n_sweeps, spl, measure=measure)
    tot_l = sum(x[1] for x in ref)
    tot_c = sum(x[2] for x in ref)
    assert cnt.levy_acc == tot_l and cnt.com_acc == tot_c
    assert cnt.bead_moves == n_sweeps * sum(O.sweep_bead_moves(d.N, d.M, d.ell) for d in descs)
    for c, d in enumerate(descs):
        assert np.max(np.abs(ctx.get_paths(c) - ref[c][0])) < 1e-9, c
        assert abs(ctx.com_step(c) - ref[c][3]) < 1e-12
        if measure:
            E, Ev, _ = ctx.blocks(c)
            Eo = np.array([e[0] for e in ref[c][4]])
            Evo = np.array([e[1] for e in ref[c][4]])
            blocks = [Eo[i:i + spl].mean() for i in range(0, n_sweeps, spl)]
            blocks_v = [Evo[i:i + spl].mean() for i in range(0, n_sweeps, spl)]
            assert np.allclose(E, blocks, rtol=1e-10, atol=1e-8)
            assert np.allclose(Ev, blocks_v, rtol=1e-10, atol=1e-8)


def test_measure_every_fp64():
    """measure_every = 3: the estimator blocks hold the energies after sweeps
    3, 6, ... (global count), as in the oracle; the paths are unchanged."""
    rng = np.random.default_rng(41)
    descs, inits = [], []
    for N, M, ell in [(3, 16, 4), (5, 24, 8)]:
        descs.append(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=100.0, beta=1.0, lam=0.5, ell=ell,
                                com_step=1.0))
        inits.append(rand_paths(rng, N, M, 100.0))
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed, opt.measure_every = A.FP64, 6, 19, 3
    ctx = A.Context(descs, opt)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    ctx.run(12, measure=True)
    ref = sweep_oracle(descs, inits, 19, 12, 6, measure=True, measure_every=3)
    for c in range(len(descs)):
        assert np.max(np.abs(ctx.get_paths(c) - ref[c][0])) < 1e-9
        E, _, _ = ctx.blocks(c)
        Eo = np.array([e[0] for e in ref[c][4]])  # 4 measurements: sweeps 3, 6 | 9, 12
        assert len(Eo) == 4 and len(E) == 2
        assert np.allclose(E, [Eo[:2].mean(), Eo[2:].mean()], rtol=1e-10, atol=1e-8)
    with pytest.raises(A.PimcError):
        ctx.run(4, measure=True)  # not a multiple of measure_every
    opt.sweeps_per_launch = 5
    with pytest.raises(A.PimcError):
        A.Context(descs, opt)  # sweeps_per_launch not a multiple of measure_every


@pytest.mark.parametrize("dim", [1, 2])
def test_sweep_trajectory_edge_cases_fp64(dim):
    """Tile kernel == oracle, FP64, at the edges of the schedule: one particle
    with ell = M = 5 (one bridge around the whole ring, from a bead back to
    itself), two-link segments (one interior bead), ell not dividing M, and 1D
    chains.  (ell > M is rejected by pimc_create.)"""
    rng = np.random.default_rng(31)
    descs, inits = [], []
    for N, M, ell in [(1, 5, 5), (2, 12, 2), (3, 16, 4), (4, 20, 6)]:
        descs.append(chain_desc(HarmonicTrap(1.0), dim=dim, M=M, N=N, L=100.0, beta=1.0, lam=0.5, ell=ell,
                                com_step=1.0))
        inits.append(np.clip(rng.normal(0, 1.0, size=(N, dim, M)), -99.0, 99.0))
    seed, n_sweeps, spl = 13, 6, 3
    ctx = make_ctx(descs, A.FP64, seed=seed, spl=spl)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    cnt = ctx.run(n_sweeps, measure=True)
    ref = sweep_oracle(descs, inits, seed, n_sweeps, spl, measure=True)
    assert cnt.levy_acc == sum(x[1] for x in ref) and cnt.com_acc == sum(x[2] for x in ref)
    assert cnt.bead_moves == n_sweeps * sum(O.sweep_bead_moves(d.N, d.M, d.ell) for d in descs)
    for c, d in enumerate(descs):
        assert np.max(np.abs(ctx.get_paths(c) - ref[c][0])) < 1e-9, c
        E, Ev, _ = ctx.blocks(c)
        Eo = np.array([e[0] for e in ref[c][4]])
        assert np.allclose(E, [Eo[i:i + spl].mean() for i in range(0, n_sweeps, spl)], rtol=1e-10, atol=1e-8), c


@pytest.mark.parametrize("pair_kernel", [1, 2])
@pytest.mark.parametrize("N,L,ra", [(30, 3.0, 1.5), (110, 20.0, 2.0)])
def test_sweep_trajectory_pair_1d_fp64(N, L, ra, pair_kernel):
    """1D pair chains == oracle, FP64: the whole-chain kernel (all-pairs with
    the path in LDS for N = 30, cell lists for N = 110) and the window kernels."""
    rng = np.random.default_rng(37)
    pair = GaussianPair(1.0, 0.5)
    d = chain_desc(HarmonicTrap(0.3), dim=1, M=16, N=N, L=L, beta=2.0, lam=0.5, pair=pair, ell=8, r_a=ra,
                   com_step=0.5)
    r = rng.uniform(-L, L, size=(N, 1, 16))
    seed, n_sweeps, spl = 6, 3, 3
    ctx = make_ctx([d], A.FP64, seed=seed, spl=spl, pair_kernel=pair_kernel)
    ctx.set_paths(0, r)
    cnt = ctx.run(n_sweeps, measure=True)
    ref = sweep_oracle([d], [r], seed, n_sweeps, spl, measure=True)
    assert cnt.levy_acc == ref[0][1] and cnt.com_acc == ref[0][2]
    assert np.max(np.abs(ctx.get_paths(0) - ref[0][0])) < 1e-9
    assert_blocks(ctx, 0, ref[0][4], n_sweeps)


def assert_blocks(ctx, c, meas, per_block, rtol=1e-9):
    """The chain's estimator blocks (E, Ev, Epair) equal the block means of the
    oracle's per-sweep energies (oracle_meas) to rtol of max(1, |value|)."""
    got = ctx.blocks(c)
    meas = np.asarray(meas, float)
    assert len(got[0]) == len(meas) // per_block, (len(got[0]), len(meas))
    for k, name in enumerate(("E", "Ev", "Epair")):
        want = block_means(meas[:, k], per_block)
        assert np.allclose(got[k], want, rtol=rtol, atol=rtol), (c, name, got[k], want)


@pytest.mark.parametrize("case", ["trap", "trap_box", "lattice_box", "optical"])
def test_sweep_trajectory_fp32(case):
    """FP32 tile kernel == oracle sweep schedule up to FP32 rounding.

    Covers both FP32 code paths of the tile kernel: segments that are one
    skew block of the LDS rows (ell a power of two dividing M: compile-time
    bridge constants, wrap only for beads outside the box) and the general
    segments (ell = 6, M = 40), in a wide trap, in a trap whose box of
    half-width 1.5 the beads leave (the harmonic sums of staging and COM then
    take the per-bead form with wrapped beads) and in a periodic lattice box
    of half-width 1.5."""
    rng = np.random.default_rng(23)
    descs, inits = [], []
    if case == "trap":
        for N, M, beta, ell in [(5, 32, 1.0, 8), (4, 64, 2.0, 16), (3, 40, 0.7, 6), (6, 24, 3.0, 4)]:
            descs.append(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=100.0, beta=beta, lam=0.5, ell=ell,
                                    com_step=1.0))
            inits.append(rand_paths(rng, N, M, 100.0))
    elif case == "trap_box":
        for N, M, ell in [(4, 32, 8), (3, 64, 16), (3, 40, 6)]:
            descs.append(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=1.5, beta=2.0, lam=0.5, ell=ell,
                                    com_step=1.0))
            inits.append(rng.uniform(-1.5, 1.5, size=(N, 2, M)))
    elif case == "optical":
        from pimc.potentials import BEAMS, OpticalLattice
        pot = OpticalLattice(BEAMS["l25"], 1.0, -4.0)
        for N, M, ell in [(4, 32, 8), (3, 64, 16), (3, 40, 6)]:
            descs.append(chain_desc(pot, dim=2, M=M, N=N, L=6.0, beta=2.0, lam=0.5, ell=ell, com_step=0.5))
            inits.append(rng.uniform(-3.0, 3.0, size=(N, 2, M)))
    else:
        for N, M, ell in [(4, 32, 8), (3, 64, 16), (3, 40, 6)]:
            descs.append(chain_desc(SquareLattice(1.0, 1.0), dim=2, M=M, N=N, L=1.5, beta=2.0, lam=0.5, ell=ell,
                                    com_step=1.0))
            inits.append(rng.uniform(-1.5, 1.5, size=(N, 2, M)))
    inits = [r.astype(np.float32).astype(np.float64) for r in inits]  # the FP32 start the kernel sees
    seed, n_sweeps, spl = 7, 2, 2
    ctx = make_ctx(descs, A.FP32, seed=seed, spl=spl)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    cnt = ctx.run(n_sweeps, measure=True)
    ref = sweep_oracle(descs, inits, seed, n_sweeps, spl, measure=True)
    assert cnt.levy_acc == sum(x[1] for x in ref) and cnt.com_acc == sum(x[2] for x in ref)
    wrapped = 0
    for c, d in enumerate(descs):
        diff = ctx.get_paths(c) - ref[c][0]
        wrapped += int(np.sum(np.abs(diff) > d.L))
        diff -= 2 * d.L * np.round(diff / (2 * d.L))  # a bead on the box edge may sit at -L in one, +L in the other
        rel = float(np.max(np.abs(diff) / np.maximum(1.0, np.abs(ref[c][0]))))
        print(case, c, "max |r32 - r64| =", float(np.max(np.abs(diff))), "relative", rel, flush=True)
        # the fast path (ell a power of two dividing M) leaves in-box beads unwrapped: FP32
        # rounding of the bridge only.  The general path applies the reference's teleport,
        # whose FP32 form ((x + L) - ...) - L rounds a bead to the spacing of L.
        fast = d.ell in (8, 16, 32) and d.M % d.ell == 0
        bound = 1e-5 if fast else 1e-5 + 2 * float(np.spacing(np.float32(d.L)))
        assert np.all(np.abs(diff) <= bound * np.maximum(1.0, np.abs(ref[c][0]))), (c, rel, bound)
        E, Ev, _ = ctx.blocks(c)
        Eo = np.mean([e[0] for e in ref[c][4]])
        assert abs(E[0] - Eo) <= 1e-4 * max(1.0, abs(Eo)), (c, E[0], Eo)
    assert wrapped <= 2
    if case in ("lattice_box", "trap_box"):
        moved_out = sum(int(np.sum(np.abs(r0 - x[0]) > 1.5)) for r0, x in zip(inits, ref))
        assert moved_out > 0  # the wrap branch ran


@pytest.mark.parametrize("pair_kernel", [1, 2])
@pytest.mark.parametrize("N,L,ra", [(110, 5.0, 2.0), (30, 3.0, 1.5)])
def test_sweep_trajectory_pair_fp64(N, L, ra, pair_kernel):
    """2D pair chains == oracle, FP64: the whole-chain kernel (cell lists for
    N > 96, all-pairs scan otherwise) and the window kernels."""
    rng = np.random.default_rng(19)
    pair = GaussianPair(1.0, 0.5)
    d = chain_desc(HarmonicTrap(0.3), dim=2, M=16, N=N, L=L, beta=2.0, lam=0.5, pair=pair, ell=8, r_a=ra,
                   com_step=0.5)
    r = rng.uniform(-L, L, size=(N, 2, 16))
    seed, n_sweeps, spl = 5, 3, 3
    ctx = make_ctx([d], A.FP64, seed=seed, spl=spl, pair_kernel=pair_kernel)
    ctx.set_paths(0, r)
    cnt = ctx.run(n_sweeps, measure=True)
    ref = sweep_oracle([d], [r], seed, n_sweeps, spl, measure=True)
    assert cnt.levy_acc == ref[0][1] and cnt.com_acc == ref[0][2]
    assert np.max(np.abs(ctx.get_paths(0) - ref[0][0])) < 1e-9
    assert_blocks(ctx, 0, ref[0][4], n_sweeps)
    assert ctx.blocks(0)[2][0] > 0.0


@pytest.mark.parametrize("precision", [A.FP64, A.FP32])
def test_window_kernels_mixed_chains(precision):
    """Several Gaussian-pair chains on the window kernels in one launch, of
    different N (not multiples of 64), M and ell: ell dividing M, ell not
    dividing M (a shorter last window), a last window of one link (no staging
    bead), M above 256 (centre-of-mass windows of M/32 slices), 1-cycle rings
    only.  FP64: the oracle's paths to 1e-9, its accept counts and estimator
    blocks; FP32: the oracle's accept counts and paths to FP32 rounding of the
    box, on the FP32-rounded start."""
    rng = np.random.default_rng(71)
    pair = GaussianPair(1.0, 0.5)
    specs = [(110, 16, 8, 5.0, 2.0), (130, 20, 8, 6.0, 2.0), (100, 17, 8, 5.0, 1.5), (97, 264, 6, 8.0, 2.0)]
    descs, inits = [], []
    for N, M, ell, L, ra in specs:
        descs.append(chain_desc(HarmonicTrap(0.3), dim=2, M=M, N=N, L=L, beta=2.0, lam=0.5, pair=pair, ell=ell,
                                r_a=ra, com_step=0.5))
        r = rng.uniform(-L, L, size=(N, 2, M))
        if precision == A.FP32:
            r = r.astype(np.float32).astype(np.float64)
        inits.append(r)
    seed, n_sweeps, spl = 13, 2, 2
    ctx = make_ctx(descs, precision, seed=seed, spl=spl, pair_kernel=2)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    cnt = ctx.run(n_sweeps, measure=True)
    ref = sweep_oracle(descs, inits, seed, n_sweeps, spl, measure=True)
    assert cnt.levy_acc == sum(x[1] for x in ref) and cnt.com_acc == sum(x[2] for x in ref)
    assert cnt.bead_moves == n_sweeps * sum(O.sweep_bead_moves(d.N, d.M, d.ell) for d in descs)
    for c, d in enumerate(descs):
        diff = ctx.get_paths(c) - ref[c][0]
        diff -= 2 * d.L * np.round(diff / (2 * d.L))
        if precision == A.FP64:
            assert np.max(np.abs(diff)) < 1e-9, c
            assert_blocks(ctx, c, ref[c][4], spl)
        else:
            assert np.max(np.abs(diff)) <= 1e-5 * d.L, (c, float(np.max(np.abs(diff))))


def test_batch_fp32_runs_and_counts():
    from pimc.batch import trapped_batch, Batch
    wl = trapped_batch(n_chains=64, seed=3)
    b = Batch(wl, sweeps_per_launch=10)
    c1 = b.run(20, measure=False)
    c2 = b.run(20, measure=True)
    assert c1.levy_acc > 0 and c2.com_acc > 0
    assert 0.3 < c2.levy_acc / c2.levy_tries <= 1.0
    e = b.estimates(0)
    assert e.n_blocks == 2 and np.isfinite(e.E)
    allc = b.ctx.all_counters()
    assert allc.shape == (64, 4)
    for c in (0, 17, 63):
        assert np.array_equal(allc[c], b.ctx.counters(c))
    tot = allc.sum(axis=0)
    assert tot[0] == c1.levy_acc + c2.levy_acc and tot[2] == c1.com_acc + c2.com_acc


def test_graph_replay_matches_eager_launches():
    """The captured launch step replayed from a hipGraph gives the same paths,
    counters and estimator blocks as eager launches, replay after replay, with
    the program's own copies and synchronisations in between."""
    from pimc.batch import trapped_batch
    wl = trapped_batch(n_chains=48, seed=21, precision=A.FP32)
    ctxs = []
    for g in (0, 1):
        opt = A.Options()
        opt.precision = A.FP32
        opt.sweeps_per_launch = 5
        opt.seed = 77
        opt.radial_bins = 16
        opt.radial_rmax = 4.0
        opt.use_graph = g
        ctx = A.Context(wl.descs, opt)
        for c, r in enumerate(wl.init):
            ctx.set_paths(c, r)
        ctxs.append(ctx)
    for rep in range(4):
        measure = rep % 2 == 1
        cnts = [ctx.run(15, measure=measure) for ctx in ctxs]
        assert cnts[0].levy_acc == cnts[1].levy_acc and cnts[0].com_acc == cnts[1].com_acc, rep
        for c in (0, 17, 47):
            assert np.array_equal(ctxs[0].get_paths(c), ctxs[1].get_paths(c)), (rep, c)
            assert np.array_equal(ctxs[0].counters(c), ctxs[1].counters(c))
    for c in (0, 17, 47):
        e0, e1 = ctxs[0].blocks(c), ctxs[1].blocks(c)
        assert all(np.array_equal(x, y) for x, y in zip(e0, e1))
        assert np.array_equal(ctxs[0].radial(c)[0], ctxs[1].radial(c)[0])
    d = wl.descs[0]
    segs = sum(1 for k in range(-(-d.M // d.ell)) if min(d.ell, d.M - k * d.ell) >= 2)
    assert ctxs[0].counters(0)[1] == 60 * d.N * segs and ctxs[0].counters(0)[3] == 60 * d.N


def test_graph_replay_with_split_estimators_and_tables():
    """Graph replay of measuring steps split into one-sweep launches plus the
    estimator kernels (a cell-list chain large enough for the automatic split),
    next to a tabulated-propagator chain with hard spheres: bit-identical to
    eager launches over several replays."""
    from pimc.propagator import PropInt
    rng = np.random.default_rng(61)
    tau = 2.0 / 16
    terms = O.prop_int_terms(64, 1e-20, 8.0, 2.0, tau)
    p = PropInt(terms, 1e-20, 8.0, 2.0, tau)
    descs = [chain_desc(NoPotential(), dim=2, M=96, N=110, L=5.0, beta=2.0, lam=0.5, pair=GaussianPair(1.0, 0.5),
                        ell=8, r_a=2.0, com_step=0.5),
             chain_desc(HarmonicTrap(0.5), dim=2, M=16, N=20, L=5.0, beta=2.0, lam=0.5, pair=p, r_a=1.0, ell=8,
                        com_step=0.5, hs_a=0.05)]
    inits = [rng.uniform(-5.0, 5.0, size=(110, 2, 96)), rng.uniform(-4.0, 4.0, size=(20, 2, 16))]
    ctxs = []
    for g in (0, 1):
        opt = A.Options()
        opt.precision = A.FP32
        opt.sweeps_per_launch = 3
        opt.seed = 5
        opt.use_graph = g
        ctx = A.Context(descs, opt)
        for c, r in enumerate(inits):
            ctx.set_paths(c, r)
        ctx.set_pair_table(1, terms, 1e-20, 8.0, 0.0)
        ctxs.append(ctx)
    for rep in range(4):
        measure = rep % 2 == 1
        cnts = [ctx.run(6, measure=measure) for ctx in ctxs]
        assert (cnts[0].levy_acc, cnts[0].com_acc) == (cnts[1].levy_acc, cnts[1].com_acc), rep
        for c in range(2):
            assert np.array_equal(ctxs[0].get_paths(c), ctxs[1].get_paths(c)), (rep, c)
    for c in range(2):
        assert all(np.array_equal(x, y) for x, y in zip(ctxs[0].blocks(c), ctxs[1].blocks(c)))


def test_graph_replay_with_two_window_groups():
    """Three window-kernel chains (two groups on two streams, the second half a
    sweep behind the first): graph replays of thermalising and measuring steps
    are bit-identical to eager launches, paths, counters, estimator blocks and
    both histograms, replay after replay."""
    rng = np.random.default_rng(62)
    pair = GaussianPair(1.0, 0.5)
    descs, inits = [], []
    for N, M, ell in [(110, 32, 8), (130, 24, 6), (120, 16, 5)]:
        descs.append(chain_desc(HarmonicTrap(0.3), dim=2, M=M, N=N, L=5.0, beta=2.0, lam=0.5, pair=pair, ell=ell,
                                r_a=2.0, com_step=0.5))
        inits.append(rng.uniform(-5.0, 5.0, size=(N, 2, M)))
    ctxs = []
    for g in (0, 1):
        opt = A.Options()
        opt.precision, opt.sweeps_per_launch, opt.seed, opt.use_graph = A.FP32, 3, 9, g
        opt.radial_bins, opt.radial_rmax, opt.cart_bins = 24, 4.0, 16
        ctx = A.Context(descs, opt)
        for c, r in enumerate(inits):
            ctx.set_paths(c, r)
        ctxs.append(ctx)
    for rep in range(4):
        measure = rep % 2 == 1
        cnts = [ctx.run(9, measure=measure) for ctx in ctxs]
        assert (cnts[0].levy_acc, cnts[0].com_acc) == (cnts[1].levy_acc, cnts[1].com_acc), rep
        for c in range(3):
            assert np.array_equal(ctxs[0].get_paths(c), ctxs[1].get_paths(c)), (rep, c)
            assert np.array_equal(ctxs[0].counters(c), ctxs[1].counters(c)), (rep, c)
    for c in range(3):
        assert all(np.array_equal(x, y) for x, y in zip(ctxs[0].blocks(c), ctxs[1].blocks(c)))
        assert np.array_equal(ctxs[0].radial(c)[0], ctxs[1].radial(c)[0])
        assert np.array_equal(ctxs[0].cartesian(c)[0], ctxs[1].cartesian(c)[0])
        assert ctxs[0].counters(c)[3] == 36 * descs[c].N


def test_cell_list_overflow_is_reported():
    """All particles of a pair chain in one cell at every slice: set_paths
    reports PIMC_ERR_CAPACITY for that chain instead of sampling with a
    truncated list."""
    pair = GaussianPair(1.0, 0.5)
    d = chain_desc(NoPotential(), dim=2, M=8, N=200, L=8.0, beta=2.0, lam=0.5, pair=pair, ell=4)
    ctx = make_ctx([d], A.FP32)
    r = np.zeros((200, 2, 8)) + 0.1
    with pytest.raises(A.PimcError) as e:
        ctx.set_paths(0, r)
    assert "CAPACITY" in str(e.value)
    ok = np.random.default_rng(1).uniform(-8, 8, size=(200, 2, 8))
    ctx.set_paths(0, ok)   # a valid configuration clears the flag
    ctx.run(2, measure=False)


@pytest.mark.parametrize("precision", [A.FP64, A.FP32])
def test_init_paths_matches_oracle(precision):
    """pimc_init_paths == the oracle's init_world restatement for every chain:
    tile chains (particle-major) and pair chains (slice-major, with cell lists
    rebuilt afterwards, so a sweep runs)."""
    pair = GaussianPair(1.0, 0.5)
    descs = [harmonic_desc(N=5, M=16, L=6.0), harmonic_desc(N=3, M=24, beta=2.0, L=100.0),
             chain_desc(NoPotential(), dim=2, M=16, N=120, L=6.0, beta=2.0, lam=0.5, pair=pair, ell=8, r_a=2.0),
             chain_desc(NoPotential(), dim=2, M=12, N=20, L=3.0, beta=1.0, lam=0.5, pair=pair, ell=4, r_a=1.5)]
    ctx = make_ctx(descs, precision, seed=3)
    ctx.init_paths(-1, 99)
    tol = 1e-12 if precision == A.FP64 else 1e-5
    for c, d in enumerate(descs):
        ref = O.init_world(99, c, d.N, d.M, d.dim, d.L, d.beta / d.M, d.lam)
        got = ctx.get_paths(c)
        diff = np.abs(got - ref)
        diff = np.minimum(diff, 2 * d.L - diff)  # a bead at the wall may round to the other side in FP32
        assert np.max(diff) <= tol * max(1.0, d.L), (c, np.max(diff))
    before = ctx.get_paths(0)
    ctx.init_paths(1, 5)  # one chain: the others keep their paths
    assert np.array_equal(ctx.get_paths(0), before)
    ref1 = O.init_world(5, 1, 3, 24, 2, 100.0, 2.0 / 24, 0.5)
    assert np.max(np.abs(ctx.get_paths(1) - ref1)) <= tol * 100.0
    cnt = ctx.run(4, measure=False)
    assert cnt.levy_acc > 0


def test_swap_sweeps_match_oracle_fp64():
    """Whole-chain kernel with permutation cycles and swap tries == the
    oracle's sweep_perm, FP64: paths, next[], all counters, energy blocks
    (swap windows inside the ring and across slice M - 1, both rules)."""
    rng = np.random.default_rng(43)
    specs = [(4, 12, 4, 3, 4, 0, 2.0), (3, 16, 8, 2, 10, 0, 1.0), (5, 8, 4, 4, 3, 1, 3.0)]
    descs, inits = [], []
    for N, M, ell, nsw, m, rule, beta in specs:
        descs.append(chain_desc(HarmonicTrap(1.0), dim=2, M=M, N=N, L=100.0, beta=beta, lam=0.5, ell=ell,
                                com_step=0.8, swaps=nsw, swap_m=m, swap_rule=rule))
        inits.append(rand_paths(rng, N, M, 100.0, scale=0.6))
    seed, n_sweeps = 29, 12
    ctx = make_ctx(descs, A.FP64, seed=seed, spl=n_sweeps)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    cnt = ctx.run(n_sweeps, measure=True)
    tot = dict(levy_acc=0, com_acc=0, com_tries=0, swap_acc=0, swap_tries=0)
    for c, (d, r) in enumerate(zip(descs, inits)):
        s = oracle_sys(d, r)
        Es = []
        for sw in range(n_sweeps):
            st = s.sweep_perm(c, sw, seed, d.ell, d.com_step, d.swaps, d.swap_m, d.swap_rule)
            for k in tot:
                tot[k] += getattr(st, k)
            Es.append(s.energy()[0])
        assert np.array_equal(ctx.get_next(c), s.next), c
        assert np.max(np.abs(ctx.get_paths(c) - s.r)) < 1e-9, c
        E, _, _ = ctx.blocks(c)
        assert abs(E[0] - np.mean(Es)) < 1e-9 * max(1.0, abs(E[0])), (c, E[0], np.mean(Es))
    assert tot["swap_acc"] > 3 and tot["swap_tries"] == n_sweeps * sum(d.swaps for d in descs)
    for k in tot:
        assert getattr(cnt, k) == tot[k], (k, getattr(cnt, k), tot[k])


@pytest.mark.parametrize("use_graph", [0, 1])
def test_adapt_ell_matches_oracle_fp64(use_graph):
    """adapt_ell: whole chains (swap cycles, pair action) change their staging
    length between thermalising launches by adjust!(::NumbOfSlices), on each
    launch's staging acceptance, in step with the oracle's sweep schedule with
    o_adapt_ell between launches: paths, next[], counters, the lengths read
    back by pimc_get_ell.  The tile chains adapt too, from one shared tile
    into tiles of their own lengths (rebuilt between launches); every chain
    keeps its length during a measuring run."""
    rng = np.random.default_rng(47)
    pair = GaussianPair(3.0, 0.5)
    descs = [
        chain_desc(HarmonicTrap(1.0), dim=2, M=24, N=4, L=100.0, beta=1.0, lam=0.5, ell=4, com_step=0.8,
                   swaps=2, swap_m=4),
        chain_desc(HarmonicTrap(0.3), dim=2, M=20, N=24, L=3.0, beta=8.0, lam=0.5, pair=pair, ell=16, r_a=1.5,
                   com_step=0.5),
        harmonic_desc(N=5, M=16, beta=1.0, ell=4),
        harmonic_desc(N=6, M=16, beta=6.0, ell=4),
        chain_desc(HarmonicTrap(1.0), dim=2, M=40, N=3, L=100.0, beta=2.0, lam=0.5, ell=16, com_step=0.8),
    ]
    inits = [rand_paths(rng, 4, 24, 100.0, scale=0.6), rng.uniform(-3.0, 3.0, size=(24, 2, 20)),
             rand_paths(rng, 5, 16, 100.0, scale=0.6), rand_paths(rng, 6, 16, 100.0, scale=0.6),
             rand_paths(rng, 3, 40, 100.0, scale=0.6)]
    seed, spl, n_launch = 31, 2, 6
    ctx = make_ctx(descs, A.FP64, seed=seed, spl=spl, adapt_ell=1, use_graph=use_graph)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    cnt_therm = ctx.run(spl * n_launch, measure=False)
    ells_after = [ctx.ell(c) for c in range(len(descs))]
    cnt_meas = ctx.run(spl * 2, measure=True)
    assert [ctx.ell(c) for c in range(len(descs))] == ells_after
    tot = dict(levy_acc=0, levy_tries=0, com_acc=0, com_tries=0, swap_acc=0, swap_tries=0)
    ells_oracle, changed = [], False
    for c, (d, r) in enumerate(zip(descs, inits)):
        s = oracle_sys(d, r)
        step, ell, sweep = d.com_step, d.ell, 0
        for launch in range(n_launch + 2):
            measure = launch >= n_launch
            la = lt = ca = ct = 0
            for _ in range(spl):
                if d.swaps:
                    st = s.sweep_perm(c, sweep, seed, ell, step, d.swaps, d.swap_m, d.swap_rule)
                else:
                    st = s.sweep(c, sweep, seed, ell, step)
                la += st.levy_acc
                lt += st.levy_tries
                ca += st.com_acc
                ct += st.com_tries
                for k in tot:
                    tot[k] += getattr(st, k)
                sweep += 1
            if not measure:
                step = O.adapt_step(step, ca, ct, d.L)
                ell = O.adapt_ell(ell, la, lt, d.M)
        ells_oracle.append(ell)
        changed |= ell != d.ell
        if d.swaps:
            assert np.array_equal(ctx.get_next(c), s.next), c
        assert np.max(np.abs(ctx.get_paths(c) - s.r)) < 1e-9, c
        assert ctx.com_step(c) == pytest.approx(step, rel=1e-12), c
    assert ells_after == ells_oracle, (ells_after, ells_oracle)
    # the tile chains left their shared length, in different directions
    assert changed and ells_after[2] != descs[2].ell and ells_after[2] != ells_after[3]
    for k in tot:
        assert getattr(cnt_therm, k) + getattr(cnt_meas, k) == tot[k], (k, getattr(cnt_therm, k), tot[k])


def test_set_next_validation():
    d_swap = chain_desc(HarmonicTrap(1.0), dim=2, M=8, N=3, L=100.0, beta=1.0, lam=0.5, ell=4, swaps=1)
    d_plain = harmonic_desc(N=3, M=8)
    ctx = make_ctx([d_swap, d_plain], A.FP64)
    ctx.set_next(0, [1, 2, 0])
    assert list(ctx.get_next(0)) == [1, 2, 0]
    with pytest.raises(A.PimcError):
        ctx.set_next(0, [1, 1, 0])  # not a permutation
    with pytest.raises(A.PimcError):
        ctx.set_next(1, [1, 0, 2])  # cycles need swap moves
    pair = GaussianPair(1.0, 0.5)  # swaps with the Gaussian pair are accepted
    c2 = make_ctx([chain_desc(NoPotential(), dim=2, M=8, N=3, L=4.0, beta=1.0, lam=0.5, pair=pair, ell=4, swaps=1)],
                  A.FP64)
    c2.set_next(0, [2, 0, 1])
    assert list(c2.get_next(0)) == [2, 0, 1]


def test_split_estimators_match_in_kernel_fp64():
    """measure_kernel 2 (one sweep per launch, estimators in a wide kernel of
    their own) against 1 (estimators inside the sweep kernel), FP64, for cell-list,
    all-pairs and permutation-cycle chains, with measure_every = 2: identical
    paths, next[] and counters, estimator blocks to 1e-12, identical histograms."""
    rng = np.random.default_rng(53)
    pair = GaussianPair(1.0, 0.5)
    descs = [
        chain_desc(HarmonicTrap(0.3), dim=2, M=16, N=110, L=5.0, beta=2.0, lam=0.5, pair=pair, ell=8, r_a=2.0,
                   com_step=0.5),
        chain_desc(HarmonicTrap(0.3), dim=2, M=16, N=30, L=3.0, beta=2.0, lam=0.5, pair=pair, ell=8, r_a=1.5,
                   com_step=0.5),
        chain_desc(HarmonicTrap(1.0), dim=2, M=12, N=4, L=100.0, beta=2.0, lam=0.5, ell=4, com_step=0.8, swaps=3,
                   swap_m=4),
    ]
    inits = [rng.uniform(-5.0, 5.0, size=(110, 2, 16)), rng.uniform(-3.0, 3.0, size=(30, 2, 16)),
             rand_paths(rng, 4, 12, 100.0, scale=0.6)]
    out = []
    for mk in (1, 2):
        opt = A.Options()
        opt.precision = A.FP64
        opt.sweeps_per_launch = 4
        opt.seed = 19
        opt.measure_every = 2
        opt.radial_bins = 16
        opt.radial_rmax = 4.0
        opt.measure_kernel = mk
        ctx = A.Context(descs, opt)
        for c, r in enumerate(inits):
            ctx.set_paths(c, r)
        ctx.run(4, measure=False)
        cnt = ctx.run(8, measure=True)
        out.append(dict(cnt=(cnt.levy_acc, cnt.com_acc, cnt.swap_acc, cnt.com_tries),
                        paths=[ctx.get_paths(c) for c in range(3)], next=ctx.get_next(2),
                        blocks=[np.array(ctx.blocks(c)) for c in range(3)],
                        radial=[ctx.radial(c) for c in range(3)]))
        ctx.close()
    a, b = out
    assert a["cnt"] == b["cnt"]
    for c in range(3):
        assert np.array_equal(a["paths"][c], b["paths"][c]), c
        assert np.allclose(a["blocks"][c], b["blocks"][c], rtol=1e-12, atol=1e-12), c
        assert np.array_equal(a["radial"][c][0], b["radial"][c][0]), c
    assert np.array_equal(a["next"], b["next"])
    assert a["blocks"][0].shape[1] == 2  # two blocks of 4 sweeps, each with 2 measured sweeps


@pytest.mark.parametrize("precision", [A.FP64, A.FP32])
@pytest.mark.parametrize("helical", [False, True])
def test_optical_lattice_moves_match_oracle(precision, helical):
    """normalized_intensity potentials (the 12-beam l25 superlattice of generate_V
    and helical cubic beams) through the staging and COM hooks: bridge beads and
    Delta V equal the oracle's."""
    from pimc.potentials import BEAMS, OpticalLattice
    rng = np.random.default_rng(81)
    pot = OpticalLattice(BEAMS["cubic"] if helical else BEAMS["l25"], 1.0, -4.0, helical=helical)
    d = chain_desc(pot, dim=2, M=24, N=4, L=8.0, beta=2.0, lam=0.5, ell=8)
    r = rng.uniform(-3.0, 3.0, size=(4, 2, 24))
    ctx = make_ctx([d], precision)
    ctx.set_paths(0, r)
    rr = ctx.get_paths(0)
    osys = oracle_sys(d, rr)
    tol = 1e-11 if precision == A.FP64 else 3e-5
    for trial in range(20):
        n, j0, m = int(rng.integers(4)), int(rng.integers(24)), int(rng.integers(2, 12))
        nrm = rng.standard_normal((m - 1, 2))
        r_gpu, w_gpu = ctx.levy_propose(0, n, j0, m, nrm)
        r_or, w_or = osys.reshape_propose(n, j0, m, nrm)
        assert np.max(np.abs(r_gpu - r_or)) <= tol * 8.0, trial
        scale = abs(w_or.wV_updated) + abs(w_or.wV_initial) + 1e-300
        assert abs((w_gpu[1] - w_gpu[0]) - (w_or.wV_updated - w_or.wV_initial)) <= 10 * tol * scale, trial
    for trial in range(10):
        n = int(rng.integers(4))
        disp = rng.uniform(-0.5, 0.5, size=2)
        r_gpu, w_gpu = ctx.com_propose(0, n, disp)
        r_or, w_or = osys.com_propose(n, disp)
        scale = abs(w_or.wV_updated) + abs(w_or.wV_initial) + 1e-300
        assert abs((w_gpu[1] - w_gpu[0]) - (w_or.wV_updated - w_or.wV_initial)) <= 10 * tol * scale, trial


def test_optical_lattice_sweeps_match_oracle_fp64():
    """Tile kernel (no pair action) and whole-chain kernel (Gaussian pair) with
    the optical potential: paths, counters and energy blocks equal the oracle's
    sweep schedule."""
    from pimc.potentials import BEAMS, OpticalLattice
    rng = np.random.default_rng(83)
    pot = OpticalLattice(BEAMS["l25"], 1.0, -4.0)
    descs = [chain_desc(pot, dim=2, M=32, N=6, L=8.0, beta=2.0, lam=0.5, ell=8, com_step=0.5),
             chain_desc(pot, dim=2, M=16, N=20, L=4.0, beta=2.0, lam=0.5, ell=8, com_step=0.5,
                        pair=GaussianPair(1.0, 0.3), r_a=1.2)]
    inits = [rng.uniform(-3.0, 3.0, size=(6, 2, 32)), rng.uniform(-4.0, 4.0, size=(20, 2, 16))]
    seed, n_sweeps = 23, 4
    ctx = make_ctx(descs, A.FP64, seed=seed, spl=n_sweeps)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    cnt = ctx.run(n_sweeps, measure=True)
    ref = sweep_oracle(descs, inits, seed, n_sweeps, n_sweeps, measure=True)
    assert cnt.levy_acc == sum(x[1] for x in ref) and cnt.com_acc == sum(x[2] for x in ref)
    for c in range(2):
        assert np.max(np.abs(ctx.get_paths(c) - ref[c][0])) < 1e-9, c
        assert_blocks(ctx, c, ref[c][4], n_sweeps)
    assert ctx.blocks(1)[2][0] > 0.0


def test_resume_from_checkpoint_is_exact_fp64():
    """Checkpoint and resume: a run of 6 + 4 thermalising sweeps in one context
    equals 6 sweeps, then a new context given the saved paths, permutation,
    COM steps and Philox sweep counter and run for 4 sweeps, bit for bit
    (tile chain, cell-list pair chain, permutation-cycle chain); without the
    counter the resumed run replays the first sweeps' random streams."""
    rng = np.random.default_rng(401)
    pair = GaussianPair(1.0, 0.5)
    descs = [harmonic_desc(N=5, M=32, beta=2.0, ell=8),
             chain_desc(HarmonicTrap(0.3), dim=2, M=16, N=110, L=5.0, beta=2.0, lam=0.5, pair=pair, ell=8, r_a=2.0,
                        com_step=0.5),
             chain_desc(HarmonicTrap(1.0), dim=2, M=12, N=4, L=100.0, beta=2.0, lam=0.5, ell=4, com_step=0.8, swaps=3,
                        swap_m=4)]
    inits = [rand_paths(rng, 5, 32, 100.0), rng.uniform(-5.0, 5.0, size=(110, 2, 16)),
             rand_paths(rng, 4, 12, 100.0, scale=0.6)]
    full = make_ctx(descs, A.FP64, seed=33, spl=2)
    for c, r in enumerate(inits):
        full.set_paths(c, r)
    full.run(6, measure=False)
    assert full.sweep() == 6
    saved = dict(paths=[full.get_paths(c) for c in range(3)], steps=[full.com_step(c) for c in range(3)],
                 nxt=full.get_next(2), sweep=full.sweep())
    full.run(4, measure=False)
    for use_counter in (True, False):
        res = make_ctx(descs, A.FP64, seed=33, spl=2)
        for c in range(3):
            res.set_paths(c, saved["paths"][c])
            res.set_com_step(c, saved["steps"][c])
        res.set_next(2, saved["nxt"])
        if use_counter:
            res.set_sweep(saved["sweep"])
        res.run(4, measure=False)
        same = all(np.array_equal(res.get_paths(c), full.get_paths(c)) for c in range(3))
        assert same == use_counter
        if use_counter:
            assert np.array_equal(res.get_next(2), full.get_next(2))
            assert all(res.com_step(c) == full.com_step(c) for c in range(3))
            assert res.sweep() == 10
        res.close()
    with pytest.raises(A.PimcError):
        full.set_sweep(-1)


def test_set_pair_table_keeps_the_adapted_staging_length():
    """pimc_set_pair_table replaces only the table of a chain: a staging
    length adapted on the device (adapt_ell) survives a new table."""
    from pimc.propagator import PropInt
    rng = np.random.default_rng(403)
    tau = 2.0 / 24
    terms = O.prop_int_terms(48, 1e-20, 6.0, 2.0, tau)
    p = PropInt(terms, 1e-20, 6.0, 2.0, tau)
    d = chain_desc(HarmonicTrap(0.5), dim=2, M=24, N=16, L=4.0, beta=2.0, lam=0.5, pair=p, r_a=1.0, ell=4,
                   com_step=0.5)
    ctx = make_ctx([d], A.FP64, seed=9, spl=2, adapt_ell=1)
    ctx.set_paths(0, rng.uniform(-3.0, 3.0, size=(16, 2, 24)))
    ctx.set_pair_table(0, terms, 1e-20, 6.0, 0.0)
    for _ in range(40):  # thermalising launches until the length has moved off its start
        ctx.run(2, measure=False)
        if ctx.ell(0) != 4:
            break
    ell = ctx.ell(0)
    assert ell != 4
    ctx.set_pair_table(0, terms, 1e-20, 6.0, 0.0)
    assert ctx.ell(0) == ell
    cnt = ctx.run(2, measure=False)
    assert cnt.levy_tries > 0


@pytest.mark.parametrize("N,L,ra", [(9, 2.0, 1.0), (12, 1.5, 1.5), (110, 5.0, 2.0)])
@pytest.mark.parametrize("pia", [0, 1])
def test_swap_sweeps_with_pairs_match_oracle_fp64(N, L, ra, pia):
    """Swap tries on chains with a Gaussian pair action == the oracle, FP64:
    the pair link sums along the permutation cycles in the staging and COM
    moves, the swap's pair sums as the reference writes them (pair_in_accept
    = 0) or as the pair action's difference (1), cell lists kept in step
    through the swaps (N = 110), the multiplicities of nbins = 2; paths,
    next[], every counter and the (E, Ev, Epair) blocks."""
    rng = np.random.default_rng(47 + N)
    pair = GaussianPair(1.0, 0.5)
    descs, inits = [], []
    for rule, m in ((0, 6), (1, 9)):
        descs.append(chain_desc(HarmonicTrap(0.5), dim=2, M=16, N=N, L=L, beta=1.0, lam=0.5, pair=pair, ell=8, r_a=ra,
                                com_step=0.4, swaps=6, swap_m=m, swap_rule=rule, pair_in_accept=bool(pia)))
        inits.append(rng.uniform(-L, L, size=(N, 1, 1)).repeat(2, axis=1).repeat(16, axis=2)
                     + rng.normal(0, 0.15, size=(N, 2, 16)))
        inits[-1] = (inits[-1] + L) % (2 * L) - L
    seed, n_sweeps = 37, 8
    ctx = make_ctx(descs, A.FP64, seed=seed, spl=n_sweeps)
    for c, r in enumerate(inits):
        ctx.set_paths(c, r)
    cnt = ctx.run(n_sweeps, measure=True)
    tot = dict(levy_acc=0, com_acc=0, com_tries=0, swap_acc=0, swap_tries=0)
    for c, (d, r) in enumerate(zip(descs, inits)):
        s = oracle_sys(d, r)
        meas = []
        for sw in range(n_sweeps):
            st = s.sweep_perm(c, sw, seed, d.ell, d.com_step, d.swaps, d.swap_m, d.swap_rule)
            for k in tot:
                tot[k] += getattr(st, k)
            meas.append(oracle_meas(s))
        assert np.array_equal(ctx.get_next(c), s.next), c
        assert np.max(np.abs(ctx.get_paths(c) - s.r)) < 1e-9, c
        assert_blocks(ctx, c, meas, n_sweeps)
        # the Energy hook on the permuted configuration (links along the cycles)
        E, Ev, Ep, Evp = ctx.energy(c)
        (Eo, Evo), (Po, Pvo) = s.energy(), s.pair_energy()
        big = d.dim * N / (2 * d.beta / d.M)
        assert abs(E - Eo) < 1e-10 * big and abs(Ev - Evo) < 1e-9 * max(1.0, abs(Evo)), c
        assert abs(Ep - Po) < 1e-9 * max(1.0, abs(Po)) and abs(Evp - Pvo) < 1e-9 * max(1.0, abs(Pvo)), c
    for k in tot:
        assert getattr(cnt, k) == tot[k], (k, getattr(cnt, k), tot[k])
    assert tot["swap_acc"] > 0


def test_swap_sweeps_with_pairs_fp32_run():
    """FP32 swap tries with a pair action (cell lists): the permutation stays a
    permutation, the estimators are finite, every try is counted."""
    rng = np.random.default_rng(48)
    d = chain_desc(HarmonicTrap(0.5), dim=2, M=32, N=120, L=5.0, beta=1.0, lam=0.5, pair=GaussianPair(1.0, 0.5),
                   ell=8, r_a=2.0, com_step=0.4, swaps=16, swap_m=8)
    ctx = make_ctx([d], A.FP32, seed=3, spl=20)
    ctx.set_paths(0, rng.uniform(-5.0, 5.0, size=(120, 1, 1)).repeat(2, axis=1).repeat(32, axis=2))
    ctx.run(40, measure=False)
    cnt = ctx.run(40, measure=True)
    assert sorted(ctx.get_next(0)) == list(range(120))
    assert cnt.swap_tries == 40 * 16 and cnt.swap_acc > 0
    E, Ev, Ep = ctx.blocks(0)
    assert np.all(np.isfinite(E)) and np.all(np.isfinite(Ep)) and np.all(Ep > 0)


def test_adapt_ell_keeps_tiles_the_workgroup_holds():
    """A tile chain of M = 1024 slices with a low staging acceptance shortens
    its segments launch by launch until the next length would give a
    particle more segments than a workgroup holds (256 items): that length
    is refused, the chain keeps ell = 4 on the host and the device, and the
    runs go on."""
    d = harmonic_desc(N=1, M=1024, beta=1000.0, ell=8)
    ctx = make_ctx([d], A.FP32, seed=3, spl=1, adapt_ell=1)
    ctx.set_paths(0, np.zeros((1, 2, 1024)))
    ells = []
    for _ in range(8):
        ctx.run(1, measure=False)
        ells.append(ctx.ell(0))
    assert ells[0] < 8 and min(ells) == 4 and ells[-1] == 4, ells
    cnt = ctx.run(2, measure=True)
    assert cnt.levy_tries == 2 * (1024 // 4)


@pytest.mark.parametrize("dim", [1, 2])
def test_estimator_kernels_agree_with_derivatives_fp64(dim):
    """Reweighting derivatives on window-size Gaussian-pair chains (cell
    lists, N = 110 and 130): the same samples measured three ways (the
    whole-chain kernel's in-kernel estimators, its wide k_meas_beads kernel,
    and the window kernels' k_meas_win) give the same E, E_virial, E_pair,
    dS/d omega, dS/dg, d<E>/d omega and d<E>/dg to 1e-9, block by block for
    the energies, and the same radial and Cartesian histograms bin for bin;
    the paths equal the oracle's."""
    rng = np.random.default_rng(71)
    pair = GaussianPair(1.0, 0.5)
    descs, inits = [], []
    for N, L, omega in [(110, 5.0 if dim == 2 else 20.0, 0.3), (130, 6.0 if dim == 2 else 24.0, 0.5)]:
        descs.append(chain_desc(HarmonicTrap(omega), dim=dim, M=16, N=N, L=L, beta=2.0, lam=0.5, pair=pair, ell=6,
                                r_a=2.0, com_step=0.5))
        inits.append(rng.uniform(-L, L, size=(N, dim, 16)))
    seed, spl = 29, 3
    fields = ("E", "Ev", "Epair", "dSdomega", "dSdg", "dE_domega", "dE_dg")
    runs = []
    for pk, mk in [(1, 1), (1, 2), (2, 0)]:
        opt = A.Options()
        opt.precision, opt.sweeps_per_launch, opt.seed = A.FP64, spl, seed
        opt.derivatives, opt.pair_kernel, opt.measure_kernel = 1, pk, mk
        opt.radial_bins, opt.radial_rmax, opt.cart_bins = 40, 4.0, 32
        ctx = A.Context(descs, opt)
        for c, r in enumerate(inits):
            ctx.set_paths(c, r)
        ctx.run(2 * spl, measure=False)
        therm = [ctx.get_paths(c) for c in range(len(descs))]
        ctx.run(4 * spl, measure=True)
        runs.append([(ctx.get_paths(c), ctx.blocks(c), ctx.estimates(c), therm[c], ctx.radial(c), ctx.cartesian(c))
                     for c in range(len(descs))])
        ctx.close()
    ref = sweep_oracle(descs, inits, seed, 2 * spl, spl)
    for c in range(len(descs)):
        assert np.max(np.abs(runs[0][c][3] - ref[c][0])) < 1e-9
        for k in (1, 2):
            assert np.max(np.abs(runs[k][c][0] - runs[0][c][0])) < 1e-9, (k, c)
            for a, b in zip(runs[k][c][1], runs[0][c][1]):
                assert np.allclose(a, b, rtol=1e-9, atol=1e-9), (k, c)
            for f in fields:
                x, y = getattr(runs[k][c][2], f), getattr(runs[0][c][2], f)
                assert abs(x - y) <= 1e-9 * max(1.0, abs(y)), (k, c, f, x, y)
            for h in (4, 5):
                assert runs[k][c][h][1] == runs[0][c][h][1] and np.array_equal(runs[k][c][h][0], runs[0][c][h][0]), (k, c, h)
        assert runs[0][c][2].dSdg != 0.0 and runs[0][c][2].dSdomega != 0.0
        assert runs[0][c][4][0].sum() > 0 and runs[0][c][5][0].sum() > 0
