"""Pins of the FP64 oracle (CPU).

The reference is Julia and cannot run here, and its tests hold no vectors for
the sampling path (its only numeric golden value is the N=1 oscillator energy
of examples/energy_2d_harmonically_trapped_bose_gas.jl:28).  Each oracle
function is therefore pinned by something the mathematics fixes: published
known-answer vectors, closed forms, brute force on small inputs, invariants,
chosen so that a dropped term, a wrong sign or index fails.
"""
import json
import math
import os

import numpy as np
import pytest

import analytic as AN
import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def blockerr(x, nb=40):
    x = np.asarray(x, float)
    b = x[: len(x) // nb * nb].reshape(nb, -1).mean(1)
    return b.mean(), b.std(ddof=1) / math.sqrt(nb)


def harmonic_sys(N, M, beta, L=100.0, seed=0, omega=1.0, scale=1.0):
    r = np.random.default_rng(seed).normal(0, scale, size=(N, 2, M))
    return O.System(r, 0.5, L, beta, O.make_pot(O.POT_HARMONIC, omega))


# ---------------------------------------------------------------- RNG
def test_philox_known_answers():
    kat = json.load(open(os.path.join(GOLD, "philox_kat.json")))["vectors"]
    for v in kat:
        out = O.philox([int(x, 16) for x in v["ctr"]], [int(x, 16) for x in v["key"]])
        assert [f"{x:08x}" for x in out] == v["out"]


def test_box_muller_closed_form_and_moments():
    # u0 = 1/2 exactly when x0 = 2^31 - 1/2 is not representable; take x0 = 2^31:
    # u0 = (2^31 + 0.5) 2^-32, u1 = 0.25 + 0.5/2^32 -> n0 = rad cos(pi/2 + eps)
    n0, n1 = O.box_muller(2 ** 31, 2 ** 30)
    u0 = (2 ** 31 + 0.5) / 2 ** 32
    u1 = (2 ** 30 + 0.5) / 2 ** 32
    rad = math.sqrt(-2 * math.log(u0))
    assert n0 == pytest.approx(rad * math.cos(2 * math.pi * u1), abs=1e-15)
    assert n1 == pytest.approx(rad * math.sin(2 * math.pi * u1), abs=1e-15)
    rng = np.random.default_rng(1)
    xs = rng.integers(0, 2 ** 32, size=(20000, 2), dtype=np.uint64)
    z = np.array([O.box_muller(int(a), int(b)) for a, b in xs]).ravel()
    assert abs(z.mean()) < 4 / math.sqrt(z.size)
    assert abs(z.var() - 1.0) < 4 * math.sqrt(2 / z.size)


# ---------------------------------------------------------------- primitives
def test_distance_and_teleport():
    L = 3.0
    assert O.distance(0.9 * L, -0.9 * L, L) == pytest.approx(0.2 * L)
    assert O.distance(1.0, 1.0, L) == 0.0
    rng = np.random.default_rng(2)
    for a, b in rng.uniform(-L, L, size=(200, 2)):
        d = O.distance(a, b, L)
        assert d == pytest.approx(O.distance(b, a, L))
        assert 0.0 <= d <= L + 1e-12
        # equals the minimum over periodic images
        assert d == pytest.approx(min(abs(a - b + k * 2 * L) for k in (-1, 0, 1)), abs=1e-12)
    assert O.teleport(L, L) == -L
    assert O.teleport(-L, L) == -L
    for x in rng.uniform(-L, L, size=100):
        for k in (-2, -1, 1, 3):
            assert O.teleport(x + 2 * L * k, L) == pytest.approx(x, abs=1e-12)
        assert O.teleport(x, L) == pytest.approx(x, abs=1e-15)


def test_levy_one_interior_bead_by_hand():
    tau, lam, L = 0.1, 0.5, 50.0
    r = np.array([[1.0, -2.0], [0.0, 0.0], [3.0, 4.0]])
    xi = np.array([0.7, -1.3])
    out = O.levy(r, tau, L, lam, xi)
    # m = 1 interior bead: alpha = 1/2, sd = sqrt(2 lambda tau / 2)
    expect = 0.5 * r[0] + 0.5 * r[2] + xi * math.sqrt(lam * tau)
    # teleport's (x + L) - L rounds at the scale of L
    assert np.allclose(out[1], expect, rtol=0, atol=4 * np.spacing(L))
    assert np.allclose(out[[0, 2]], r[[0, 2]])


def test_levy_end_shift_across_boundary():
    tau, lam, L = 0.1, 0.5, 2.0
    r = np.array([[1.8, 0.0], [0.0, 0.0], [-1.8, 0.0]])
    out = O.levy(r, tau, L, lam, np.zeros(2))
    # the far end moves by +2L to 2.2, the midpoint 2.0 wraps to -2.0
    assert out[1, 0] == pytest.approx(-2.0)
    assert out[2, 0] == pytest.approx(-1.8)
    # sign(0) = 0: an end more than L away from a start at 0 is not shifted
    r = np.array([[0.0, 0.0], [0.0, 0.0], [2.5, 0.0]])
    out = O.levy(r, tau, L, lam, np.zeros(2))
    assert out[1, 0] == pytest.approx(O.teleport(1.25, L))


def test_levy_bridge_moments():
    """Bead k of a K-interval bridge: mean r0 + k/K (rK - r0), variance 2 lambda tau k (K-k)/K."""
    tau, lam, L, K = 0.05, 0.5, 100.0, 6
    r = np.zeros((K + 1, 1))
    r[0, 0], r[K, 0] = -1.0, 2.0
    rng = np.random.default_rng(3)
    n = 20000
    out = np.array([O.levy(r, tau, L, lam, rng.standard_normal(K - 1))[:, 0] for _ in range(n)])
    for k in range(1, K):
        mean = -1.0 + k / K * 3.0
        var = 2 * lam * tau * k * (K - k) / K
        assert abs(out[:, k].mean() - mean) < 4 * math.sqrt(var / n)
        assert abs(out[:, k].var() - var) < 4 * var * math.sqrt(2 / n)


# ---------------------------------------------------------------- moves vs brute force
def test_reshape_delta_equals_full_action_difference():
    s = harmonic_sys(5, 24, 2.0, seed=4, scale=1.3)
    rng = np.random.default_rng(4)
    for _ in range(40):
        n, j0, m = int(rng.integers(5)), int(rng.integers(24)), int(rng.integers(2, 12))
        before = s.total_lnV()
        acc, w = s.reshape_apply(n, j0, m, rng.standard_normal((m - 1, 2)), 0.0)
        assert acc  # u = 0 accepts every move
        after = s.total_lnV()
        scale = abs(w.wV_updated) + abs(w.wV_initial)
        assert (w.wV_updated - w.wV_initial) == pytest.approx(after - before, abs=1e-12 * scale + 1e-13)


def test_com_keeps_kinetic_action_and_matches_potential_difference():
    s = harmonic_sys(3, 20, 1.0, L=5.0, seed=5, scale=2.0)
    rng = np.random.default_rng(5)
    for _ in range(20):
        n = int(rng.integers(3))
        d = rng.uniform(-3, 3, size=2)
        k0, v0 = s.total_lnK(), s.total_lnV()
        acc, w = s.com_apply(n, d, 0.0)
        assert acc
        assert s.total_lnK() == pytest.approx(k0, rel=1e-10)  # relative positions unchanged, minimum image
        assert (w.wV_updated - w.wV_initial) == pytest.approx(s.total_lnV() - v0, abs=1e-10)


def pair_sys(N, L, M, ra, seed, g=1.0, sigma=0.5):
    r = np.random.default_rng(seed).uniform(-L, L, size=(N, 2, M))
    return O.System(r, 0.5, L, 2.0, O.make_pot(), O.make_pair(O.PAIR_GAUSSIAN, g, sigma), ra, True)


def brute_neighbours(s, target, j, exclude, nbins):
    """Independent restatement of the 9-cell stencil with multiplicities."""
    r = s.r
    w = 2 * s._s.L / nbins
    def cell(p):
        return tuple(int(math.floor((p[d] + s._s.L) / w)) for d in range(2))
    ct = cell(target)
    out = []
    for k in range(s.N):
        if k == exclude:
            continue
        pk = r[k, :, j]
        ck = cell(pk)
        mult = sum(1 for dx in (-1, 0, 1) for dy in (-1, 0, 1)
                   if ((ct[0] + dx) % nbins, (ct[1] + dy) % nbins) == ck)
        dist = math.sqrt(sum(O.distance(pk[d], target[d], s._s.L) ** 2 for d in range(2)))
        if mult and dist <= s.h:
            out += [k] * mult
    return sorted(out)


@pytest.mark.parametrize("N,L,ra", [(40, 4.0, 2.0), (30, 1.5, 1.5), (25, 2.0, 3.0), (60, 6.0, 1.0)])
def test_find_nns_against_brute_force(N, L, ra):
    s = pair_sys(N, L, 4, ra, seed=6)
    rng = np.random.default_rng(6)
    for _ in range(30):
        n, j = int(rng.integers(N)), int(rng.integers(4))
        t = s.r[n, :, j]
        assert sorted(s.find_nns(t, j, n)) == brute_neighbours(s, t, j, n, s.nbins)


def test_pair_sums_against_brute_force_total():
    s = pair_sys(50, 4.0, 6, 2.0, seed=7)   # nbins = 4: no repeated cells
    tot = sum(s.interaction_old(n, j) for n in range(50) for j in range(6))
    assert tot == pytest.approx(2 * s.total_lnU(), rel=1e-12)
    rng = np.random.default_rng(7)
    for _ in range(20):
        n, j0, m = int(rng.integers(50)), int(rng.integers(6)), int(rng.integers(2, 5))
        before = s.total_lnU()
        acc, w = s.reshape_apply(n, j0, m, rng.standard_normal((m - 1, 2)), 0.0)
        assert acc
        assert (w.wU_updated - w.wU_initial) == pytest.approx(s.total_lnU() - before, abs=1e-12)
    for _ in range(10):
        n = int(rng.integers(50))
        before = s.total_lnU()
        acc, w = s.com_apply(n, rng.uniform(-1, 1, size=2), 0.0)
        assert acc
        assert (w.wU_updated - w.wU_initial) == pytest.approx(s.total_lnU() - before, abs=1e-12)


def test_pair_energy_matches_total_action():
    s = pair_sys(40, 4.0, 6, 2.0, seed=8)
    Ep, _ = s.pair_energy()
    # E_pair = <S_U> / beta with S_U = -sum lnU
    assert Ep == pytest.approx(-s.total_lnU() / 2.0, rel=1e-12)


# ---------------------------------------------------------------- estimators
def test_density_bin_convention():
    L, nb = 2.0, 4
    w = 2 * L / nb
    # beads at the centres of bins (bx, by), 0-based; the reference stores bin b at index b-1
    # and drops b = 0 (src/measurement.jl:48-50)
    pts = [(0, 0), (1, 2), (3, 3), (2, 0), (1, 1)]
    r = np.zeros((1, 2, len(pts)))
    for j, (bx, by) in enumerate(pts):
        r[0, :, j] = (-L + (bx + 0.5) * w, -L + (by + 0.5) * w)
    s = O.System(r, 0.5, L, 1.0, O.make_pot())
    dens, nd = s.density(nb)
    assert nd == len(pts)
    expect = np.zeros((nb, nb))
    for bx, by in pts:
        if bx > 0 and by > 0:
            expect[bx - 1, by - 1] += 1
    assert np.array_equal(dens, expect)
    assert dens.sum() == 3  # (0,0) and (2,0) dropped


def test_radial_histogram():
    s = harmonic_sys(6, 16, 1.0, seed=9)
    h, nd = s.radial(8, 3.0)
    rr = np.sqrt((s.r ** 2).sum(axis=1)).ravel()
    expect, _ = np.histogram(rr, bins=8, range=(0, 3.0))
    assert np.array_equal(h, expect)
    assert nd == 16


def test_reference_example_energy():
    """examples/energy_2d_harmonically_trapped_bose_gas.jl: N=1, M=5, beta=1,
    COM + ReshapeLinear(s, 2).  The sampled thermodynamic and virial energies
    reach the M=5 primitive-action value (the example's `exact`, 2.16395...,
    is the continuum limit it approaches as M grows)."""
    gold = json.load(open(os.path.join(GOLD, "analytic_energies.json")))
    assert AN.energy_boltzmann_continuum(1, 1.0) == pytest.approx(gold["reference_example_exact"], rel=1e-14)
    assert AN.energy_boltzmann_primitive(1, 4096, 1.0) == pytest.approx(gold["reference_example_exact"], rel=1e-6)
    s = harmonic_sys(1, 5, 1.0, seed=10)
    run = O.RefRunner(s, cap=2, step=1.0, seed=3)
    run.run(20000)
    E, Ev = run.run(400000, ncycle=10, measure=True)
    exact5 = AN.energy_boltzmann_primitive(1, 5, 1.0)
    m, e = blockerr(E)
    assert abs(m - exact5) < 3.5 * e, (m, e, exact5)
    m, e = blockerr(Ev)
    assert abs(m - exact5) < 3.5 * e + 0.01, (m, e, exact5)


def test_sweep_schedule_energy_and_x2():
    """The sweep schedule of the kernels samples the primitive-action
    distribution: <E>, <E_virial> and <x^2> against the normal-mode closed form."""
    N, M, beta = 4, 16, 2.0
    s = harmonic_sys(N, M, beta, seed=11)
    E, Ev, x2 = [], [], []
    for sw in range(30000):
        s.sweep(0, sw, 77, 4, 1.0)
        if sw >= 500:
            e = s.energy()
            E.append(e[0])
            Ev.append(e[1])
            x2.append(np.mean(s.r ** 2))
    for vals, exact in ((E, AN.energy_boltzmann_primitive(N, M, beta)),
                        (Ev, AN.energy_boltzmann_primitive(N, M, beta)),
                        (x2, AN.x2_primitive(M, beta))):
        m, e = blockerr(vals)
        assert abs(m - exact) < 3.5 * e, (m, e, exact)


def test_reference_schedule_counts_and_adaptation():
    s = harmonic_sys(3, 32, 1.0, seed=12)
    run = O.RefRunner(s, cap=8, step=1.0, seed=4)
    run.run(5000)
    st = run.st
    assert st.levy_tries + st.com_tries == 5000
    assert st.bead_moves >= st.com_tries * 32 + st.levy_tries
    assert 2 <= st.cap <= 30 and 0.1 <= st.step <= 50.0


def test_sweep_bead_move_count_and_step_rule():
    # M = 10, ell = 4: segments of 4, 4, 2 links -> 3 + 3 + 1 interior beads, + M for COM
    assert O.sweep_bead_moves(1, 10, 4) == 7 + 10
    assert O.sweep_bead_moves(3, 64, 8) == 3 * (8 * 7 + 64)
    assert O.adapt_step(1.0, 1, 10, 100.0) == pytest.approx(0.9)
    assert O.adapt_step(1.0, 9, 10, 100.0) == pytest.approx(1.1)
    assert O.adapt_step(1.0, 5, 10, 100.0) == 1.0
    assert O.adapt_step(0.1, 0, 10, 100.0) == 0.1
    assert O.adapt_step(49.0, 10, 10, 100.0) == 50.0


def test_staging_length_rule():
    """adjust!(::NumbOfSlices) (src/updates/helper.jl:42-52): strict
    comparisons at 0.6 and 0.8, one step per call, clamped to [2, M - 2]
    (ReshapeLinear's defaults) and to 32 links."""
    assert O.adapt_ell(8, 59, 100, 64) == 7
    assert O.adapt_ell(8, 60, 100, 64) == 8   # 0.6 is not below 0.6
    assert O.adapt_ell(8, 80, 100, 64) == 8   # 0.8 is not above 0.8
    assert O.adapt_ell(8, 81, 100, 64) == 9
    assert O.adapt_ell(2, 0, 100, 64) == 2
    assert O.adapt_ell(32, 100, 100, 64) == 32
    assert O.adapt_ell(10, 100, 100, 12) == 10  # maxslices = M - 2
    assert O.adapt_ell(12, 50, 100, 12) == 10   # above maxslices: clamped, not just decremented
    assert O.adapt_ell(3, 100, 100, 4) == 2    # M - 2 = 2
    assert O.adapt_ell(2, 100, 100, 3) == 2    # M - 2 < 2: floor 2
    assert O.adapt_ell(5, 0, 0, 64) == 5       # no tries: unchanged


def test_init_world_structure_and_keys():
    """init_world (src/system.jl:36-78, a = 0): first bead == last bead, all in
    [-L, L), reproducible for (seed, chain), different for another chain."""
    r = O.init_world(11, 2, 6, 12, 2, 3.0, 0.2, 0.5)
    assert r.shape == (6, 2, 12)
    assert np.array_equal(r[:, :, 0], r[:, :, -1])
    assert np.all(r >= -3.0) and np.all(r < 3.0)
    assert np.array_equal(r, O.init_world(11, 2, 6, 12, 2, 3.0, 0.2, 0.5))
    assert not np.allclose(r, O.init_world(11, 3, 6, 12, 2, 3.0, 0.2, 0.5))


def _min_pair_distance_per_slice(r, L):
    """Smallest distance (norm of the per-component |minimum-image| offsets)
    between beads of different particles at the same slice, brute force."""
    N = r.shape[0]
    best = np.inf
    for i in range(N):
        for k in range(i):
            d = np.abs(r[i] - r[k])
            d = np.minimum(d, 2 * L - d)
            best = min(best, float(np.sqrt((d ** 2).sum(axis=0)).min()))
    return best


def test_init_world_hard_spheres():
    """init_world with hard spheres (src/system.jl:44-64): a = 0 gives the plain
    init_world; with a > 0 no two beads of a slice end closer than a; a
    particle placed on its first try keeps the plain path (try 0 draws), and a
    particle that needed more tries had a first-try path that came closer than
    a to an earlier particle's beads; an impossible a reports failure."""
    N, M, dim, L, tau, lam = 12, 16, 2, 1.5, 0.1, 0.5
    plain = O.init_world(1, 0, N, M, dim, L, tau, lam)
    r0, t0, rc0 = O.init_world_hs(1, 0, N, M, dim, L, tau, lam, 0.0)
    assert rc0 == 0 and np.array_equal(r0, plain) and np.all(t0 == 1)
    a = 0.3
    r, tries, rc = O.init_world_hs(1, 0, N, M, dim, L, tau, lam, a)
    assert rc == 0
    assert _min_pair_distance_per_slice(r, L) >= a
    assert np.array_equal(r[:, :, 0], r[:, :, -1]) and np.all(r >= -L) and np.all(r < L)
    assert tries[0] == 1 and tries.max() > 1
    for p in range(N):
        if tries[p] == 1:
            assert np.array_equal(r[p], plain[p])
        else:
            # the first try's path overlaps a particle placed before p
            assert _min_pair_distance_per_slice(np.concatenate([r[:p], plain[p:p + 1]]), L) < a, p
    _, _, rc_bad = O.init_world_hs(1, 0, 6, 8, dim, 1.0, tau, lam, 2.5)  # a > box diagonal
    assert rc_bad == -1


def test_init_world_statistics():
    """Start beads uniform in the box (Kolmogorov-Smirnov), interior beads a
    Brownian bridge from the start bead back to itself: mean 0 and variance
    2 lambda tau i (K - i) / K per coordinate relative to the start, K = M - 1
    links (a wide box, so no wrap)."""
    from scipy import stats
    N, M, L, tau, lam = 20000, 9, 1000.0, 0.3, 0.5
    r = O.init_world(5, 0, N, M, 2, L, tau, lam)
    start = r[:, :, 0].ravel()
    assert stats.kstest((start + L) / (2 * L), "uniform").pvalue > 1e-3
    K = M - 1
    rel = r - r[:, :, :1]
    rel -= 2 * L * np.round(rel / (2 * L))  # minimum image: paths that start near the wall wrap
    for i in range(1, M - 1):
        var = 2 * lam * tau * i * (K - i) / K
        x = rel[:, :, i].ravel()
        assert abs(x.mean()) < 5 * math.sqrt(var / x.size)
        # variance of a variance estimate: 2 var^2 / n
        assert abs(x.var() - var) < 5 * var * math.sqrt(2 / x.size), (i, x.var(), var)


def _oracle_trap(N, M, beta, seed=3):
    rng = np.random.default_rng(seed)
    r = rng.normal(0, 0.7, size=(N, 2, M))
    return O.System(r, 0.5, 100.0, beta, O.make_pot(1, 1.0, 0.0, 1.0), None, 0.0, False)


def _blocked(x):
    x = np.asarray(x, float)
    b, errs = x.copy(), []
    while len(b) >= 16:
        errs.append(b.std(ddof=1) / math.sqrt(len(b)))
        b = 0.5 * (b[: len(b) // 2 * 2: 2] + b[1: len(b) // 2 * 2: 2])
    return x.mean(), max(errs)


def test_swap_exchange_sector_probability():
    """Two trapped bosons: the fraction of configurations in the exchanged
    sector is z_2 / (z_1^2 + z_2), z_k the partition function of one k-cycle
    (a ring of k M beads).  The swap move with the heat-bath normalisations
    (rule 0) samples it; the reference's acceptance exp(w_updated - w_initial)
    (rule 1, src/updates/reshape.jl:266) does not."""
    M, beta, m = 8, 1.0, 4
    tau = beta / M
    lz1, lz2 = 2 * AN._ln_ring(M, tau, 1.0, 0.5), 2 * AN._ln_ring(2 * M, tau, 1.0, 0.5)
    p = math.exp(lz2) / (math.exp(2 * lz1) + math.exp(lz2))
    out = {}
    for rule in (0, 1):
        s = _oracle_trap(2, M, beta)
        f = []
        for sw in range(150000):
            s.sweep_perm(0, sw, 91, 4, 1.0, 2, m, rule)
            if sw >= 2000:
                f.append(float(s.next[0] != 0))
        out[rule] = _blocked(f)
    (f0, e0), (f1, e1) = out[0], out[1]
    assert abs(f0 - p) < 3.5 * e0, (f0, e0, p)
    assert (f1 - p) / e1 > 3.0, (f1, e1, p)


def test_swap_boson_energy_closed_form():
    """Three trapped bosons, M = 8, beta = 2: the thermodynamic energy with
    swaps matches -d ln Z_3 / d beta of the primitive action (canonical
    recursion over cycles), far from the distinguishable-particle value."""
    N, M, beta = 3, 8, 2.0
    s = _oracle_trap(N, M, beta, seed=1)
    E = []
    for sw in range(150000):
        st = s.sweep_perm(0, sw, 77, 4, 1.0, 3, 4, 0)
        if sw >= 2000:
            E.append(s.energy()[0])
    nxt = s.next
    assert sorted(nxt.tolist()) == list(range(N))  # still a permutation
    e, err = _blocked(E)
    exact = AN.energy_boson_primitive(N, M, beta)
    assert abs(e - exact) < 3.5 * err, (e, err, exact)
    assert AN.energy_boltzmann_primitive(N, M, beta) - e > 20 * err


def test_optical_lattice_known_answers_and_gradient():
    """normalized_intensity of the reference's own test (test/testpotential.jl:
    19-30): three beams at 0, 2pi/3, 4pi/3 with scale 1 give intensity 1 at the
    origin, at (2/sqrt(3), 0) and at (0, 2/3); the gradient matches central
    differences, for plain and helical beams."""
    cyclic_3 = [2 * math.pi * k / 3 for k in range(3)]
    pot = O.make_pot(O.POT_OPTICAL, V0=1.0, angles=cyclic_3, scale=1.0)
    for r in ([0.0, 0.0], [2 / math.sqrt(3), 0.0], [0.0, 2 / 3]):
        assert O.pot_V(pot, r) == pytest.approx(1.0, rel=1e-12)
    # intensity is at most 1 and 0 where the beams cancel
    rng = np.random.default_rng(2)
    vals = [O.pot_V(pot, x) for x in rng.uniform(-2, 2, size=(200, 2))]
    assert max(vals) <= 1.0 + 1e-12 and min(vals) >= 0.0
    cubic = [2 * math.pi * k / 4 for k in range(4)]
    for helical in (False, True):
        p = O.make_pot(O.POT_OPTICAL, V0=-4.0, angles=cubic + [0.3], scale=0.7, helical=helical)
        for x in rng.uniform(-3, 3, size=(10, 2)):
            g = O.pot_dV(p, x)
            h = 1e-6
            fd = [(O.pot_V(p, x + h * e) - O.pot_V(p, x - h * e)) / (2 * h) for e in np.eye(2)]
            assert np.allclose(g, fd, rtol=1e-6, atol=1e-7)


# ---------------------------------------------------------------- square lattice and Gaussian pair
@pytest.mark.parametrize("V0,a", [(4.0, 1.0), (-2.5, 0.7)])
def test_square_lattice_special_values_period_and_gradient(V0, a):
    """V = V0 sum_d cos^2(pi x_d / a) (the lattice of the interacting-lattice
    workload): maxima dim V0 at the lattice points k a, zeros at (k + 1/2) a,
    V0 (1 + 1/2) at (a/4, 0) and V0 (1 + 1/4) at (a/3, 0), period a in each
    coordinate, a stationary point at every (k + 1/2) a and k a, and the
    gradient equal to central differences.  A cos(2 pi x / a), a missing
    square or a gradient off by a factor 2 fails one of these."""
    pot = O.make_pot(O.POT_LATTICE, V0=V0, a=a)
    assert O.pot_V(pot, [0.0, 0.0]) == pytest.approx(2 * V0, rel=1e-15)
    assert O.pot_V(pot, [a / 2, a / 2]) == pytest.approx(0.0, abs=1e-15 * abs(V0))
    assert O.pot_V(pot, [-a / 2, 3 * a / 2]) == pytest.approx(0.0, abs=1e-14 * abs(V0))
    assert O.pot_V(pot, [a / 4, 0.0]) == pytest.approx(1.5 * V0, rel=1e-14)
    assert O.pot_V(pot, [a / 3, 0.0]) == pytest.approx(1.25 * V0, rel=1e-14)
    assert O.pot_V(pot, [0.0, a / 4]) == pytest.approx(1.5 * V0, rel=1e-14)
    assert O.pot_V(pot, [a / 4]) == pytest.approx(0.5 * V0, rel=1e-14)  # 1D
    rng = np.random.default_rng(91)
    for x in rng.uniform(-5, 5, size=(20, 2)):
        v = O.pot_V(pot, x)
        for shift in ([a, 0.0], [0.0, -a], [3 * a, 2 * a]):
            assert O.pot_V(pot, x + np.array(shift)) == pytest.approx(v, rel=1e-12, abs=1e-12)
        g = O.pot_dV(pot, x)
        h = 1e-6
        fd = [(O.pot_V(pot, x + h * e) - O.pot_V(pot, x - h * e)) / (2 * h) for e in np.eye(2)]
        assert np.allclose(g, fd, rtol=1e-7, atol=1e-7 * abs(V0))
    for k in (-2, 0, 3):
        for pt in ([(k + 0.5) * a, (k + 0.5) * a], [k * a, (k + 0.5) * a]):
            assert np.allclose(O.pot_dV(pot, pt), 0.0, atol=1e-12 * abs(V0) / a)
    # the extremes over a fine grid of one cell are 0 and 2 V0
    xs = np.linspace(0, a, 41)
    vals = [O.pot_V(pot, [x, y]) for x in xs for y in xs]
    assert min(vals) == pytest.approx(min(0.0, 2 * V0), abs=1e-12)
    assert max(vals) == pytest.approx(max(0.0, 2 * V0), abs=1e-12)


def test_reference_sine_lattice_is_a_shifted_square_lattice():
    """The reference's tests sample v1d(x) = depth sin^2(2 pi x scale)
    (test/testsystem.jl, test/testmeasurements.jl).  That potential is
    depth + V of the square lattice with V0 = -depth and a = 1 / (2 scale),
    so a constant shift of the action: the same sampled paths."""
    depth, scale = 3.0, 0.8
    pot = O.make_pot(O.POT_LATTICE, V0=-depth, a=1.0 / (2 * scale))
    for x in np.linspace(-3, 3, 37):
        v1d = depth * math.sin(2 * math.pi * x * scale) ** 2
        assert O.pot_V(pot, [x]) + depth == pytest.approx(v1d, abs=1e-13)


def test_gaussian_pair_special_values_and_link_action():
    """v(r) = g exp(-r^2 / (2 sigma^2)): v(0) = g, v(sigma) = g e^-1/2,
    v(sqrt(2) sigma) = g / e, v(4 sigma) = g e^-8 (the cut-off of the
    interacting workloads).  The primitive link action lnU of one pair is
    -tau/2 (v(|r_j|) + v(|r_j+1|)) from the two per-component offsets."""
    g, sigma, tau = 1.5, 0.5, 0.1
    pair = O.make_pair(O.PAIR_GAUSSIAN, g, sigma)
    assert O.vpair(pair, 0.0) == g
    assert O.vpair(pair, sigma) == pytest.approx(g * math.exp(-0.5), rel=1e-15)
    assert O.vpair(pair, math.sqrt(2) * sigma) == pytest.approx(g / math.e, rel=1e-15)
    assert O.vpair(pair, 4 * sigma) == pytest.approx(g * math.exp(-8.0), rel=1e-14)
    # offsets (0.3, 0.4): |r| = 0.5 = sigma; (0, 0): v = g; (2, 0): 4 sigma; (0.5, 0.5): sqrt(2) sigma
    assert O.lnU(pair, [0.3, 0.4], [0.0, 0.0], tau) == pytest.approx(-0.05 * (g * math.exp(-0.5) + g), rel=1e-15)
    assert O.lnU(pair, [2.0, 0.0], [0.5, 0.5], tau) == pytest.approx(-0.05 * (g * math.exp(-8) + g / math.e),
                                                                     rel=1e-14)
    assert O.lnU(pair, [0.5], [0.0], tau) == pytest.approx(-0.05 * (g * math.exp(-0.5) + g), rel=1e-15)  # 1D
    assert O.vpair(O.make_pair(), 0.3) == 0.0


def test_pair_energy_of_a_hand_placed_pair():
    """Two particles, M = 4, beta = 4 (tau = 1), in a wide box: separation
    (0.3, 0.4) at slices 0 and 2 (|r| = sigma) and (0.5, 0.5) at slices 1
    and 3 (|r| = sqrt(2) sigma).  Every link finds the pair, so S_U =
    tau/2 * 4 (v0 + v1) and sum r . grad S_U = tau/2 * 4 (-(r0^2/sigma^2) v0
    - (r1^2/sigma^2) v1) = 2 (-v0 - 2 v1); E_pair = S_U / beta, the pair
    virial E_pair + that / (2 beta)."""
    g, sigma = 1.3, 0.5
    r = np.zeros((2, 2, 4))
    r[0, :, 0] = (1.0, -2.0)
    r[1, :, 0] = (1.3, -1.6)
    r[0, :, 1] = (0.2, 0.7)
    r[1, :, 1] = (0.7, 1.2)
    r[:, :, 2:] = r[:, :, :2] + np.array([0.1, -0.2])[None, :, None]
    s = O.System(r, 0.5, 10.0, 4.0, O.make_pot(), O.make_pair(O.PAIR_GAUSSIAN, g, sigma), 2.0, True)
    v0, v1 = g * math.exp(-0.5), g / math.e
    Ep, Evp = s.pair_energy()
    assert Ep == pytest.approx((v0 + v1) / 2.0, rel=1e-14)
    assert Evp == pytest.approx((v0 + v1) / 2.0 + (-v0 - 2.0 * v1) / 4.0, rel=1e-14)
    assert -s.total_lnU() == pytest.approx(2 * (v0 + v1), rel=1e-14)


def test_pair_virial_against_scaling_derivative():
    """sum r . grad S_U = d/ds S_U(s r) at s = 1 (the positions scaled about
    the origin, inside a wide box so that no offset wraps and the neighbour
    sets do not change): the pair virial of o_pair_energy against a central
    difference of the brute-force total action."""
    N, M, L, beta = 12, 6, 20.0, 2.0
    rng = np.random.default_rng(92)
    r = rng.uniform(-2.0, 2.0, size=(N, 2, M))
    pair = O.make_pair(O.PAIR_GAUSSIAN, 0.7, 0.5)
    # r_a = 8: nbins = 5, the stencil covers every particle and the cut-off 8 exceeds every distance
    s = O.System(r, 0.5, L, beta, O.make_pot(), pair, 8.0, True)
    Ep, Evp = s.pair_energy()
    SU = -s.total_lnU()
    assert Ep == pytest.approx(SU / beta, rel=1e-12)
    h = 1e-5
    s.r = r * (1 + h)
    up = -s.total_lnU()
    s.r = r * (1 - h)
    dn = -s.total_lnU()
    rgrad = (up - dn) / (2 * h)
    assert (Evp - Ep) * 2 * beta == pytest.approx(rgrad, rel=1e-7)


def test_thermodynamic_equals_virial_energy_with_pair():
    """For a confined system the primitive-action thermodynamic and virial
    estimators have the same mean at every M (scaling identity of the
    discretised path integral), the pair terms included: E_pair thermodynamic
    S_U / beta against the virial S_U / beta + (r . grad S_U) / (2 beta).
    Trapped 2D particles with the Gaussian repulsion (the differentiation
    workload's interaction, g = 0.5, sigma = 0.5), in the oracle's sweep
    schedule.  A wrong sign or factor in grad V or in the pair virial breaks
    the equality."""
    N, M, beta = 4, 12, 2.0
    rng = np.random.default_rng(93)
    r = rng.normal(0, 0.7, size=(N, 2, M))
    pair = O.make_pair(O.PAIR_GAUSSIAN, 0.5, 0.5)
    s = O.System(r, 0.5, 100.0, beta, O.make_pot(O.POT_HARMONIC, 1.0), pair, 2.0, True)
    E, Ev, Ep = [], [], []
    for sw in range(16000):
        s.sweep(0, sw, 5, 4, 0.8)
        if sw >= 400:
            e0, ev0 = s.energy()
            p0, pv0 = s.pair_energy()
            E.append(e0 + p0)
            Ev.append(ev0 + pv0)
            Ep.append(p0)
    d = np.array(E) - np.array(Ev)
    m, e = blockerr(d)
    assert abs(m) < 3.5 * e, (m, e)
    # the pair terms matter at this size: the pair energy is far from zero
    mp, ep = blockerr(Ep)
    assert mp > 20 * ep and mp > 0.05 * np.mean(Ev)


# ---------------------------------------------------------------- move weights
def _swap_trap(seed=5):
    rng = np.random.default_rng(seed)
    return O.System(rng.normal(0, 0.7, size=(4, 2, 16)), 0.5, 100.0, 1.0, O.make_pot(1, 1.0, 0.0, 1.0), None, 0.0,
                    False)


def test_move_weights_every_one_is_the_plain_sweep():
    """every = (1, 1, 1) is o_sweep_perm bit for bit, swaps included."""
    a, b = _swap_trap(), _swap_trap()
    for sw in range(40):
        sa = a.sweep_perm(0, sw, 7, 4, 0.8, 2, 5)
        sb = b.sweep_every(0, sw, 7, 4, 0.8, (1, 1, 1), 2, 5)
        assert (sa.levy_acc, sa.levy_tries, sa.com_acc, sa.com_tries, sa.swap_acc, sa.swap_tries) == \
            (sb.levy_acc, sb.levy_tries, sb.com_acc, sb.com_tries, sb.swap_acc, sb.swap_tries)
    assert np.array_equal(a.r, b.r) and np.array_equal(a.next, b.next)


def test_move_weights_passes_switched_off():
    """A sweep without its staging pass moves each ring rigidly (every bead
    by one shift, modulo the box) or not at all; without its COM pass it is the
    sweep whose COM shifts are zero (com_step = 0: accepted, nothing moves);
    without its swap tries it is the sweep with no swap tries."""
    L = 100.0
    # staging off (sweeps 1..29 of staging_every = 1000)
    s = harmonic_sys(5, 24, 1.5, seed=21)
    for sw in range(1, 30):
        r0 = s.r.copy()
        st = s.sweep_every(0, sw, 3, 6, 1.0, (1000, 1, 1))
        assert st.levy_tries == 0 and st.com_tries == 5
        d = (s.r - r0 + L) % (2 * L) - L
        assert np.allclose(d, d[:, :, :1], atol=1e-12)
        moved = np.abs(d[:, :, 0]).sum(1) > 0
        assert moved.sum() == st.com_acc
    # COM off on odd sweeps of com_every = 2 == com_step 0
    a, b = harmonic_sys(5, 24, 1.5, seed=22), harmonic_sys(5, 24, 1.5, seed=22)
    for sw in range(1, 40, 2):
        sa = a.sweep_every(0, sw, 9, 6, 1.0, (1, 2, 1))
        sb = b.sweep(0, sw, 9, 6, 0.0)
        assert sa.com_tries == 0 and sb.com_tries == 5 and sb.com_acc == 5
        assert (sa.levy_acc, sa.levy_tries) == (sb.levy_acc, sb.levy_tries)
    # the zero shift still passes through teleport, which rounds: equal to rounding
    assert np.allclose(a.r, b.r, rtol=0, atol=1e-12)
    # swap tries off on sweeps with sweep % 3 != 0
    a, b = _swap_trap(8), _swap_trap(8)
    for sw in range(0, 30):
        sa = a.sweep_every(0, sw, 11, 4, 0.8, (1, 1, 3), 2, 5)
        sb = b.sweep_perm(0, sw, 11, 4, 0.8, 2 if sw % 3 == 0 else 0, 5)
        assert sa.swap_tries == (2 if sw % 3 == 0 else 0) == sb.swap_tries
    assert np.array_equal(a.r, b.r) and np.array_equal(a.next, b.next)


def test_move_weights_keep_the_sampled_distribution():
    """Passes at weights 1/2 (staging) and 1/3 (COM) still sample the
    primitive-action distribution: <E> and <x^2> against the closed form."""
    N, M, beta = 4, 16, 2.0
    s = harmonic_sys(N, M, beta, seed=31)
    E, x2 = [], []
    for sw in range(40000):
        s.sweep_every(0, sw, 78, 4, 1.0, (2, 3, 1))
        if sw >= 600:
            E.append(s.energy()[0])
            x2.append(np.mean(s.r ** 2))
    for vals, exact in ((E, AN.energy_boltzmann_primitive(N, M, beta)), (x2, AN.x2_primitive(M, beta))):
        m, e = blockerr(vals)
        assert abs(m - exact) < 3.5 * e, (m, e, exact)


# ---------------------------------------------------------------- swaps with a pair action
def _apply_swap(r, nxt, n1, n2, j0, m, rp1, rp2):
    """The commit of an accepted swap: the new beads along the new cycles, the
    tails after the window exchanged, next[] of n1 and n2 exchanged."""
    M = r.shape[2]
    q1, q2 = int(nxt[n1]), int(nxt[n2])
    for nw, qn, rp in ((n1, q2, rp1), (n2, q1, rp2)):
        for i in range(1, m):
            u = j0 + i
            r[nw if u < M else qn, :, u % M] = rp[i]
    if j0 + m < M:
        r[[n1, n2], :, j0 + m:] = r[[n2, n1], :, j0 + m:]
    nxt[n1], nxt[n2] = q2, q1


def _swap_case(rng, N, M, L, nxt):
    r = rng.uniform(-L, L, size=(N, 2, M))
    s = O.System(r, 0.5, L, 1.0, O.make_pot(O.POT_NONE), O.make_pair(O.PAIR_GAUSSIAN, 1.0, 0.5), 1.0, True)
    s.next = nxt
    return s, r


def _swap_paths(rng, r, nxt, n1, n2, j0, m, L):
    """New paths of a swap try: the start beads of n1, n2 at j0, the ends of n2,
    n1 at j0 + m along the cycles, interior beads anywhere in the box."""
    M = r.shape[2]
    jm = j0 + m
    end = lambda n: r[n if jm < M else nxt[n], :, jm % M]
    rp1 = np.vstack([r[n1, :, j0], rng.uniform(-L, L, size=(m - 1, 2)), end(n2)])
    rp2 = np.vstack([r[n2, :, j0], rng.uniform(-L, L, size=(m - 1, 2)), end(n1)])
    return rp1, rp2


def test_swap_pair_sums_equal_the_pair_action_change():
    """The swap's pair log ratio with each pair once (pair_in_accept = 1) is
    the change of the total pair action (o_total_lnU, along the cycles) from
    before the swap to after its commit: windows inside the ring and across its
    end, particles on cycles of several lengths, nbins = 4."""
    rng = np.random.default_rng(71)
    N, M, L = 7, 12, 2.0
    for trial in range(24):
        nxt = np.array([1, 2, 0, 3, 5, 4, 6]) if trial % 2 else np.arange(N)
        s, r = _swap_case(rng, N, M, L, nxt)
        n1, n2 = rng.choice(N, size=2, replace=False)
        j0, m = int(rng.integers(0, M)), int(rng.integers(2, M))
        rp1, rp2 = _swap_paths(rng, r, nxt, n1, n2, j0, m, L)
        lw = s.swap_pair_lw(int(n1), int(n2), j0, m, rp1, rp2, 0)
        u0 = s.total_lnU()
        r2, nx2 = r.copy(), nxt.copy()
        _apply_swap(r2, nx2, n1, n2, j0, m, rp1, rp2)
        s2, _ = _swap_case(rng, N, M, L, nx2)
        s2.r = r2
        assert abs(lw - (s2.total_lnU() - u0)) < 1e-11 * max(1.0, abs(u0)), trial


def test_swap_pair_sums_as_the_reference_writes_them():
    """Two particles (pair_in_accept = 0, ReshapeSwapLinear's sums): the old
    sums list the pair once from each particle, the new sums exclude both, so
    the log ratio is minus twice the old pair sums of the window's links, each
    the reference's interaction_action! of the bead's holder."""
    rng = np.random.default_rng(72)
    N, M, L = 2, 10, 1.0
    nonzero = 0
    for trial in range(12):
        nxt = np.array([1, 0]) if trial % 2 else np.arange(N)
        s, r = _swap_case(rng, N, M, L, nxt)
        j0, m = int(rng.integers(0, M)), int(rng.integers(2, M))
        rp1, rp2 = _swap_paths(rng, r, nxt, 0, 1, j0, m, L)
        lw = s.swap_pair_lw(0, 1, j0, m, rp1, rp2, 1)
        old = sum(s.interaction_old(0 if j0 + t < M else int(nxt[0]), (j0 + t) % M) for t in range(m))
        assert abs(lw + 2.0 * old) < 1e-12 * max(1.0, abs(old)), trial
        nonzero += old != 0.0
    assert nonzero >= 6


def _table_swap_sys(seed, hs_a, hs_ctr=10000, N=12, M=12, L=2.0):
    rng = np.random.default_rng(seed)
    tau = 2.0 / M
    r1 = float(math.ceil(math.sqrt(2.0) * L))
    tab = O.prop_int_terms(48, 1e-20, r1, 4.0, tau)
    pair = O.make_pair(O.PAIR_TABLE, table=tab, r0=1e-20, r1=r1, mu=0.4)
    r = rng.uniform(-L, L, size=(N, 2, 1)).repeat(M, axis=2) + rng.normal(0, 0.1, size=(N, 2, M))
    r = (r + L) % (2 * L) - L
    s = O.System(r, 0.5, L, 2.0, O.make_pot(O.POT_HARMONIC, 0.3), pair, 1.0, True, hs_a=hs_a, hs_ctr=hs_ctr)
    s._keep = tab
    return s


def test_hard_sphere_swaps_reduce_to_plain_swaps_and_reject():
    """Swap tries with the tabulated propagator: a hard-sphere radius no bead
    comes near gives the sweeps without hard spheres (paths to rounding, the
    same counters and permutation); a radius that every redraw violates with
    one draw allowed rejects every swap try while the plain sweeps accept some."""
    a, b = _table_swap_sys(3, 0.0), _table_swap_sys(3, 1e-9)
    for sw in range(20):
        sa = a.sweep_perm(0, sw, 5, 4, 0.5, 6, 5)
        sb = b.sweep_perm(0, sw, 5, 4, 0.5, 6, 5)
        assert (sa.levy_acc, sa.com_acc, sa.swap_acc, sa.swap_tries) == (sb.levy_acc, sb.com_acc, sb.swap_acc,
                                                                         sb.swap_tries)
    assert np.allclose(a.r, b.r, atol=1e-12) and np.array_equal(a.next, b.next)
    plain, blocked = _table_swap_sys(4, 0.0), _table_swap_sys(4, 1.0, hs_ctr=1)
    acc_p = acc_b = tries_b = 0
    for sw in range(20):
        acc_p += plain.sweep_perm(0, sw, 6, 4, 0.5, 6, 5).swap_acc
        st = blocked.sweep_perm(0, sw, 6, 4, 0.5, 6, 5)
        acc_b += st.swap_acc
        tries_b += st.swap_tries
    assert acc_p > 0 and acc_b == 0 and tries_b == 20 * 6
    assert np.array_equal(blocked.next, np.arange(12))
