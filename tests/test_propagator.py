"""Pins of the oracle's pair-product propagator (CPU).

prop_rel_interpolate_terms (src/propagator.jl:36-70) integrates three Bessel
products over k with quadgk; the oracle uses a fixed Gauss-Legendre rule.  The
rule is pinned against an independent adaptive quadrature (scipy's QUADPACK).
The integrand is pinned by what the physics fixes: the s-wave relative
propagator free + terms is a heat kernel, so it composes, rho_2tau = rho_tau *
rho_tau over the plane, which fails for a wrong sign or weight of any term.
The table interpolation and lnU are pinned by closed forms of their
definitions.
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

import oracle as O


def terms_scipy(r1, r2, g0, tau, cross_sign=1.0):
    """The integrand of src/propagator.jl:43-63 by adaptive quadrature (cross_sign
    = -1 plants a wrong sign on term2, to show the composition test catches it)."""
    def f(k):
        d = (2 / np.pi) * (np.euler_gamma + np.log(k / 2)) - 4 / g0
        a, b = 1 / (1 + d * d), d / (1 + d * d)
        j1, j2, y1, y2 = special.j0(k * r1), special.j0(k * r2), special.y0(k * r1), special.y0(k * r2)
        return k * np.exp(-tau * k * k) / (2 * np.pi) * (-a * j1 * j2 - cross_sign * b * (j1 * y2 + j2 * y1) + a * y1 * y2)
    K = math.sqrt(40 / tau)
    return (integrate.quad(f, 0, 1, epsabs=1e-14, epsrel=1e-13, limit=2000)[0] +
            integrate.quad(f, 1, K, epsabs=1e-14, epsrel=1e-13, limit=5000)[0])


def free_s(a, b, tau):
    """s-wave part of exp(-|r - r'|^2 / (4 tau)) / (4 pi tau) (angle average)."""
    return special.ive(0, a * b / (2 * tau)) * np.exp(-(a - b) ** 2 / (4 * tau)) / (4 * np.pi * tau)


@pytest.mark.parametrize("g0,tau", [(4.0, 0.05), (20.0, 0.01), (1.0, 0.1)])
def test_quadrature_rule_matches_adaptive_quadrature(g0, tau):
    rmax = 4.0
    for r1, r2 in [(0.1, 0.1), (0.3, 0.7), (1e-20, 0.5), (1.5, 2.0), (0.05, 1.0)]:
        a = O.prop_int_term(r1, r2, g0, tau, rmax)
        b = terms_scipy(r1, r2, g0, tau)
        assert abs(a - b) <= 1e-11 * max(1.0, abs(b)), (r1, r2, a, b)


def test_s_wave_propagator_composes():
    """(free + terms)_2tau = (free + terms)_tau * (free + terms)_tau, the s-wave
    kernel composed over the plane (2 pi r3 dr3); the free part composes by
    itself, which leaves terms_2tau = F*T + T*F + T*T."""
    g0, tau, rmax = 4.0, 0.05, 4.0
    T = lambda a, b, t=tau: O.prop_int_term(a, b, g0, t, rmax)
    for r1, r2 in [(0.2, 0.3), (0.5, 0.5), (0.1, 0.8)]:
        f = lambda r3: 2 * np.pi * r3 * (free_s(r1, r3, tau) * T(r3, r2) + T(r1, r3) * free_s(r3, r2, tau)
                                         + T(r1, r3) * T(r3, r2))
        rhs = integrate.quad(f, 0, rmax, epsabs=1e-12, epsrel=1e-10, limit=500, points=[r1, r2])[0]
        lhs = T(r1, r2, 2 * tau)
        assert abs(lhs - rhs) < 1e-9, (r1, r2, lhs, rhs)
    # the same identity with term2's sign planted wrong fails by far
    r1, r2 = 0.2, 0.3
    Tw = lambda a, b, t=tau: terms_scipy(a, b, g0, t, cross_sign=-1.0)
    grid = np.linspace(1e-6, rmax, 400)
    vals = [2 * np.pi * x * (free_s(r1, x, tau) * Tw(x, r2) + Tw(r1, x) * free_s(x, r2, tau) + Tw(r1, x) * Tw(x, r2))
            for x in grid]
    assert abs(Tw(r1, r2, 2 * tau) - np.trapezoid(vals, grid)) > 1e-2


def test_weak_coupling_limit():
    """g0 -> 0+: tan(delta) -> 0, so the terms vanish and prop_int -> 1."""
    for g0 in (1e-2, 1e-3):
        t = O.prop_int_term(0.3, 0.4, g0, 0.05, 4.0)
        assert abs(t) < 0.05 * g0 * 4.0


def test_table_builder_matches_point_values():
    n, r0, r1, g0, tau = 7, 1e-20, 3.0, 4.0, 0.05
    tab = O.prop_int_terms(n, r0, r1, g0, tau)
    grid = r0 + np.arange(n) * ((r1 - r0) / (n - 1))
    for i in range(n):
        for k in range(n):
            assert tab[i, k] == pytest.approx(O.prop_int_term(grid[i], grid[k], g0, tau, r1), rel=1e-13, abs=1e-15)
    assert np.allclose(tab, tab.T, rtol=1e-13, atol=1e-15)


def test_table_eval_is_bilinear_and_clamps():
    n, r0, r1 = 9, 0.5, 2.5
    grid = r0 + np.arange(n) * ((r1 - r0) / (n - 1))
    f = lambda x, y: 2.0 + 3.0 * x - y + 0.5 * x * y  # bilinear: reproduced exactly
    tab = f(grid[:, None], grid[None, :])
    pp = O.make_pair(O.PAIR_TABLE, table=tab, r0=r0, r1=r1)
    rng = np.random.default_rng(3)
    for x, y in rng.uniform(r0, r1, size=(50, 2)):
        assert O.prop_table_eval(pp, x, y) == pytest.approx(f(x, y), rel=1e-13)
    assert O.prop_table_eval(pp, 0.0, 1.0) == pytest.approx(f(r0, 1.0), rel=1e-13)
    assert O.prop_table_eval(pp, 9.0, 9.0) == pytest.approx(f(r1, r1), rel=1e-13)
    assert O.prop_table_eval(pp, r1, r0) == pytest.approx(f(r1, r0), rel=1e-13)


def test_prop_int_and_lnU_definitions():
    """prop_int = 1 + c / prop_rel0 for a constant table c, prop_rel0 taken between
    the per-component absolute offsets; lnU = log(prop_int), or -mu below zero."""
    tau, mu = 0.1, 0.7
    for c in (0.05, -0.02, -5.0):
        pp = O.make_pair(O.PAIR_TABLE, table=np.full((4, 4), c), r0=0.0, r1=3.0, mu=mu)
        r1, r2 = np.array([0.3, 0.4]), np.array([0.1, 0.9])
        p0 = math.exp(-((0.3 - 0.1) ** 2 + (0.4 - 0.9) ** 2) / (4 * tau)) / (4 * math.pi * tau)
        p = 1 + c / p0
        assert O.prop_int(pp, r1, r2, tau) == pytest.approx(p, rel=1e-14)
        assert O.lnU(pp, r1, r2, tau) == pytest.approx(math.log(p) if p >= 0 else -mu, rel=1e-14)


def table_pair_sys(N, L, M, ra, seed, g0=4.0, mu=0.3):
    r = np.random.default_rng(seed).uniform(-L, L, size=(N, 2, M))
    tau = 2.0 / M
    r1 = math.ceil(math.sqrt(2) * L)
    tab = O.prop_int_terms(24, 1e-20, r1, g0, tau)
    pair = O.make_pair(O.PAIR_TABLE, table=tab, r0=1e-20, r1=r1, mu=mu)
    return O.System(r, 0.5, L, 2.0, O.make_pot(), pair, ra, True)


def test_table_pair_sums_against_brute_force_total():
    """The cell-list link sums of the tabulated lnU (interaction_action!) sum to
    twice the all-pairs total, and the Delta of a move equals the change of the
    total."""
    s = table_pair_sys(40, 4.0, 6, 2.0, seed=9)
    tot = sum(s.interaction_old(n, j) for n in range(40) for j in range(6))
    assert tot == pytest.approx(2 * s.total_lnU(), rel=1e-12)
    assert s.total_lnU() != 0.0
    rng = np.random.default_rng(9)
    for _ in range(10):
        n, j0, m = int(rng.integers(40)), int(rng.integers(6)), int(rng.integers(2, 5))
        before = s.total_lnU()
        acc, w = s.reshape_apply(n, j0, m, rng.standard_normal((m - 1, 2)), 0.0)
        assert acc
        assert (w.wU_updated - w.wU_initial) == pytest.approx(s.total_lnU() - before, abs=1e-11)
    Ep, _ = s.pair_energy()
    assert Ep == 0.0  # no pair energy for the tabulated propagator
