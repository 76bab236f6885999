"""Pins of the oracle's hard-sphere rejection (CPU): hardspherelevy! and
move_polymer! (src/updates/helper.jl:141-181, 368-395) with a > 0.

find_nn against an independent nearest-bead search over the stencil cells;
a = 0 leaves the sweep bit-identical to the plain one; with a > 0 no sweep
ever places two beads of one slice closer than a when the start has none
(the property the redraws exist for), the redraw limit turns crowded tries
into rejections, and the accept counts move with the limit as they must.
"""
import math

import numpy as np
import pytest

import oracle as O


def grid_paths(N, M, L, spacing, jitter, seed):
    rng = np.random.default_rng(seed)
    side = int(math.ceil(math.sqrt(N)))
    pts = np.array([[((i % side) + 0.5) * spacing - L, ((i // side) + 0.5) * spacing - L] for i in range(N)])
    r = np.repeat(pts[:, :, None], M, axis=2) + rng.normal(0, jitter, size=(N, 2, M))
    return np.clip(r, -L, L - 1e-9)


def min_same_slice_distance(s):
    r, L = s.r, s._s.L
    best = np.inf
    for j in range(s.M):
        x = r[:, :, j]
        d = np.abs(x[:, None, :] - x[None, :, :])
        d = np.minimum(d, 2 * L - d)
        dd = np.sqrt((d ** 2).sum(-1)) + np.eye(len(x)) * 1e9
        best = min(best, dd.min())
    return best


def test_find_nn_against_stencil_search():
    rng = np.random.default_rng(4)
    for nbins_ra, L in [(1.5, 3.0), (3.0, 3.0), (1.0, 4.0)]:
        r = rng.uniform(-L, L, size=(30, 2, 3))
        s = O.System(r, 0.5, L, 1.0, O.make_pot(), None, nbins_ra)
        nb, w = s.nbins, 2 * L / s.nbins
        cell = lambda p: tuple(int(math.floor((p[d] + L) / w)) % nb for d in range(2))
        for _ in range(40):
            j, n = int(rng.integers(3)), int(rng.integers(30))
            t = rng.uniform(-L, L, size=2)
            ct = cell(t)
            best, bd = -1, np.inf
            for k in range(30):
                if k == n:
                    continue
                ck = cell(r[k, :, j])
                if all(((ck[d] - ct[d]) % nb) in (0, 1, nb - 1) for d in range(2)):
                    dv = np.abs(r[k, :, j] - t)
                    dv = np.minimum(dv, 2 * L - dv)
                    dd = math.sqrt((dv ** 2).sum())
                    if dd < bd:
                        best, bd = k, dd
            assert s.find_nn(t, j, n) == best


def test_zero_radius_is_the_plain_sweep():
    r = grid_paths(12, 8, 3.0, 1.5, 0.2, seed=1)
    a = O.System(r, 0.5, 3.0, 1.0, O.make_pot(O.POT_HARMONIC, 0.5), None, 1.5)
    b = O.System(r, 0.5, 3.0, 1.0, O.make_pot(O.POT_HARMONIC, 0.5), None, 1.5, hs_a=0.0, hs_ctr=1)
    for sw in range(5):
        sa, sb = a.sweep(0, sw, 9, 4, 0.5), b.sweep(0, sw, 9, 4, 0.5)
        assert (sa.levy_acc, sa.com_acc) == (sb.levy_acc, sb.com_acc)
    assert np.array_equal(a.r, b.r)


@pytest.mark.parametrize("hs_a", [0.3, 0.6])
def test_hard_spheres_stay_apart(hs_a):
    r = grid_paths(12, 8, 3.0, 1.5, 0.1, seed=2)
    s = O.System(r, 0.5, 3.0, 1.0, O.make_pot(O.POT_HARMONIC, 0.5), None, 1.5, hs_a=hs_a)
    assert min_same_slice_distance(s) >= hs_a
    plain = O.System(r, 0.5, 3.0, 1.0, O.make_pot(O.POT_HARMONIC, 0.5), None, 1.5)
    closest_plain = np.inf
    for sw in range(60):
        s.sweep(0, sw, 13, 4, 0.8)
        plain.sweep(0, sw, 13, 4, 0.8)
        assert min_same_slice_distance(s) >= hs_a, sw
        closest_plain = min(closest_plain, min_same_slice_distance(plain))
    assert closest_plain < hs_a  # without the rejection, beads do come closer


def test_redraw_limit_rejects_crowded_tries():
    r = grid_paths(16, 8, 2.0, 1.0, 0.05, seed=3)
    acc = {}
    for ctr in (1, 10000):
        s = O.System(r, 0.5, 2.0, 2.0, O.make_pot(O.POT_HARMONIC, 0.5), None, 1.0, hs_a=0.7, hs_ctr=ctr)
        la = lt = ca = 0
        for sw in range(20):
            st = s.sweep(0, sw, 17, 4, 0.4)
            la += st.levy_acc
            lt += st.levy_tries
            ca += st.com_acc
        assert min_same_slice_distance(s) >= 0.7
        acc[ctr] = (la, ca, lt)
    assert acc[1][2] == acc[10000][2]
    assert acc[1][0] < acc[10000][0]
