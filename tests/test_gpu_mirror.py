"""The reference's own tests (test/testsystem.jl, test/testmeasurements.jl),
restated on the Python mirror of its API over the C ABI, plus the swap moves
of the batched sweep seen through the same API."""
import math

import numpy as np
import pytest

from pimc import (Density, Energy, HIPSweep, HarmonicTrap, NoPotential, OpticalLattice, ReshapeLinear,
                  SingleCenterOfMass, SquareLattice, System, run, sweep_weights)

pytestmark = pytest.mark.gpu


def test_system_initialisation():
    """testsystem.jl: System(V) has two particles; System(V; dim=2, M=100,
    N=5, L=4.0, T=1.0) builds five worldlines of 100 slices in the box."""
    assert System(SquareLattice(8.0, 2.0), seed=1).N == 2
    s = System(SquareLattice(8.0, 2.0), dim=2, M=100, N=5, L=4.0, T=1.0, seed=2)
    w = s.world
    assert len(w) == 5
    for n, p in enumerate(w):
        assert p.r.shape == (100, 2) and p.next == n
        assert np.all(p.r >= -4.0) and np.all(p.r < 4.0)
        assert np.array_equal(p.r[0], p.r[-1])  # init_world closes the path on its first bead


def test_system_with_the_references_optical_potential():
    """testsystem.jl Initialization2d with its own potential: generate_V(cyclic_4,
    scale 0.5, depth 8) of normalized_intensity, System(V; dim=2, M=100, N=5,
    L=4.0, T=1.0); testpotential.jl's intensity 1 at the lattice maxima; and the
    periodic-bounds check of testsystem.jl after run! (no bead outside [-L, L))."""
    from pimc import OpticalLattice
    cyclic_4 = [2 * math.pi * k / 4 for k in range(4)]
    V = OpticalLattice(cyclic_4, 0.5, 8.0)
    assert V([0.0, 0.0]) == pytest.approx(8.0)
    cyclic_3 = [2 * math.pi * k / 3 for k in range(3)]
    I3 = OpticalLattice(cyclic_3, 1.0, 1.0)
    for r in ([0.0, 0.0], [2 / math.sqrt(3), 0.0], [0.0, 2 / 3]):
        assert I3(r) == pytest.approx(1.0)
    s = System(V, dim=2, M=100, N=5, L=4.0, T=1.0, seed=7)
    assert s.N == 5
    updates = [(2, SingleCenterOfMass(s, 3.0)), (1, ReshapeLinear(s, 20))]
    run(s, 2000, updates)
    r = s.paths
    assert np.all(r >= -4.0) and np.all(r < 4.0)


def test_density_integrates_to_particle_number():
    """testmeasurements.jl: after run! with SingleCenterOfMass and
    ReshapeLinear, the Density histogram integrates to N (up to the beads in
    the dropped first bin)."""
    s = System(SquareLattice(8.0, 2.0), dim=2, N=2, M=100, L=4.0, T=1.0, seed=3)
    updates = [(2, SingleCenterOfMass(s, 3.0)), (1, ReshapeLinear(s, 20))]
    d = Density(s)
    run(s, 3000, updates, Zmeasurements=[d])
    numb = d.dens.sum() / d.ndata
    assert d.ndata > 0
    assert abs(numb - s.N) < 1e-2, numb


def test_swaps_through_the_sweep_functor():
    """Swap moves (ReshapeSwapLinear) run inside HIPSweep: the permutation
    changes, the Energy estimator follows the cycles, the density still
    integrates to N, and the single-move functors refuse the chain."""
    s = System(HarmonicTrap(1.0), dim=2, N=4, M=32, L=20.0, T=0.5, lam=0.5, seed=4, ell=8, swaps=4, swap_m=8)
    sweep = HIPSweep(s)
    e, d = Energy(s), Density(s, nbins=200)
    seen = set()
    for _ in range(200):
        sweep(s)
        seen.add(tuple(s.next))
    run(s, 300, [(1, sweep)], Zmeasurements=[e, d])
    assert len(seen) > 1  # permutations visited
    assert sorted(s.next) == [0, 1, 2, 3]
    assert np.isfinite(e.energy[4]).all() and len(e.energy[4]) > 0
    assert abs(d.dens.sum() / d.ndata - 4) < 1e-2
    with pytest.raises(Exception):
        ReshapeLinear(s, 8)(s)


# the update mix of the reference's own tests (test/testsystem.jl:39-43, test/testmeasurements.jl:13-17)
REF_MIX = [(2, "SingleCenterOfMass"), (1, "ReshapeLinear"), (1, "ReshapeSwapLinear")]


def v1d_2d(depth=8.0, scale=0.5):
    """The reference tests' v1d(x) = depth sin^2(2 pi x[1] scale) in the 2D
    System, a potential of x alone: two counter-propagating beams (angles
    +-pi/2) give normalized_intensity = cos^2(2 pi scale x), so -depth times it
    is v1d - depth (a constant shift: same paths, same density)."""
    return OpticalLattice([math.pi / 2, -math.pi / 2], scale, -depth)


def mix_system(potential, seed):
    """System(potential) with the reference's defaults (dim 2, M 100, N 2, L 4,
    T 1, lambda 1) sampled by the batched sweep with the reference mix:
    ReshapeLinear(s, 20) as ell = 20, ReshapeSwapLinear(s, 20) as swap tries
    of 20 links, as many per sweep as staging tries (equal weights), COM
    tries every second sweep (weight 1/2)."""
    w = sweep_weights(REF_MIX)
    assert w == {"com_every": 2, "staging_every": 1, "swap_every": 1}
    return System(potential, seed=seed, ell=20, swaps=2 * 5, swap_m=20, **w)


def test_reference_v1d_is_an_optical_lattice_of_two_beams():
    V = v1d_2d()
    rng = np.random.default_rng(3)
    for r in rng.uniform(-4, 4, size=(50, 2)):
        assert V(r) + 8.0 == pytest.approx(8.0 * math.sin(2 * math.pi * r[0] * 0.5) ** 2, abs=1e-12)


def test_reference_periodic_bounds_as_written():
    """testsystem.jl "Periodic Bounds": System(x -> 0), the reference mix
    with swap tries, run! 10 000 times; no bead lies outside the box (the
    reference counts beads with every coordinate outside; here no coordinate
    is)."""
    s = mix_system(NoPotential(), seed=11)
    sweep = HIPSweep(s)
    seen = set()
    for _ in range(100):
        sweep(s)
        seen.add(tuple(s.next))
    run(s, 10_000, [(1, sweep)])
    r = s.paths
    assert np.all(r >= -s.L) and np.all(r < s.L)
    assert len(seen) > 1  # the swap tries exchange the two particles
    c = s.ctx.counters(0)
    assert c[3] > 0 and c[1] > 0


def test_reference_density_with_v1d_as_written():
    """testmeasurements.jl "Number of particles with potential": System(v1d),
    the reference mix with swap tries, Density measured while running; the
    density integrates to N within 1e-2."""
    s = mix_system(v1d_2d(), seed=12)
    sweep = HIPSweep(s)
    run(s, 2000, [(1, sweep)])
    d = Density(s)
    run(s, 10_000, [(1, sweep)], Zmeasurements=[d])
    numb = d.dens.sum() / d.ndata
    assert d.ndata > 0
    assert abs(numb - s.N) < 1e-2, numb
    assert np.all(s.paths >= -s.L) and np.all(s.paths < s.L)


def test_reference_srl_example_with_swaps_small():
    """examples/density_SRL_lattice.jl's system at a small size: the l25
    superlattice (generate_V(1, 4, :l25; attractive)), interactions=true with
    build_prop_int(ceil(sqrt(2) L), g, 1/(T M)) and g = 2 (hard spheres of
    radius exp(-pi)), lambda = 1/pi^2, the update mix (1, SingleCenterOfMass),
    (1, ReshapeLinear(s, 5)), (1, ReshapeSwapLinear(s, 20)) as move weights and
    swap tries; the permutation changes and the Density integrates to N up to
    the beads in its dropped first bins (a strip of 2/500 of the box here,
    where this lattice leaves beads at the edges)."""
    from pimc import build_prop_int, generate_V
    L, M, N, T, g = 4.0, 40, 6, 0.2, 2.0
    propint = build_prop_int(float(math.ceil(math.sqrt(2 * L * L))), g, 1.0 / (T * M), n=200)
    w = sweep_weights([(1, "SingleCenterOfMass"), (1, "ReshapeLinear"), (1, "ReshapeSwapLinear")])
    s = System(generate_V(1.0, 4.0, "l25", attractive=True), lam=1 / math.pi ** 2, M=M, N=N, L=L, T=T,
               propint=propint, interactions=True, g=g, seed=21, ell=5, swaps=N * (M // 5), swap_m=20, **w)
    assert s.a == pytest.approx(math.exp(-math.pi))
    sweep = HIPSweep(s)
    seen = set()
    for _ in range(200):
        sweep(s)
        seen.add(tuple(s.next))
    d = Density(s)
    run(s, 2000, [(1, sweep)], Zmeasurements=[d])
    assert len(seen) > 1 and sorted(s.next) == list(range(N))
    assert d.ndata > 0 and abs(d.dens.sum() / d.ndata - N) < 0.01 * N
    r = s.paths
    assert np.all(r >= -L) and np.all(r < L)
