"""The reference's own tests (test/testsystem.jl, test/testmeasurements.jl),
restated on the Python mirror of its API over the C ABI, plus the swap moves
of the batched sweep seen through the same API."""
import numpy as np
import pytest

from pimc import (Density, Energy, HIPSweep, HarmonicTrap, ReshapeLinear, SingleCenterOfMass, SquareLattice, System,
                  run)

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
    import math
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
