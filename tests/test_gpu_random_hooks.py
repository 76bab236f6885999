"""Seeded random chains through the single-move hooks (the reference's
ReshapeLinear and SingleCenterOfMass called one move at a time, with the
draws injected): 1D or 2D; harmonic trap, square or optical lattice, or no
potential; no pair action, a Gaussian pair or the tabulated propagator, in
either pair mode, with cut-offs from one cell a side to six and N up to 129
(cell lists); 150 moves each.  Every accept decision equals the oracle's and
the paths agree to 1e-10 in FP64."""
import math

import numpy as np
import pytest

import oracle as O
from pimc import _capi as A
from pimc.potentials import GaussianPair, HarmonicTrap, NoPotential, OpticalLattice, SquareLattice
from pimc.propagator import PropInt
from pimc.system import chain_desc
from test_gpu_parity import oracle_sys

pytestmark = pytest.mark.gpu


def random_chain(rng):
    dim = int(rng.integers(1, 3))
    M = int(rng.choice([6, 8, 16, 24, 40]))
    N = int(rng.choice([1, 2, 5, 20, 60, 110, 129]))
    L = float(rng.uniform(2.0, 6.0))
    beta = float(rng.uniform(0.5, 3.0))
    kind = str(rng.choice(["none", "gauss", "table"]))
    u = rng.random()
    if u < 0.4:
        pot = HarmonicTrap(float(rng.uniform(0.3, 1.5)))
    elif u < 0.6:
        pot = SquareLattice(2.0, float(rng.uniform(0.5, 1.5)))
    elif u < 0.8 or dim == 1:
        pot = NoPotential()
    else:
        pot = OpticalLattice([0.0, math.pi / 2], 0.5, 2.0)
    if kind == "gauss":
        pair = GaussianPair(float(rng.uniform(0.5, 2.0)), float(rng.uniform(0.2, 0.5)))
    elif kind == "table":
        pair = PropInt(np.zeros((2, 2)), 0.0, 1.0, 4.0, beta / M)  # kind marker; the table is set separately
    else:
        pair = None
    ra = 2 * L / int(rng.integers(1, 7)) if pair else 0.0
    d = chain_desc(pot, dim=dim, M=M, N=N, L=L, beta=beta, lam=0.5, pair=pair, ell=min(8, M - 1), r_a=ra,
                   com_step=0.5, pair_in_accept=bool(rng.integers(0, 2)))
    r = rng.uniform(-L, L, size=(N, dim, 1)).repeat(M, axis=2) + rng.normal(0, 0.1, size=(N, dim, M))
    return d, (r + L) % (2 * L) - L, kind


@pytest.mark.parametrize("trial", range(16))
def test_random_chain_single_moves_match_oracle_fp64(trial):
    rng = np.random.default_rng(40000 + trial)
    d, r, kind = random_chain(rng)
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed = A.FP64, 4, 7
    ctx = A.Context([d], opt)
    ctx.set_paths(0, r)
    rr = ctx.get_paths(0)
    if kind == "table":
        r1 = float(math.ceil(math.sqrt(2.0) * d.L))
        tab = O.prop_int_terms(48, 1e-20, r1, 4.0, d.beta / d.M)
        ctx.set_pair_table(0, tab, 1e-20, r1, 0.4)
        s = O.System(rr, d.lam, d.L, d.beta, O.pot_from_desc(d),
                     O.make_pair(O.PAIR_TABLE, table=tab, r0=1e-20, r1=r1, mu=0.4), d.r_a, bool(d.pair_in_accept))
    else:
        s = oracle_sys(d, rr)
    for mv in range(150):
        u = float(rng.random())
        if mv % 3 == 0:
            n, disp = int(rng.integers(d.N)), rng.uniform(-0.5, 0.5, size=d.dim)
            got, (want, _) = ctx.center_of_mass(0, n, disp, u), s.com_apply(n, disp, u)
        else:
            n, j0, m = int(rng.integers(d.N)), int(rng.integers(d.M)), int(rng.integers(2, min(d.M - 1, 16) + 1))
            nrm = rng.standard_normal((m - 1, d.dim))
            got, (want, _) = ctx.reshape_linear(0, n, j0, m, nrm, u), s.reshape_apply(n, j0, m, nrm, u)
        assert got == want, mv
    assert np.max(np.abs(ctx.get_paths(0) - s.r)) < 1e-10
