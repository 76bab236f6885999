"""System and Particle, mirroring src/system.jl over one GPU chain."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ._capi import FP32, FP64, PAIR_NONE, PAIR_TABLE, POT_OPTICAL, ChainDesc, Context, Options
from .potentials import GaussianPair, HarmonicTrap, NoPotential, SquareLattice


@dataclass
class Particle:
    """src/system.jl:1-6: r is M x dim; V holds the link actions lnV(r_j, r_j+1)."""
    r: np.ndarray
    V: np.ndarray
    next: int


def _teleport(x, L):
    # src/propagator.jl:30-32
    return (x + L) - np.floor(x / (2 * L) + 0.5) * (2 * L) - L


def _levy_host(r: np.ndarray, tau: float, L: float, lam: float, rng: np.random.Generator) -> np.ndarray:
    """levy! (src/updates/helper.jl:118-139) on a rows x dim array, host side,
    used only to build initial paths the way init_world does."""
    r = r.copy()
    for d in range(r.shape[1]):
        if abs(r[0, d] - r[-1, d]) > L:
            r[-1, d] += np.sign(r[0, d]) * 2 * L
    m = r.shape[0] - 2
    for j in range(1, m + 1):
        alpha = (m + 1 - j) / (m + 2 - j)
        r[j] = alpha * r[j - 1] + (1 - alpha) * r[-1] + rng.standard_normal(r.shape[1]) * math.sqrt(2 * lam * alpha * tau)
    return _teleport(r, L)


def init_world(N: int, M: int, dim: int, L: float, tau: float, lam: float, rng: np.random.Generator) -> np.ndarray:
    """init_world (src/system.jl:36-78) without hard spheres (a = 0): each
    particle starts uniformly in the box, r[end] = r[begin], and levy! closes
    the path.  Returns N x dim x M (the C ABI layout)."""
    out = np.zeros((N, dim, M))
    for n in range(N):
        r = np.zeros((M, dim))
        r[0] = 2 * L * (rng.random(dim) - 0.5)
        r[-1] = r[0]
        out[n] = _levy_host(r, tau, L, lam, rng).T
    return out


def chain_desc(potential, *, dim, M, N, L, beta, lam, pair=None, r_a=0.0, pair_in_accept=True, ell=0,
               com_step=1.0, swaps=0, swap_m=0, swap_rule=0, hs_a=0.0, staging_every=1, com_every=1,
               swap_every=1) -> ChainDesc:
    """pimc_chain_desc of one chain.  swaps > 0 adds that many swap tries per
    sweep (ReshapeSwapLinear) of swap_m links (0: ell); swap_rule 1 keeps the
    reference's acceptance exp(w_updated - w_initial).  The staging pass, the
    centre-of-mass pass and the swap tries run on the sweeps s with
    s % every == 0 (the move weights 1/every of run!, see sweep_weights)."""
    d = ChainDesc()
    d.N, d.M, d.dim = int(N), int(M), int(dim)
    d.potential = getattr(potential, "kind", 0)
    d.beta, d.lam, d.L = float(beta), float(lam), float(L)
    d.omega = getattr(potential, "omega", 1.0)
    d.V0 = getattr(potential, "V0", 0.0)
    d.a = getattr(potential, "a", 1.0)
    if getattr(potential, "kind", 0) == POT_OPTICAL:
        d.beam_scale, d.n_beams, d.helical = potential.scale, len(potential.angles), int(potential.helical)
        for i, a in enumerate(potential.angles):
            d.beam_angle[i] = a
    if pair is not None and pair.kind == PAIR_TABLE:
        d.pair = PAIR_TABLE
        d.g, d.sigma = pair.g0, 1.0
        d.r_a = r_a if r_a > 0 else pair.nn_range()  # determine_nnrange (src/system.jl:29-31)
    elif pair is not None:
        d.pair = pair.kind
        d.g, d.sigma = pair.g, pair.sigma
        d.r_a = r_a if r_a > 0 else pair.cutoff
    else:
        d.pair = PAIR_NONE
        d.sigma = 1.0
        d.r_a = r_a
    d.pair_in_accept = 1 if pair_in_accept else 0
    d.ell = int(ell)
    d.com_step = float(com_step)
    d.swaps, d.swap_m, d.swap_rule = int(swaps), int(swap_m), int(swap_rule)
    d.hs_a = float(hs_a)
    d.staging_every, d.com_every, d.swap_every = int(staging_every), int(com_every), int(swap_every)
    return d


class System:
    """System(potential; kwargs...) (src/system.jl:129-167) on one GPU chain.

    swaps > 0 adds that many swap tries of swap_m links to every HIPSweep
    sweep (ReshapeSwapLinear, src/updates/reshape.jl:99-283; the chain then
    carries permutation cycles, see `next`); swap_rule = 1 keeps the
    reference's acceptance.

    Keyword names follow the reference (lam for lambda, mu for mu, r_a for
    r_a).  `pair` is a device-side pair action: GaussianPair, or the
    tabulated pair-product propagator of build_prop_int (PropInt), which
    `interactions=True, propint=...` also selects as in the reference;
    `reference_pair_semantics=True` keeps the reference's behaviour of
    computing the pair sums but leaving them out of the Metropolis ratio.
    With interactions, a coupling g != 0 sets the reference's hard-sphere
    radius a = exp(-2 pi / g) (src/system.jl:148): bridge beads and COM shifts
    are redrawn on the device while they come closer than a to another
    particle (hardspherelevy!, move_polymer!, src/updates/helper.jl:141-181,
    368-395).  g = 0, the reference's default, gives a = 0.
    staging_every, com_every and swap_every set the move weights of the
    batched sweep (HIPSweep): `System(..., **sweep_weights(updates))` takes
    them from a reference-style update list.
    """

    def __init__(self, potential=None, *, dV=None, dim: int = 2, M: int = 100, N: int = 2, mu: float = 0.0,
                 L: float = 4.0, T: float = 1.0, lam: float = 1.0, pair: GaussianPair | None = None,
                 r_a: float = 0.0, length_measurement_cycle: int = 10, precision: str = "fp64",
                 seed: int | None = None, r0: np.ndarray | None = None, reference_pair_semantics: bool = False,
                 ell: int = 8, swaps: int = 0, swap_m: int = 0, swap_rule: int = 0, interactions: bool = False,
                 propint=None, g: float = 0.0, staging_every: int = 1, com_every: int = 1, swap_every: int = 1):
        hs_a = 0.0
        if interactions:
            if propint is None:
                raise ValueError("interactions=True needs propint (build_prop_int)")
            pair = propint
            hs_a = math.exp(-2 * math.pi / g) if g != 0 else 0.0  # exp(-2 pi / 0.0) = 0 in the reference
        self.a = hs_a
        self.potential = potential if potential is not None else NoPotential()
        self.dV = dV if dV is not None else getattr(self.potential, "dV", None)
        self.dim, self.M, self.N, self.Ninit = dim, M, N, N
        self.mu, self.lam, self.L = mu, lam, L
        self.vol = (2 * L) ** dim
        self.beta = 1.0 / T
        self.tau = self.beta / M
        self.pair = pair
        self.Ncycle = length_measurement_cycle
        self.N_MC = {N: 0}
        self.Nctr = {N: 0}
        self.rng = np.random.default_rng(seed)
        desc = chain_desc(self.potential, dim=dim, M=M, N=N, L=L, beta=self.beta, lam=lam, pair=pair, r_a=r_a,
                          pair_in_accept=not reference_pair_semantics, ell=min(ell, M), swaps=swaps, swap_m=swap_m,
                          swap_rule=swap_rule, hs_a=hs_a, staging_every=staging_every, com_every=com_every,
                          swap_every=swap_every)
        opt = Options()
        opt.precision = FP64 if precision == "fp64" else FP32
        opt.sweeps_per_launch = 1
        opt.seed = int(self.rng.integers(0, 2 ** 63))
        self.ctx = Context([desc], opt)
        self.r_a = desc.r_a
        if desc.pair == PAIR_TABLE:
            self.ctx.set_pair_table(0, pair.terms, pair.r0, pair.r1, mu)
        self.nbins = max(1, int(math.floor(2 * L / (desc.r_a if desc.r_a > 0 else L / 4))))
        if r0 is None:
            self.ctx.init_paths(0, int(self.rng.integers(0, 2 ** 63)))  # init_world on the device
        else:
            self.ctx.set_paths(0, r0)

    # paths in the C ABI layout (N, dim, M)
    @property
    def paths(self) -> np.ndarray:
        return self.ctx.get_paths(0)

    @paths.setter
    def paths(self, r):
        self.ctx.set_paths(0, r)

    @property
    def world(self) -> list[Particle]:
        """s.world with Particle.r as M x dim matrices and Particle.next (0-based)
        read back from the device; V[j] is the link action to the next bead
        along the permutation cycle."""
        r = self.paths
        nxt = self.next
        out = []
        for n in range(self.N):
            rn = r[n].T.copy()
            after = r[nxt[n]].T[0]  # slice 0 of the particle that follows across slice M - 1
            V = np.array([-0.5 * self.tau * (self.potential(rn[j]) +
                                             self.potential(rn[j + 1] if j + 1 < self.M else after))
                          for j in range(self.M)])
            out.append(Particle(rn, V, int(nxt[n])))
        return out

    @property
    def next(self) -> np.ndarray:
        """The permutation: next[p] follows particle p across slice M - 1."""
        return self.ctx.get_next(0)
