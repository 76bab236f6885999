"""ctypes front end of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only the
test-suite, the smoke check and the CPU-baseline leg of the benchmark import
this module; the product never does.

The C library restates PIMC.jl's sampling path (file:line citations are in
pimc_oracle.c); this module only marshals numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

POT_NONE, POT_HARMONIC, POT_LATTICE, POT_OPTICAL = 0, 1, 2, 3
PAIR_NONE, PAIR_GAUSSIAN, PAIR_TABLE = 0, 1, 2


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (make rebuilds it when a
    source is newer; force rebuilds it regardless)."""
    if force or not os.path.exists(_LIB_PATH) or \
            any(os.path.getmtime(os.path.join(_HERE, f)) > os.path.getmtime(_LIB_PATH)
                for f in ("pimc_oracle.c", "pimc_oracle.h", "Makefile") if os.path.exists(os.path.join(_HERE, f))):
        subprocess.run(["make", "-s", "-B" if force else "liboracle.so"], cwd=_HERE, check=True)
    return _LIB_PATH


class OPot(C.Structure):
    _fields_ = [("kind", C.c_int), ("omega", C.c_double), ("V0", C.c_double), ("a", C.c_double),
                ("nbeam", C.c_int), ("helical", C.c_int), ("scale", C.c_double), ("angle", C.c_double * 16)]


class OPair(C.Structure):
    _fields_ = [("kind", C.c_int), ("g", C.c_double), ("sigma", C.c_double),
                ("tab", C.POINTER(C.c_double)), ("tab_n", C.c_int), ("tab_r0", C.c_double),
                ("tab_r1", C.c_double), ("mu", C.c_double)]


class OSys(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("M", C.c_int), ("N", C.c_int),
        ("lam", C.c_double), ("L", C.c_double), ("beta", C.c_double), ("tau", C.c_double),
        ("pot", OPot), ("pair", OPair),
        ("nbins", C.c_int), ("h", C.c_double), ("pair_in_accept", C.c_int),
        ("r", C.POINTER(C.c_double)), ("V", C.POINTER(C.c_double)), ("next", C.POINTER(C.c_int64)),
        ("hs_a", C.c_double), ("hs_ctr", C.c_int), ("cells", C.c_void_p),
    ]


class OWeights(C.Structure):
    _fields_ = [("wV_initial", C.c_double), ("wV_updated", C.c_double),
                ("wU_initial", C.c_double), ("wU_updated", C.c_double)]


class ORefState(C.Structure):
    _fields_ = [
        ("cap", C.c_int), ("step", C.c_double), ("bead_moves", C.c_int64),
        ("levy_tries", C.c_int64), ("levy_acc", C.c_int64), ("com_tries", C.c_int64), ("com_acc", C.c_int64),
        ("rng_counter", C.c_uint64), ("n_meas", C.c_int64), ("win_range", C.c_int), ("adj", C.c_int),
    ]


class OSweepStats(C.Structure):
    _fields_ = [("levy_acc", C.c_int64), ("levy_tries", C.c_int64), ("com_acc", C.c_int64),
                ("com_tries", C.c_int64), ("bead_moves", C.c_int64), ("swap_acc", C.c_int64),
                ("swap_tries", C.c_int64)]


_D = C.POINTER(C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        u32p = C.POINTER(C.c_uint32)
        L.o_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.o_u01.argtypes = [C.c_uint32]; L.o_u01.restype = C.c_double
        L.o_sweep_perm.argtypes = [C.POINTER(OSys), C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, C.c_double,
                                   C.c_int, C.c_int, C.c_int, C.POINTER(OSweepStats)]
        L.o_swap_pair_lw.argtypes = [C.POINTER(OSys), C.c_int, C.c_int, C.c_int, C.c_int, _D, _D, C.c_int]
        L.o_swap_pair_lw.restype = C.c_double
        L.o_sweep_every.argtypes = [C.POINTER(OSys), C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, C.c_double,
                                    C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(OSweepStats)]
        L.o_init_world.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_double, _D]
        L.o_init_world_hs.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_double, _D, C.POINTER(C.c_int)]
        L.o_init_world_hs.restype = C.c_int
        L.o_box_muller.argtypes = [C.c_uint32, C.c_uint32, _D, _D]
        L.o_distance.argtypes = [C.c_double] * 3; L.o_distance.restype = C.c_double
        L.o_teleport.argtypes = [C.c_double] * 2; L.o_teleport.restype = C.c_double
        L.o_lnK.argtypes = [_D, _D, C.c_int, C.c_double, C.c_double, C.c_double]; L.o_lnK.restype = C.c_double
        L.o_V.argtypes = [C.POINTER(OPot), _D, C.c_int]; L.o_V.restype = C.c_double
        L.o_dV.argtypes = [C.POINTER(OPot), _D, C.c_int, _D]
        L.o_lnV.argtypes = [_D, _D, C.c_int, C.c_double, C.POINTER(OPot)]; L.o_lnV.restype = C.c_double
        L.o_lnU.argtypes = [_D, _D, C.c_int, C.c_double, C.POINTER(OPair)]; L.o_lnU.restype = C.c_double
        L.o_vpair.argtypes = [C.POINTER(OPair), C.c_double]; L.o_vpair.restype = C.c_double
        L.o_metropolis.argtypes = [C.c_double, C.c_double]; L.o_metropolis.restype = C.c_int
        L.o_prop_quad_nodes.argtypes = [C.c_double, C.c_double, _D, _D, C.c_int]; L.o_prop_quad_nodes.restype = C.c_int
        L.o_prop_int_term.argtypes = [C.c_double] * 5; L.o_prop_int_term.restype = C.c_double
        L.o_prop_int_terms.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _D]
        L.o_prop_table_eval.argtypes = [C.POINTER(OPair), C.c_double, C.c_double]
        L.o_prop_table_eval.restype = C.c_double
        L.o_prop_int.argtypes = [C.POINTER(OPair), _D, _D, C.c_int, C.c_double]; L.o_prop_int.restype = C.c_double
        L.o_levy.argtypes = [_D, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _D]
        L.o_sys_init.argtypes = [C.POINTER(OSys), C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                 OPot, OPair, C.c_double, C.c_int, _D]
        L.o_sys_init.restype = C.c_int
        L.o_sys_free.argtypes = [C.POINTER(OSys)]
        L.o_sys_refresh_V.argtypes = [C.POINTER(OSys)]
        L.o_bin.argtypes = [C.POINTER(OSys), _D]; L.o_bin.restype = C.c_int
        L.o_find_nns.argtypes = [C.POINTER(OSys), _D, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int]
        L.o_find_nns.restype = C.c_int
        L.o_find_nn.argtypes = [C.POINTER(OSys), _D, C.c_int, C.c_int]; L.o_find_nn.restype = C.c_int
        L.o_interaction_old.argtypes = [C.POINTER(OSys), C.c_int, C.c_int]; L.o_interaction_old.restype = C.c_double
        L.o_interaction_new.argtypes = [C.POINTER(OSys), _D, C.c_int, C.c_int, C.c_int, C.c_int]
        L.o_interaction_new.restype = C.c_double
        L.o_interaction_com.argtypes = [C.POINTER(OSys), _D, C.c_int, C.c_int]; L.o_interaction_com.restype = C.c_double
        L.o_reshape_propose.argtypes = [C.POINTER(OSys), C.c_int, C.c_int, C.c_int, _D, _D, C.POINTER(OWeights)]
        L.o_reshape_apply.argtypes = [C.POINTER(OSys), C.c_int, C.c_int, C.c_int, _D, C.c_double,
                                      C.POINTER(C.c_int), C.POINTER(OWeights)]
        L.o_com_propose.argtypes = [C.POINTER(OSys), C.c_int, _D, _D, C.POINTER(OWeights)]
        L.o_com_apply.argtypes = [C.POINTER(OSys), C.c_int, _D, C.c_double, C.POINTER(C.c_int), C.POINTER(OWeights)]
        L.o_energy.argtypes = [C.POINTER(OSys), _D, _D]
        L.o_pair_energy.argtypes = [C.POINTER(OSys), _D, _D]
        L.o_density.argtypes = [C.POINTER(OSys), C.c_int, _D, C.POINTER(C.c_int64)]
        L.o_radial.argtypes = [C.POINTER(OSys), C.c_int, C.c_double, _D, C.POINTER(C.c_int64)]
        for f in ("o_total_lnV", "o_total_lnU", "o_total_lnK"):
            getattr(L, f).argtypes = [C.POINTER(OSys)]
            getattr(L, f).restype = C.c_double
        L.o_refstate_init.argtypes = [C.POINTER(ORefState), C.POINTER(OSys), C.c_int, C.c_double]
        L.o_run_reference.argtypes = [C.POINTER(OSys), C.POINTER(ORefState), C.c_int64, C.c_uint64, C.c_int,
                                      _D, _D, C.c_int64]
        L.o_sweep.argtypes = [C.POINTER(OSys), C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, C.c_double,
                              C.POINTER(OSweepStats)]
        L.o_adapt_step.argtypes = [C.c_double, C.c_int64, C.c_int64, C.c_double]; L.o_adapt_step.restype = C.c_double
        L.o_adapt_ell.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_int]; L.o_adapt_ell.restype = C.c_int
        L.o_sweep_bead_moves.argtypes = [C.c_int, C.c_int, C.c_int]; L.o_sweep_bead_moves.restype = C.c_int64
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_D)


# ---------------------------------------------------------------- primitives
def philox(ctr, key) -> np.ndarray:
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().o_philox4x32_10(c, k, o)
    return np.array(list(o), dtype=np.uint32)


def box_muller(x0: int, x1: int):
    a, b = C.c_double(), C.c_double()
    lib().o_box_muller(x0, x1, C.byref(a), C.byref(b))
    return a.value, b.value


def distance(x1, x2, L):
    return lib().o_distance(x1, x2, L)


def teleport(x, L):
    return lib().o_teleport(x, L)


def metropolis(delta, u):
    return bool(lib().o_metropolis(delta, u))


def levy(rp: np.ndarray, tau: float, L: float, lam: float, normals: np.ndarray) -> np.ndarray:
    """levy! on a (rows x dim) array; returns the new array."""
    rp = np.asarray(rp, dtype=np.float64)
    nrows, dim = rp.shape
    buf = np.ascontiguousarray(rp.T.copy())          # column-major (dim, rows)
    nrm = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(nrows - 2, dim))
    lib().o_levy(_dp(buf), nrows, dim, tau, L, lam, _dp(nrm))
    return buf.T.copy()


def make_pot(kind=POT_NONE, omega=1.0, V0=0.0, a=1.0, angles=(), scale=1.0, helical=False) -> OPot:
    p = OPot(kind, omega, V0, a)
    p.nbeam, p.helical, p.scale = len(angles), int(bool(helical)), scale
    for i, x in enumerate(angles):
        p.angle[i] = x
    return p


def make_pair(kind=PAIR_NONE, g=0.0, sigma=1.0, table=None, r0=0.0, r1=0.0, mu=0.0) -> OPair:
    """table: the n x n integral terms on the grid r0..r1 (kind PAIR_TABLE); the
    array is kept alive by the returned structure."""
    p = OPair(kind, g, sigma)
    if table is not None:
        t = np.ascontiguousarray(np.asarray(table, dtype=np.float64))
        p.tab, p.tab_n, p.tab_r0, p.tab_r1, p.mu = _dp(t), t.shape[0], r0, r1, mu
        p._keep = t
    return p


def pot_from_desc(d) -> OPot:
    """The oracle potential of a chain descriptor (any object with the
    pimc_chain_desc field names)."""
    angles = [d.beam_angle[i] for i in range(d.n_beams)] if d.potential == POT_OPTICAL else ()
    return make_pot(d.potential, d.omega, d.V0, d.a, angles=angles, scale=d.beam_scale, helical=bool(d.helical))


def pot_V(pot: OPot, r):
    a = np.ascontiguousarray(r, dtype=np.float64)
    return lib().o_V(C.byref(pot), _dp(a), len(a))


def pot_dV(pot: OPot, r):
    a = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(len(a))
    lib().o_dV(C.byref(pot), _dp(a), len(a), _dp(out))
    return out


def prop_quad_nodes(rmax, tau):
    n = lib().o_prop_quad_nodes(rmax, tau, None, None, 0)
    k, w = np.zeros(n), np.zeros(n)
    lib().o_prop_quad_nodes(rmax, tau, _dp(k), _dp(w), n)
    return k, w


def prop_int_term(r1, r2, g0, tau, rmax):
    """int_terms(r1, r2) of prop_rel_interpolate_terms by the oracle's quadrature rule."""
    return lib().o_prop_int_term(r1, r2, g0, tau, rmax)


def prop_int_terms(n, r0, r1, g0, tau):
    out = np.zeros((n, n))
    lib().o_prop_int_terms(n, r0, r1, g0, tau, _dp(out))
    return out


def prop_table_eval(pair: OPair, x, y):
    return lib().o_prop_table_eval(C.byref(pair), x, y)


def prop_int(pair: OPair, r1rel, r2rel, tau):
    a = np.ascontiguousarray(r1rel, dtype=np.float64)
    b = np.ascontiguousarray(r2rel, dtype=np.float64)
    return lib().o_prop_int(C.byref(pair), _dp(a), _dp(b), len(a), tau)


def vpair(pair: OPair, r):
    """The pair potential v(r) of a Gaussian pair (0 for other kinds)."""
    return lib().o_vpair(C.byref(pair), float(r))


def lnU(pair: OPair, r1rel, r2rel, tau):
    a = np.ascontiguousarray(r1rel, dtype=np.float64)
    b = np.ascontiguousarray(r2rel, dtype=np.float64)
    return lib().o_lnU(_dp(a), _dp(b), len(a), tau, C.byref(pair))


class System:
    """Oracle system; r has shape (N, dim, M) (the reference's per-particle M x dim, column-major)."""

    def __init__(self, r, lam, L, beta, pot: OPot, pair: OPair | None = None, r_a: float = 0.0,
                 pair_in_accept: bool = False, hs_a: float = 0.0, hs_ctr: int = 10000):
        r = np.ascontiguousarray(np.asarray(r, dtype=np.float64))
        N, dim, M = r.shape
        self._s = OSys()
        rc = lib().o_sys_init(C.byref(self._s), dim, M, N, lam, L, beta, pot,
                              pair if pair is not None else make_pair(), r_a, int(pair_in_accept), _dp(r))
        if rc != 0:
            raise ValueError(f"o_sys_init failed ({rc})")
        self._pair = pair  # keeps a pair table alive while the system holds its pointer
        self._s.hs_a, self._s.hs_ctr = hs_a, hs_ctr
        self.N, self.dim, self.M = N, dim, M

    @property
    def hs_ctr(self):
        return self._s.hs_ctr

    @hs_ctr.setter
    def hs_ctr(self, v):
        self._s.hs_ctr = int(v)

    def find_nn(self, target, j, exclude):
        t = np.ascontiguousarray(target, dtype=np.float64)
        return lib().o_find_nn(C.byref(self._s), _dp(t), j, exclude)

    def __del__(self):
        try:
            lib().o_sys_free(C.byref(self._s))
        except Exception:
            pass

    @property
    def ref(self):
        return C.byref(self._s)

    @property
    def tau(self):
        return self._s.tau

    @property
    def nbins(self):
        return self._s.nbins

    @property
    def h(self):
        return self._s.h

    @property
    def r(self) -> np.ndarray:
        n = self.N * self.dim * self.M
        return np.ctypeslib.as_array(self._s.r, shape=(n,)).reshape(self.N, self.dim, self.M).copy()

    @r.setter
    def r(self, val):
        val = np.ascontiguousarray(np.asarray(val, dtype=np.float64)).reshape(-1)
        arr = np.ctypeslib.as_array(self._s.r, shape=(val.size,))
        arr[:] = val
        lib().o_sys_refresh_V(self.ref)

    @property
    def V(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._s.V, shape=(self.N * self.M,)).reshape(self.N, self.M).copy()

    # -- hot-path pieces
    def reshape_propose(self, n, j0, m, normals):
        rp = np.zeros((self.dim, m + 1))
        w = OWeights()
        nrm = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(m - 1, self.dim))
        ok = lib().o_reshape_propose(self.ref, n, j0, m, _dp(nrm), _dp(rp), C.byref(w))
        if not ok:
            raise ValueError("bad staging length")
        return rp.T.copy(), w

    def reshape_apply(self, n, j0, m, normals, u):
        acc = C.c_int()
        w = OWeights()
        nrm = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(m - 1, self.dim))
        lib().o_reshape_apply(self.ref, n, j0, m, _dp(nrm), u, C.byref(acc), C.byref(w))
        return bool(acc.value), w

    def com_propose(self, n, d):
        rp = np.zeros((self.dim, self.M))
        w = OWeights()
        dd = np.ascontiguousarray(np.asarray(d, dtype=np.float64))
        lib().o_com_propose(self.ref, n, _dp(dd), _dp(rp), C.byref(w))
        return rp.T.copy(), w

    def com_apply(self, n, d, u):
        acc = C.c_int()
        w = OWeights()
        dd = np.ascontiguousarray(np.asarray(d, dtype=np.float64))
        lib().o_com_apply(self.ref, n, _dp(dd), u, C.byref(acc), C.byref(w))
        return bool(acc.value), w

    def find_nns(self, target, j, exclude):
        out = (C.c_int * 8192)()
        t = np.ascontiguousarray(np.asarray(target, dtype=np.float64))
        cnt = lib().o_find_nns(self.ref, _dp(t), j, exclude, out, 8192)
        return [out[i] for i in range(min(cnt, 8192))]

    def interaction_old(self, n, j):
        return lib().o_interaction_old(self.ref, n, j)

    def energy(self):
        E, Ev = C.c_double(), C.c_double()
        lib().o_energy(self.ref, C.byref(E), C.byref(Ev))
        return E.value, Ev.value

    def pair_energy(self):
        E, Ev = C.c_double(), C.c_double()
        lib().o_pair_energy(self.ref, C.byref(E), C.byref(Ev))
        return E.value, Ev.value

    def density(self, nbins, dens=None, ndata=0):
        if dens is None:
            dens = np.zeros((nbins,) * self.dim)
        buf = np.asfortranarray(dens)
        nd = C.c_int64(ndata)
        lib().o_density(self.ref, nbins, buf.ctypes.data_as(_D), C.byref(nd))
        return buf, nd.value

    def radial(self, nb, rmax, hist=None, ndata=0):
        if hist is None:
            hist = np.zeros(nb)
        hist = np.ascontiguousarray(hist)
        nd = C.c_int64(ndata)
        lib().o_radial(self.ref, nb, rmax, _dp(hist), C.byref(nd))
        return hist, nd.value

    def total_lnV(self):
        return lib().o_total_lnV(self.ref)

    def total_lnU(self):
        return lib().o_total_lnU(self.ref)

    def total_lnK(self):
        return lib().o_total_lnK(self.ref)

    def sweep(self, chain, sweep, seed, ell, com_step):
        st = OSweepStats()
        lib().o_sweep(self.ref, chain, sweep, seed, ell, com_step, C.byref(st))
        return st

    def sweep_perm(self, chain, sweep, seed, ell, com_step, n_swap, swap_m, rule=0):
        """o_sweep_perm: the sweep with permutation cycles and n_swap swap tries."""
        st = OSweepStats()
        lib().o_sweep_perm(self.ref, chain, sweep, seed, ell, com_step, n_swap, swap_m, rule, C.byref(st))
        return st

    def swap_pair_lw(self, n1, n2, j0, m, rp1, rp2, reference):
        """o_swap_pair_lw: rp1, rp2 are (m + 1) x dim new paths (bead-major)."""
        a = np.ascontiguousarray(np.asarray(rp1, float).T)
        b = np.ascontiguousarray(np.asarray(rp2, float).T)
        return lib().o_swap_pair_lw(self.ref, n1, n2, j0, m, _dp(a), _dp(b), int(reference))

    def sweep_every(self, chain, sweep, seed, ell, com_step, every=(1, 1, 1), n_swap=0, swap_m=0, rule=0):
        """o_sweep_every: the sweep whose staging pass, COM pass and swap tries
        run only on the sweeps with sweep % every == 0 (every = (staging, com, swap))."""
        st = OSweepStats()
        lib().o_sweep_every(self.ref, chain, sweep, seed, ell, com_step, n_swap, swap_m, rule,
                            int(every[0]), int(every[1]), int(every[2]), C.byref(st))
        return st

    @property
    def next(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._s.next, shape=(self.N,)).copy()

    @next.setter
    def next(self, nxt):
        a = np.ctypeslib.as_array(self._s.next, shape=(self.N,))
        a[:] = np.asarray(nxt, dtype=np.int64)


class RefRunner:
    """run!(s, n, [(1, SingleCenterOfMass(s, step)), (1, ReshapeLinear(s, cap))]) on the oracle."""

    def __init__(self, system: System, cap: int = 8, step: float = 1.0, seed: int = 1):
        self.s = system
        self.st = ORefState()
        self.seed = seed
        lib().o_refstate_init(C.byref(self.st), system.ref, cap, step)

    def run(self, n_iter: int, ncycle: int = 10, measure: bool = False):
        if measure:
            nE = n_iter // ncycle + 1
            E = np.zeros(nE)
            Ev = np.zeros(nE)
            n0 = self.st.n_meas
            lib().o_run_reference(self.s.ref, C.byref(self.st), n_iter, self.seed, ncycle, _dp(E), _dp(Ev), nE)
            k = self.st.n_meas - n0
            self.st.n_meas = 0
            return E[:k], Ev[:k]
        lib().o_run_reference(self.s.ref, C.byref(self.st), n_iter, self.seed, ncycle, None, None, 0)
        return None


def sweep_bead_moves(N, M, ell):
    return lib().o_sweep_bead_moves(N, M, ell)


def init_world(seed, chain, N, M, dim, L, tau, lam):
    """o_init_world: N x dim x M initial paths (init_world with a = 0, Philox draws)."""
    out = np.zeros((N, dim, M))
    lib().o_init_world(seed, chain, N, M, dim, L, tau, lam, out.ctypes.data_as(_D))
    return out


def init_world_hs(seed, chain, N, M, dim, L, tau, lam, a):
    """o_init_world_hs: (paths N x dim x M, tries per particle, rc) of init_world
    with hard spheres of radius a (rc -1: a particle found no place)."""
    out = np.zeros((N, dim, M))
    tries = np.zeros(N, dtype=np.int32)
    rc = lib().o_init_world_hs(seed, chain, N, M, dim, L, tau, lam, a, out.ctypes.data_as(_D),
                               tries.ctypes.data_as(C.POINTER(C.c_int)))
    return out, tries, rc


def adapt_step(step, acc, tries, L):
    return lib().o_adapt_step(step, acc, tries, L)


def adapt_ell(ell, acc, tries, M):
    return lib().o_adapt_ell(ell, acc, tries, M)
