"""ctypes binding of libpimc_hip.so (C ABI in include/pimc_hip.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libpimc_hip.so")

PIMC_OK = 0
ERRORS = {-1: "PIMC_ERR_ARG", -2: "PIMC_ERR_HIP", -3: "PIMC_ERR_STATE", -4: "PIMC_ERR_CAPACITY", -5: "PIMC_ERR_NOMEM"}
POT_NONE, POT_HARMONIC, POT_LATTICE, POT_OPTICAL = 0, 1, 2, 3
MAX_BEAMS = 16
PAIR_NONE, PAIR_GAUSSIAN, PAIR_TABLE = 0, 1, 2
FP32, FP64 = 0, 1

EXPORTS = [
    "pimc_create", "pimc_destroy", "pimc_last_error", "pimc_device_info",
    "pimc_set_paths", "pimc_get_paths", "pimc_init_paths", "pimc_set_next", "pimc_get_next",
    "pimc_get_com_step", "pimc_set_com_step", "pimc_get_ell", "pimc_get_sweep", "pimc_set_sweep",
    "pimc_prop_int_terms", "pimc_set_pair_table",
    "pimc_run", "pimc_get_estimates", "pimc_get_counters", "pimc_get_blocks", "pimc_get_radial", "pimc_get_cartesian",
    "pimc_reset_estimators", "pimc_levy_propose", "pimc_reshape_linear", "pimc_com_propose",
    "pimc_center_of_mass", "pimc_interaction", "pimc_energy", "pimc_density", "pimc_density_noalloc",
    "pimc_adjoint", "pimc_action",
]


class ChainDesc(C.Structure):
    _fields_ = [
        ("N", C.c_int32), ("M", C.c_int32), ("dim", C.c_int32),
        ("potential", C.c_int32), ("pair", C.c_int32), ("pair_in_accept", C.c_int32),
        ("ell", C.c_int32), ("swaps", C.c_int32),
        ("beta", C.c_double), ("lam", C.c_double), ("L", C.c_double),
        ("omega", C.c_double), ("V0", C.c_double), ("a", C.c_double),
        ("g", C.c_double), ("sigma", C.c_double), ("r_a", C.c_double), ("com_step", C.c_double),
        ("swap_m", C.c_int32), ("swap_rule", C.c_int32), ("hs_a", C.c_double),
        ("beam_scale", C.c_double), ("n_beams", C.c_int32), ("helical", C.c_int32),
        ("beam_angle", C.c_double * 16),
        ("staging_every", C.c_int32), ("com_every", C.c_int32), ("swap_every", C.c_int32),
    ]


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("sweeps_per_launch", C.c_int32),
        ("radial_bins", C.c_int32), ("cart_bins", C.c_int32), ("radial_rmax", C.c_double),
        ("tile_beads", C.c_int32), ("block_threads", C.c_int32),
        ("use_graph", C.c_int32), ("derivatives", C.c_int32), ("seed", C.c_uint64),
        ("measure_every", C.c_int32), ("adapt_ell", C.c_int32), ("measure_kernel", C.c_int32),
        ("pair_kernel", C.c_int32),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("sweeps", C.c_int64), ("bead_moves", C.c_int64),
        ("levy_tries", C.c_int64), ("levy_acc", C.c_int64), ("com_tries", C.c_int64), ("com_acc", C.c_int64),
        ("wall_ms", C.c_double), ("kernel_ms", C.c_double), ("kernel_launches", C.c_int64),
        ("swap_tries", C.c_int64), ("swap_acc", C.c_int64),
    ]


class Estimates(C.Structure):
    _fields_ = [
        ("n_blocks", C.c_int64), ("n_meas", C.c_int64),
        ("E", C.c_double), ("E_err", C.c_double), ("Ev", C.c_double), ("Ev_err", C.c_double),
        ("Epair", C.c_double), ("Epair_err", C.c_double),
        ("dSdomega", C.c_double), ("dSdg", C.c_double),
        ("dE_domega", C.c_double), ("dE_domega_err", C.c_double),
        ("dE_dg", C.c_double), ("dE_dg_err", C.c_double),
    ]


_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_lib = None


class PimcError(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PimcError(f"{path} is missing: build it with `python pimc.jl_amd/build.py` "
                        "(there is no CPU fallback)")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.pimc_create.argtypes = [C.POINTER(vp), C.POINTER(ChainDesc), C.c_int32, C.POINTER(Options)]
    L.pimc_destroy.argtypes = [vp]
    L.pimc_last_error.argtypes = [vp]
    L.pimc_last_error.restype = C.c_char_p
    L.pimc_device_info.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_int32)]
    L.pimc_set_paths.argtypes = [vp, C.c_int32, _D]
    L.pimc_get_paths.argtypes = [vp, C.c_int32, _D]
    L.pimc_init_paths.argtypes = [vp, C.c_int32, C.c_uint64]
    L.pimc_set_next.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
    L.pimc_get_next.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
    L.pimc_get_com_step.argtypes = [vp, C.c_int32, _D]
    L.pimc_set_com_step.argtypes = [vp, C.c_int32, C.c_double]
    L.pimc_get_ell.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
    L.pimc_get_sweep.argtypes = [vp, _I64]
    L.pimc_set_sweep.argtypes = [vp, C.c_int64]
    L.pimc_prop_int_terms.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, _D]
    L.pimc_set_pair_table.argtypes = [vp, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, _D]
    L.pimc_run.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(Counters)]
    L.pimc_get_estimates.argtypes = [vp, C.c_int32, C.POINTER(Estimates)]
    L.pimc_get_counters.argtypes = [vp, C.c_int32, _I64]
    L.pimc_get_blocks.argtypes = [vp, C.c_int32, _D, _D, _D, C.c_int64, _I64]
    L.pimc_get_radial.argtypes = [vp, C.c_int32, _D, _I64]
    L.pimc_get_cartesian.argtypes = [vp, C.c_int32, _D, _I64]
    L.pimc_reset_estimators.argtypes = [vp]
    L.pimc_levy_propose.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _D, _D, _D]
    L.pimc_reshape_linear.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _D, C.c_double,
                                      C.POINTER(C.c_int32)]
    L.pimc_com_propose.argtypes = [vp, C.c_int32, C.c_int32, _D, _D, _D]
    L.pimc_center_of_mass.argtypes = [vp, C.c_int32, C.c_int32, _D, C.c_double, C.POINTER(C.c_int32)]
    L.pimc_interaction.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, _D]
    L.pimc_energy.argtypes = [vp, C.c_int32, _D]
    L.pimc_density.argtypes = [vp, C.c_int32, C.c_int32, _D, _I64]
    L.pimc_density_noalloc.argtypes = [vp, C.c_int32, C.c_int32, _D, _I64]
    L.pimc_adjoint.argtypes = [vp, C.c_int32, _D, _D, _D]
    L.pimc_action.argtypes = [vp, C.c_int32, _D]
    for name in EXPORTS:
        if name not in ("pimc_last_error",):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def dp(a: np.ndarray):
    return a.ctypes.data_as(_D)


class Context:
    """Owns one pimc_ctx (device paths, stream, kernels) for a batch of chains."""

    def __init__(self, chains: list[ChainDesc], options: Options):
        self.lib = load()
        arr = (ChainDesc * len(chains))(*chains)
        self.h = C.c_void_p()
        rc = self.lib.pimc_create(C.byref(self.h), arr, len(chains), C.byref(options))
        if rc != PIMC_OK:
            msg = self.lib.pimc_last_error(self.h).decode() if self.h else ""
            if self.h:
                self.lib.pimc_destroy(self.h)
                self.h = C.c_void_p()
            raise PimcError(f"pimc_create: {ERRORS.get(rc, rc)} {msg}")
        self.chains = list(chains)
        self.options = options

    def check(self, rc: int, what: str):
        if rc != PIMC_OK:
            raise PimcError(f"{what}: {ERRORS.get(rc, rc)} {self.lib.pimc_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.lib.pimc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- paths
    def set_next(self, chain: int, nxt):
        """The permutation of one chain: nxt[p] follows particle p across slice M - 1."""
        a = np.ascontiguousarray(np.asarray(nxt, dtype=np.int32))
        self.check(self.lib.pimc_set_next(self.h, chain, a.ctypes.data_as(C.POINTER(C.c_int32))), "pimc_set_next")

    def get_next(self, chain: int) -> np.ndarray:
        a = np.zeros(self.chains[chain].N, dtype=np.int32)
        self.check(self.lib.pimc_get_next(self.h, chain, a.ctypes.data_as(C.POINTER(C.c_int32))), "pimc_get_next")
        return a

    def init_paths(self, chain: int = -1, seed: int = 0):
        """init_world on the device for one chain (or every chain: -1)."""
        self.check(self.lib.pimc_init_paths(self.h, chain, seed), "pimc_init_paths")

    def set_paths(self, chain: int, r: np.ndarray):
        d = self.chains[chain]
        r = np.ascontiguousarray(np.asarray(r, dtype=np.float64)).reshape(d.N, d.dim, d.M)
        self.check(self.lib.pimc_set_paths(self.h, chain, dp(r)), "pimc_set_paths")

    def get_paths(self, chain: int) -> np.ndarray:
        d = self.chains[chain]
        r = np.zeros((d.N, d.dim, d.M))
        self.check(self.lib.pimc_get_paths(self.h, chain, dp(r)), "pimc_get_paths")
        return r

    def com_step(self, chain: int) -> float:
        x = C.c_double()
        self.check(self.lib.pimc_get_com_step(self.h, chain, C.byref(x)), "pimc_get_com_step")
        return x.value

    def set_com_step(self, chain: int, step: float):
        self.check(self.lib.pimc_set_com_step(self.h, chain, step), "pimc_set_com_step")

    def set_pair_table(self, chain: int, terms: np.ndarray, r0: float, r1: float, mu: float = 0.0):
        """Table of a PIMC_PAIR_TABLE chain: terms[i, k] on the grid r0..r1."""
        t = np.ascontiguousarray(np.asarray(terms, dtype=np.float64))
        if t.ndim != 2 or t.shape[0] != t.shape[1]:
            raise PimcError("the pair table must be square")
        self.check(self.lib.pimc_set_pair_table(self.h, chain, t.shape[0], r0, r1, mu, dp(t)), "pimc_set_pair_table")

    def sweep(self) -> int:
        """The Philox sweep counter (pimc_get_sweep)."""
        v = C.c_int64()
        self.check(self.lib.pimc_get_sweep(self.h, C.byref(v)), "pimc_get_sweep")
        return v.value

    def set_sweep(self, sweep: int):
        self.check(self.lib.pimc_set_sweep(self.h, int(sweep)), "pimc_set_sweep")

    def ell(self, chain: int) -> int:
        x = C.c_int32()
        self.check(self.lib.pimc_get_ell(self.h, chain, C.byref(x)), "pimc_get_ell")
        return x.value

    # -- batched sweeps
    def run(self, n_sweeps: int, measure: bool) -> Counters:
        cnt = Counters()
        self.check(self.lib.pimc_run(self.h, n_sweeps, int(measure), C.byref(cnt)), "pimc_run")
        return cnt

    def estimates(self, chain: int) -> Estimates:
        e = Estimates()
        self.check(self.lib.pimc_get_estimates(self.h, chain, C.byref(e)), "pimc_get_estimates")
        return e

    def counters(self, chain: int) -> np.ndarray:
        """{staging accepted, staging tries, COM accepted, COM tries} of one chain."""
        out = np.zeros(4, dtype=np.int64)
        self.check(self.lib.pimc_get_counters(self.h, chain, out.ctypes.data_as(_I64)), "pimc_get_counters")
        return out

    def all_counters(self) -> np.ndarray:
        """counters() of every chain, as an (n_chains, 4) array, in one copy."""
        out = np.zeros((len(self.chains), 4), dtype=np.int64)
        self.check(self.lib.pimc_get_counters(self.h, -1, out.ctypes.data_as(_I64)), "pimc_get_counters")
        return out

    def blocks(self, chain: int):
        n = C.c_int64()
        self.check(self.lib.pimc_get_blocks(self.h, chain, None, None, None, 0, C.byref(n)), "pimc_get_blocks")
        E, Ev, Ep = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value)
        self.check(self.lib.pimc_get_blocks(self.h, chain, dp(E), dp(Ev), dp(Ep), n.value, C.byref(n)),
                   "pimc_get_blocks")
        return E, Ev, Ep

    def radial(self, chain: int):
        h = np.zeros(self.options.radial_bins)
        nd = C.c_int64()
        self.check(self.lib.pimc_get_radial(self.h, chain, dp(h), C.byref(nd)), "pimc_get_radial")
        return h, nd.value

    def cartesian(self, chain: int):
        d = self.chains[chain]
        nb = self.options.cart_bins
        h = np.zeros(nb ** d.dim)
        nd = C.c_int64()
        self.check(self.lib.pimc_get_cartesian(self.h, chain, dp(h), C.byref(nd)), "pimc_get_cartesian")
        return h.reshape((nb,) * d.dim, order="F"), nd.value

    def reset_estimators(self):
        self.check(self.lib.pimc_reset_estimators(self.h), "pimc_reset_estimators")

    # -- single moves and estimators (the reference functors)
    def levy_propose(self, chain, n, j0, m, normals):
        d = self.chains[chain]
        nrm = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(m - 1, d.dim))
        r = np.zeros((d.dim, m + 1))
        w = np.zeros(4)
        self.check(self.lib.pimc_levy_propose(self.h, chain, n, j0, m, dp(nrm), dp(r), dp(w)), "pimc_levy_propose")
        return r.T.copy(), w

    def reshape_linear(self, chain, n, j0, m, normals, u) -> bool:
        d = self.chains[chain]
        nrm = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(m - 1, d.dim))
        acc = C.c_int32()
        self.check(self.lib.pimc_reshape_linear(self.h, chain, n, j0, m, dp(nrm), u, C.byref(acc)),
                   "pimc_reshape_linear")
        return bool(acc.value)

    def com_propose(self, chain, n, disp):
        d = self.chains[chain]
        dd = np.ascontiguousarray(np.asarray(disp, dtype=np.float64).reshape(d.dim))
        r = np.zeros((d.dim, d.M))
        w = np.zeros(4)
        self.check(self.lib.pimc_com_propose(self.h, chain, n, dp(dd), dp(r), dp(w)), "pimc_com_propose")
        return r.T.copy(), w

    def center_of_mass(self, chain, n, disp, u) -> bool:
        d = self.chains[chain]
        dd = np.ascontiguousarray(np.asarray(disp, dtype=np.float64).reshape(d.dim))
        acc = C.c_int32()
        self.check(self.lib.pimc_center_of_mass(self.h, chain, n, dp(dd), u, C.byref(acc)), "pimc_center_of_mass")
        return bool(acc.value)

    def interaction(self, chain, n, j) -> float:
        w = C.c_double()
        self.check(self.lib.pimc_interaction(self.h, chain, n, j, C.byref(w)), "pimc_interaction")
        return w.value

    def energy(self, chain) -> np.ndarray:
        out = np.zeros(4)
        self.check(self.lib.pimc_energy(self.h, chain, dp(out)), "pimc_energy")
        return out

    def density(self, chain, nbins, dens=None, ndata=0):
        d = self.chains[chain]
        if dens is None:
            dens = np.zeros((nbins,) * d.dim, order="F")
        buf = np.asfortranarray(dens, dtype=np.float64)
        nd = C.c_int64(ndata)
        self.check(self.lib.pimc_density(self.h, chain, nbins, buf.ctypes.data_as(_D), C.byref(nd)), "pimc_density")
        return buf, nd.value

    def density_noalloc(self, chain, nbins, dens=None, ndata=0):
        """pimc_density_noalloc: Density(s) with the counts in the library's fixed device array."""
        d = self.chains[chain]
        if dens is None:
            dens = np.zeros((nbins,) * d.dim, order="F")
        buf = np.asfortranarray(dens, dtype=np.float64)
        nd = C.c_int64(ndata)
        self.check(self.lib.pimc_density_noalloc(self.h, chain, nbins, buf.ctypes.data_as(_D), C.byref(nd)),
                   "pimc_density_noalloc")
        return buf, nd.value

    def adjoint(self, chain):
        d = self.chains[chain]
        g = np.zeros((d.N, d.dim, d.M))
        par = np.zeros(2)
        S = C.c_double()
        self.check(self.lib.pimc_adjoint(self.h, chain, dp(g), dp(par), C.byref(S)), "pimc_adjoint")
        return S.value, g, par

    def action(self, chain) -> float:
        S = C.c_double()
        self.check(self.lib.pimc_action(self.h, chain, C.byref(S)), "pimc_action")
        return S.value


def prop_int_terms(n: int, r0: float, r1: float, g0: float, tau: float) -> np.ndarray:
    """pimc_prop_int_terms: the n x n integral terms of the pair-product
    propagator on the grid r0..r1, computed on the GPU."""
    L = load()
    out = np.zeros((n, n))
    rc = L.pimc_prop_int_terms(n, r0, r1, g0, tau, dp(out))
    if rc != PIMC_OK:
        raise PimcError(f"pimc_prop_int_terms: {ERRORS.get(rc, rc)} {L.pimc_last_error(None).decode()}")
    return out


def device_info():
    L = load()
    buf = C.create_string_buffer(256)
    cu = C.c_int32()
    rc = L.pimc_device_info(buf, 256, C.byref(cu))
    if rc != PIMC_OK:
        raise PimcError("no HIP device")
    return buf.value.decode(), cu.value
