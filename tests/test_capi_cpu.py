"""CPU checks of the C ABI and of the host-side mirror (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from pimc import _capi as A
from pimc.potentials import GaussianPair, HarmonicTrap, SquareLattice
from pimc.system import chain_desc
from pimc.updates import (Counter, NumbOfSlices, ReshapeLinear, SingleCenterOfMass, Step, acceptance, adjust,
                          queue, sweep_weights)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pimc_hip.h")).read()
    return sorted(set(re.findall(r"\b(pimc_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(A.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(A.EXPORTS)


C_NAMES = {"lam": "lambda"}  # ctypes field -> header field where the names differ


@pytest.mark.parametrize("cstruct,pystruct", [("pimc_chain_desc", A.ChainDesc), ("pimc_options", A.Options),
                                              ("pimc_counters", A.Counters), ("pimc_estimates", A.Estimates)])
def test_struct_layouts_match_header(tmp_path, cstruct, pystruct):
    """sizeof and every field offset of the ctypes mirrors equal the C header's,
    as gcc lays the header out."""
    fields = [f[0] for f in pystruct._fields_]
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pimc_hip.h"', "int main(void) {",
             f'printf("%zu\\n", sizeof({cstruct}));']
    lines += [f'printf("%zu\\n", offsetof({cstruct}, {C_NAMES.get(f, f)}));' for f in fields]
    lines += ["return 0; }"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(pystruct)
    for f, off in zip(fields, out[1:]):
        assert getattr(pystruct, f).offset == off, f


@pytest.mark.parametrize("field,value,msg", [
    ("dim", 3, "dim"), ("N", 0, "N"), ("M", 2, "M"), ("beta", -1.0, "beta"), ("ell", 1, "ell"),
    ("ell", 40, "ell"), ("potential", 7, "potential"), ("sigma", 0.0, "sigma"),
    ("com_every", -1, "every"), ("staging_every", 70000, "every"), ("swap_every", -3, "every"),
])
def test_invalid_descriptors_are_rejected_before_any_device_call(field, value, msg):
    d = chain_desc(HarmonicTrap(1.0), dim=2, M=32, N=4, L=10.0, beta=1.0, lam=0.5, pair=GaussianPair(1.0, 0.5))
    setattr(d, field, value)
    opt = A.Options()
    opt.precision = A.FP32
    with pytest.raises(A.PimcError) as e:
        A.Context([d], opt)
    assert "PIMC_ERR_ARG" in str(e.value)
    assert msg in str(e.value)


def test_invalid_options_are_rejected():
    d = chain_desc(SquareLattice(4.0, 1.0), dim=2, M=32, N=4, L=10.0, beta=1.0, lam=0.5)
    opt = A.Options()
    opt.precision = 5
    with pytest.raises(A.PimcError):
        A.Context([d], opt)
    opt.precision = A.FP64
    opt.radial_bins = 16
    opt.radial_rmax = 0.0
    with pytest.raises(A.PimcError):
        A.Context([d], opt)


def test_counter_and_adjust_follow_the_reference():
    # queue! keeps a sliding window of `range` flags (src/simulation.jl:3-10)
    c = Counter(adj=10, range=4)
    for f in (1, 0, 1, 1, 1, 1):
        queue(c, bool(f))
    assert c.tries == 6 and list(c.queue) == [1, 1, 1, 1] and acceptance(c.queue) == 1.0
    # adjust!(::Step) x0.9 / x1.1 within [minstep, maxstep] (src/updates/helper.jl:22-32)
    s = Step(1.0, 0.1, 2.0, 0.4, 0.6)
    adjust(s, 0.3)
    assert s.size == pytest.approx(0.9)
    adjust(s, 0.7)
    assert s.size == pytest.approx(0.99)
    adjust(s, 0.5)
    assert s.size == pytest.approx(0.99)
    s.size = 1.95
    adjust(s, 0.9)
    assert s.size == 2.0
    # adjust!(::NumbOfSlices) -1 / +1 within [minslices, maxslices] (:42-52)
    n = NumbOfSlices(3, 2, 4, 0.6, 0.8)
    adjust(n, 0.9)
    adjust(n, 0.9)
    assert n.m == 4
    for _ in range(5):
        adjust(n, 0.1)
    assert n.m == 2


@pytest.mark.parametrize("case,msg", [
    ("table_without_ra", "r_a"), ("hs_without_table", "hard spheres"), ("hs_above_cell_side", "cell side"),
    ("negative_hs", "hs_a"), ("table_with_derivatives", "derivatives"), ("measure_kernel", "measure_kernel"),
    ("pair_kernel", "pair_kernel"), ("pair_kind", "pair"),
])
def test_table_and_hard_sphere_descriptors_are_checked_before_any_device_call(case, msg):
    from pimc.propagator import PropInt
    p = PropInt(np.zeros((4, 4)), 0.0, 3.0, 4.0, 0.05)
    d = chain_desc(HarmonicTrap(1.0), dim=2, M=32, N=4, L=4.0, beta=1.0, lam=0.5, pair=p, r_a=1.0)
    opt = A.Options()
    opt.precision = A.FP32
    if case == "table_without_ra":
        d.r_a = 0.0
    elif case == "hs_without_table":
        d = chain_desc(HarmonicTrap(1.0), dim=2, M=32, N=4, L=4.0, beta=1.0, lam=0.5, pair=GaussianPair(1.0, 0.5),
                       hs_a=0.1)
    elif case == "hs_above_cell_side":
        d.hs_a = 1.5  # nbins = floor(8 / 1.0) = 8, cell side 1.0
    elif case == "negative_hs":
        d.hs_a = -0.1
    elif case == "table_with_derivatives":
        opt.derivatives = 1
    elif case == "measure_kernel":
        opt.measure_kernel = 3
    elif case == "pair_kernel":
        opt.pair_kernel = 3
    elif case == "pair_kind":
        d.pair = 3
    with pytest.raises(A.PimcError) as e:
        A.Context([d], opt)
    assert "PIMC_ERR_ARG" in str(e.value)
    assert msg in str(e.value)


def test_prop_int_terms_checks_its_arguments_without_a_device():
    for args in [(1, 0.0, 1.0, 4.0, 0.1), (8, 1.0, 0.5, 4.0, 0.1), (8, 0.0, 1.0, 0.0, 0.1), (8, 0.0, 1.0, 4.0, 0.0),
                 (8, -1.0, 1.0, 4.0, 0.1)]:
        with pytest.raises(A.PimcError) as e:
            A.prop_int_terms(*args)
        assert "PIMC_ERR_ARG" in str(e.value)


def test_sweep_weights_from_reference_update_lists():
    """run!'s weights 1/every (src/simulation.jl:33-36) as pass periods: the
    reference's test mix (test/testsystem.jl:39-43), a scaled copy of it (the
    weights are relative), names as strings, and the lists refused."""
    mix = [(2, SingleCenterOfMass), (1, ReshapeLinear), (1, "ReshapeSwapLinear")]
    assert sweep_weights(mix) == {"com_every": 2, "staging_every": 1, "swap_every": 1}
    assert sweep_weights([(e * 3, f) for e, f in mix]) == sweep_weights(mix)
    assert sweep_weights([(1, "SingleCenterOfMass"), (1, "ReshapeLinear")]) == \
        {"com_every": 1, "staging_every": 1, "swap_every": 1}
    assert sweep_weights([(6, "SingleCenterOfMass"), (4, "ReshapeLinear")]) == \
        {"com_every": 3, "staging_every": 2, "swap_every": 1}
    for bad in ([(1, "ReshapeLinear")], [(1, "ReshapeLinear"), (1, "Energy"), (1, "SingleCenterOfMass")],
                [(0, "ReshapeLinear"), (1, "SingleCenterOfMass")],
                [(1, "ReshapeLinear"), (1, "ReshapeLinear"), (1, "SingleCenterOfMass")]):
        with pytest.raises(ValueError):
            sweep_weights(bad)
    d = chain_desc(HarmonicTrap(1.0), dim=2, M=32, N=4, L=10.0, beta=1.0, lam=0.5, **sweep_weights(mix))
    assert (d.staging_every, d.com_every, d.swap_every) == (1, 2, 1)

