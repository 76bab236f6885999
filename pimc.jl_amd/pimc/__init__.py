"""Python mirror of PIMC.jl's System / update / measurement API over the HIP path.

The reference (Julia package `Pimc`) drives the sampler through callable
structs: updates `<: UpdateC` called as `u(s)::Bool` inside `run!`, and
measurements `<: ZMeasurement` called as `m(s)` (src/types.jl:1-13,
src/simulation.jl:12-42).  This module keeps that shape and these names, with
every move and estimator executed by libpimc_hip.so on the GPU:

    s = System(HarmonicTrap(1.0), dV=..., lam=0.5, M=64, N=10, L=100.0, T=1.0)
    updates = [(1, SingleCenterOfMass(s, 1.0)), (1, ReshapeLinear(s, 8))]
    run(s, 10_000, updates)                               # thermalise
    run(s, 10_000, updates, Zmeasurements=[Energy(s)])    # measure

`Batch` is the batched form (thousands of chains in one launch, `HIPSweep`
granularity) used for throughput.
"""
from __future__ import annotations

from ._capi import (FP32, FP64, PAIR_GAUSSIAN, PAIR_NONE, PAIR_TABLE, POT_HARMONIC, POT_LATTICE, POT_NONE, ChainDesc,
                    Context, Counters, Estimates, Options, PimcError, device_info, load, prop_int_terms)
from .potentials import GaussianPair, HarmonicTrap, NoPotential, OpticalLattice, SquareLattice, generate_V
from .propagator import PropInt, build_prop_int
from .system import Particle, System, init_world
from .updates import (Counter, NumbOfSlices, ReshapeLinear, SingleCenterOfMass, HIPSweep, Step, Update, UpdateC,
                      acceptance, adjust, queue, sweep_weights)
from .measurement import Density, Energy, Measurement, ZMeasurement, measurement_Z_sector
from .simulation import apply, run
from .batch import Batch

run_bang = run  # run! in the reference

__all__ = [
    "System", "Particle", "init_world", "run", "run_bang", "apply",
    "ReshapeLinear", "SingleCenterOfMass", "HIPSweep", "Update", "UpdateC", "sweep_weights",
    "Counter", "Step", "NumbOfSlices", "adjust", "queue", "acceptance",
    "Energy", "Density", "Measurement", "ZMeasurement", "measurement_Z_sector",
    "HarmonicTrap", "SquareLattice", "OpticalLattice", "generate_V", "NoPotential", "GaussianPair", "PropInt", "build_prop_int", "prop_int_terms",
    "Batch", "Context", "ChainDesc", "Options", "Counters", "Estimates", "PimcError", "device_info", "load",
    "FP32", "FP64", "POT_NONE", "POT_HARMONIC", "POT_LATTICE", "PAIR_NONE", "PAIR_GAUSSIAN", "PAIR_TABLE",
]
