"""Writes the golden fixtures of tests/golden/ (run from the repo root).

* philox_kat.json: the Philox4x32-10 known-answer vectors published with
  Random123 (kat_vectors: counter, key -> output), the RNG that keys every
  draw of the sampling path.
* analytic_energies.json: closed-form energies of the trapped ideal gas for
  the workloads (normal modes of the discrete primitive-action ring and the
  continuum / bosonic forms of examples/tools/energytools.jl:1-21), and the
  value the reference's example quotes for N=1, beta=1
  (examples/energy_2d_harmonically_trapped_bose_gas.jl:28).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import analytic as A  # noqa: E402

PHILOX_KAT = [
    {"ctr": ["00000000", "00000000", "00000000", "00000000"], "key": ["00000000", "00000000"],
     "out": ["6627e8d5", "e169c58d", "bc57ac4c", "9b00dbd8"]},
    {"ctr": ["ffffffff", "ffffffff", "ffffffff", "ffffffff"], "key": ["ffffffff", "ffffffff"],
     "out": ["408f276d", "41c83b0e", "a20bc7c6", "6d5451fd"]},
    {"ctr": ["243f6a88", "85a308d3", "13198a2e", "03707344"], "key": ["a4093822", "299f31d0"],
     "out": ["d16cfe09", "94fdcceb", "5001e420", "24126ea1"]},
]


def main():
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump({"source": "Random123 kat_vectors, philox4x32 10 rounds", "vectors": PHILOX_KAT}, f, indent=1)
    cases = []
    for N, M, beta in [(10, 64, 1.0), (1, 5, 1.0), (1, 64, 1.0), (4, 16, 2.0), (1, 4, 1.0), (64, 256, 2.0),
                       (4, 32, 0.5), (64, 256, 4.0)]:
        cases.append({
            "N": N, "M": M, "beta": beta, "dim": 2, "omega": 1.0, "lambda": 0.5,
            "E_boltzmann_primitive": A.energy_boltzmann_primitive(N, M, beta),
            "x2_primitive": A.x2_primitive(M, beta),
            "E_boltzmann_continuum": A.energy_boltzmann_continuum(N, beta),
        })
    out = {
        "reference_example_exact": 2.1639534137386534,
        "reference_example_source": "examples/energy_2d_harmonically_trapped_bose_gas.jl:28 (N=1, beta=1, 2D)",
        "E_boson_continuum_N10_beta1": A.energy_boson_continuum(10, 1.0),
        "dE_domega_boltzmann_continuum_N10_beta1": A.denergy_domega_continuum(10, 1.0),
        "cases": cases,
    }
    with open(os.path.join(HERE, "analytic_energies.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
