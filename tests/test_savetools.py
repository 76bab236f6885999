"""Output formats of examples/tools/savetools.jl (CPU, no device): column
layout, the closing row along the permutation cycle, the density
normalisation (sum(den) * bin = sum(dens) / ndata, the invariant of
test/testmeasurements.jl:24-29) and the matrix orientation of a 2D density."""
import csv
import types

import numpy as np

from pimc import savetools as ST


def fake_system(N=3, M=4, dim=2, beta=2.0, L=3.0, nxt=(1, 0, 2), seed=0):
    r = np.random.default_rng(seed).uniform(-L, L, size=(N, dim, M))
    return types.SimpleNamespace(paths=r, next=np.array(nxt), N=N, M=M, dim=dim, beta=beta, L=L)


def test_paths_df_columns_and_cycle_closure(tmp_path):
    s = fake_system()
    names, cols = ST.paths_df(s)
    assert names == ["tau", "p1 x", "p1 y", "p2 x", "p2 y", "p3 x", "p3 y"]
    assert np.allclose(cols[0], [0.0, 0.5, 1.0, 1.5, 2.0])
    # particles 1 and 2 form a 2-cycle: each column closes on the other's slice 0
    assert np.array_equal(cols[1][:4], s.paths[0, 0]) and cols[1][4] == s.paths[1, 0, 0]
    assert cols[4][4] == s.paths[0, 1, 0]
    assert cols[6][4] == s.paths[2, 1, 0]  # a 1-cycle closes on itself
    path = ST.save_paths(s, str(tmp_path / "paths.csv"))
    rows = list(csv.reader(open(path)))
    assert rows[0] == names and len(rows) == s.M + 2
    assert float(rows[3][2]) == cols[2][2]


def test_paths_df_one_dimension():
    s = fake_system(N=2, dim=1, nxt=(0, 1))
    names, _ = ST.paths_df(s)
    assert names == ["tau", "p1 x", "p2 x"]


def test_density_df_normalisation_and_orientation(tmp_path):
    s = fake_system(L=2.0)
    d = types.SimpleNamespace(nbins=5, bin=4.0 / 5, ndata=12, dens=np.arange(25.0).reshape(5, 5))
    names, cols = ST.density_df(s, d)
    assert names == ["pos", "x1", "x2", "x3", "x4", "x5"]
    assert np.allclose(cols[0], np.linspace(-2.0, 2.0, 5))
    den = np.array(cols[1:]).T  # back to the matrix
    assert np.allclose(den, d.dens / (d.bin * d.ndata))
    assert np.isclose(den.sum() * d.bin, d.dens.sum() / d.ndata)
    d1 = types.SimpleNamespace(nbins=4, bin=1.0, ndata=2, dens=np.array([0.0, 2.0, 4.0, 2.0]))
    n1, c1 = ST.density_df(s, d1)
    assert n1 == ["pos", "Density"] and np.allclose(c1[1], [0.0, 1.0, 2.0, 1.0])
    p = ST.save_density(s, d1, 1.5, "trap", 4, directory=str(tmp_path))
    assert p.endswith("densitytrapM4beta2.0N3g1.5V4.csv")
    rows = list(csv.reader(open(p)))
    assert rows[0] == ["pos", "Density"] and float(rows[2][1]) == 1.0
