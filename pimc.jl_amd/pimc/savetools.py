"""Output formats of the reference's example tools (examples/tools/savetools.jl:4-69).

paths_df / save_paths write the worldlines as a table with a `tau` column
(j beta / M for j = 0..M) and one column per particle and coordinate
("p1 x", "p1 y", ...): M beads of the particle plus, as row M + 1, slice 0 of
the particle that follows it along its permutation cycle, so each column
closes its link across slice M - 1.  density_df / save_density write the
Cartesian Density normalised by bin * ndata next to a `pos` column
range(-L, L, length = nbins).  CSV text as CSV.jl writes it for these tables:
a header line, comma separators, numbers in shortest round-trip form.
"""
from __future__ import annotations

import numpy as np


def _fmt(x) -> str:
    return repr(float(x))


def paths_df(s) -> tuple[list[str], list[np.ndarray]]:
    """paths_df(s) (savetools.jl:4-31): (column names, columns).  s needs
    paths (N x dim x M), next (0-based), M, N, dim and beta, as System has."""
    r, nxt = s.paths, s.next
    names, cols = ["tau"], [np.array([j * s.beta / s.M for j in range(s.M + 1)])]
    axes = "xy"
    for n in range(s.N):
        for d in range(s.dim):
            names.append(f"p{n + 1} {axes[d]}")
            cols.append(np.append(r[n, d, :], r[nxt[n], d, 0]))
    return names, cols


def _write(path, names, cols):
    with open(path, "w") as f:
        f.write(",".join(names) + "\n")
        for i in range(len(cols[0])):
            f.write(",".join(_fmt(c[i]) for c in cols) + "\n")


def save_paths(s, path: str = "paths.csv") -> str:
    """save_paths(s) (savetools.jl:33-37)."""
    names, cols = paths_df(s)
    _write(path, names, cols)
    return path


def density_df(s, d) -> tuple[list[str], list[np.ndarray]]:
    """density_df(s, d) (savetools.jl:39-49): `pos` and the normalised
    density; a 2D density becomes columns x1 .. x_nbins (the matrix columns)."""
    den = np.asarray(d.dens, dtype=float) / (d.bin * d.ndata)
    pos = np.linspace(-s.L, s.L, d.nbins)
    if den.ndim == 1:
        return ["pos", "Density"], [pos, den]
    return ["pos"] + [f"x{k + 1}" for k in range(den.shape[1])], [pos] + [den[:, k] for k in range(den.shape[1])]


def save_density(s, d, g, name, V, directory: str = ".") -> str:
    """save_density(s, d, g, name, V) (savetools.jl:61-69): the file name
    carries M, beta, N, g and V as the reference's does."""
    path = f"{directory}/density{name}M{s.M}beta{float(s.beta)}N{s.N}g{g}V{V}.csv"
    names, cols = density_df(s, d)
    _write(path, names, cols)
    return path
