"""Synthetic systems of the benchmark and parity workloads.

Cell types:

* type A: h0 = 20.7 Angstrom * I, cubic C11/C12/C44 = 40/20/10 GPa, one state;
* type B: h0 = 20.9 Angstrom * I, C11/C12/C44 = 30/18/7 GPa, and a second state with h0 scaled
  by 0.97 and C scaled by 0.8; states combined at T_eff = 300 K.

Both carry a cell mass of 8 x 6000 amu, so a node surrounded by eight cells weighs 6000 amu
(node masses are one eighth of each adjacent cell's mass, micmec/utils.py:217). Systems are
built with :func:`micmec_amd.topology.build_system` (micmec/utils.py:164-263); node positions
get seeded Gaussian noise on top of the lattice.
"""

import numpy as np

from . import units
from .system import System
from .topology import build_system, build_type

__all__ = ["cubic_elasticity", "type_a", "type_b", "make_grid", "make_system", "workload",
           "batch_replicas", "WORKLOADS"]

_V = {0: (0, 0), 1: (1, 1), 2: (2, 2), 3: (1, 2), 4: (0, 2), 5: (0, 1)}


def voigt_to_tensor(matrix):
    """Full 3x3x3x3 elasticity tensor from a 6x6 Voigt matrix (micmec/analysis/tensor.py:188-197)."""
    tensor = np.zeros((3, 3, 3, 3))
    inv = {v: k for k, v in _V.items()}
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    I = inv[tuple(sorted((i, j)))]
                    J = inv[tuple(sorted((k, l)))]
                    tensor[i, j, k, l] = matrix[I, J]
    return tensor


def cubic_elasticity(c11, c12, c44):
    """Cubic elasticity tensor (atomic units) from C11, C12, C44 in GPa."""
    gpa = units.gigapascal
    m = np.zeros((6, 6))
    m[:3, :3] = c12 * gpa
    for i in range(3):
        m[i, i] = c11 * gpa
        m[3 + i, 3 + i] = c44 * gpa
    return voigt_to_tensor(m)


CELL_MASS = 8 * 6000 * units.amu


def type_a():
    h0 = 20.7 * units.angstrom * np.identity(3)
    return build_type("A", CELL_MASS, [h0], [cubic_elasticity(40.0, 20.0, 10.0)],
                      free_energy=[0.0], effective_temp=300.0 * units.kelvin)


def type_b():
    h0 = 20.9 * units.angstrom * np.identity(3)
    C = cubic_elasticity(30.0, 18.0, 7.0)
    return build_type("B", CELL_MASS, [h0, 0.97 * h0], [C, 0.8 * C], free_energy=[0.0, 0.0],
                      effective_temp=300.0 * units.kelvin)


def make_grid(shape, p_b=0.0, p_hole=0.0, seed=0):
    """Type grid: 1 = A, 2 = B with probability p_b, 0 = empty with probability p_hole."""
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in shape)
    grid = np.where(rng.random(shape) < p_b, 2, 1).astype(np.int64)
    if p_hole > 0:
        holes = rng.random(shape) < p_hole
        grid[holes] = 0
    if not (grid != 0).any():
        grid.flat[0] = 1
    return grid


def make_system(shape, p_b=0.0, p_hole=0.0, seed=0, noise=0.05 * units.angstrom, pbc=True,
                strain=0.0):
    """A System on a seeded grid with Gaussian position noise (standard deviation ``noise``)
    and, if ``strain`` > 0, a seeded uniform random symmetric strain of that amplitude applied
    to positions and domain vectors."""
    grid = make_grid(shape, p_b, p_hole, seed)
    data = {1: type_a(), 2: type_b()}
    used = {int(t) for t in np.unique(grid) if t != 0}
    out = build_system({t: data[t] for t in used}, grid, pbc=pbc)
    rng = np.random.default_rng(seed + 7919)
    pos = out["pos"] + rng.normal(0.0, 1.0, out["pos"].shape) * noise
    rvecs = out["rvecs"]
    if strain > 0:
        e = rng.uniform(-strain, strain, (3, 3))
        e = 0.5 * (e + e.T)
        F = np.identity(3) + e
        pos = pos @ F
        rvecs = rvecs @ F if len(rvecs) else rvecs
    out["pos"] = pos
    out["rvecs"] = rvecs
    return System.from_dict(out)


def batch_replicas(count=512, seed=4000, nmin=4, nmax=24):
    """Replica set of the batched workload: grid edge n uniform in {nmin..nmax}, a seeded type
    mix (probability of type B uniform in [0, 1]), a defect fraction uniform in [0, 0.3] and an
    initial temperature uniform in [100, 500] K. Returns (system, temperature, seed) triples;
    replica r uses seed + r for its grid and its velocities."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        p_b = float(rng.uniform(0.0, 1.0))
        p_hole = float(rng.uniform(0.0, 0.3))
        temp = float(rng.uniform(100.0, 500.0))
        out.append((make_system((n, n, n), p_b=p_b, p_hole=p_hole, seed=seed + r), temp,
                    seed + r))
    return out


def workload(name, **overrides):
    """Description dict of a named workload (sizes, ensemble, seeds)."""
    w = dict(WORKLOADS[name])
    w.update(overrides)
    return w


WORKLOADS = {
    # plain method, CPU-runnable: 3^3 periodic, type A only, NVE
    "nve3": dict(shape=(3, 3, 3), p_b=0.0, p_hole=0.0, seed=0, ensemble="nve", steps=1000,
                 timestep_fs=1.0, temp=300.0),
    # plain method, mid size: 32^3, 20% type B, NVT with a Nose-Hoover chain
    "nhc32": dict(shape=(32, 32, 32), p_b=0.2, p_hole=0.0, seed=1, ensemble="nvt-nhc",
                  steps=2000, timestep_fs=1.0, temp=300.0, chain=3, timecon_fs=100.0),
    # plain method, headline size: 128^3, 20% B, 5% empty cells, NPT with NHC + MTK at 1 bar
    "npt128": dict(shape=(128, 128, 128), p_b=0.2, p_hole=0.05, seed=2, ensemble="npt-mtk",
                   steps=500, timestep_fs=1.0, temp=300.0, press_bar=1.0, chain=3,
                   timecon_fs=100.0, baro_timecon_fs=1000.0),
}
