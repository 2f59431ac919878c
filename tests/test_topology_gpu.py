"""Topology built on the GPU (mm_topology_build, micmec/utils.py:55-263) against the host
builder, which is pinned to the reference's CHK fixtures: every table equal entry for entry
(indices exactly; masses and positions bit for bit, since the device sums the masses in the
same (cell, corner) order) on random grids with holes, several types, every periodicity and
axes of one or two cells; the reference's fixtures themselves; and a large grid."""

import glob
import os

import numpy as np
import pytest

from micmec_amd import units
from micmec_amd.chk import load_chk
from micmec_amd.topology import (Grid, build_system, build_system_device, build_type,
                                 periodic_wrap_bits)

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CHK_FILES = sorted(glob.glob(os.path.join(HERE, "golden", "chk", "*_micmec.chk")))
KEYS_INT = ("surrounding_nodes", "surrounding_cells", "boundary_nodes", "types")
KEYS_FLOAT = ("masses", "pos", "rvecs")


def types_of(d):
    out = {}
    for key in d:
        if key.startswith("type") and "/" in key:
            t = int(key.split("/")[0][4:])
            if t == 0:
                continue
            out.setdefault(t, {})[key.split("/", 1)[1]] = d[key]
    return out


def random_types(rng, ntypes):
    data = {}
    for t in range(1, ntypes + 1):
        h = np.diag(rng.uniform(15.0, 25.0, 3)) * units.angstrom
        data[t] = build_type("t%d" % t, rng.uniform(1e3, 9e3) * units.amu, h,
                             np.zeros((3, 3, 3, 3)))
    return data


def compare(dev, host):
    for key in KEYS_INT:
        a, b = np.asarray(dev[key]), np.asarray(host[key])
        assert a.shape == b.shape and np.array_equal(a, b), key
    for key in KEYS_FLOAT:
        a, b = np.asarray(dev[key]), np.asarray(host[key])
        assert a.shape == b.shape and np.array_equal(a, b), key


@pytest.mark.parametrize("seed", range(16))
def test_device_topology_equals_host_builder(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(x) for x in rng.integers(1, 9, size=3))
    if seed == 0:
        shape = (1, 1, 1)
    ntypes = int(rng.integers(1, 4))
    grid = rng.integers(1, ntypes + 1, size=shape)
    grid[rng.random(shape) < rng.uniform(0.0, 0.4)] = 0
    if not grid.any():
        grid.flat[0] = 1
    pbc = [bool(b) for b in rng.integers(0, 2, size=3)]
    data = random_types(rng, ntypes)
    print("topology", seed, shape, pbc, flush=True)
    host = build_system(data, grid, pbc=pbc)
    dev = build_system_device(data, grid, pbc=pbc, return_extra=True)
    compare(dev, host)
    g = Grid(grid, pbc=pbc)
    cell_ijk = np.array(g.cells, dtype=np.int64).reshape(-1, 3)
    assert np.array_equal(dev["wrap_bits"], periodic_wrap_bits(cell_ijk, grid.shape, g.pbc))
    assert np.array_equal(dev["cell_point"], np.flatnonzero(grid))


@pytest.mark.parametrize("fn", CHK_FILES, ids=[os.path.basename(f) for f in CHK_FILES])
def test_device_topology_reproduces_fixture(fn):
    d = load_chk(fn)
    grid = np.rint(d["grid"]).astype(np.int64)
    pbc = [bool(p) for p in d["pbc"]]
    data = types_of(d)
    dev = build_system_device(data, grid, pbc=pbc)
    compare(dev, build_system(data, grid, pbc=pbc))
    for key in KEYS_INT:
        assert np.array_equal(np.asarray(dev[key]), np.rint(d[key]).astype(np.int64)), key


def test_large_grid():
    rng = np.random.default_rng(5)
    grid = rng.integers(1, 3, size=(96, 80, 64))
    grid[rng.random(grid.shape) < 0.05] = 0
    data = random_types(rng, 2)
    compare(build_system_device(data, grid), build_system(data, grid))
