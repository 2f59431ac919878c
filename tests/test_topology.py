"""Topology, units and CHK I/O against the reference's fixtures and a brute-force restatement."""

import glob
import os

import numpy as np
import pytest

from micmec_amd import units
from micmec_amd.chk import dump_chk, load_chk
from micmec_amd.system import Domain, System
from micmec_amd.topology import Grid, build_system, periodic_wrap_bits
from oracle import constants as oc
from oracle import topology as otop

HERE = os.path.dirname(os.path.abspath(__file__))
CHK_FILES = sorted(glob.glob(os.path.join(HERE, "golden", "chk", "*_micmec.chk")))


def _types_of(d):
    out = {}
    for key in d:
        if key.startswith("type") and "/" in key:
            t = int(key.split("/")[0][4:])
            if t == 0:
                continue
            out.setdefault(t, {})[key.split("/", 1)[1]] = d[key]
    return out


def test_units_pinned_by_fixture():
    d = load_chk(os.path.join(HERE, "golden", "chk", "2x2x2_test_micmec.chk"))
    # 10 Angstrom lattice spacing and the C44 = 10 GPa tensor entry, as printed in the file
    assert "%.15e" % (10.0 * units.angstrom) == "%.15e" % d["type1/cell"][0, 0, 0]
    assert "%.15e" % (10e9 * units.pascal) == "%.15e" % d["type1/elasticity"][0, 1, 2, 1, 2]
    assert d["type1/effective_temp"] == 300.0 * units.kelvin
    for name in ("angstrom", "pascal", "bar", "femtosecond", "amu", "boltzmann", "kelvin"):
        assert getattr(units, name) == getattr(oc, name), name


def test_fixture_count():
    assert len(CHK_FILES) == 19


@pytest.mark.parametrize("fn", CHK_FILES, ids=[os.path.basename(f) for f in CHK_FILES])
def test_build_system_reproduces_fixture(fn):
    d = load_chk(fn)
    grid = np.rint(d["grid"]).astype(np.int64)
    out = build_system(_types_of(d), grid, pbc=[bool(p) for p in d["pbc"]])
    for key in ("surrounding_nodes", "surrounding_cells", "boundary_nodes", "types"):
        assert np.array_equal(np.asarray(out[key]), np.rint(d[key]).astype(np.int64)), key
    # Floats are stored with 16 significant digits; two fixtures were written from a type cell
    # that is itself rounded in the file, hence one unit of the 16th digit.
    for key in ("masses", "pos", "rvecs"):
        a, b = np.asarray(out[key]), d[key]
        assert a.shape == b.shape
        assert np.all(np.abs(a - b) <= 1e-15 * np.abs(b) + 1e-300), key


@pytest.mark.parametrize("seed", range(12))
def test_vectorised_topology_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(rng.integers(1, 5, size=3))
    pbc = [bool(b) for b in rng.integers(0, 2, size=3)] if seed % 3 else [True] * 3
    if any(p and n < 2 for p, n in zip(pbc, shape)):
        pbc = [p and n >= 2 for p, n in zip(pbc, shape)]
    grid = rng.integers(0, 3, size=shape)
    if not (grid != 0).any():
        grid[0, 0, 0] = 1
    masses = {1: 3.0, 2: 5.0}
    sn0, sc0, m0, nodes0, cells0, types0 = otop.surrounding_tables(grid, masses, pbc)
    g = Grid(grid, pbc=pbc)
    sn, sc = g.surrounding()
    assert g.nodes == nodes0 and g.cells == cells0 and list(g.types) == types0
    assert np.array_equal(sn, sn0) and np.array_equal(sc, sc0)
    data = {t: {"mass": masses[t], "cell": [np.identity(3) * (1 + t)]} for t in (1, 2)}
    out = build_system(data, grid, pbc=pbc)
    assert np.array_equal(out["masses"], m0)


@pytest.mark.parametrize("shape,hole_seed", [((2, 2, 2), None), ((3, 3, 3), None),
                                              ((4, 3, 5), None), ((2, 3, 4), 1),
                                              ((3, 3, 3), 2), ((5, 5, 5), 3), ((4, 4, 4), 4)])
def test_wrap_bits_equal_reference_image_table(shape, hole_seed):
    grid = np.ones(shape, dtype=np.int64)
    if hole_seed is not None:
        rng = np.random.default_rng(hole_seed)
        grid[rng.random(shape) < 0.4] = 0
        grid[0, 0, 0] = 1
    mic = otop.get_mic(grid, True)
    g = Grid(grid, pbc=True)
    sn, _ = g.surrounding()
    bits = g.wrap_bits()
    offs = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
                     (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)])
    for c in range(len(sn)):
        for k in range(1, 8):
            expect = np.array([((bits[c] >> d) & 1) * offs[k, d] for d in range(3)], float)
            assert np.array_equal(mic[sn[c, 0], sn[c, k]], expect), (c, k)


def test_chk_roundtrip(tmp_path):
    d = load_chk(CHK_FILES[0])
    fn = str(tmp_path / "x.chk")
    dump_chk(fn, d)
    d2 = load_chk(fn)
    assert set(d) == set(d2)
    for k in d:
        if isinstance(d[k], np.ndarray):
            assert np.array_equal(d[k], d2[k]), k
        else:
            assert d[k] == d2[k], k


def test_system_file_roundtrip(tmp_path):
    s = System.from_file(CHK_FILES[3])
    fn = str(tmp_path / "s.chk")
    s.to_file(fn)
    s2 = System.from_file(fn)
    assert np.array_equal(s.pos, s2.pos)
    assert np.array_equal(s.surrounding_cells, s2.surrounding_cells)
    assert s2.domain.volume == s.domain.volume


def test_domain_volume_and_gvecs():
    r = np.array([[3.0, 0.1, 0.0], [0.2, 4.0, 0.3], [0.0, 0.5, 5.0]])
    dom = Domain(r)
    assert dom.nvec == 3
    assert abs(dom.volume - abs(np.linalg.det(r))) < 1e-12
    assert np.allclose(dom.gvecs @ r.T, np.identity(3), atol=1e-14)
    assert Domain(np.zeros((0, 3))).volume == 0.0
    with pytest.raises(TypeError):
        Domain(np.zeros((4, 3)))


def test_periodic_wrap_bits_nonperiodic_axes():
    ijk = np.array([[0, 0, 0], [1, 2, 3], [1, 0, 3]])
    assert list(periodic_wrap_bits(ijk, (2, 3, 4), [True, False, True])) == [0, 1 | 4, 1 | 4]
