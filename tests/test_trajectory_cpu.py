"""On-disk formats on the host (micmec/sampling/trajectory.py:25-126, micmec/system.py:261-370):
the HDF5 system group round trip, the HDF5 trajectory of an optimisation (one row per hook
call, every state item), the XYZ writer read back, and the h5py gate."""

import numpy as np
import pytest

from h5mem import Group
from micmec_amd import units
from micmec_amd.sampling import CartesianDOF, CGOptimizer, HDF5Writer, XYZWriter
from micmec_amd.system import System
from micmec_amd.workloads import make_system


def read_xyz(fn):
    frames = []
    with open(fn) as fh:
        lines = fh.read().splitlines()
    i = 0
    while i < len(lines):
        n = int(lines[i])
        title = lines[i + 1]
        rows = [ln.split() for ln in lines[i + 2:i + 2 + n]]
        frames.append((title, [r[0] for r in rows],
                       np.array([[float(x) for x in r[1:]] for r in rows])))
        i += 2 + n
    return frames


def test_system_hdf5_round_trip():
    s = make_system((3, 2, 4), p_b=0.4, p_hole=0.2, seed=3)
    f = Group()
    s.to_hdf5(f)
    with pytest.raises(ValueError):
        s.to_hdf5(f)
    s2 = System.from_hdf5(f)
    for key in ("pos", "masses", "surrounding_cells", "surrounding_nodes", "boundary_nodes",
                "grid", "types"):
        assert np.array_equal(np.asarray(getattr(s2, key)), np.asarray(getattr(s, key))), key
    assert np.array_equal(s2.domain.rvecs, s.domain.rvecs)
    assert sorted(s2.params) == sorted(s.params)
    for k in s.params:
        assert np.array_equal(np.asarray(s2.params[k]), np.asarray(s.params[k])), k


class _OracleForceField(object):
    def __init__(self, system):
        from micmec_amd.pes.mmff import system_wrap_bits
        self.system = system
        self.wrap_bits = system_wrap_bits(system)
        self.parts = []

    def update_rvecs(self, rvecs):
        self.system.domain.update_rvecs(np.asarray(rvecs, dtype=float))

    def update_pos(self, pos):
        self.system.pos[:] = pos

    def compute(self, gpos=None, vtens=None):
        from oracle.mmff import compute_vectorised
        s = self.system
        e, g, v = compute_vectorised(s.pos, s.domain.rvecs, s.types, s.surrounding_nodes,
                                     s.surrounding_cells, self.wrap_bits, s.params)
        if gpos is not None:
            gpos += g
        if vtens is not None:
            vtens += v
        return e


def test_hdf5_and_xyz_trajectories_of_an_optimisation(tmp_path):
    s = make_system((3, 3, 3), p_b=0.3, seed=4, strain=0.02, noise=0.05 * units.angstrom)
    f = Group()
    fn = str(tmp_path / "opt.xyz")
    opt = CGOptimizer(CartesianDOF(_OracleForceField(s)),
                      hooks=[HDF5Writer(f, step=2), XYZWriter(fn, step=3)])
    seen = {}
    for _ in range(7):
        opt.propagate()
        seen[opt.counter] = (opt.epot, np.array(s.pos))
    traj = f["trajectory"]
    assert "system" in f and np.array_equal(f["system"]["masses"][:], s.masses)
    assert list(traj["counter"][:]) == [0, 2, 4, 6]
    assert traj["pos"].shape == (4, s.nnodes, 3)
    for row, c in enumerate([2, 4, 6]):
        assert np.array_equal(traj["pos"][row + 1], seen[c][1])
        assert traj["epot"][row + 1] == seen[c][0]
    frames = read_xyz(fn)
    assert [int(t.split()[0]) for t, _, _ in frames] == [0, 3, 6]
    assert all(sym == ["Cs"] * s.nnodes for _, sym, _ in frames)
    for (title, _, xyz), c in zip(frames[1:], [3, 6]):
        assert np.abs(xyz - seen[c][1] / units.angstrom).max() < 1e-9
        assert abs(float(title.split("=")[1]) - seen[c][0]) < 1e-9


def test_system_file_formats(tmp_path):
    s = make_system((2, 2, 2), p_b=0.0, seed=1)
    s.to_file(str(tmp_path / "s.xyz"))
    (_, sym, xyz), = read_xyz(str(tmp_path / "s.xyz"))
    assert sym == ["Cs"] * s.nnodes
    assert np.abs(xyz - s.pos / units.angstrom).max() < 1e-9
    with pytest.raises(NotImplementedError):
        s.to_file(str(tmp_path / "s.pdb"))
    try:
        import h5py  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError):
            s.to_file(str(tmp_path / "s.h5"))
    else:
        s.to_file(str(tmp_path / "s.h5"))
        s2 = System.from_file(str(tmp_path / "s.h5"))
        assert np.array_equal(s2.pos, s.pos)
