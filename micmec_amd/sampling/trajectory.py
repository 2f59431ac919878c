"""Trajectory writers (drop-in for micmec/sampling/trajectory.py:25-126).

``HDF5Writer`` and ``BaseHDF5Writer`` keep the reference's layout: a ``trajectory`` group with
one resizable dataset per state item (first axis = frame), the items' attributes on the group,
and a ``system`` group written once (``System.to_hdf5``). They take an open file object with the
h5py interface (``h5py.File`` where h5py is installed; any object offering ``create_group``,
``create_dataset(name, shape, maxshape=, dtype=)``, ``resize`` and ``attrs`` works).

``XYZWriter`` writes the node positions as an XYZ trajectory, every node labelled as a caesium
atom like the reference (trajectory.py:108-126), coordinates in angstrom, without molmod.
"""

import numpy as np

from ..units import angstrom
from .iterative import Hook

__all__ = ["BaseHDF5Writer", "HDF5Writer", "XYZWriter", "write_xyz_frame"]


def _recorded_items(iterative):
    """The (key, item) pairs a trajectory stores: items that hold a value and whose shape has
    no empty axis."""
    for key, item in iterative.state.items():
        if item.value is None or 0 in tuple(item.shape):
            continue
        yield key, item


class BaseHDF5Writer(Hook):
    """Append the iterative's state items, one frame per call, to ``f['trajectory']``
    (layout of trajectory.py:31-80)."""

    def __init__(self, f, start=0, step=1):
        self.f = f
        Hook.__init__(self, start, step)

    def __call__(self, iterative):
        if "trajectory" not in self.f:
            self.init_trajectory(iterative)
        group = self.f["trajectory"]
        frame = self._next_frame(group, iterative)
        for key, item in _recorded_items(iterative):
            dataset = group[key]
            if dataset.shape[0] < frame + 1:
                dataset.resize(frame + 1, axis=0)
            dataset[frame] = item.value

    @staticmethod
    def _next_frame(group, iterative):
        """Index of the frame to write: the length of the shortest dataset of this iterative's
        state, so that a frame an interrupted call left incomplete is written again."""
        lengths = [group[key].shape[0] for key in iterative.state if key in group]
        return min(lengths)

    def dump_system(self, system, grp):
        system.to_hdf5(grp)

    def init_trajectory(self, iterative):
        """Create one empty, extensible dataset per recorded item (frames along axis 0) and
        store the items' attributes on the group."""
        group = self.f.create_group("trajectory")
        for key, item in _recorded_items(iterative):
            group.create_dataset(key, (0,) + tuple(item.shape),
                                 maxshape=(None,) + tuple(item.shape), dtype=item.dtype)
            for name, value in item.iter_attrs(iterative):
                group.attrs[name] = value


class HDF5Writer(BaseHDF5Writer):
    """``BaseHDF5Writer`` that stores the system in the file before the first frame
    (trajectory.py:83-87)."""

    def __call__(self, iterative):
        if "system" not in self.f:
            self.dump_system(iterative.mmf.system, self.f)
        BaseHDF5Writer.__call__(self, iterative)


def write_xyz_frame(fh, symbols, title, pos):
    """One XYZ frame: atom count, title line, one 'symbol x y z' line per atom (angstrom)."""
    fh.write("%i\n%s\n" % (len(symbols), title))
    for sym, p in zip(symbols, np.asarray(pos) / angstrom):
        fh.write("%2s % 15.10f % 15.10f % 15.10f\n" % (sym, p[0], p[1], p[2]))


class XYZWriter(Hook):
    """Node positions of every due iteration appended to ``fn_xyz`` (trajectory.py:89-126)."""

    def __init__(self, fn_xyz, select=None, start=0, step=1):
        self.fn_xyz = fn_xyz
        self.select = select
        self._symbols = None
        Hook.__init__(self, start, step)

    def __call__(self, iterative):
        pos = np.asarray(iterative.mmf.system.pos)
        if self.select is not None:
            pos = pos[self.select]
        mode = "a"
        if self._symbols is None:
            self._symbols = ["Cs"] * len(pos)
            mode = "w"
        title = "%7i E_pot = %.10f     " % (iterative.counter, iterative.epot)
        with open(self.fn_xyz, mode) as fh:
            write_xyz_frame(fh, self._symbols, title, pos)
