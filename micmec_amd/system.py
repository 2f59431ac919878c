"""The micromechanical system container and its periodic domain.

``System`` mirrors micmec/system.py:38-370: the same constructor arguments, ``from_file`` and
``to_file`` for CHK, HDF5 (where h5py is installed) and XYZ files, ``to_hdf5`` / ``from_hdf5``
on open h5py-style files, and attribute names read by the force field and integrators.
``Domain`` mirrors the Cython extension type of micmec/pes/ext.pyx:36-123 and its C helper
micmec/pes/domain.c:13-49 (rvecs, gvecs from the SVD pseudo-inverse, volume, nvec).
"""

import numpy as np

from .chk import load_chk, dump_chk

__all__ = ["System", "Domain"]


class Domain(object):
    """Periodic boundary conditions of the simulation domain (ext.pyx:36-123)."""

    def __init__(self, rvecs):
        self.update_rvecs(rvecs)

    def update_rvecs(self, rvecs):
        """Store new domain vectors; recompute gvecs and volume (ext.pyx:52-72, domain.c:13-49)."""
        if rvecs is None or np.asarray(rvecs).size == 0:
            mod_rvecs = np.identity(3)
            gvecs = mod_rvecs
            nvec = 0
        else:
            rvecs = np.asarray(rvecs, dtype=np.float64)
            if rvecs.ndim != 2 or rvecs.shape[0] > 3 or rvecs.shape[1] != 3:
                raise TypeError("rvecs must be an array with three columns and at most three rows.")
            nvec = len(rvecs)
            Up, Sp, Vt = np.linalg.svd(rvecs, full_matrices=True)
            S = np.ones(3)
            S[:nvec] = Sp
            U = np.identity(3)
            U[:nvec, :nvec] = Up
            mod_rvecs = np.dot(U * S, Vt)
            mod_rvecs[:nvec] = rvecs
            gvecs = np.dot(U / S, Vt)
        self._nvec = nvec
        self._rvecs_full = np.array(mod_rvecs, dtype=np.float64).reshape(3, 3)
        self._gvecs_full = np.array(gvecs, dtype=np.float64).reshape(3, 3)
        r = self._rvecs_full.ravel()
        if nvec == 0:
            vol = 0.0
        elif nvec == 1:
            vol = np.sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])
        elif nvec == 2:
            tmp = r[0] * r[3] + r[1] * r[4] + r[2] * r[5]
            tmp = (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) * \
                  (r[3] * r[3] + r[4] * r[4] + r[5] * r[5]) - tmp * tmp
            vol = np.sqrt(tmp) if tmp > 0 else 0.0
        else:
            vol = abs(r[0] * (r[4] * r[8] - r[5] * r[7])
                      + r[1] * (r[5] * r[6] - r[3] * r[8])
                      + r[2] * (r[3] * r[7] - r[4] * r[6]))
        self._volume = float(vol)

    @property
    def nvec(self):
        return self._nvec

    @property
    def volume(self):
        return self._volume

    def _get_rvecs(self, full=False):
        """Domain vectors; with ``full`` the 3x3 matrix completed by orthonormal directions
        (ext.pyx:84-94)."""
        out = (self._rvecs_full if full else self._rvecs_full[: self._nvec]).copy()
        out.setflags(write=False)
        return out

    def _get_gvecs(self, full=False):
        """Reciprocal vectors (ext.pyx:96-106)."""
        out = (self._gvecs_full if full else self._gvecs_full[: self._nvec]).copy()
        out.setflags(write=False)
        return out

    @property
    def rvecs(self):
        return self._get_rvecs()

    @property
    def gvecs(self):
        return self._get_gvecs()

    @property
    def parameters(self):
        """Lengths of the domain vectors and the angles between them (ext.pyx:107-121)."""
        rvecs = self.rvecs
        tmp = np.dot(rvecs, rvecs.T)
        lengths = np.sqrt(np.diag(tmp))
        tmp = tmp / lengths
        tmp = tmp / lengths.reshape((-1, 1))
        if len(rvecs) < 2:
            cosines = np.array([])
        elif len(rvecs) == 2:
            cosines = np.array([tmp[0, 1]])
        else:
            cosines = np.array([tmp[1, 2], tmp[2, 0], tmp[0, 1]])
        return lengths, np.arccos(np.clip(cosines, -1, 1))


def _import_h5py():
    try:
        import h5py
    except ImportError as exc:
        raise ImportError("HDF5 files need h5py, which is not installed") from exc
    return h5py


class System(object):
    """A micromechanical system (system.py:38-107); all quantities in atomic units."""

    def __init__(self, pos, masses, rvecs, surrounding_cells, surrounding_nodes,
                 boundary_nodes=None, grid=None, types=None, params=None):
        self._pos_source = None
        self.masses = masses
        self.pos = pos
        self.domain = Domain(rvecs)
        self.grid = grid
        self.types = types
        self.params = params
        self.surrounding_cells = surrounding_cells
        self.surrounding_nodes = surrounding_nodes
        self.boundary_nodes = boundary_nodes
        self.nnodes = len(self.surrounding_cells)
        self.ncells = len(self.surrounding_nodes)

    @property
    def pos(self):
        """Node positions [N, 3]. After a device MD run they are fetched from the device on the
        first read (see ``defer_pos``)."""
        source = self._pos_source
        if source is not None:
            self._pos_source = None
            source(self._pos)
        return self._pos

    @pos.setter
    def pos(self, value):
        self._pos_source = None
        self._pos = value

    def defer_pos(self, fill):
        """Mark the positions as held elsewhere: the next read of ``pos`` first calls
        ``fill(array)``, which must write the current positions into the array in place.
        Assigning ``pos`` cancels the pending fill. The device integrator uses this to write its
        positions through to the system (the reference does so every step, verlet.py:147 via
        mmff.py:195-197) without a device-to-host copy per run that nobody reads."""
        self._pos_source = fill

    def update_params(self, new_params, type_=1):
        for key, val in new_params.items():
            self.params["type%i/%s" % (int(type_), key)] = val

    _HDF5_KEYS = ("pos", "masses", "rvecs", "surrounding_cells", "surrounding_nodes",
                  "boundary_nodes", "grid", "types")

    @classmethod
    def from_file(cls, fn, **user_kwargs):
        """Construct a system from a CHK or HDF5 file (system.py:213-259)."""
        if fn.endswith(".h5"):
            h5py = _import_h5py()
            with h5py.File(fn, "r") as f:
                return cls.from_hdf5(f, **user_kwargs)
        if not fn.endswith(".chk"):
            raise IOError('Cannot read from file "%s".' % fn)
        allowed = ["pos", "masses", "rvecs", "surrounding_cells", "surrounding_nodes",
                   "boundary_nodes", "grid", "types"]
        kwargs, params = {}, {}
        for key, value in load_chk(fn).items():
            if key in allowed:
                kwargs[key] = value
            if "type" in key and key != "types":
                params[key] = value
        kwargs["params"] = params
        kwargs.update(user_kwargs)
        return cls(**kwargs)

    @classmethod
    def from_dict(cls, output):
        """Construct a system from the dict returned by :func:`topology.build_system`."""
        params = {k: v for k, v in output.items() if k.startswith("type") and k != "types"}
        return cls(
            pos=np.array(output["pos"], dtype=np.float64),
            masses=np.array(output["masses"], dtype=np.float64),
            rvecs=np.array(output["rvecs"], dtype=np.float64),
            surrounding_cells=np.asarray(output["surrounding_cells"]),
            surrounding_nodes=np.asarray(output["surrounding_nodes"]),
            boundary_nodes=np.asarray(output["boundary_nodes"]),
            grid=np.asarray(output["grid"]),
            types=np.asarray(output["types"]),
            params=params,
        )

    @classmethod
    def from_hdf5(cls, f, **user_kwargs):
        """Construct a system from the ``system`` group of an open HDF5 file (system.py:261-296).
        Type parameters stored as ``params/<name>`` datasets are read back as well."""
        sgrp = f["system"]
        kwargs = {"pos": np.asarray(sgrp["pos"][:])}
        for key in cls._HDF5_KEYS:
            if key in sgrp:
                kwargs[key] = np.asarray(sgrp[key][:])
        if "params" in sgrp:
            pgrp = sgrp["params"]
            kwargs["params"] = {k.replace(":", "/"): np.asarray(pgrp[k][()])
                                for k in pgrp.keys()}
        kwargs.update(user_kwargs)
        for key in ("surrounding_cells", "surrounding_nodes"):
            kwargs.setdefault(key, np.zeros((0, 8), dtype=int))
        kwargs.setdefault("masses", None)
        kwargs.setdefault("rvecs", np.zeros((0, 3)))
        return cls(**kwargs)

    def to_hdf5(self, f):
        """Write the system to the ``system`` group of an open HDF5 file (system.py:344-370).

        As the reference: positions and masses, and a ValueError if the file already holds a
        system. In addition the domain, topology and type parameters (under ``params``, with
        ':' for '/') are stored, so that ``from_hdf5`` gives back the whole system."""
        if "system" in f:
            raise ValueError("The HDF5 file already contains a system description.")
        sgrp = f.create_group("system")
        sgrp.create_dataset("pos", data=np.asarray(self.pos))
        if self.masses is not None:
            sgrp.create_dataset("masses", data=np.asarray(self.masses))
        for key in self._HDF5_KEYS[2:]:
            value = np.asarray(self.domain.rvecs) if key == "rvecs" else getattr(self, key)
            if value is not None:
                sgrp.create_dataset(key, data=np.asarray(value))
        if self.params:
            pgrp = sgrp.create_group("params")
            for k, v in self.params.items():
                pgrp.create_dataset(k.replace("/", ":"), data=np.asarray(v))

    def to_file(self, fn):
        """Write the system to a CHK, HDF5 or XYZ file (system.py:298-342)."""
        if fn.endswith(".h5"):
            h5py = _import_h5py()
            with h5py.File(fn, "w") as f:
                self.to_hdf5(f)
            return
        if fn.endswith(".xyz"):
            from .sampling.trajectory import write_xyz_frame
            with open(fn, "w") as fh:
                write_xyz_frame(fh, ["Cs"] * len(self.pos), str(self), self.pos)
            return
        if not fn.endswith(".chk"):
            raise NotImplementedError(
                "The extension of %s does not correspond to any known format." % fn)
        output = {
            "pos": np.asarray(self.pos),
            "masses": np.asarray(self.masses),
            "rvecs": np.asarray(self.domain.rvecs),
            "surrounding_cells": np.asarray(self.surrounding_cells),
            "surrounding_nodes": np.asarray(self.surrounding_nodes),
            "boundary_nodes": np.asarray(self.boundary_nodes),
            "grid": np.asarray(self.grid),
            "types": np.asarray(self.types),
        }
        output.update(self.params)
        dump_chk(fn, output)
