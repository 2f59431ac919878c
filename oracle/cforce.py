"""ctypes front end of the C restatement ``oracle/c/mm_oracle.c`` (test infrastructure only).

``compute_c`` has the signature and results of ``oracle.mmff.compute_vectorised`` and runs the
same reference arithmetic (micmec/pes/mmff.py:299-403, nanocell.py:40-132,
nanocell_original.py:38-132) cell by cell in C on all host cores (OpenMP). Build it with
``make -C oracle/c``; the library is ``oracle/libmm_oracle.so``.
"""

import ctypes
import os

import numpy as np

from .constants import boltzmann
from .mmff import parse_params

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmm_oracle.so")
_lib = None


def available():
    return os.path.exists(_PATH)


def _load():
    global _lib
    if _lib is None:
        if not available():
            raise ImportError("oracle/libmm_oracle.so is not built (make -C oracle/c)")
        L = ctypes.CDLL(_PATH)
        P, i64, i32, f64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
        L.mmo_compute.argtypes = [ctypes.c_int, i64, i64, P, P, ctypes.c_int, P, P, P, P,
                                  ctypes.c_int, ctypes.c_int, P, P, P, P, P, f64, P, P, P]
        L.mmo_compute.restype = ctypes.c_int
        L.mmo_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def num_threads():
    return int(_load().mmo_num_threads())


class ParamTable(object):
    """Per-type parameter rows of the ``type<t>/<key>`` dict (mmff.py:218-231)."""

    def __init__(self, params, types):
        p1, p2, p3, p4 = parse_params(params)
        ids = sorted(set(int(round(float(t))) for t in np.asarray(types).ravel()))
        smax = max([len(np.atleast_1d(p2[t])) for t in ids] + [1])
        T = max(len(ids), 1)
        self.h0 = np.tile(np.identity(3).ravel(), (T, smax, 1))
        self.c0 = np.zeros((T, smax, 81))
        self.free = np.zeros((T, smax))
        self.temp = np.ones(T)
        self.nstates = np.ones(T, dtype=np.int32)
        for r, t in enumerate(ids):
            f = np.atleast_1d(np.asarray(p2[t], dtype=float))
            S = len(f)
            self.nstates[r] = S
            self.h0[r, :S] = np.asarray(p3[t], dtype=float).reshape(S, 9)
            self.c0[r, :S] = np.asarray(p4[t], dtype=float).reshape(S, 81)
            self.free[r, :S] = f
            self.temp[r] = float(p1[t])
        row_of = {t: r for r, t in enumerate(ids)}
        self.cell_row = np.array([row_of[int(round(float(t)))] for t in np.asarray(types)],
                                 dtype=np.int32)
        self.smax = smax


def compute_c(pos, rvecs, types, surrounding_nodes, surrounding_cells, wrap_bits, params,
              model="corner8", table=None):
    """(energy, gpos [N, 3], vtens [3, 3]); ``table`` may hold a prebuilt ``ParamTable``."""
    L = _load()
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    rv = np.ascontiguousarray(np.asarray(rvecs, dtype=np.float64).reshape(-1, 3))
    sn = np.ascontiguousarray(surrounding_nodes, dtype=np.int64).reshape(-1, 8)
    sc = np.ascontiguousarray(surrounding_cells, dtype=np.int64).reshape(-1, 8)
    wb = np.ascontiguousarray(wrap_bits, dtype=np.uint8)
    tab = table if table is not None else ParamTable(params, types)
    N, C = len(pos), len(sn)
    gpos = np.zeros((N, 3))
    vtens = np.zeros(9)
    energy = ctypes.c_double(0.0)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = L.mmo_compute({"corner8": 0, "original": 1}[model], C, N, vp(pos), vp(rv), len(rv),
                       vp(sn), vp(sc), vp(wb), vp(tab.cell_row), len(tab.nstates), tab.smax,
                       vp(tab.nstates), vp(tab.h0), vp(tab.c0), vp(tab.free), vp(tab.temp),
                       boltzmann, ctypes.byref(energy), vp(gpos), vp(vtens))
    if rc != 0:
        raise MemoryError("mmo_compute: out of memory")
    return energy.value, gpos, vtens.reshape(3, 3)


def make_force(system, model="corner8", wrap_bits=None):
    """``force(pos, rvecs) -> (energy, gpos, vtens)`` for ``OracleMD`` on a fixed topology."""
    from micmec_amd.topology import Grid
    if wrap_bits is None:
        rv = np.asarray(system.domain.rvecs)
        wrap_bits = Grid(np.rint(system.grid).astype(int), pbc=len(rv) > 0).wrap_bits()
    tab = ParamTable(system.params, system.types)
    sn, sc = system.surrounding_nodes, system.surrounding_cells

    def force(pos, rvecs):
        return compute_c(pos, rvecs, system.types, sn, sc, wrap_bits, system.params, model,
                         table=tab)
    return force
