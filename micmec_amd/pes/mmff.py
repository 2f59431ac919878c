"""Force-field plugin API and the micromechanical force part on MI355X.

Drop-in for micmec/pes/mmff.py:39-323:

* ``ForcePart`` keeps the reference contract of ``compute(gpos=None, vtens=None)``: private
  buffers are zeroed, ``_internal_compute`` fills them, a NaN raises ``ValueError`` and the
  results are then *added* to the caller's arrays (mmff.py:87-149). The arrays may be numpy
  arrays (host) or torch tensors (device); the arithmetic of the mechanical part always runs
  in the HIP kernels of ``libmicmec_amd.so``.
* ``MicMecForceField`` sums its parts (mmff.py:156-201).
* ``ForcePartMechanical`` evaluates the nanocell energy, node gradient and virial
  (mmff.py:204-403) on the GPU. ``model="corner8"`` is the eight-corner model of
  micmec/pes/nanocell.py; ``model="original"`` the averaged-matrix model that mmff.py:27-30
  imports as shipped.
* ``ForcePartPressure`` adds the external-pressure term P.V (the Yaff force part that MicMec's
  NPT runs combine with the MTK barostat).
"""

import ctypes

import numpy as np
import torch

from .. import units
from .. import _native
from ..topology import Grid, periodic_wrap_bits

__all__ = ["ForcePart", "MicMecForceField", "ForcePartMechanical", "ForcePartPressure",
           "parse_type_params"]


def _is_torch(x):
    return isinstance(x, torch.Tensor)


def _zero(buf):
    if _is_torch(buf):
        buf.zero_()
    else:
        buf[:] = 0.0


def _any_nan(buf):
    if _is_torch(buf):
        return bool(torch.isnan(buf).any().item())
    return bool(np.isnan(buf).any())


def _private_like(cache, arr):
    """A private buffer of the caller's kind (numpy or torch on the caller's device)."""
    if _is_torch(arr):
        if cache is None or not _is_torch(cache) or cache.device != arr.device or \
                cache.shape != arr.shape:
            return torch.zeros(arr.shape, dtype=torch.float64, device=arr.device)
        return cache
    if cache is None or _is_torch(cache) or cache.shape != np.shape(arr):
        return np.zeros(np.shape(arr), float)
    return cache


class ForcePart(object):
    """Base class of anything that computes an energy, and optionally the node gradient and the
    virial, for a ``System`` (mmff.py:39-153)."""

    def __init__(self, name, system):
        self.name = name
        self.energy = 0.0
        self._stale = set()
        self.gpos = np.zeros((system.nnodes, 3), float)
        self.vtens = np.zeros((3, 3), float)
        self.clear()

    def clear(self):
        """Mark the cached results invalid (mmff.py:61-65): the energy becomes NaN at once, the
        cached ``gpos`` and ``vtens`` read as NaN from then on. The arrays are filled on their
        next read rather than here, so that invalidating a large system's cache every MD batch
        costs nothing unless the cache is looked at."""
        self.energy = np.nan
        self._stale = {"gpos", "vtens"}

    def _cache(name):
        key = "_" + name

        def get(self):
            buf = self.__dict__[key]
            if name in self._stale:
                self._stale.discard(name)
                if _is_torch(buf):
                    buf.fill_(float("nan"))
                else:
                    buf[...] = np.nan
            return buf

        def put(self, value):
            self._stale.discard(name)
            self.__dict__[key] = value
        return property(get, put)

    gpos = _cache("gpos")
    vtens = _cache("vtens")
    del _cache

    def update_rvecs(self, rvecs):
        self.clear()

    def update_pos(self, pos):
        self.clear()

    def compute(self, gpos=None, vtens=None):
        """Compute the energy; optionally add the gradient and virial to ``gpos``/``vtens``.

        Same contract as mmff.py:87-149: outputs are Fortran-style, results are *added*, and a
        not-a-number energy, gradient or virial raises ``ValueError``.
        """
        if gpos is None:
            my_gpos = None
        else:
            my_gpos = self.gpos = _private_like(self._gpos, gpos)
            _zero(my_gpos)
        if vtens is None:
            my_vtens = None
        else:
            my_vtens = self.vtens = _private_like(self._vtens, vtens)
            _zero(my_vtens)
        self.energy = self._internal_compute(my_gpos, my_vtens)
        if np.isnan(self.energy):
            raise ValueError("The energy is not-a-number (``nan``).")
        if gpos is not None:
            if _any_nan(my_gpos):
                raise ValueError("Some ``gpos`` element(s) is/are not-a-number (``nan``).")
            gpos += my_gpos
        if vtens is not None:
            if _any_nan(my_vtens):
                raise ValueError("Some ``vtens`` element(s) is/are not-a-number (``nan``).")
            vtens += my_vtens
        return self.energy

    def _internal_compute(self, gpos, vtens):
        raise NotImplementedError


class MicMecForceField(ForcePart):
    """A complete micromechanical force field: the sum of its parts (mmff.py:156-201)."""

    def __init__(self, system, parts):
        ForcePart.__init__(self, "all", system)
        self.system = system
        self.parts = []
        for part in parts:
            self.add_part(part)

    def add_part(self, part):
        self.parts.append(part)
        name = "part_%s" % part.name
        if name in self.__dict__:
            raise ValueError("The part %s occurs twice in the micromechanical force field." % name)
        self.__dict__[name] = part

    def update_rvecs(self, rvecs):
        ForcePart.update_rvecs(self, rvecs)
        self.system.domain.update_rvecs(rvecs)

    def update_pos(self, pos):
        ForcePart.update_pos(self, pos)
        if _is_torch(pos) and not _is_torch(self.system.pos):
            self.system.pos[:] = pos.detach().cpu().numpy()
        else:
            self.system.pos[:] = pos

    def _internal_compute(self, gpos, vtens):
        return sum([part.compute(gpos, vtens) for part in self.parts])


def parse_type_params(params, types):
    """Per-type parameter tables from the ``type<t>/<key>`` dict (mmff.py:218-231).

    Returns ``(type_ids, nstates, h0 [T,S,3,3], c0 [T,S,3,3,3,3], free [T,S], temp [T])``
    for the types that occur in ``types``, rows in increasing type id; unused state slots are
    filled with the identity cell and zeros.
    """
    p1, p2, p3, p4 = {}, {}, {}, {}
    for key, item in params.items():
        if not key.startswith("type") or "/" not in key:
            continue
        try:
            t = int(key.split("/")[0][4:])
        except ValueError:
            continue
        if "effective_temp" in key:
            p1[t] = item
        if "free_energy" in key:
            p2[t] = item
        if "cell" in key:
            p3[t] = item
        if "elasticity" in key:
            p4[t] = item
    type_ids = sorted(set(int(round(float(t))) for t in np.asarray(types).ravel()))
    ns = []
    for t in type_ids:
        if t not in p3 or t not in p4 or t not in p2 or t not in p1:
            raise ValueError("Parameters of cell type %i are incomplete." % t)
        h = np.asarray(p3[t], dtype=np.float64).reshape(-1, 3, 3)
        c = np.asarray(p4[t], dtype=np.float64).reshape(-1, 3, 3, 3, 3)
        f = np.atleast_1d(np.asarray(p2[t], dtype=np.float64))
        if not (len(h) == len(c) == len(f)):
            raise ValueError("Cell type %i: cell, elasticity and free_energy disagree on the "
                             "number of states." % t)
        ns.append(len(h))
    T, S = len(type_ids), max(ns) if ns else 1
    h0 = np.tile(np.identity(3), (T, S, 1, 1))
    c0 = np.zeros((T, S, 3, 3, 3, 3))
    free = np.zeros((T, S))
    temp = np.zeros(T)
    for r, t in enumerate(type_ids):
        h0[r, : ns[r]] = np.asarray(p3[t], dtype=np.float64).reshape(-1, 3, 3)
        c0[r, : ns[r]] = np.asarray(p4[t], dtype=np.float64).reshape(-1, 3, 3, 3, 3)
        free[r, : ns[r]] = np.atleast_1d(np.asarray(p2[t], dtype=np.float64))
        temp[r] = float(p1[t])
    return type_ids, np.array(ns, dtype=np.int32), h0, c0, free, temp


class DeviceField(object):
    """Owner of one ``mm_field`` object of the HIP library (device topology and parameters)."""

    def __init__(self, surrounding_nodes, surrounding_cells, wrap_bits, cell_row, nstates, h0,
                 c0, free, temp, model="corner8", device=None):
        if model not in _native.MODELS:
            raise ValueError("Unknown model %r; choose one of %s" % (model, list(_native.MODELS)))
        L = _native.lib()
        self.device = torch.device(device if device is not None else "cuda")
        dev_index = self.device.index if self.device.index is not None else \
            torch.cuda.current_device()
        self.device = torch.device("cuda", dev_index)
        sn = np.ascontiguousarray(surrounding_nodes, dtype=np.int64).reshape(-1, 8)
        sc = np.ascontiguousarray(surrounding_cells, dtype=np.int64).reshape(-1, 8)
        wb = np.ascontiguousarray(wrap_bits, dtype=np.uint8)
        cr = np.ascontiguousarray(cell_row, dtype=np.int32)
        ns = np.ascontiguousarray(nstates, dtype=np.int32)
        h0 = np.ascontiguousarray(h0, dtype=np.float64)
        c0 = np.ascontiguousarray(c0, dtype=np.float64)
        free = np.ascontiguousarray(free, dtype=np.float64)
        temp = np.ascontiguousarray(temp, dtype=np.float64)
        self.ncells, self.nnodes = len(sn), len(sc)
        self.model = model
        handle = ctypes.c_void_p()
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        rc = L.mm_field_create(ctypes.byref(handle), dev_index, _native.MODELS[model],
                               self.ncells, self.nnodes, vp(sn), vp(sc), vp(wb), vp(cr),
                               int(len(ns)), int(h0.shape[1]), vp(ns), vp(h0), vp(c0), vp(free),
                               vp(temp), float(units.boltzmann))
        _native.check(rc, "mm_field_create")
        self._handle = handle
        self._lib = L

    def compute(self, pos, rvecs, energy, gpos=None, vtens=None, accumulate=False, stream=None):
        """Launch the force kernels (asynchronous): all arguments are float64 device tensors."""
        rc = self._lib.mm_field_compute(self._handle, _native.ptr(pos), _native.ptr(rvecs),
                                        _native.ptr(energy), _native.ptr(gpos),
                                        _native.ptr(vtens), int(bool(accumulate)),
                                        _native.stream_ptr(stream))
        _native.check(rc, "mm_field_compute")

    def cg_trial(self, x, g, d, beta, set_dir, alpha, rvecs, xa, ga, out, host_out=None,
                 stream=None):
        """One line-search trial: with ``set_dir`` first d <- beta d - g, then xa <- x + alpha d,
        energy and gradient ga at xa; out[:7] = energy, ga.d, ga.ga, ga.g, max_n |ga_n|^2,
        sum_n |xa_n - x_n|^2, max_n |xa_n - x_n|^2. Asynchronous unless ``host_out`` (a ctypes
        array of 7 doubles) is given, which then receives the results."""
        rc = self._lib.mm_cg_trial(self._handle, _native.ptr(x), _native.ptr(g), _native.ptr(d),
                                   float(beta), int(bool(set_dir)), float(alpha),
                                   _native.ptr(rvecs), _native.ptr(xa), _native.ptr(ga),
                                   _native.ptr(out), host_out, _native.stream_ptr(stream))
        _native.check(rc, "mm_cg_trial")

    def cg_stats(self, x, x_old, g, g_old, out, host_out=None, stream=None):
        """out[:5] = sum |g_n|^2, sum g_n . g_old_n, sum |x_n - x_old_n|^2, max |g_n|^2,
        max |x_n - x_old_n|^2 over the nodes; ``host_out`` as for ``cg_trial``."""
        rc = self._lib.mm_cg_stats(self._handle, _native.ptr(x), _native.ptr(x_old),
                                   _native.ptr(g), _native.ptr(g_old), _native.ptr(out),
                                   host_out, _native.stream_ptr(stream))
        _native.check(rc, "mm_cg_stats")

    def check(self, stream=None):
        """Synchronise and raise ``ValueError`` if any compute since the last check saw NaN."""
        flags = ctypes.c_int32(0)
        rc = self._lib.mm_field_check(self._handle, _native.stream_ptr(stream),
                                      ctypes.byref(flags))
        _native.check(rc, "mm_field_check")
        return flags.value

    @property
    def handle(self):
        return self._handle

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                self._lib.mm_field_destroy(h)
            except Exception:
                pass
            self._handle = None


def system_wrap_bits(system):
    """Image-shift bits of every cell of ``system`` (replaces the table of mmff.py:258-286)."""
    nper = np.asarray(system.domain.rvecs).shape[0]
    if nper == 0:
        return np.zeros(system.ncells, dtype=np.uint8)
    grid = np.rint(np.asarray(system.grid)).astype(np.int64)
    g = Grid(grid, pbc=True)
    if min(grid.shape) < 2:
        raise ValueError("Periodic grids need at least two cells along every axis "
                         "(the reference image table is undefined below that).")
    sn, _ = g.surrounding()
    if sn.shape != np.shape(system.surrounding_nodes) or \
            not np.array_equal(sn, np.asarray(system.surrounding_nodes)):
        raise ValueError("surrounding_nodes does not follow the node and cell order of `grid`.")
    return periodic_wrap_bits(g.cell_ijk, grid.shape, [True] * 3)


class ForcePartMechanical(ForcePart):
    """The micromechanical part of the force field, evaluated by HIP kernels (mmff.py:204-323).

    Parameters
    ----------
    system : micmec_amd.system.System
    model : {"corner8", "original"}
        Nanocell energy model; see the module docstring.
    device : torch device, optional
        GPU that holds the topology and parameter tables (default: the current device).
    """

    def __init__(self, system, model="corner8", device=None):
        ForcePart.__init__(self, "micmec", system)
        self.system = system
        self.model = model
        self.pbc = self.get_pbc(np.asarray(system.domain.rvecs))
        wrap = system_wrap_bits(system)
        type_ids, ns, h0, c0, free, temp = parse_type_params(system.params, system.types)
        row_of = {t: r for r, t in enumerate(type_ids)}
        types = np.rint(np.asarray(system.types, dtype=np.float64)).astype(np.int64)
        cell_row = np.array([row_of[int(t)] for t in types], dtype=np.int32) if len(types) \
            else np.zeros(0, np.int32)
        self.type_ids = type_ids
        self.field = DeviceField(system.surrounding_nodes, system.surrounding_cells, wrap,
                                 cell_row, ns, h0, c0, free, temp, model=model, device=device)
        self.device = self.field.device
        N = system.nnodes
        kw = dict(dtype=torch.float64, device=self.device)
        self._pos_d = torch.zeros((N, 3), **kw)
        self._gpos_d = torch.zeros((N, 3), **kw)
        self._vtens_d = torch.zeros((3, 3), **kw)
        self._energy_d = torch.zeros(1, **kw)
        self._rvecs_d = torch.zeros((3, 3), **kw)

    @staticmethod
    def get_pbc(rvecs):
        nper = rvecs.shape[0]
        pbc = nper > 0
        if pbc and nper != 3:
            raise ValueError("Attribute `rvecs` only supports finite systems or 3D periodic "
                             "systems, not %iD periodic systems." % nper)
        return pbc

    def rvecs_device(self, rvecs=None):
        rv = np.zeros((3, 3))
        r = np.asarray(self.system.domain.rvecs if rvecs is None else rvecs, dtype=np.float64)
        if r.size:
            rv[: len(r)] = r
        self._rvecs_d.copy_(torch.from_numpy(rv))
        return self._rvecs_d

    def _internal_compute(self, gpos, vtens):
        pos = self.system.pos
        if _is_torch(pos):
            pos_d = pos.to(device=self.device, dtype=torch.float64).contiguous()
        else:
            self._pos_d.copy_(torch.from_numpy(np.ascontiguousarray(pos, dtype=np.float64)))
            pos_d = self._pos_d
        rv = self.rvecs_device()
        want_g = gpos is not None
        want_v = vtens is not None
        g_d = (gpos if (_is_torch(gpos) and gpos.device == self.device and gpos.is_contiguous())
               else self._gpos_d) if want_g else None
        v_d = (vtens if (_is_torch(vtens) and vtens.device == self.device and
                         vtens.is_contiguous()) else self._vtens_d) if want_v else None
        self.field.compute(pos_d, rv, self._energy_d, g_d, v_d, accumulate=False)
        energy = float(self._energy_d.item())
        if want_g and g_d is not gpos:
            if _is_torch(gpos):
                gpos.copy_(g_d)
            else:
                gpos[:] = g_d.cpu().numpy()
        if want_v and v_d is not vtens:
            if _is_torch(vtens):
                vtens.copy_(v_d)
            else:
                vtens[:] = v_d.cpu().numpy()
        return energy

    def compute_device(self, pos, rvecs, energy, gpos=None, vtens=None, accumulate=False,
                       stream=None):
        """Asynchronous device entry point: float64 device tensors in and out, no host sync."""
        self.field.compute(pos, rvecs, energy, gpos, vtens, accumulate, stream)


class ForcePartPressure(ForcePart):
    """External pressure term: energy P.V and virial P.V.I for a 3D periodic domain.

    MicMec's NPT runs add it to the force field next to ``ForcePartMechanical``. Used together
    with an MTK barostat at pressure 0 it is equivalent to the barostat at pressure P
    (micmec/sampling/npt.py:665-672, 645-648); never combine it with a barostat at P.
    """

    def __init__(self, system, pext):
        ForcePart.__init__(self, "press", system)
        self.system = system
        self.pext = pext

    def _internal_compute(self, gpos, vtens):
        domain = self.system.domain
        if domain.nvec != 3:
            raise ValueError("The pressure term needs a 3D periodic domain.")
        if vtens is not None:
            pv = self.pext * domain.volume
            if _is_torch(vtens):
                vtens += pv * torch.eye(3, dtype=vtens.dtype, device=vtens.device)
            else:
                vtens += pv * np.identity(3)
        return domain.volume * self.pext
