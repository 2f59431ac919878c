"""Batched form: many independent systems of uneven size integrated together on one GPU.

``BatchVerlet`` runs NVE or NVT (Nose-Hoover chain) velocity Verlet for a list of systems in a
single device engine: the systems' cells and nodes are concatenated into one batched force-field
object (``mm_batch_field_create``), every kernel of a step covers all systems in one launch, each
system keeps its own scalar block (domain, chain, energies, conserved-quantity tracking), and the
step is captured once in a HIP graph and replayed. Small systems that cannot fill the GPU on
their own (a 4^3 grid is 64 cells) thereby share it.

Each system is initialised exactly as a lone ``VerletIntegrator`` with an ``NHCThermostat`` would
be when numpy's global random stream is seeded with that system's seed: Maxwell-Boltzmann
velocities (micmec/sampling/utils.py:28-65), momentum cleaning (142-174) in the integrator and in
the thermostat's ``init``, then the chain's random velocities (micmec/sampling/nvt.py:402-407).
The reference has no batched form; per-system results equal one-at-a-time runs.
"""

import ctypes

import numpy as np
import torch

from . import _layout as L
from . import _native, units
from .pes.mmff import parse_type_params, system_wrap_bits
from .sampling.utils import clean_momenta, get_ndof_internal_md, get_random_vel

__all__ = ["BatchVerlet", "NODE_ALIGN", "concat_systems"]

NODE_ALIGN = 256   # node block of the device engine; each system starts on a block boundary

_P = ctypes.c_void_p
_i32, _i64, _f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
_native.register_signature("mm_batch_field_create",
                           [ctypes.POINTER(_P), ctypes.c_int, ctypes.c_int, _i32, _i64, _i64, _P,
                            _P, _P, _P, _P, _i32, _i32, _P, _P, _P, _P, _P, _f64])
_native.register_signature("mm_batch_compute", [_P, _P, _P, _P, _P, _P, _P])
_native.register_signature("mm_md_create_batch", [ctypes.POINTER(_P), _P, _i32, _P, _P, _P, _P,
                                                  _P, _P, _P, ctypes.c_int, ctypes.c_int])


def _type_table(systems):
    """One parameter table for all systems; a type id must mean the same parameters in each."""
    ref = {}
    for s in systems:
        ids, ns, h0, c0, free, temp = parse_type_params(s.params, s.types)
        for r, t in enumerate(ids):
            entry = (ns[r], h0[r, : ns[r]], c0[r, : ns[r]], free[r, : ns[r]], temp[r])
            if t in ref:
                o = ref[t]
                same = o[0] == entry[0] and all(np.array_equal(a, b) for a, b in
                                                zip(o[1:], entry[1:]))
                if not same:
                    raise ValueError("type %i has different parameters in different systems" % t)
            else:
                ref[t] = entry
    ids = sorted(ref)
    S = max(ref[t][0] for t in ids)
    T = len(ids)
    h0 = np.tile(np.identity(3), (T, S, 1, 1))
    c0 = np.zeros((T, S, 3, 3, 3, 3))
    free = np.zeros((T, S))
    temp = np.zeros(T)
    ns = np.zeros(T, dtype=np.int32)
    for r, t in enumerate(ids):
        n, h, c, f, tt = ref[t]
        ns[r] = n
        h0[r, :n], c0[r, :n], free[r, :n], temp[r] = h, c, f, tt
    return ids, ns, h0, c0, free, temp


def concat_systems(systems, node_align=1):
    """Concatenated topology of several systems for one batched device object.

    System k's nodes start at ``node_start[k]`` (a multiple of ``node_align``; the gap up to the
    next system holds padding nodes without cells), its cells follow those of system k - 1.
    Returns a dict with ``surrounding_nodes``, ``surrounding_cells``, ``wrap_bits``,
    ``cell_row``, ``cell_sys``, ``node_start`` [nsys + 1], ``ncells``, ``nnodes`` and the shared
    parameter table ``types`` = (ids, nstates, h0, c0, free, temp) of :func:`_type_table`."""
    ids, ns, h0, c0, free, temp = _type_table(systems)
    row_of = {t: r for r, t in enumerate(ids)}
    starts = [0]
    for s in systems:
        starts.append(starts[-1] + -(-s.nnodes // node_align) * node_align)
    sn_all, sc_all, wb_all, row_all, sys_all = [], [], [], [], []
    cell_off = 0
    for k, s in enumerate(systems):
        sn = np.asarray(s.surrounding_nodes, dtype=np.int64) + starts[k]
        sc = np.asarray(s.surrounding_cells, dtype=np.int64)
        sc = np.where(sc >= 0, sc + cell_off, -1)
        pad = np.full((starts[k + 1] - starts[k] - s.nnodes, 8), -1, dtype=np.int64)
        sn_all.append(sn)
        sc_all.extend([sc, pad])
        wb_all.append(system_wrap_bits(s))
        types = np.rint(np.asarray(s.types, dtype=float)).astype(int)
        row_all.append(np.array([row_of[int(t)] for t in types], dtype=np.int32))
        sys_all.append(np.full(s.ncells, k, dtype=np.int32))
        cell_off += s.ncells
    sn = np.ascontiguousarray(np.concatenate(sn_all))
    return dict(
        surrounding_nodes=sn,
        surrounding_cells=np.ascontiguousarray(np.concatenate(sc_all)),
        wrap_bits=np.ascontiguousarray(np.concatenate(wb_all), dtype=np.uint8),
        cell_row=np.ascontiguousarray(np.concatenate(row_all)),
        cell_sys=np.ascontiguousarray(np.concatenate(sys_all)),
        node_start=np.array(starts, dtype=np.int64),
        ncells=len(sn), nnodes=int(starts[-1]),
        types=(ids, np.ascontiguousarray(ns, dtype=np.int32), np.ascontiguousarray(h0),
               np.ascontiguousarray(c0), np.ascontiguousarray(free),
               np.ascontiguousarray(temp)))


class BatchVerlet(object):
    """NVE / NVT-NHC velocity Verlet of many systems in one device engine.

    Parameters
    ----------
    systems : list of micmec_amd.system.System
    timestep : float
    temps : sequence of float
        Initial (and thermostat) temperature of every system.
    seeds : sequence of int
        numpy legacy-stream seed of every system's random initial velocities.
    thermostat : bool
        Nose-Hoover chain (length ``chainlength``, time constant ``timecon``) on every system.
    """

    def __init__(self, systems, timestep, temps, seeds, model="corner8", thermostat=True,
                 timecon=100 * units.femtosecond, chainlength=3, device=None, use_graph=True):
        self.systems = list(systems)
        self.nsys = n = len(self.systems)
        if n < 1:
            raise ValueError("need at least one system")
        self.timestep = timestep
        self.temps = [float(t) for t in temps]
        self.use_graph = use_graph
        self.thermostat = thermostat
        self.device = torch.device(device if device is not None else "cuda")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        # ---------------------------------------------------------- concatenated topology
        topo = concat_systems(self.systems, NODE_ALIGN)
        starts = topo["node_start"]
        self.node_start = starts
        Ntot = topo["nnodes"]
        self.ncells_total = topo["ncells"]
        self.lib = _native.lib()
        handle = ctypes.c_void_p()
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        ids, ns, h0, c0, free, temp = topo["types"]
        rc = self.lib.mm_batch_field_create(
            ctypes.byref(handle), self.device.index, _native.MODELS[model], n, topo["ncells"],
            Ntot, vp(topo["surrounding_nodes"]), vp(topo["surrounding_cells"]),
            vp(topo["wrap_bits"]), vp(topo["cell_row"]), vp(topo["cell_sys"]), len(ids),
            int(h0.shape[1]), vp(ns), vp(h0), vp(c0), vp(free), vp(temp), float(units.boltzmann))
        _native.check(rc, "mm_batch_field_create")
        self.field = handle
        # ---------------------------------------------------------- host initialisation
        pos = np.zeros((Ntot, 3))
        vel = np.zeros((Ntot, 3))
        masses = np.ones(Ntot)
        S = np.zeros((n, L.S_COUNT))
        self.chains = []
        for k, s in enumerate(self.systems):
            a, N = starts[k], s.nnodes
            p = np.array(s.pos, dtype=np.float64)
            m = np.array(s.masses, dtype=np.float64)
            np.random.seed(int(seeds[k]))
            v = get_random_vel(self.temps[k], True, m)
            clean_momenta(p, v, m, s.domain)
            ndof = p.size
            Sk = S[k]
            if thermostat:
                clean_momenta(p, v, m, s.domain)
                ndof = get_ndof_internal_md(N, s.domain.nvec)
                afreq = 2 * np.pi / timecon
                cm = np.ones(chainlength) * (units.boltzmann * self.temps[k] / afreq ** 2)
                cm[0] *= ndof
                cv = np.random.normal(0, np.sqrt(cm * units.boltzmann * self.temps[k]),
                                      chainlength) / cm
                Sk[L.S_THERMO] = L.THERMO_NHC
                Sk[L.S_CHAIN_LEN] = chainlength
                Sk[L.S_CHAIN_NDOF] = ndof
                Sk[L.S_THERMO_TEMP] = self.temps[k]
                Sk[L.S_CHAIN_MASS:L.S_CHAIN_MASS + chainlength] = cm
                Sk[L.S_CHAIN_VEL:L.S_CHAIN_VEL + chainlength] = cv
            pos[a:a + N] = p
            vel[a:a + N] = v
            masses[a:a + N] = m
            rv = np.zeros((3, 3))
            r = np.asarray(s.domain.rvecs)
            rv[: len(r)] = r
            Sk[L.S_TIMESTEP] = timestep
            Sk[L.S_NDOF] = ndof
            Sk[L.S_BOLTZMANN] = units.boltzmann
            Sk[L.S_NVEC] = s.domain.nvec
            Sk[L.S_RVECS:L.S_RVECS + 9] = rv.ravel()
            Sk[L.S_VOLUME] = s.domain.volume
            Sk[L.S_NNODES] = N
        kw = dict(dtype=torch.float64, device=self.device)
        self.pos = torch.tensor(pos, **kw)
        self.vel = torch.tensor(vel, **kw)
        self.gpos = torch.zeros((Ntot, 3), **kw)
        self.masses = torch.tensor(masses, **kw)
        self.posold = self.pos.clone()
        self.S = torch.tensor(S, **kw)
        md = ctypes.c_void_p()
        rc = self.lib.mm_md_create_batch(
            ctypes.byref(md), self.field, n, vp(self.node_start), _native.ptr(self.pos),
            _native.ptr(self.vel), _native.ptr(self.gpos), _native.ptr(self.masses),
            _native.ptr(self.posold), _native.ptr(self.S),
            L.THERMO_NHC if thermostat else L.THERMO_NONE, L.BARO_NONE)
        _native.check(rc, "mm_md_create_batch")
        self.md = md
        self.counter = 0
        self._call("mm_md_init_force", 0)
        self._call("mm_md_init_properties")
        self._check()

    def _stream(self):
        return _native.stream_ptr(torch.cuda.current_stream(self.device))

    def _call(self, name, *args):
        _native.check(getattr(self.lib, name)(self.md, *args, self._stream()), name)

    def _check(self):
        self._S = self.S.cpu().numpy()
        bad = np.nonzero(self._S[:, L.S_NAN] != 0)[0]
        if len(bad):
            raise ValueError("not-a-number in systems %s" % list(bad))

    def run(self, nsteps):
        """Advance every system by nsteps (one graph replay per step)."""
        self._call("mm_md_step", int(nsteps), int(bool(self.use_graph)))
        self._call("mm_md_sync")
        self.counter += nsteps
        self._check()

    def scalars(self, name):
        """Per-system value of one scalar (e.g. ``"EPOT"``, ``"EKIN"``, ``"ECONS"``, ``"TEMP"``)."""
        return self._S[:, getattr(L, "S_" + name)].copy()

    def positions(self, k):
        a = int(self.node_start[k])
        return self.pos[a:a + self.systems[k].nnodes].cpu().numpy()

    def velocities(self, k):
        a = int(self.node_start[k])
        return self.vel[a:a + self.systems[k].nnodes].cpu().numpy()

    @property
    def grid_cells(self):
        """Grid cells of every system (empty cells included), the unit of cell-steps."""
        return np.array([int(np.prod(np.shape(s.grid))) for s in self.systems])

    def __del__(self):
        for name, fn in (("md", "mm_md_destroy"), ("field", "mm_field_destroy")):
            h = getattr(self, name, None)
            if h is not None and h.value:
                try:
                    torch.cuda.synchronize(self.device)
                    getattr(self.lib, fn)(h)
                except Exception:
                    pass
                setattr(self, name, None)
