"""Velocity Verlet on the GPU (drop-in for micmec/sampling/verlet.py:49-374).

``VerletIntegrator`` keeps the reference constructor, state items, hook protocol and
attributes. Initialisation runs on the host exactly as in the reference (random velocities from
numpy's global stream, momentum cleaning, the hooks' ``init``), the force evaluations included
run on the GPU, and then every step runs in the device engine of ``libmicmec_amd.so``
(``mm_md_step``): kick, drift, force, kick and the thermostat/barostat updates, with the scalar state of
the integrator and its hooks kept in device memory. Steps between two host-hook calls are issued
as one batch of HIP-graph replays; host attributes (``pos``, ``vel``, ``epot``, ``econs``, ...)
are refreshed before every hook call and after every ``run``/``propagate``.

Supported force fields: one ``ForcePartMechanical`` plus at most one ``ForcePartPressure``.
Supported device hooks: ``NHCThermostat``, ``LangevinThermostat``, ``MTKBarostat``,
``LangevinBarostat`` and their ``TBCombination``. Host hooks (screen log, writers, user hooks)
run at their own cadence.
"""

import ctypes
import time

import numpy as np
import torch

from .. import _layout as L
from .. import _native
from ..log import log
from ..units import boltzmann, kelvin
from .iterative import (AttributeStateItem, DomainStateItem, Hook, Iterative, PosStateItem,
                        TemperatureStateItem, VolumeStateItem)
from .utils import clean_momenta, get_random_vel

__all__ = ["VerletIntegrator", "VerletHook", "VerletScreenLog", "ConsErrTracker"]

_P = ctypes.c_void_p
_native.register_signature("mm_md_create", [ctypes.POINTER(_P), _P, _P, _P, _P, _P, _P, _P,
                                            ctypes.c_int, ctypes.c_int])
_native.register_signature("mm_md_init_force", [_P, ctypes.c_int, _P])
_native.register_signature("mm_md_init_properties", [_P, _P])
_native.register_signature("mm_md_step", [_P, ctypes.c_int32, ctypes.c_int, _P])
_native.register_signature("mm_md_sync", [_P, _P])
_native.register_signature("mm_md_snapshot", [_P, _P])
_native.register_signature("mm_md_rmsd", [_P, _P])


class ConsErrTracker(object):
    """Running variances of ekin and econs (verlet.py:275-307); the device keeps the live copy."""

    def __init__(self):
        self.counter = 0
        self.ekin_m = 0.0
        self.ekin_s = 0.0
        self.econs_m = 0.0
        self.econs_s = 0.0

    def update(self, ekin, econs):
        if self.counter == 0:
            self.ekin_m = ekin
            self.econs_m = econs
        else:
            t = ekin - self.ekin_m
            self.ekin_m += t / (self.counter + 1)
            self.ekin_s += t * (ekin - self.ekin_m)
            t = econs - self.econs_m
            self.econs_m += t / (self.counter + 1)
            self.econs_s += t * (econs - self.econs_m)
        self.counter += 1

    def get(self):
        if self.counter > 1:
            return np.sqrt(self.econs_s / self.ekin_s)
        return 0.0


class VerletHook(Hook):
    """Hook with init/pre/post entry points, used for thermostats and barostats
    (verlet.py:310-339). As in the reference the constructor ignores start and step."""

    def __init__(self, start=0, step=1):
        self.econs_correction = 0.0
        Hook.__init__(self, start=0, step=1)

    def __call__(self, iterative):
        pass

    def init(self, iterative):
        raise NotImplementedError

    def pre(self, iterative):
        raise NotImplementedError

    def post(self, iterative):
        raise NotImplementedError


class VerletScreenLog(Hook):
    """Screen logger (verlet.py:342-374); due only when the log level is medium or higher."""

    def __init__(self, start=0, step=1):
        Hook.__init__(self, start, step)
        self.time0 = None

    def expects_call(self, counter):
        return log.do_medium and Hook.expects_call(self, counter)

    def __call__(self, iterative):
        if self.time0 is None:
            self.time0 = time.time()
            log.hline()
            log("counter  Cons.Err.       Temp     d-RMSD     g-RMSD   Walltime")
            log.hline()
        log("%7i %10.5f %10.3f %10.3e %10.3e %10.1f" % (
            iterative.counter, iterative.cons_err, iterative.temp, iterative.rmsd_delta,
            iterative.rmsd_gpos, time.time() - self.time0))


class _DeviceEngine(object):
    """Device buffers and the ``mm_md`` handle of one integrator."""

    def __init__(self, integ, thermo, baro):
        from ..pes.mmff import ForcePartMechanical, ForcePartPressure
        mmf = integ.mmf
        parts = list(getattr(mmf, "parts", [mmf]))
        mech = [p for p in parts if isinstance(p, ForcePartMechanical)]
        press = [p for p in parts if isinstance(p, ForcePartPressure)]
        if len(mech) != 1 or len(press) > 1 or len(mech) + len(press) != len(parts):
            raise NotImplementedError(
                "The device integrator needs exactly one ForcePartMechanical and at most one "
                "ForcePartPressure; got %s" % [type(p).__name__ for p in parts])
        self.fpm = mech[0]
        self.ext_press = float(press[0].pext) if press else 0.0
        self.device = self.fpm.device
        kw = dict(dtype=torch.float64, device=self.device)
        N = mmf.system.nnodes
        self.N = N
        self.pos = torch.tensor(np.ascontiguousarray(integ._pos), **kw)
        self.vel = torch.zeros((N, 3), **kw)
        self.gpos = torch.zeros((N, 3), **kw)
        self.masses = torch.tensor(np.ascontiguousarray(integ.masses), **kw)
        self.posold = self.pos.clone()
        self.S = torch.zeros(L.S_COUNT, **kw)
        self.lib = _native.lib()
        handle = ctypes.c_void_p()
        rc = self.lib.mm_md_create(ctypes.byref(handle), self.fpm.field.handle,
                                   _native.ptr(self.pos), _native.ptr(self.vel),
                                   _native.ptr(self.gpos), _native.ptr(self.masses),
                                   _native.ptr(self.posold), _native.ptr(self.S), int(thermo),
                                   int(baro))
        _native.check(rc, "mm_md_create")
        self.handle = handle

    def stream(self):
        return _native.stream_ptr(torch.cuda.current_stream(self.device))

    def call(self, name, *args):
        _native.check(getattr(self.lib, name)(self.handle, *args, self.stream()), name)

    def put_scalars(self, S):
        self.S.copy_(torch.from_numpy(np.ascontiguousarray(S, dtype=np.float64)))

    def get_scalars(self):
        return self.S.cpu().numpy()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.device)
                self.lib.mm_md_destroy(h)
            except Exception:
                pass
            self.handle = None


class VerletIntegrator(Iterative):
    """Velocity Verlet integrator (verlet.py:49-190) running on the GPU."""

    default_state = [
        AttributeStateItem("counter"),
        AttributeStateItem("time"),
        AttributeStateItem("epot"),
        PosStateItem(),
        AttributeStateItem("vel"),
        AttributeStateItem("rmsd_delta"),
        AttributeStateItem("rmsd_gpos"),
        AttributeStateItem("ekin"),
        TemperatureStateItem(),
        AttributeStateItem("etot"),
        AttributeStateItem("econs"),
        AttributeStateItem("cons_err"),
        AttributeStateItem("ptens"),
        AttributeStateItem("vtens"),
        AttributeStateItem("press"),
        VolumeStateItem(),
        DomainStateItem(),
    ]
    log_name = "VERLET"

    def __init__(self, mmf, timestep=None, state=None, hooks=None, vel0=None,
                 temp0=300 * kelvin, scalevel0=True, time0=None, ndof=None, counter0=None,
                 use_graph=True):
        self.ndof = ndof
        self.hooks = hooks
        self.use_graph = use_graph
        if time0 is None:
            time0 = 0.0
        if counter0 is None:
            counter0 = 0
        self._engine = None
        self._time0 = time0
        self._pos = np.array(mmf.system.pos, dtype=np.float64).copy()
        self.rvecs = np.array(mmf.system.domain.rvecs, dtype=np.float64).copy()
        self.masses = np.array(mmf.system.masses, dtype=np.float64).copy()
        self.timestep = timestep
        self.mmf = mmf
        self._verify_hooks()
        if vel0 is None:
            self._vel = get_random_vel(temp0, scalevel0, self.masses)
            clean_momenta(self._pos, self._vel, self.masses, mmf.system.domain)
        else:
            self._vel = np.array(vel0, dtype=np.float64).copy()
        self._gpos = np.zeros(self._pos.shape)
        self.delta = np.zeros(self._pos.shape)
        self._cons_err_tracker = ConsErrTracker()
        self._S = np.zeros(L.S_COUNT)
        self._host_cache = {}
        Iterative.__init__(self, mmf, state, self.hooks, counter0)

    # ------------------------------------------------------------------ hook bookkeeping
    def _add_default_hooks(self):
        if not any(isinstance(hook, VerletScreenLog) for hook in self.hooks):
            self.hooks.append(VerletScreenLog())

    def _verify_hooks(self):
        """Merge a thermostat and a barostat into one TBCombination (verlet.py:216-272; the
        reference deletes ``hooks[max(index_thermo, index_thermo)]``, which removes the wrong
        hook unless the barostat follows the thermostat; here both are removed)."""
        hooks = self.hooks
        if hooks is None:
            self.hooks = []
            return
        if not hasattr(hooks, "__len__"):
            hooks = [hooks]
        thermo = baro = None
        for hook in hooks:
            if hook.method == "thermostat":
                thermo = hook
            elif hook.method == "barostat":
                baro = hook
        if thermo is not None and baro is not None:
            from .npt import TBCombination
            hooks = [h for h in hooks if h is not thermo and h is not baro]
            hooks.append(TBCombination(thermo, baro))
        self.hooks = hooks

    def _device_kinds(self):
        from .npt import LangevinBarostat, MTKBarostat, TBCombination
        from .nvt import LangevinThermostat, NHCThermostat
        thermo = baro = None
        for hook in self.hooks:
            if isinstance(hook, TBCombination):
                thermo, baro = hook.thermostat, hook.barostat
            elif isinstance(hook, (NHCThermostat, LangevinThermostat)):
                thermo = hook
            elif isinstance(hook, (MTKBarostat, LangevinBarostat)):
                baro = hook
            elif isinstance(hook, VerletHook):
                raise NotImplementedError("%s has no device implementation" %
                                          type(hook).__name__)
        return thermo, baro

    def call_verlet_hooks(self, kind):
        for hook in self.hooks:
            if isinstance(hook, VerletHook) and hook.expects_call(self.counter):
                if kind == "init":
                    hook.init(self)
                else:
                    raise NotImplementedError("pre/post hooks run inside the device engine")

    # ------------------------------------------------------------------ host <-> device
    def _live(self):
        return self._engine is not None and getattr(self, "_uploaded", False)

    def _host(self, name):
        if not self._live():
            return getattr(self, "_" + name)
        key = (name, self.counter, self._S_version)
        if key not in self._host_cache:
            self._host_cache = {k: v for k, v in self._host_cache.items() if k[1:] == key[1:]}
            t = getattr(self._engine, name)
            self._host_cache[key] = t.cpu().numpy()
        return self._host_cache[key]

    pos = property(lambda self: self._host("pos"))
    vel = property(lambda self: self._host("vel"))
    gpos = property(lambda self: self._host("gpos"))

    @pos.setter
    def pos(self, value):
        if self._live():
            self._engine.pos.copy_(torch.as_tensor(np.asarray(value, dtype=np.float64)))
            self._S_version += 1
        else:
            self._pos = value

    @vel.setter
    def vel(self, value):
        if self._live():
            self._engine.vel.copy_(torch.as_tensor(np.asarray(value, dtype=np.float64)))
            self._S_version += 1
        else:
            self._vel = value

    @gpos.setter
    def gpos(self, value):
        if self._live():
            self._engine.gpos.copy_(torch.as_tensor(np.asarray(value, dtype=np.float64)))
            self._S_version += 1
        else:
            self._gpos = value

    def _pull(self):
        self._S = self._engine.get_scalars()
        self._S_version = getattr(self, "_S_version", 0) + 1
        if self._S[L.S_NAN] != 0.0:
            raise ValueError("The energy or the gradient became not-a-number (``nan``).")
        if getattr(self, "_uploaded", False):
            if self._baro is not None:
                self.rvecs = self._S[L.S_RVECS:L.S_RVECS + 9].reshape(3, 3).copy()
                self.mmf.update_rvecs(self.rvecs)
            self._mirror_hooks()

    def _write_through(self):
        """Publish the state after a batch of device steps as the reference does after every
        step (verlet.py:147-151 via mmff.py:191-197): the system's positions follow the
        integrator (copied from the device on their first read, ``System.defer_pos``), the
        domain is already passed on by ``_pull``, and the force-field caches are invalidated,
        keeping the energies of the last evaluation."""
        pos_t = self._engine.pos

        def fill(arr):
            arr[...] = pos_t.cpu().numpy().reshape(np.shape(arr))
        mmf = self.mmf
        parts = list(getattr(mmf, "parts", []))
        for part in [mmf] + parts:
            part.clear()
        mmf.system.defer_pos(fill)
        from ..pes.mmff import ForcePartMechanical
        mmf.energy = self.epot
        mech = self._s(L.S_EPOT_MECH)
        for part in parts:
            part.energy = mech if isinstance(part, ForcePartMechanical) else self.epot - mech

    def _s(self, i, n=None):
        if n is None:
            return float(self._S[i])
        return self._S[i:i + n].copy()

    epot = property(lambda self: self._s(L.S_EPOT))
    ekin = property(lambda self: self._s(L.S_EKIN))
    temp = property(lambda self: self._s(L.S_TEMP))
    etot = property(lambda self: self._s(L.S_ETOT))
    econs = property(lambda self: self._s(L.S_ECONS))
    cons_err = property(lambda self: self._s(L.S_CONS_ERR))
    vtens = property(lambda self: self._s(L.S_VTENS, 9).reshape(3, 3))
    ptens = property(lambda self: self._s(L.S_PTENS, 9).reshape(3, 3))
    press = property(lambda self: self._s(L.S_PRESS))
    time = property(lambda self: self._s(L.S_TIME))
    rmsd_gpos = property(lambda self: self._s(L.S_RMSD_GPOS))
    rmsd_delta = property(lambda self: self._s(L.S_RMSD_DELTA))

    def _init_force(self, with_vtens):
        """Force evaluation during initialisation (verlet.py:123-125, npt.py:611-614)."""
        self._engine.call("mm_md_init_force", int(bool(with_vtens)))
        self._pull()

    # ------------------------------------------------------------------ initialisation
    def initialize(self):
        from .npt import LangevinBarostat
        from .nvt import LangevinThermostat, draw_seed
        thermo, baro = self._device_kinds()
        self._thermo, self._baro = thermo, baro
        lt = isinstance(thermo, LangevinThermostat)
        lb = isinstance(baro, LangevinBarostat)
        self._engine = _DeviceEngine(
            self, L.THERMO_NONE if thermo is None else (L.THERMO_LANGEVIN if lt else L.THERMO_NHC),
            L.BARO_NONE if baro is None else (L.BARO_LANGEVIN if lb else L.BARO_MTK))
        self._uploaded = False
        self._S_version = 0
        S = np.zeros(L.S_COUNT)
        S[L.S_TIMESTEP] = self.timestep
        S[L.S_BOLTZMANN] = boltzmann
        S[L.S_NVEC] = self.mmf.system.domain.nvec
        rv = np.zeros((3, 3))
        rv[: len(self.rvecs)] = self.rvecs
        S[L.S_RVECS:L.S_RVECS + 9] = rv.ravel()
        S[L.S_VOLUME] = self.mmf.system.domain.volume
        S[L.S_EXT_PRESS] = self._engine.ext_press
        S[L.S_TIME] = self._time0
        S[L.S_NDOF] = 1.0
        S[L.S_NNODES] = self.mmf.system.nnodes
        self._engine.put_scalars(S)
        self.mmf.update_pos(self._pos)
        self._init_force(with_vtens=False)
        self._gpos = self._engine.gpos.cpu().numpy()
        self.call_verlet_hooks("init")
        if self.ndof is None:
            self.ndof = np.size(self._pos)
        S = self._S.copy()
        S[L.S_NDOF] = self.ndof
        S[L.S_COUNTER] = self.counter
        # keys of the device noise, drawn after every hook's init
        for hook in (thermo if lt else None, baro if lb else None):
            if hook is not None and hook.seed is None:
                hook.seed = draw_seed()

        def put_seed(slot, seed):
            seed = int(seed) & ((1 << 64) - 1)
            S[slot] = float(seed & 0xFFFFFFFF)
            S[slot + 1] = float(seed >> 32)

        if lt:
            c1, amp = thermo.device_scalars(self.timestep)
            S[L.S_THERMO] = L.THERMO_LANGEVIN
            S[L.S_THERMO_TEMP] = thermo.temp
            S[L.S_LT_C1] = c1
            S[L.S_LT_A] = amp
            put_seed(L.S_LT_SEED, thermo.seed)
        elif thermo is not None:
            ch = thermo.chain
            S[L.S_THERMO] = L.THERMO_NHC
            S[L.S_CHAIN_LEN] = ch.length
            S[L.S_CHAIN_NDOF] = ch.ndof
            S[L.S_THERMO_TEMP] = thermo.temp
            S[L.S_CHAIN_MASS:L.S_CHAIN_MASS + ch.length] = ch.masses
            S[L.S_CHAIN_POS:L.S_CHAIN_POS + ch.length] = ch.pos
            S[L.S_CHAIN_VEL:L.S_CHAIN_VEL + ch.length] = ch.vel
        if baro is not None:
            S[L.S_BARO] = L.BARO_LANGEVIN if lb else L.BARO_MTK
            S[L.S_BARO_TEMP] = baro.temp
            S[L.S_BARO_PRESS] = baro.press
            S[L.S_MASS_PRESS] = baro.mass_press
            S[L.S_BARO_NDOF] = baro.baro_ndof
            S[L.S_BARO_ANISO] = 1.0 if baro.anisotropic else 0.0
            S[L.S_BARO_VOLCON] = 1.0 if baro.vol_constraint else 0.0
            vp = np.zeros(9)
            vp[:np.size(baro.vel_press)] = np.asarray(baro.vel_press).ravel()
            S[L.S_VEL_PRESS:L.S_VEL_PRESS + 9] = vp
            bt = getattr(baro, "baro_thermo", None)
            if bt is not None:
                ch = bt.chain
                S[L.S_BCHAIN_ON] = 1.0
                S[L.S_BCHAIN_LEN] = ch.length
                S[L.S_BCHAIN_TEMP] = bt.temp
                S[L.S_BCHAIN_NDOF] = ch.ndof
                S[L.S_BCHAIN_MASS:L.S_BCHAIN_MASS + ch.length] = ch.masses
                S[L.S_BCHAIN_POS:L.S_BCHAIN_POS + ch.length] = ch.pos
                S[L.S_BCHAIN_VEL:L.S_BCHAIN_VEL + ch.length] = ch.vel
            if lb:
                S[L.S_LB_TIMECON] = baro.timecon
                S[L.S_LB_RSCALE] = baro.random_force_scale()
                put_seed(L.S_LB_SEED, baro.seed)
        self._engine.put_scalars(S)
        self._engine.pos.copy_(torch.from_numpy(np.ascontiguousarray(self._pos)))
        self._engine.vel.copy_(torch.from_numpy(np.ascontiguousarray(self._vel)))
        self._engine.posold.copy_(self._engine.pos)
        self._uploaded = True
        self._engine.call("mm_md_init_properties")
        self._pull()
        self._mirror_hooks()
        Iterative.initialize(self)

    # ------------------------------------------------------------------ stepping
    def _next_hook_distance(self, limit):
        best = limit
        for hook in self.hooks:
            if getattr(hook, "_device_hook", False):
                continue
            if isinstance(hook, VerletScreenLog) and not log.do_medium:
                continue
            c0 = max(self.counter + 1, hook.start)
            r = (c0 - hook.start) % hook.step
            c = c0 if r == 0 else c0 + hook.step - r
            best = min(best, c - self.counter)
        return max(1, best)

    def _advance(self, k, snapshot_last):
        if snapshot_last:
            if k > 1:
                self._engine.call("mm_md_step", k - 1, int(bool(self.use_graph)))
            self._engine.call("mm_md_snapshot")
            self._engine.call("mm_md_step", 1, int(bool(self.use_graph)))
        else:
            self._engine.call("mm_md_step", k, int(bool(self.use_graph)))
        # a thermostat factor may still be pending on the velocities; apply it so that the
        # device arrays hold the reference state between batches
        self._engine.call("mm_md_sync")
        self.counter += k
        self._pull()

    def propagate(self):
        self._advance(1, True)
        self._write_through()
        self.call_hooks()
        return False

    def run(self, nsteps=None):
        remaining = float("inf") if nsteps is None else int(nsteps)
        while remaining > 0:
            limit = int(min(remaining, 1 << 30))
            k = self._next_hook_distance(limit)
            due = self._hooks_due(self.counter + k)
            self._advance(k, due)
            self._write_through()
            if due:
                self.call_hooks()
            remaining -= k
        self.finalize()

    def _hooks_due(self, counter):
        for hook in self.hooks:
            if getattr(hook, "_device_hook", False):
                continue
            if hook.expects_call(counter):
                return True
        return False

    def _sync_for_hooks(self):
        """Apply pending transforms, compute rmsd values, mirror state for the hooks."""
        self._engine.call("mm_md_sync")
        self._engine.call("mm_md_rmsd")
        self._pull()
        self._write_through()
        self._mirror_hooks()

    def _mirror_hooks(self):
        S = self._S
        if self._thermo is not None:
            ch = getattr(self._thermo, "chain", None)
            if ch is not None:
                ch.pos = S[L.S_CHAIN_POS:L.S_CHAIN_POS + ch.length].copy()
                ch.vel = S[L.S_CHAIN_VEL:L.S_CHAIN_VEL + ch.length].copy()
            self._thermo.econs_correction = float(S[L.S_ECONS_THERMO])
        if self._baro is not None:
            if self._baro.anisotropic:
                self._baro.vel_press = S[L.S_VEL_PRESS:L.S_VEL_PRESS + 9].reshape(3, 3).copy()
            else:
                self._baro.vel_press = float(S[L.S_VEL_PRESS])
            self._baro.econs_correction = float(S[L.S_ECONS_BARO])
            bt = getattr(self._baro, "baro_thermo", None)
            if bt is not None:
                bt.chain.pos = S[L.S_BCHAIN_POS:L.S_BCHAIN_POS + bt.chain.length].copy()
                bt.chain.vel = S[L.S_BCHAIN_VEL:L.S_BCHAIN_VEL + bt.chain.length].copy()
        for hook in self.hooks:
            if getattr(hook, "name", None) == "TBCombination":
                hook.econs_correction = float(S[L.S_ECONS_CORR])
        t = self._cons_err_tracker
        t.counter = int(S[L.S_CE])
        t.ekin_m, t.ekin_s, t.econs_m, t.econs_s = (float(x) for x in S[L.S_CE + 1:L.S_CE + 5])

    def sync(self):
        """Apply pending device transforms and refresh all host attributes."""
        self._sync_for_hooks()

    def compute_properties(self):
        self._engine.call("mm_md_rmsd")
        self._pull()

    def finalize(self):
        if log.do_medium:
            log.hline()
