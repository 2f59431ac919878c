"""Barostats and the thermostat-barostat combination (micmec/sampling/npt.py:52-177, 295-787).

``MTKBarostat`` (anisotropic full 3x3 or isotropic cell fluctuations, with or without volume
constraint, with or without its own Nose-Hoover chain ``baro_thermo``), ``LangevinBarostat``
(anisotropic or isotropic, with or without volume constraint) and ``TBCombination`` keep the
reference constructors. Host-side ``init`` mirrors the reference
(domain symmetrisation at construction, momentum cleaning, barostat mass, random symmetric
barostat velocity); the per-step barostat updates, including the extra force evaluation of each
``baro`` call and the Langevin barostat's random force, run inside the device integrator.
"""

import numpy as np

from ..units import boltzmann, femtosecond
from .iterative import StateItem
from .nvt import LangevinThermostat, NHCThermostat
from .utils import (clean_momenta, domain_symmetrize, get_ndof_baro, get_ndof_internal_md,
                    get_random_vel_press)
from .verlet import VerletHook

__all__ = ["TBCombination", "MTKBarostat", "LangevinBarostat", "MTKAttributeStateItem"]


class TBCombination(VerletHook):
    """Thermostat and barostat called in the order of npt.py:99-148."""
    name = "TBCombination"
    _device_hook = True

    def __init__(self, thermostat, barostat, start=0):
        self.thermostat = thermostat
        self.barostat = barostat
        self.start = start
        if not self.verify():
            self.barostat = thermostat
            self.thermostat = barostat
            if not self.verify():
                raise TypeError("The Thermostat or Barostat instance is not supported (yet).")
        self.step_thermo = self.thermostat.step
        self.step_baro = self.barostat.step
        VerletHook.__init__(self, start, min(self.step_thermo, self.step_baro))

    def init(self, iterative):
        set_ndof = iterative.ndof is not None
        self.thermostat.init(iterative)
        self.barostat.init(iterative)
        # ndof = 3N when a stochastic hook lets the centre-of-mass momentum fluctuate and the
        # caller did not fix ndof (npt.py:87-95)
        p_cm_fluct = isinstance(self.thermostat, LangevinThermostat) or \
            isinstance(self.barostat, LangevinBarostat)
        if not set_ndof and p_cm_fluct:
            iterative.ndof = iterative.pos.size
        self.chainvel0 = None
        self.G1_add = None

    def verify(self):
        return isinstance(self.thermostat, (NHCThermostat, LangevinThermostat)) and \
            isinstance(self.barostat, (MTKBarostat, LangevinBarostat))

    def pre(self, iterative):
        raise NotImplementedError("TBC updates run inside the device integrator.")

    post = pre


class MTKBarostat(VerletHook):
    """Martyna-Tobias-Klein barostat (npt.py:513-757): anisotropic (3x3 symmetric barostat
    velocity) or isotropic (one scalar), optionally volume-preserving, optionally with its own
    Nose-Hoover chain ``baro_thermo`` (an ``NHCThermostat`` that is not itself a hook of the
    integrator; it thermostats the barostat alone and decouples it from a particle thermostat).
    An isotropic barostat with a volume constraint has no degree of freedom and raises
    ``AssertionError`` as in the reference (utils.py:498-500)."""
    name = "MTTK"
    kind = "deterministic"
    method = "barostat"
    _device_hook = True

    def __init__(self, mmf, temp, press, start=0, step=1, timecon=1000 * femtosecond,
                 anisotropic=True, vol_constraint=False, baro_thermo=None, vel_press0=None,
                 restart=False):
        if baro_thermo is not None and not isinstance(baro_thermo, NHCThermostat):
            raise TypeError("baro_thermo must be an NHCThermostat.")
        if baro_thermo is not None and baro_thermo.chain.length > 8:
            raise ValueError("The barostat thermostat's chain holds at most 8 beads.")
        self.temp = temp
        self.press = press
        self.timecon_press = timecon
        self.anisotropic = anisotropic
        self.vol_constraint = vol_constraint
        self.baro_thermo = baro_thermo
        self.dim = mmf.system.domain.nvec
        if self.dim != 3:
            raise ValueError("The MTK barostat needs a 3D periodic domain.")
        self.restart = restart
        self.baro_ndof = get_ndof_baro(self.dim, self.anisotropic, self.vol_constraint)
        if self.anisotropic and not self.restart:
            domain_symmetrize(mmf)
        self.domain = np.asarray(mmf.system.domain.rvecs).copy()
        self.vel_press = None if vel_press0 is None else np.array(vel_press0, dtype=float)
        self.econs_correction = 0.0
        VerletHook.__init__(self, start, step)

    def init(self, iterative):
        """MTKBarostat.init (npt.py:580-614), random draws in the reference's order: the
        barostat velocity, then the barostat chain's velocities."""
        self.timestep_press = iterative.timestep
        if not self.restart:
            clean_momenta(iterative.pos, iterative.vel, iterative.masses,
                          iterative.mmf.system.domain)
        if iterative.ndof is None:
            iterative.ndof = get_ndof_internal_md(iterative.mmf.system.nnodes,
                                                  iterative.mmf.system.domain.nvec)
        angfreq = 2 * np.pi / self.timecon_press
        self.mass_press = (iterative.ndof + self.dim ** 2) * boltzmann * self.temp / angfreq ** 2
        if self.vel_press is None:
            self.vel_press = get_random_vel_press(self.mass_press, self.temp)
            if not self.anisotropic:
                self.vel_press = self.vel_press[0][0]
        if self.baro_thermo is not None:
            self.baro_thermo.chain.timestep = iterative.timestep
            self.baro_thermo.chain.set_ndof(self.baro_ndof)
        if self.vol_constraint:
            self.vel_press = self.vel_press - np.trace(self.vel_press) / 3 * np.eye(3)
        # gpos and vtens after the symmetrisation (npt.py:611-614)
        iterative._init_force(with_vtens=True)

    def pre(self, iterative, chainvel0=None):
        raise NotImplementedError("MTK updates run inside the device integrator.")

    post = pre

    def add_press_cont(self):
        if self.baro_thermo is not None:
            return 0
        return 2 * self._compute_ekin_baro() - self.baro_ndof * (self.temp * boltzmann)

    def _compute_ekin_baro(self):
        if self.anisotropic:
            return 0.5 * self.mass_press * np.trace(np.dot(self.vel_press.T, self.vel_press))
        return 0.5 * self.mass_press * self.vel_press ** 2


class LangevinBarostat(VerletHook):
    """Langevin barostat (npt.py:295-510; Feller, Zhang, Pastor and Brooks, J. Chem. Phys. 103,
    4613, 1995): the MTK cell update with a friction exp(-dt / (8 timecon)) on both sides of
    every barostat-velocity update and a random symmetric force R of variance
    2 W k_B T / (dt timecon), drawn on the device (Philox4x64-10 keyed by ``seed``)."""
    name = "Langevin"
    kind = "stochastic"
    method = "barostat"
    _device_hook = True

    def __init__(self, mmf, temp, press, start=0, step=1, timecon=1000 * femtosecond,
                 anisotropic=True, vol_constraint=False, seed=None):
        self.temp = temp
        self.press = press
        self.timecon = timecon
        self.anisotropic = anisotropic
        self.vol_constraint = vol_constraint
        self.seed = seed
        self.dim = mmf.system.domain.nvec
        if self.dim != 3:
            raise ValueError("The Langevin barostat needs a 3D periodic domain.")
        self.baro_ndof = get_ndof_baro(self.dim, self.anisotropic, self.vol_constraint)
        if self.anisotropic:
            domain_symmetrize(mmf)
        self.domain = np.asarray(mmf.system.domain.rvecs).copy()
        self.econs_correction = 0.0
        VerletHook.__init__(self, start, step)

    def init(self, iterative):
        self.timestep_press = iterative.timestep
        clean_momenta(iterative.pos, iterative.vel, iterative.masses,
                      iterative.mmf.system.domain)
        if iterative.ndof is None:
            iterative.ndof = iterative.pos.size
        self.mass_press = (iterative.ndof + 3) / 3 * boltzmann * self.temp * (
            self.timecon / (2 * np.pi)) ** 2
        self.vel_press = get_random_vel_press(self.mass_press, self.temp)
        if self.vol_constraint:
            self.vel_press -= np.trace(self.vel_press) / 3 * np.eye(3)
        if not self.anisotropic:
            self.vel_press = self.vel_press[0][0]
        # gpos and vtens after the symmetrisation (npt.py:368-372)
        iterative._init_force(with_vtens=True)

    def random_force_scale(self):
        """sqrt(2 W k_B T / (dt timecon)) in the reference's operation order (npt.py:492-497)."""
        return np.sqrt(2 * self.mass_press * boltzmann * self.temp /
                       (self.timestep_press * self.timecon))

    def pre(self, iterative, chainvel0=None):
        raise NotImplementedError("Langevin barostat updates run inside the device integrator.")

    post = pre


class MTKAttributeStateItem(StateItem):
    def __init__(self, attr):
        StateItem.__init__(self, "baro_" + attr)
        self.attr = attr

    def get_value(self, iterative):
        baro = None
        for hook in iterative.hooks:
            if isinstance(hook, MTKBarostat):
                baro = hook
                break
            if isinstance(hook, TBCombination):
                if isinstance(hook.barostat, MTKBarostat):
                    baro = hook.barostat
                break
        if baro is None:
            raise TypeError("Iterative does not contain an MTKBarostat hook.")
        if self.key.startswith("baro_chain_"):
            if baro.baro_thermo is None:
                return 0
            return getattr(baro.baro_thermo.chain, self.key.split("_")[2])
        return getattr(baro, self.attr)

    def copy(self):
        return self.__class__(self.attr)
