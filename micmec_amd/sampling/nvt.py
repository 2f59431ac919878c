"""Thermostats: Nose-Hoover chain (micmec/sampling/nvt.py:361-557) and Langevin
(micmec/sampling/nvt.py:165-218).

``NHCThermostat`` and ``LangevinThermostat`` keep the reference constructors and attributes.
Their ``init`` runs on the host with numpy (momentum cleaning, chain masses, random chain
velocities in the reference's draw order); their per-step ``pre``/``post`` updates run inside
the device integrator (``micmec_amd/csrc/mm_md.hip``), whose state is mirrored back into the
hook (``chain.pos`` / ``chain.vel``, ``econs_correction``) whenever host hooks run.

The Langevin noise is drawn on the device from a counter-based generator (Philox4x64-10 keyed
by the hook's ``seed``) instead of numpy's global stream: every node draws its own normals for
every hook call, in the node passes that already update the velocities. A seeded run is
reproducible; its noise differs from the reference's draws, so agreement with the reference
is statistical (the same update applied to the same distribution of noise).
"""

import numpy as np

from ..units import boltzmann, femtosecond
from .iterative import StateItem
from .utils import clean_momenta, get_ndof_internal_md
from .verlet import VerletHook

__all__ = ["NHChain", "NHCThermostat", "LangevinThermostat", "NHCAttributeStateItem"]


class NHChain(object):
    """Chain of Nose-Hoover thermostats (nvt.py:361-458)."""

    def __init__(self, length, timestep, temp, ndof, pos0, vel0, timecon=100 * femtosecond):
        self.length = length
        self.timestep = timestep
        self.temp = temp
        self.timecon = timecon
        self.restart_pos = pos0 is not None
        self.restart_vel = vel0 is not None
        if ndof > 0:
            self.set_ndof(ndof)
        self.pos = pos0.copy() if self.restart_pos else np.zeros(length)
        # as in nvt.py:388-391, a draw made by set_ndof above is overwritten here
        self.vel = np.array(vel0, dtype=float).copy() if self.restart_vel else np.zeros(length)

    def set_ndof(self, ndof):
        self.ndof = ndof
        afreq = 2 * np.pi / self.timecon
        self.masses = np.ones(self.length) * (boltzmann * self.temp / afreq ** 2)
        self.masses[0] *= ndof
        if not self.restart_vel:
            self.vel = self.get_random_vel_therm()

    def get_random_vel_therm(self):
        return np.random.normal(0, np.sqrt(self.masses * boltzmann * self.temp),
                                self.length) / self.masses

    def __call__(self, ekin, vel, G1_add):
        """Host evaluation of one half step of the chain (nvt.py:410-451)."""
        def do_bead(k, ekin):
            if k == 0:
                g = 2 * ekin - self.ndof * self.temp * boltzmann
                if G1_add is not None:
                    g += G1_add
            else:
                g = self.masses[k - 1] * self.vel[k - 1] ** 2 - self.temp * boltzmann
            g /= self.masses[k]
            if k == self.length - 1:
                self.vel[k] += g * self.timestep / 4
            else:
                self.vel[k] *= np.exp(-self.vel[k + 1] * self.timestep / 8)
                self.vel[k] += g * self.timestep / 4
                self.vel[k] *= np.exp(-self.vel[k + 1] * self.timestep / 8)

        for k in range(self.length - 1, -1, -1):
            do_bead(k, ekin)
        self.pos += self.vel * self.timestep / 2
        factor = np.exp(-self.vel[0] * self.timestep / 2)
        vel *= factor
        ekin *= factor ** 2
        for k in range(self.length):
            do_bead(k, ekin)
        return vel, ekin

    def get_econs_correction(self):
        kt = boltzmann * self.temp
        return 0.5 * (self.vel ** 2 * self.masses).sum() + kt * (
            self.ndof * self.pos[0] + self.pos[1:].sum())


class NHCThermostat(VerletHook):
    """Nose-Hoover chain thermostat (nvt.py:461-532)."""
    name = "NHC"
    kind = "deterministic"
    method = "thermostat"
    _device_hook = True

    def __init__(self, temp, start=0, timecon=100 * femtosecond, chainlength=3, chain_pos0=None,
                 chain_vel0=None, restart=False):
        if chainlength < 1 or chainlength > 8:
            raise ValueError("The device integrator supports chains of 1 to 8 beads.")
        self.temp = temp
        self.restart = restart
        self.chain = NHChain(chainlength, 0.0, temp, 0, chain_pos0, chain_vel0, timecon)
        VerletHook.__init__(self, start, 1)

    def init(self, iterative):
        if not self.restart:
            clean_momenta(iterative.pos, iterative.vel, iterative.masses,
                          iterative.mmf.system.domain)
        if iterative.ndof is None:
            iterative.ndof = get_ndof_internal_md(iterative.pos.shape[0],
                                                  iterative.mmf.system.domain.nvec)
        self.chain.timestep = iterative.timestep
        self.chain.set_ndof(iterative.ndof)

    def pre(self, iterative, G1_add=None):
        raise NotImplementedError("NHC updates run inside the device integrator.")

    post = pre


def draw_seed():
    """A 62-bit key for the device noise, from numpy's global stream (drawn after every hook's
    init, so the reference's draws at initialisation keep their order)."""
    return int(np.random.randint(0, 2 ** 62, dtype=np.int64))


class LangevinThermostat(VerletHook):
    """Langevin thermostat (nvt.py:165-218): at every pre and post call
    v <- c1 v + sqrt((1 - c1^2) T k_B / m) xi with c1 = exp(-dt / (2 timecon)) and xi standard
    normals, and econs_correction += ekin before - ekin after (Bussi and Parrinello,
    Phys. Rev. E 75, 056707, 2007)."""
    name = "Langevin"
    kind = "stochastic"
    method = "thermostat"
    _device_hook = True

    def __init__(self, temp, start=0, timecon=100 * femtosecond, seed=None):
        self.temp = temp
        self.timecon = timecon
        self.seed = seed
        VerletHook.__init__(self, start, 1)

    def init(self, iterative):
        # the external momenta must be zero (nvt.py:193-196)
        clean_momenta(iterative.pos, iterative.vel, iterative.masses,
                      iterative.mmf.system.domain)

    def device_scalars(self, timestep):
        """c1 and (1 - c1^2) T k_B in the reference's operation order (nvt.py:212-215)."""
        c1 = np.exp(-timestep / self.timecon / 2)
        return c1, (1.0 - c1 ** 2) * self.temp * boltzmann

    def pre(self, iterative, G1_add=None):
        raise NotImplementedError("Langevin updates run inside the device integrator.")

    post = pre


class NHCAttributeStateItem(StateItem):
    def __init__(self, attr):
        StateItem.__init__(self, "thermo_" + attr)
        self.attr = attr

    def get_value(self, iterative):
        from .npt import TBCombination
        chain = None
        for hook in iterative.hooks:
            if isinstance(hook, NHCThermostat):
                chain = hook.chain
                break
            if isinstance(hook, TBCombination):
                if isinstance(hook.thermostat, NHCThermostat):
                    chain = hook.thermostat.chain
                break
        if chain is None:
            raise TypeError("Iterative does not contain a NHCThermostat hook.")
        return getattr(chain, self.attr)

    def copy(self):
        return self.__class__(self.attr)
