"""CPU restatement of MicMec's velocity Verlet with its thermostat and barostat hooks (test
infrastructure only).

Follows, in the same order of operations and random draws:
  micmec/sampling/verlet.py:71-190        VerletIntegrator (init, propagate, properties)
  micmec/sampling/verlet.py:275-307       ConsErrTracker
  micmec/sampling/nvt.py:361-532          NHChain and NHCThermostat
  micmec/sampling/npt.py:52-177           TBCombination
  micmec/sampling/npt.py:513-757          MTKBarostat (anisotropic or isotropic, with or
                                          without volume constraint, with or without a
                                          barostat thermostat)
  micmec/sampling/nvt.py:165-218          LangevinThermostat
  micmec/sampling/npt.py:295-510          LangevinBarostat (anisotropic or isotropic, with or
                                          without volume constraint)
  micmec/sampling/utils.py:28-174, 322-383, 450-501  init helpers
The force field is any callable ``force(pos, rvecs) -> (epot, gpos, vtens)``.
"""

import numpy as np

from .constants import boltzmann, femtosecond
from .philox import barostat_noise, thermostat_noise


# ---------------------------------------------------------------- sampling/utils.py


def get_random_vel(temp0, scalevel0, masses):
    vel0 = np.random.normal(0, 1, (len(masses), 3)) * np.sqrt(
        boltzmann * temp0 / masses).reshape(-1, 1)
    if scalevel0 and temp0 > 0:
        temp = np.mean((vel0 ** 2 * masses.reshape(-1, 1))) / boltzmann
        vel0 *= np.sqrt(temp0 / temp)
    return vel0


def remove_com_moment(vel, masses):
    velcom = np.dot(masses, vel) / np.sum(masses)
    vel[:] -= velcom


def remove_angular_moment(pos, vel, masses):
    pos = pos.copy()
    pos -= np.dot(masses, pos) / np.sum(masses)
    itens = np.identity(3) * (masses.reshape(-1, 1) * pos ** 2).sum() - np.dot(
        pos.T, masses.reshape(-1, 1) * pos)
    lin = masses.reshape(-1, 1) * vel
    amom = np.zeros(3)
    amom[0] = np.sum(pos[:, 1] * lin[:, 2] - pos[:, 2] * lin[:, 1])
    amom[1] = np.sum(pos[:, 2] * lin[:, 0] - pos[:, 0] * lin[:, 2])
    amom[2] = np.sum(pos[:, 0] * lin[:, 1] - pos[:, 1] * lin[:, 0])
    evals, evecs = np.linalg.eigh(itens)
    mask = evals > 1e-10
    avel = np.dot(evecs[:, mask], np.dot(evecs[:, mask].T, amom) / evals[mask])
    rb = np.zeros(pos.shape)
    rb[:, 0] = avel[1] * pos[:, 2] - avel[2] * pos[:, 1]
    rb[:, 1] = avel[2] * pos[:, 0] - avel[0] * pos[:, 2]
    rb[:, 2] = avel[0] * pos[:, 1] - avel[1] * pos[:, 0]
    vel[:] -= rb


def clean_momenta(pos, vel, masses, nvec):
    remove_com_moment(vel, masses)
    if nvec == 0:
        remove_angular_moment(pos, vel, masses)
    elif nvec == 1:
        raise NotImplementedError


def get_ndof_internal_md(nnodes, nper):
    if nper == 0:
        return 3 * nnodes - 6
    if nper == 1:
        return 3 * nnodes - 4
    return 3 * nnodes - 3


def domain_symmetrize(pos, rvecs):
    U, s, Vt = np.linalg.svd(rvecs)
    rot = np.dot(Vt.T, U.T)
    return np.dot(pos, rot), np.dot(rvecs, rot)


def get_random_vel_press(mass, temp):
    rand = np.random.normal(0, np.sqrt(mass * boltzmann * temp), (3, 3)) / mass
    vp = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            vp[i, j] = rand[i, j] if i >= j else rand[j, i]
            if i != j:
                vp[i, j] /= np.sqrt(2)
    return vp


def volume_of(rvecs):
    r = np.zeros(9)
    r[: rvecs.size] = rvecs.ravel()
    if rvecs.shape[0] == 0:
        return 0.0
    return abs(r[0] * (r[4] * r[8] - r[5] * r[7]) + r[1] * (r[5] * r[6] - r[3] * r[8])
               + r[2] * (r[3] * r[7] - r[4] * r[6]))


# ---------------------------------------------------------------- nvt.py NHChain


class NHChain(object):
    def __init__(self, length, timestep, temp, timecon=100 * femtosecond):
        self.length = length
        self.timestep = timestep
        self.temp = temp
        self.timecon = timecon
        self.pos = np.zeros(length)
        self.vel = np.zeros(length)

    def set_ndof(self, ndof, draw=True):
        self.ndof = ndof
        afreq = 2 * np.pi / self.timecon
        self.masses = np.ones(self.length) * (boltzmann * self.temp / afreq ** 2)
        self.masses[0] *= ndof
        if draw:
            self.vel = np.random.normal(0, np.sqrt(self.masses * boltzmann * self.temp),
                                        self.length) / self.masses

    def __call__(self, ekin, vel, G1_add):
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

    def econs_correction(self):
        kt = boltzmann * self.temp
        return 0.5 * (self.vel ** 2 * self.masses).sum() + kt * (
            self.ndof * self.pos[0] + self.pos[1:].sum())


# ---------------------------------------------------------------- the integrator


class OracleMD(object):
    """Velocity Verlet with an optional thermostat (NHC or Langevin) and barostat (MTK or
    Langevin), combined as in TBC.

    Parameters mirror the reference constructors: ``thermo = dict(temp, timecon, chainlength)``
    for NHC or ``dict(kind="langevin", temp, timecon, seed)``; ``baro = dict(temp, press,
    timecon, anisotropic, vol_constraint, baro_thermo)`` for MTK (``baro_thermo = dict(temp,
    timecon, chainlength)`` couples a Nose-Hoover chain to the barostat alone) or
    ``dict(kind="langevin", temp, press, timecon, anisotropic, vol_constraint, seed)``.
    ``vel0``/``chain_vel0``/``vel_press0`` skip the corresponding random draws exactly as the
    reference does. The Langevin noise is the counter-based noise of
    ``oracle.philox`` for the given seeds (the reference draws it from numpy's global stream).
    """

    def __init__(self, force, pos, rvecs, masses, timestep, temp0=300.0, vel0=None,
                 thermo=None, baro=None, chain_vel0=None, vel_press0=None, ndof=None,
                 ext_press=0.0, scalevel0=True):
        self.force = force
        self.timestep = timestep
        self.masses = masses.copy()
        self.nvec = rvecs.shape[0]
        self.ext_press = ext_press
        pos = pos.copy()
        rvecs = rvecs.copy()
        self.thermo = None if thermo is None else dict(thermo)
        self.baro = None if baro is None else dict(baro)
        self.thermo_kind = None if thermo is None else self.thermo.get("kind", "nhc")
        self.baro_kind = None if baro is None else self.baro.get("kind", "mtk")
        if self.baro is not None:
            self.aniso = self.baro.get("anisotropic", True)
            self.volcon = self.baro.get("vol_constraint", False)
            # the barostat constructors symmetrise the domain before the integrator copies pos
            # (npt.py:566-568, 347-349 via utils.py:346-383)
            if self.aniso:
                pos, rvecs = domain_symmetrize(pos, rvecs)
            # get_ndof_baro (utils.py:478-501)
            self.baro_ndof = (3 * 4 // 2 if self.aniso else 1) - (1 if self.volcon else 0)
            if self.baro_ndof == 0:
                raise AssertionError("Isotropic barostat called with a volume constraint.")
        self.pos = pos
        self.rvecs = rvecs
        self.time = 0.0
        self.counter = 0
        self.ndof = ndof
        if vel0 is None:
            self.vel = get_random_vel(temp0, scalevel0, self.masses)
            clean_momenta(self.pos, self.vel, self.masses, self.nvec)
        else:
            self.vel = vel0.copy()
        self.vtens = np.zeros((3, 3))
        # initialize(): force without virial (verlet.py:119-137)
        self.epot, self.gpos, _ = self._force(self.pos, self.rvecs)
        self.posold = self.pos.copy()
        self.chain = None
        self.econs_thermo = 0.0
        self.econs_baro = 0.0
        set_ndof = self.ndof is not None
        if self.thermo_kind == "nhc":
            clean_momenta(self.pos, self.vel, self.masses, self.nvec)
            if self.ndof is None:
                self.ndof = get_ndof_internal_md(len(self.pos), self.nvec)
            self.chain = NHChain(thermo.get("chainlength", 3), timestep, thermo["temp"],
                                 thermo.get("timecon", 100 * femtosecond))
            if chain_vel0 is not None:
                self.chain.set_ndof(self.ndof, draw=False)
                self.chain.vel = np.array(chain_vel0, dtype=float)
            else:
                self.chain.set_ndof(self.ndof)
        elif self.thermo_kind == "langevin":
            # LangevinThermostat.init (nvt.py:193-196)
            clean_momenta(self.pos, self.vel, self.masses, self.nvec)
        self.bchain = None
        if self.baro_kind == "mtk":
            # MTKBarostat.init (npt.py:580-614)
            clean_momenta(self.pos, self.vel, self.masses, self.nvec)
            if self.ndof is None:
                self.ndof = get_ndof_internal_md(len(self.pos), self.nvec)
            angfreq = 2 * np.pi / self.baro.get("timecon", 1000 * femtosecond)
            self.mass_press = (self.ndof + 9) * boltzmann * self.baro["temp"] / angfreq ** 2
            if vel_press0 is None:
                self.vel_press = get_random_vel_press(self.mass_press, self.baro["temp"])
                if not self.aniso:
                    self.vel_press = self.vel_press[0][0]
            else:
                self.vel_press = np.array(vel_press0, dtype=float)
            bt = self.baro.get("baro_thermo")
            if bt is not None:
                self.bchain = NHChain(bt.get("chainlength", 3), timestep, bt["temp"],
                                      bt.get("timecon", 100 * femtosecond))
                if bt.get("chain_vel0") is not None:
                    self.bchain.set_ndof(self.baro_ndof, draw=False)
                    self.bchain.vel = np.array(bt["chain_vel0"], dtype=float)
                else:
                    self.bchain.set_ndof(self.baro_ndof)
            if self.volcon:
                self.vel_press -= np.trace(self.vel_press) / 3 * np.eye(3)
            self.epot, self.gpos, self.vtens = self._force(self.pos, self.rvecs)
        elif self.baro_kind == "langevin":
            # LangevinBarostat.init (npt.py:351-372)
            clean_momenta(self.pos, self.vel, self.masses, self.nvec)
            if self.ndof is None:
                self.ndof = self.pos.size
            timecon = self.baro.get("timecon", 1000 * femtosecond)
            self.mass_press = (self.ndof + 3) / 3 * boltzmann * self.baro["temp"] * (
                timecon / (2 * np.pi)) ** 2
            if vel_press0 is None:
                self.vel_press = get_random_vel_press(self.mass_press, self.baro["temp"])
            else:
                self.vel_press = np.array(vel_press0, dtype=float)
            if self.volcon:
                self.vel_press -= np.trace(self.vel_press) / 3 * np.eye(3)
            if not self.aniso:
                self.vel_press = self.vel_press[0][0]
            self.epot, self.gpos, self.vtens = self._force(self.pos, self.rvecs)
        # TBCombination.init (npt.py:81-95): stochastic hooks let the momentum fluctuate
        if self.thermo is not None and self.baro is not None and not set_ndof and \
                "langevin" in (self.thermo_kind, self.baro_kind):
            self.ndof = self.pos.size
        if self.ndof is None:
            self.ndof = self.pos.size
        self.cons = [0, 0.0, 0.0, 0.0, 0.0]
        self.compute_properties()

    def _force(self, pos, rvecs):
        e, g, v = self.force(pos, rvecs)
        if self.ext_press:
            V = volume_of(rvecs)
            e = e + V * self.ext_press
            v = v + self.ext_press * V * np.identity(3)
        return e, g, v

    def ekin_now(self):
        return np.sum(0.5 * (self.vel ** 2.0 * self.masses.reshape(-1, 1)))

    # ------------------------------------------------------------ Langevin thermostat
    def _langevin_thermo(self, post):
        """LangevinThermostat.pre/post (nvt.py:197-218)."""
        th = self.thermo
        ekin0 = self.ekin
        c1 = np.exp(-self.timestep / th.get("timecon", 100 * femtosecond) / 2)
        c2 = np.sqrt((1.0 - c1 ** 2) * th["temp"] * boltzmann / self.masses).reshape(-1, 1)
        xi = thermostat_noise(th["seed"], self.counter, post, len(self.masses))
        self.vel = c1 * self.vel + c2 * xi
        self.ekin = self.ekin_now()
        self.econs_thermo += ekin0 - self.ekin

    # ------------------------------------------------------------ MTK (npt.py:616-757)
    def _baro_ekin(self):
        if self.aniso:
            return 0.5 * self.mass_press * np.trace(np.dot(self.vel_press.T, self.vel_press))
        return 0.5 * self.mass_press * self.vel_press ** 2

    def _update_baro_vel(self, chainvel0, post=0, sub=0):
        dt = self.timestep
        lb = self.baro_kind == "langevin"
        if lb:
            tau = self.baro.get("timecon", 1000 * femtosecond)
            self.vel_press *= np.exp(-dt / (8 * tau))
        if chainvel0 is not None:
            self.vel_press *= np.exp(-dt * chainvel0 / 8)
        ptens_vol = np.dot(self.vel.T * self.masses, self.vel) - self.vtens
        ptens_vol = 0.5 * (ptens_vol.T + ptens_vol)
        G = (ptens_vol + (2.0 * self.ekin / self.ndof
                          - self.baro["press"] * volume_of(self.rvecs)) * np.eye(3)) \
            / self.mass_press
        if not lb:
            # MTK (npt.py:671-676): the isotropic form keeps the trace; the volume constraint
            # removes it (the two together are refused at construction)
            if not self.aniso:
                G = np.trace(G)
            if self.volcon:
                G -= np.trace(G) / self.nvec * np.eye(self.nvec)
            self.vel_press += G * dt / 4
        else:
            # LangevinBarostat.getR (npt.py:489-505) with the counter-based draw
            rand = barostat_noise(self.baro["seed"], self.counter, post, sub) * np.sqrt(
                2 * self.mass_press * boltzmann * self.baro["temp"] / (dt * tau))
            R = np.zeros((3, 3))
            for i in range(3):
                for j in range(3):
                    R[i, j] = rand[i, j] if i >= j else rand[j, i]
            if self.volcon:
                G -= np.trace(G) / self.nvec * np.eye(self.nvec)
                R -= np.trace(R) / self.nvec * np.eye(self.nvec)
            if not self.aniso:
                G = np.trace(G)
                R = R[0][0]
            self.vel_press += (G - R / self.mass_press) * dt / 4
            self.vel_press *= np.exp(-dt / (8 * tau))
        if chainvel0 is not None:
            self.vel_press *= np.exp(-dt * chainvel0 / 8)

    def _baro(self, chainvel0, post=0):
        dt = self.timestep
        epot0, ekin0 = self.epot, self.ekin
        self._update_baro_vel(chainvel0, post, 0)
        if self.aniso:
            Dr, Qg = np.linalg.eigh(self.vel_press)
            rot = np.dot(np.dot(Qg, np.diagflat(np.exp(Dr * dt / 2))), Qg.T)
            self.pos = np.dot(self.pos, rot)
            self.rvecs = np.dot(self.rvecs, rot)
        else:
            c = np.exp(self.vel_press * dt / 2)
            self.pos = c * self.pos
            self.rvecs = c * self.rvecs
        self.epot, self.gpos, self.vtens = self._force(self.pos, self.rvecs)
        if self.aniso:
            if self.volcon:
                Dg, Eg = np.linalg.eigh(self.vel_press)
            else:
                Dg, Eg = np.linalg.eigh(self.vel_press + (np.trace(self.vel_press) / self.ndof)
                                        * np.eye(3))
            rot = np.dot(np.dot(Eg, np.diagflat(np.exp(-Dg * dt / 2))), Eg.T)
            self.vel = np.dot(self.vel, rot)
        else:
            self.vel = np.exp(-((1.0 + 3.0 / self.ndof) * self.vel_press) * dt / 2) * self.vel
        self.ekin = self.ekin_now()
        self._update_baro_vel(chainvel0, post, 1)
        if self.baro_kind == "langevin":
            self.econs_baro += epot0 - self.epot + ekin0 - self.ekin

    def _add_press_cont(self):
        # MTKBarostat.add_press_cont (npt.py:738-746): a barostat thermostat decouples them
        if self.bchain is not None:
            return 0
        return 2 * self._baro_ekin() - self.baro_ndof * (self.baro["temp"] * boltzmann)

    def _bchain_call(self):
        """The barostat thermostat's half step on the barostat kinetic energy (npt.py:625-629,
        634-638); the barostat velocity itself is scaled inside update_baro_vel."""
        vp = np.zeros(np.shape(self.vel_press))
        vp[...] = self.vel_press
        self.bchain(self._baro_ekin(), vp, 0)

    # ------------------------------------------------------------ hooks in TBC order
    def _thermo_call(self, post):
        if self.thermo_kind == "nhc":
            # G1_add only from an MTK barostat (npt.py:109-113)
            g1 = self._add_press_cont() if self.baro_kind == "mtk" else None
            self.vel, self.ekin = self.chain(self.ekin, self.vel, g1)
            if post:
                self.econs_thermo = self.chain.econs_correction()
        elif self.thermo_kind == "langevin":
            self._langevin_thermo(post)

    def _pre(self):
        if self.baro is not None:
            chainvel0 = self.chain.vel[0] if self.chain is not None else None
            if self.bchain is not None:
                chainvel0 = self.bchain.vel[0]
            self._baro(chainvel0, 0)
            if self.bchain is not None:
                self._bchain_call()
        self._thermo_call(0)

    def _post(self):
        self._thermo_call(1)
        if self.baro is not None:
            chainvel0 = self.chain.vel[0] if self.chain is not None else None
            if self.bchain is not None:
                self._bchain_call()
                chainvel0 = self.bchain.vel[0]
            self._baro(chainvel0, 1)
            if self.baro_kind == "mtk":
                # MTKBarostat.post (npt.py:640-651)
                self.econs_baro = self._baro_ekin()
                if not self.volcon:
                    self.econs_baro += self.baro["press"] * volume_of(self.rvecs)
                if self.bchain is not None:
                    self.econs_baro += self.bchain.econs_correction()

    def econs_correction(self):
        corr = 0.0
        if self.thermo is not None and self.baro is not None:
            corr = self.econs_thermo + self.econs_baro
            if self.chain is not None and self.bchain is None:
                corr += self.baro_ndof * (boltzmann * self.chain.temp) * self.chain.pos[0]
        elif self.thermo is not None:
            corr = self.econs_thermo
        elif self.baro is not None:
            corr = self.econs_baro
        return corr

    def propagate(self):
        self._pre()
        dt = self.timestep
        acc = -self.gpos / self.masses.reshape(-1, 1)
        self.vel += 0.5 * acc * dt
        self.pos += dt * self.vel
        self.epot, self.gpos, self.vtens = self._force(self.pos, self.rvecs)
        acc = -self.gpos / self.masses.reshape(-1, 1)
        self.vel += 0.5 * acc * dt
        self.ekin = self.ekin_now()
        self._post()
        self.delta = self.pos - self.posold
        self.posold = self.pos.copy()
        self.time += dt
        self.counter += 1
        self.compute_properties()

    def compute_properties(self):
        self.rmsd_gpos = np.sqrt(np.mean(self.gpos ** 2))
        self.rmsd_delta = np.sqrt(np.mean(getattr(self, "delta", np.zeros(1)) ** 2))
        self.ekin = self.ekin_now()
        self.temp = (self.ekin / self.ndof) * (2.0 / boltzmann)
        self.etot = self.ekin + self.epot
        self.econs = self.etot + self.econs_correction()
        c = self.cons
        if c[0] == 0:
            c[1], c[3] = self.ekin, self.econs
        else:
            t = self.ekin - c[1]
            c[1] += t / (c[0] + 1)
            c[2] += t * (self.ekin - c[1])
            t = self.econs - c[3]
            c[3] += t / (c[0] + 1)
            c[4] += t * (self.econs - c[3])
        c[0] += 1
        self.cons_err = np.sqrt(c[4] / c[2]) if c[0] > 1 else 0.0
        if self.nvec > 0:
            self.ptens = (np.dot(self.vel.T * self.masses, self.vel) - self.vtens) \
                / volume_of(self.rvecs)
            self.press = np.trace(self.ptens) / 3.0

    def run(self, nsteps):
        for _ in range(nsteps):
            self.propagate()
