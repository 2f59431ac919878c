"""Langevin hooks on the host (no GPU): the counter-based generator and the oracle's Langevin
thermostat and barostat (micmec/sampling/nvt.py:165-218, npt.py:295-510).

The generator is pinned against numpy's own Philox4x64-10 bit generator and the published
known-answer vector; its normals are checked for their distribution. The oracle's hooks are
pinned by properties the mathematics fixes: the exact stationary velocity variance k_B T / m of
the Ornstein-Uhlenbeck update, the telescoping of the econs correction, the constant volume
under the volume constraint, the fixed cell shape of the isotropic barostat, and the target
pressure of the barostat's stationary state."""

import numpy as np
import pytest
import scipy.stats

from oracle import philox
from oracle.constants import boltzmann, femtosecond
from oracle.integrators import OracleMD


def test_philox_matches_numpy_bit_generator():
    rng = np.random.default_rng(7)
    for _ in range(5):
        key = rng.integers(0, 2 ** 63, size=2, dtype=np.uint64)
        ctr = rng.integers(0, 2 ** 62, size=4, dtype=np.uint64)
        bg = np.random.Philox(key=key, counter=ctr)
        raw = bg.random_raw(4 * 40)
        # numpy increments the 256-bit counter before each block
        ctrs = np.tile(ctr, (40, 1))
        ctrs[:, 0] += np.arange(1, 41, dtype=np.uint64)
        got = philox.philox4x64_10(ctrs, (int(key[0]), int(key[1]))).ravel()
        assert np.array_equal(got, raw)


def test_philox_known_answer_vector():
    out = philox.philox4x64_10(np.zeros((1, 4), dtype=np.uint64), (0, 0))[0]
    assert [int(x) for x in out] == [0x16554D9ECA36314C, 0xDB20FE9D672D0FDC,
                                     0xD7E772CEE186176B, 0x7E68B68AEC7BA23B]


def test_normals_are_independent_standard_normals():
    n = 100000
    z = philox.thermostat_noise(12345, 7, 1, n)
    flat = z.ravel()
    m = flat.size
    assert abs(flat.mean()) < 5 / np.sqrt(m)
    assert abs(flat.var() - 1.0) < 5 * np.sqrt(2.0 / m)
    assert scipy.stats.kstest(flat, "norm").pvalue > 1e-4
    # components of one node, neighbouring nodes, the pre and post draws and the two streams
    pairs = [(z[:, 0], z[:, 1]), (z[:, 1], z[:, 2]), (z[:-1, 0], z[1:, 0]),
             (z[:, 0], philox.thermostat_noise(12345, 7, 0, n)[:, 0]),
             (z[:, 0], philox.thermostat_noise(12346, 7, 1, n)[:, 0])]
    for a, b in pairs:
        assert abs(np.corrcoef(a, b)[0, 1]) < 5 / np.sqrt(n)
    # the uniforms stay inside (0, 1) at the extremes of the 64-bit words
    w = np.array([0, 2 ** 64 - 1], dtype=np.uint64)
    u = philox.open_uniform(w)
    assert 0.0 < u[0] < 1e-15 and 1.0 - 1e-15 < u[1] < 1.0


def _free_particles(n, seed, mass=1000.0):
    rng = np.random.default_rng(seed)
    rvecs = np.identity(3) * 50.0
    pos = rng.uniform(0.0, 50.0, (n, 3))
    masses = mass * rng.uniform(0.5, 2.0, n)

    def force(p, r):
        return 0.0, np.zeros_like(p), np.zeros((3, 3))
    return pos, rvecs, masses, force


def test_langevin_thermostat_reaches_the_stationary_velocity_variance():
    # Zero force: each velocity component is an Ornstein-Uhlenbeck chain v <- c1 v + c2 xi
    # (twice per step) whose stationary variance is c2^2 / (1 - c1^2) = k_B T / m exactly.
    temp = 300.0
    pos, rvecs, masses, force = _free_particles(400, 1)
    np.random.seed(2)
    md = OracleMD(force, pos, rvecs, masses, 1.0 * femtosecond, temp0=30.0,
                  thermo=dict(kind="langevin", temp=temp, timecon=5 * femtosecond, seed=11))
    md.run(60)  # 12 time constants
    acc = []
    for _ in range(300):
        md.run(1)
        acc.append((md.vel ** 2 * masses[:, None]).mean() / (boltzmann * temp))
    acc = np.array(acc)
    # about 30 independent samples of 1200 squared normals each: relative error ~ 0.7 %
    assert abs(acc.mean() - 1.0) < 0.03, acc.mean()
    assert md.ndof == pos.size


def test_langevin_thermostat_econs_correction_telescopes():
    # With zero force the energy changes only through the thermostat, so etot plus the
    # accumulated correction (ekin before - ekin after of every call) stays at its start value.
    pos, rvecs, masses, force = _free_particles(50, 3)
    np.random.seed(4)
    md = OracleMD(force, pos, rvecs, masses, 2.0 * femtosecond, temp0=300.0,
                  thermo=dict(kind="langevin", temp=600.0, timecon=20 * femtosecond, seed=5))
    e0 = md.econs
    ekin0 = md.ekin
    md.run(200)
    assert md.ekin > 1.5 * ekin0  # the thermostat did heat the system
    assert abs(md.econs - e0) <= 1e-12 * abs(e0)


# The barostat mass (ndof + 3) / 3 k_B T (timecon / 2 pi)^2 (npt.py:358-365) does not grow
# with the node masses; for these nanocells the cell oscillates with a period of about
# timecon / 1000, so timecon = 20 ps gives about 20 steps of 2 fs per period.
BARO_TIMECON = 20000 * femtosecond


def _small_solid(seed=3, strain=0.0, p_b=0.3):
    from oracle import mmff as omm
    from micmec_amd.topology import Grid
    from micmec_amd.workloads import make_system

    s = make_system((3, 3, 3), p_b=p_b, p_hole=0.0, seed=seed, strain=strain)
    wb = Grid(np.rint(s.grid).astype(int), pbc=True).wrap_bits()

    def force(p, r):
        return omm.compute_vectorised(p, r, s.types, s.surrounding_nodes,
                                      s.surrounding_cells, wb, s.params)
    return s, force


@pytest.mark.parametrize("aniso,volcon", [(True, True), (False, False), (True, False)])
def test_langevin_barostat_invariants(aniso, volcon):
    from micmec_amd import units

    s, force = _small_solid(seed=6, strain=0.01)
    np.random.seed(8)
    md = OracleMD(force, s.pos, np.array(s.domain.rvecs), s.masses, 2.0 * femtosecond,
                  temp0=300.0,
                  thermo=dict(kind="langevin", temp=300.0, timecon=50 * femtosecond, seed=1),
                  baro=dict(kind="langevin", temp=300.0, press=0.1 * units.gigapascal,
                            timecon=BARO_TIMECON, anisotropic=aniso, vol_constraint=volcon,
                            seed=2))
    rv0 = md.rvecs.copy()
    v0 = abs(np.linalg.det(rv0))
    e0 = md.econs
    dev = []
    for _ in range(20):
        md.run(5)
        dev.append(abs(md.econs - e0))
    v = abs(np.linalg.det(md.rvecs))
    assert md.ndof == md.pos.size  # TBCombination.init with stochastic hooks
    if volcon:
        # tr(v_g) = 0 after the projection: det exp(v_g dt / 2) = 1
        assert abs(v - v0) <= 1e-12 * v0
    else:
        assert abs(v - v0) > 1e-6 * v0  # the cell did move
    if not aniso:
        # isotropic: rvecs stay proportional to the initial ones
        ratio = md.rvecs / rv0
        ratio = ratio[np.abs(rv0) > 1e-8 * np.abs(rv0).max()]
        assert np.ptp(ratio) <= 1e-12 * abs(ratio.mean())
    # econs drifts only by the Verlet integration error (the hooks' corrections cancel the energy
    # they add), far below the energy exchanged with the hooks
    exchanged = abs(md.econs_thermo) + abs(md.econs_baro)
    assert max(dev) < 1e-3 * exchanged, (max(dev), exchanged)


def test_langevin_barostat_reaches_the_target_pressure():
    # isotropic Langevin piston at P: the time average of the internal pressure equals P
    # (less k_B T / V, negligible here); starting from a strained cell far from it
    from micmec_amd import units

    s, force = _small_solid(seed=9, strain=0.02, p_b=0.0)
    P = 1.0 * units.gigapascal
    np.random.seed(10)
    md = OracleMD(force, s.pos, np.array(s.domain.rvecs), s.masses, 5.0 * femtosecond,
                  temp0=300.0,
                  thermo=dict(kind="langevin", temp=300.0, timecon=100 * femtosecond, seed=3),
                  baro=dict(kind="langevin", temp=300.0, press=P, timecon=2.5 * BARO_TIMECON,
                            anisotropic=False, seed=4))
    p_start = md.press
    md.run(200)
    samples = []
    for _ in range(800):
        md.run(1)
        samples.append(md.press)
    samples = np.array(samples)
    blocks = samples.reshape(8, -1).mean(axis=1)
    err = blocks.std(ddof=1) / np.sqrt(len(blocks))
    assert abs(p_start - P) > 0.5 * P
    assert abs(samples.mean() - P) < max(5 * err, 0.05 * P), (samples.mean() / P, err / P)
