"""GPU MD parity: the device integrator against the CPU oracle integrator.

Same seeded system, same numpy random stream (velocities, chain and barostat velocities are
drawn in the reference's order by both sides), 100 steps of dt = 1 fs. Bar: every node position
within 1e-8 Angstrom, velocities, energies and hook state within tight relative tolerances.
"""

import numpy as np
import pytest
import torch

from micmec_amd import units
from micmec_amd.pes import ForcePartMechanical, ForcePartPressure, MicMecForceField
from micmec_amd.sampling import (MTKBarostat, NHCThermostat, VerletIntegrator, Hook)
from micmec_amd.topology import Grid
from micmec_amd.workloads import make_system
from oracle import mmff as omm
from oracle.integrators import OracleMD

pytestmark = pytest.mark.gpu

FS = units.femtosecond
ANG = units.angstrom


def oracle_force(system, model):
    wb = Grid(np.rint(system.grid).astype(int), pbc=len(system.domain.rvecs) > 0).wrap_bits()
    sn, sc, types, params = (system.surrounding_nodes, system.surrounding_cells, system.types,
                             system.params)

    def force(pos, rvecs):
        return omm.compute_vectorised(pos, rvecs, types, sn, sc, wb, params, model)
    return force


def build(shape, p_b, p_hole, seed, ensemble, model, ext_press=0.0, press=1.0 * units.bar,
          pbc=True):
    s = make_system(shape, p_b=p_b, p_hole=p_hole, seed=seed, pbc=pbc)
    pos0 = s.pos.copy()
    rvecs0 = np.array(s.domain.rvecs).copy()
    parts = [ForcePartMechanical(s, model=model)]
    if ext_press:
        parts.append(ForcePartPressure(s, ext_press))
    mmf = MicMecForceField(s, parts)
    np.random.seed(seed)
    hooks = []
    thermo = baro = None
    if "nhc" in ensemble:
        hooks.append(NHCThermostat(temp=300.0, timecon=100 * FS))
        thermo = dict(temp=300.0, timecon=100 * FS, chainlength=3)
    if "mtk" in ensemble:
        hooks.append(MTKBarostat(mmf, temp=300.0, press=press, timecon=1000 * FS))
        baro = dict(temp=300.0, press=press, timecon=1000 * FS)
    verlet = VerletIntegrator(mmf, timestep=1.0 * FS, hooks=hooks, temp0=300.0)
    np.random.seed(seed)
    omd = OracleMD(oracle_force(s, model), pos0, rvecs0, s.masses, 1.0 * FS, temp0=300.0,
                   thermo=thermo, baro=baro, ext_press=ext_press)
    return s, verlet, omd


def compare(verlet, omd, tag):
    dpos = np.abs(verlet.pos - omd.pos).max() / ANG
    assert dpos < 1e-8, (tag, "pos", dpos)
    vscale = np.abs(omd.vel).max()
    assert np.abs(verlet.vel - omd.vel).max() < 1e-8 * vscale, (tag, "vel")
    for name in ("epot", "ekin", "econs", "etot", "temp"):
        a, b = getattr(verlet, name), getattr(omd, name)
        assert abs(a - b) <= 1e-9 * abs(b) + 1e-15, (tag, name, a, b)
    # cons_err is a ratio of running variances; in NVE the econs variance is rounding noise
    assert abs(verlet.cons_err - omd.cons_err) <= 1e-5 * abs(omd.cons_err) + 1e-10, tag
    if omd.nvec > 0:
        assert np.abs(verlet.ptens - omd.ptens).max() <= 1e-8 * np.abs(omd.ptens).max(), tag
    assert abs(verlet.time - omd.time) < 1e-9 * omd.time + 1e-12


@pytest.mark.parametrize("ensemble", ["nve", "nvt-nhc", "npt-nhc-mtk", "nph-mtk"])
@pytest.mark.parametrize("model", ["corner8", "original"])
def test_trajectory_parity(ensemble, model):
    print("md parity", ensemble, model, flush=True)
    s, verlet, omd = build((4, 3, 5), 0.3, 0.1, 31, ensemble, model)
    compare(verlet, omd, "init")
    for chunk in range(4):
        verlet.run(25)
        omd.run(25)
        compare(verlet, omd, "step %d" % verlet.counter)
    if "nhc" in ensemble:
        from micmec_amd.sampling import TBCombination
        th = [h for h in verlet.hooks if isinstance(h, (NHCThermostat, TBCombination))][0]
        ch = th.thermostat.chain if isinstance(th, TBCombination) else th.chain
        assert np.abs(ch.vel - omd.chain.vel).max() < 1e-8 * np.abs(omd.chain.vel).max()
        assert np.abs(ch.pos - omd.chain.pos).max() < 1e-8 * np.abs(omd.chain.pos).max()
    if "mtk" in ensemble:
        assert np.abs(verlet.rvecs - omd.rvecs).max() / ANG < 1e-8
        assert np.abs(s.domain.rvecs - omd.rvecs).max() / ANG < 1e-8


@pytest.mark.parametrize("ensemble", ["nve", "nvt-nhc"])
def test_open_boundary_trajectory_parity(ensemble):
    """A free-standing piece (no domain vectors: ndof 3N - 6, no wrap bits, no pressure)."""
    print("md open boundaries", ensemble, flush=True)
    s, verlet, omd = build((4, 3, 3), 0.3, 0.05, 8, ensemble, "corner8", pbc=False)
    assert s.domain.nvec == 0
    compare(verlet, omd, "init")
    for chunk in range(3):
        verlet.run(25)
        omd.run(25)
        compare(verlet, omd, "step %d" % verlet.counter)


def test_nve_config0_1000_steps():
    """Plain method, CPU-runnable case: 3^3 periodic grid, one type, NVE for 1000 steps."""
    s, verlet, omd = build((3, 3, 3), 0.0, 0.0, 0, "nve", "corner8")
    verlet.run(1000)
    omd.run(1000)
    compare(verlet, omd, "1000 steps")


def test_pressure_part_in_md():
    s, verlet, omd = build((3, 3, 3), 0.5, 0.0, 9, "npt-nhc-mtk", "corner8",
                           ext_press=50 * units.bar, press=0.0)
    verlet.run(40)
    omd.run(40)
    compare(verlet, omd, "pressure part")


def test_graph_and_eager_agree_bitwise():
    _, v1, _ = build((4, 4, 3), 0.3, 0.05, 5, "npt-nhc-mtk", "corner8")
    _, v2, _ = build((4, 4, 3), 0.3, 0.05, 5, "npt-nhc-mtk", "corner8")
    v2.use_graph = False
    v1.run(30)
    v2.run(30)
    assert np.array_equal(v1.pos, v2.pos) and np.array_equal(v1.vel, v2.vel)
    assert v1.econs == v2.econs


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("ensemble", ["npt-nhc-mtk", "nph-mtk", "nvt-nhc"])
def test_steps_of_one_call_match_single_steps_bitwise(ensemble, use_graph):
    """Every step of an mm_md_step call but the last ends with the next step's opening scalar
    phase, and with a barostat keeps the node gradients of its main and post-barostat force
    evaluations out of HBM (nothing reads them before the next step's pre-barostat
    evaluation). One 30-step call must equal 30 one-step calls (all full steps) bit for bit:
    positions, velocities, the gradient after the run, every scalar."""
    print("lean steps", ensemble, use_graph, flush=True)
    _, v1, _ = build((4, 5, 3), 0.3, 0.05, 8, ensemble, "corner8")
    _, v2, _ = build((4, 5, 3), 0.3, 0.05, 8, ensemble, "corner8")
    v1.use_graph = v2.use_graph = use_graph
    v1.run(30)
    for _ in range(30):
        v2.propagate()
    assert v1.counter == v2.counter == 30
    assert np.array_equal(v1.pos, v2.pos) and np.array_equal(v1.vel, v2.vel)
    assert np.array_equal(v1.gpos, v2.gpos)
    assert np.array_equal(v1._engine.get_scalars(), v2._engine.get_scalars())
    assert v1.econs == v2.econs and v1.rmsd_gpos == v2.rmsd_gpos


def test_graph_recaptured_when_state_counts_change():
    """New state counts (mm_field_set_params) move the cell pass's split between the
    multistate and single-state blocks; the captured step graph is rebuilt, so graph and eager
    runs still agree bit for bit (a stale graph would skip the blocks that changed kind)."""
    import ctypes
    from micmec_amd import _native
    from micmec_amd.adjoint import voigt_from_tensor
    from micmec_amd.pes.mmff import parse_type_params
    runs = []
    for graph in (True, False):
        s, v, _ = build((4, 4, 3), 0.3, 0.05, 6, "nve", "corner8")
        v.use_graph = graph
        v.run(5)
        _, ns, h0, c0, free, temp = parse_type_params(s.params, s.types)
        assert ns.max() == 2
        cv = np.array([[voigt_from_tensor(c) for c in row] for row in c0])
        arrays = [np.ascontiguousarray(a, dtype=t) for a, t in
                  ((np.ones_like(ns), np.int32), (h0, np.float64), (cv, np.float64),
                   (free, np.float64), (temp, np.float64))]
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        rc = _native.lib().mm_field_set_params(v._engine.fpm.field._handle,
                                               *[vp(a) for a in arrays], float(units.boltzmann),
                                               _native.stream_ptr())
        assert rc == 0
        v.run(5)
        runs.append(v)
    assert np.array_equal(runs[0].pos, runs[1].pos) and np.array_equal(runs[0].vel, runs[1].vel)
    assert runs[0].epot == runs[1].epot


class Recorder(Hook):
    def __init__(self, start, step):
        Hook.__init__(self, start, step)
        self.seen = []

    def __call__(self, iterative):
        self.seen.append((iterative.counter, iterative.epot, iterative.rmsd_delta,
                          iterative.mmf.system.pos.copy()))


def test_host_hooks_cadence_and_state():
    s = make_system((3, 4, 3), p_b=0.3, seed=3)
    mmf = MicMecForceField(s, [ForcePartMechanical(s)])
    np.random.seed(3)
    rec = Recorder(start=2, step=7)
    verlet = VerletIntegrator(mmf, timestep=1.0 * FS, hooks=[rec, NHCThermostat(300.0)])
    verlet.run(30)
    assert [c for c, *_ in rec.seen] == [2, 9, 16, 23, 30]
    # the state seen by the hook is the state after that step
    np.random.seed(3)
    s2 = make_system((3, 4, 3), p_b=0.3, seed=3)
    mmf2 = MicMecForceField(s2, [ForcePartMechanical(s2)])
    v2 = VerletIntegrator(mmf2, timestep=1.0 * FS, hooks=[NHCThermostat(300.0)])
    for c, epot, rmsd_delta, pos in rec.seen:
        while v2.counter < c:
            prev = v2.pos.copy()
            v2.propagate()
        assert v2.epot == epot
        assert np.array_equal(v2.pos, pos)
        assert abs(rmsd_delta - np.sqrt(np.mean((pos - prev) ** 2))) < 1e-12 * rmsd_delta


def test_nan_in_md_raises():
    s = make_system((3, 3, 3), seed=4)
    mmf = MicMecForceField(s, [ForcePartMechanical(s)])
    np.random.seed(4)
    verlet = VerletIntegrator(mmf, timestep=1.0 * FS)
    verlet._engine.pos[3, 0] = float("nan")
    with pytest.raises(ValueError):
        verlet.run(2)


def test_hdf5_writer_records_md_frames():
    """HDF5Writer on the device integrator (trajectory.py:31-87): one row per due step, every
    state item, the values the integrator holds at that step."""
    from h5mem import Group
    from micmec_amd.sampling import HDF5Writer

    s, verlet, omd = build((3, 3, 4), 0.3, 0.0, 17, "npt-nhc-mtk", "corner8")
    f = Group()
    verlet.hooks.append(HDF5Writer(f, step=10))
    verlet.run(40)
    omd.run(40)
    traj = f["trajectory"]
    assert list(traj["counter"][:]) == [10, 20, 30, 40]
    assert np.array_equal(traj["pos"][3], verlet.pos)
    assert traj["epot"][3] == verlet.epot and traj["econs"][3] == verlet.econs
    assert np.abs(traj["pos"][3] - omd.pos).max() / ANG < 1e-8
    assert np.array_equal(traj["domain"][3], verlet.rvecs)
    assert traj["temp"].shape == (4,) and traj["ptens"].shape == (4, 3, 3)
    assert np.array_equal(f["system"]["masses"][:], s.masses)


@pytest.mark.parametrize("ensemble", ["nve", "npt-nhc-mtk"])
def test_positions_written_through_after_run(ensemble):
    """After ``run(n)`` with no host hook due, the system holds the integrator's positions and
    domain, and a fresh force evaluation gives the integrator's energy (the reference writes
    both through on every step: verlet.py:147-151, mmff.py:191-197)."""
    print("write-through", ensemble, flush=True)
    s, verlet, omd = build((4, 3, 5), 0.3, 0.1, 31, ensemble, "corner8", ext_press=(
        50 * units.bar if "mtk" in ensemble else 0.0))
    mmf = verlet.mmf
    pos_before = np.array(s.pos, copy=True)
    verlet.run(7)
    omd.run(7)
    assert not np.array_equal(s.pos, pos_before)
    assert np.array_equal(s.pos, verlet.pos)
    assert np.abs(s.pos - omd.pos).max() / ANG < 1e-8
    assert np.array_equal(np.asarray(s.domain.rvecs), verlet.rvecs)
    # caches: invalidated, energies of the last evaluation kept
    assert mmf.energy == verlet.epot
    assert np.isnan(mmf.gpos).all() and np.isnan(mmf.part_micmec.vtens).all()
    e_parts = sum(p.energy for p in mmf.parts)
    assert abs(e_parts - verlet.epot) <= 1e-12 * abs(verlet.epot)
    gpos = np.zeros_like(s.pos)
    e = mmf.compute(gpos)
    assert abs(e - verlet.epot) <= 1e-12 * abs(verlet.epot), (e, verlet.epot)
    assert np.abs(gpos - verlet.gpos).max() <= 1e-12 * np.abs(verlet.gpos).max()
    # a later run publishes again; an assignment in between is overwritten by it
    s.pos = s.pos + 1.0
    verlet.run(3)
    assert np.array_equal(s.pos, verlet.pos)


@pytest.mark.parametrize("variant", ["isotropic", "vol_constraint", "baro_thermo",
                                     "isotropic-nph", "baro_thermo-nph"])
def test_mtk_variants_trajectory_parity(variant):
    """MTKBarostat(anisotropic=False / vol_constraint=True / baro_thermo=NHCThermostat(...)),
    with or without a particle NHC thermostat: 100 steps against the oracle (npt.py:518-757)."""
    print("mtk variant", variant, flush=True)
    s = make_system((4, 3, 5), p_b=0.3, p_hole=0.1, seed=37)
    pos0, rvecs0 = s.pos.copy(), np.array(s.domain.rvecs).copy()
    mmf = MicMecForceField(s, [ForcePartMechanical(s)])
    kw, okw = {}, {}
    if variant.startswith("isotropic"):
        kw["anisotropic"] = okw["anisotropic"] = False
    if variant.startswith("vol_constraint"):
        kw["vol_constraint"] = okw["vol_constraint"] = True
    if variant.startswith("baro_thermo"):
        kw["baro_thermo"] = NHCThermostat(temp=300.0, timecon=500 * FS)
        okw["baro_thermo"] = dict(temp=300.0, timecon=500 * FS, chainlength=3)
    nph = variant.endswith("-nph")
    np.random.seed(37)
    hooks = [] if nph else [NHCThermostat(temp=300.0, timecon=100 * FS)]
    baro = MTKBarostat(mmf, temp=300.0, press=100 * units.bar, timecon=1000 * FS, **kw)
    hooks.append(baro)
    verlet = VerletIntegrator(mmf, timestep=1.0 * FS, hooks=hooks, temp0=300.0)
    np.random.seed(37)
    omd = OracleMD(oracle_force(s, "corner8"), pos0, rvecs0, s.masses, 1.0 * FS, temp0=300.0,
                   thermo=None if nph else dict(temp=300.0, timecon=100 * FS, chainlength=3),
                   baro=dict(temp=300.0, press=100 * units.bar, timecon=1000 * FS, **okw))
    compare(verlet, omd, variant + " init")
    verlet.run(100)
    omd.run(100)
    compare(verlet, omd, variant + " 100 steps")
    assert np.abs(verlet.rvecs - omd.rvecs).max() <= 1e-12 * np.abs(omd.rvecs).max()
    if variant.startswith("baro_thermo"):
        assert np.abs(baro.baro_thermo.chain.vel - omd.bchain.vel).max() <= \
            1e-9 * np.abs(omd.bchain.vel).max()
    if variant.startswith("vol_constraint"):
        assert abs(verlet.mmf.system.domain.volume / abs(np.linalg.det(rvecs0)) - 1.0) < 1e-12
    if variant.startswith("isotropic"):
        assert np.ndim(baro.vel_press) == 0
        assert abs(baro.vel_press - omd.vel_press) <= 1e-9 * abs(omd.vel_press)
