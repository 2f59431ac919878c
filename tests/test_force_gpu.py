"""GPU parity of the force field: HIP kernels vs the CPU oracle and the golden vectors.

Bar: energy, node gradient and virial within 1e-10 relative (max-norm, ||d||inf / ||ref||inf).
"""

import os

import numpy as np
import pytest
import torch

from micmec_amd import units
from micmec_amd.pes import ForcePartMechanical, MicMecForceField, ForcePartPressure
from micmec_amd.system import System
from micmec_amd.topology import Grid, build_system, build_type
from micmec_amd.workloads import CELL_MASS, cubic_elasticity, make_system, type_a, type_b
from oracle import mmff as omm

from golden.make_golden import load_system

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = np.load(os.path.join(HERE, "golden", "force_cases.npz"))
TOL = 1e-10


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b)))


def gpu_eval(system, model):
    fpm = ForcePartMechanical(system, model=model)
    mmf = MicMecForceField(system, [fpm])
    gpos = np.zeros(system.pos.shape)
    vtens = np.zeros((3, 3))
    e = mmf.compute(gpos, vtens)
    return e, gpos, vtens


def oracle_eval(system, model):
    wb = Grid(np.rint(np.asarray(system.grid, dtype=float)).astype(int),
              pbc=len(system.domain.rvecs) > 0).wrap_bits()
    return omm.compute_vectorised(system.pos, system.domain.rvecs, system.types,
                                  system.surrounding_nodes, system.surrounding_cells, wb,
                                  system.params, model)


@pytest.mark.parametrize("case", [str(c) for c in GOLDEN["cases"]])
@pytest.mark.parametrize("model", ["corner8", "original"])
def test_golden_cases(case, model):
    print("case", case, model, flush=True)
    s = load_system(GOLDEN, case + "/")
    e, g, v = gpu_eval(s, model)
    assert rel(e, GOLDEN[case + "/" + model + ":energy"]) < TOL
    assert rel(g, GOLDEN[case + "/" + model + ":gpos"]) < TOL
    assert rel(v, GOLDEN[case + "/" + model + ":vtens"]) < TOL


@pytest.mark.parametrize("fn", sorted(f for f in os.listdir(os.path.join(HERE, "golden", "chk"))))
def test_reference_fixtures(fn):
    print("fixture", fn, flush=True)
    s = System.from_file(os.path.join(HERE, "golden", "chk", fn))
    s.pos = s.pos + np.random.default_rng(17).normal(0, 0.08, s.pos.shape)
    for model in ("corner8", "original"):
        e, g, v = gpu_eval(s, model)
        e0, g0, v0 = oracle_eval(s, model)
        assert rel(e, e0) < TOL and rel(g, g0) < TOL and rel(v, v0) < TOL, (model, fn)


@pytest.mark.parametrize("shape,p_b,p_hole,seed", [
    ((3, 3, 3), 0.0, 0.0, 0),
    ((5, 4, 6), 0.3, 0.1, 21),
    ((8, 8, 8), 0.2, 0.05, 22),
    ((16, 16, 16), 0.2, 0.05, 23),
    ((2, 2, 2), 0.5, 0.0, 24),
])
def test_synthetic_multistate(shape, p_b, p_hole, seed):
    print("synthetic", shape, p_b, p_hole, seed, flush=True)
    s = make_system(shape, p_b=p_b, p_hole=p_hole, seed=seed, noise=0.1 * units.angstrom)
    for model in ("corner8", "original"):
        e, g, v = gpu_eval(s, model)
        e0, g0, v0 = oracle_eval(s, model)
        assert rel(e, e0) < TOL, (model, e, e0)
        assert rel(g, g0) < TOL, (model, rel(g, g0))
        assert rel(v, v0) < TOL, (model, rel(v, v0))


def many_state_system(shape, ns, seed):
    """Seeded grid of type A (one state), type B (two states) and a type C of ``ns`` states in
    about equal shares, 10% of the cells empty. C's states differ in h0 by under 1.5% and in
    free energy by 1e-4 Ha steps (kT at 300 K is 9.5e-4 Ha), so every state carries weight."""
    rng = np.random.default_rng(seed)
    grid = rng.integers(1, 4, size=shape).astype(np.int64)
    grid[rng.random(shape) < 0.1] = 0
    h0 = 20.8 * units.angstrom * np.identity(3)
    scales = [1.0, 0.995, 1.004, 0.99, 1.008, 0.985][:ns]
    C = cubic_elasticity(35.0, 19.0, 8.0)
    tc = build_type("C", CELL_MASS, [s * h0 for s in scales],
                    [(1.0 - 0.05 * i) * C for i in range(ns)],
                    free_energy=[1e-4 * i for i in range(ns)],
                    effective_temp=300.0 * units.kelvin)
    data = {1: type_a(), 2: type_b(), 3: tc}
    out = build_system({t: data[t] for t in sorted({int(t) for t in np.unique(grid) if t})},
                       grid, pbc=True)
    out["pos"] = out["pos"] + rng.normal(0.0, 0.1 * units.angstrom, out["pos"].shape)
    return System.from_dict(out)


@pytest.mark.parametrize("ns", [3, 4, 5])
def test_many_state_types(ns):
    """Types of three to five states beside one- and two-state types: the multistate lanes loop
    over more states (the energy-only instance keeps them all in LDS). Against the oracle
    (mmff.py:386-398)."""
    print("many states", ns, flush=True)
    s = many_state_system((9, 7, 8), ns, seed=40 + ns)
    assert len(s.params["type3/free_energy"]) == ns
    for model in ("corner8", "original"):
        e, g, v = gpu_eval(s, model)
        e0, g0, v0 = oracle_eval(s, model)
        assert rel(e, e0) < TOL, (model, e, e0)
        assert rel(g, g0) < TOL, (model, rel(g, g0))
        assert rel(v, v0) < TOL, (model, rel(v, v0))


@pytest.mark.parametrize("shape,p_b,p_hole,pbc,seed", [
    ((1, 1, 1), 0.0, 0.0, False, 30),   # one cell, open boundaries
    ((1, 1, 1), 1.0, 0.0, False, 31),   # one two-state cell
    ((6, 1, 5), 0.3, 0.1, False, 32),   # a single layer
    ((2, 3, 7), 0.4, 0.1, True, 33),    # two cells along x under periodicity
    ((7, 5, 4), 1.0, 0.0, True, 34),    # two-state cells only: no single-state blocks
    ((9, 9, 9), 0.2, 0.5, True, 35),    # half the cells empty, nodes dropped
    ((9, 9, 9), 0.2, 0.5, False, 36),
])
def test_edge_grids(shape, p_b, p_hole, pbc, seed):
    print("edge grid", shape, p_b, p_hole, pbc, seed, flush=True)
    s = make_system(shape, p_b=p_b, p_hole=p_hole, seed=seed, noise=0.1 * units.angstrom,
                    pbc=pbc)
    for model in ("corner8", "original"):
        e, g, v = gpu_eval(s, model)
        e0, g0, v0 = oracle_eval(s, model)
        assert rel(e, e0) < TOL, (model, e, e0)
        assert rel(g, g0) < TOL, (model, rel(g, g0))
        assert rel(v, v0) < TOL, (model, rel(v, v0))


def test_rest_state_is_zero():
    s = make_system((4, 4, 4), seed=1, noise=0.0)
    # At the lattice every corner matrix equals h0, so strains vanish up to the rounding of
    # h0^-1 (about 1e-16). Scales: C11 * V for energy and virial, C11 * |h0|^2 for gradients.
    e, g, v = gpu_eval(s, "corner8")
    c11v = 40.0 * units.gigapascal * s.domain.volume
    c11a = 40.0 * units.gigapascal * (20.7 * units.angstrom) ** 2
    assert abs(e) < 1e-20 * c11v
    assert np.max(np.abs(g)) < 1e-14 * c11a
    assert np.max(np.abs(v)) < 1e-14 * c11v


def test_accumulate_and_cache_semantics():
    s = make_system((3, 4, 3), p_b=0.5, seed=2)
    fpm = ForcePartMechanical(s)
    g1 = np.ones(s.pos.shape)
    v1 = np.ones((3, 3))
    e = fpm.compute(g1, v1)
    g2 = np.zeros(s.pos.shape)
    v2 = np.zeros((3, 3))
    e2 = fpm.compute(g2, v2)
    assert e == e2
    assert rel(g1 - 1.0, g2) < 1e-13
    assert np.array_equal(fpm.gpos, g2)
    assert np.array_equal(fpm.vtens, v2)
    fpm.update_pos(s.pos)
    assert np.isnan(fpm.energy) and np.isnan(fpm.gpos).all()
    # energy only
    assert fpm.compute() == e


def test_torch_device_outputs_match_numpy():
    s = make_system((4, 3, 5), p_b=0.4, p_hole=0.1, seed=3)
    fpm = ForcePartMechanical(s)
    g = torch.zeros((s.nnodes, 3), dtype=torch.float64, device="cuda")
    v = torch.zeros((3, 3), dtype=torch.float64, device="cuda")
    e = fpm.compute(g, v)
    g2 = np.zeros(s.pos.shape)
    v2 = np.zeros((3, 3))
    e2 = fpm.compute(g2, v2)
    assert e == e2
    assert np.array_equal(g.cpu().numpy(), g2)
    assert np.array_equal(v.cpu().numpy(), v2)


def test_bitwise_deterministic():
    s = make_system((12, 10, 14), p_b=0.3, p_hole=0.05, seed=4)
    fpm = ForcePartMechanical(s)
    outs = []
    for _ in range(3):
        g = np.zeros(s.pos.shape)
        v = np.zeros((3, 3))
        e = fpm.compute(g, v)
        outs.append((e, g, v))
    for e, g, v in outs[1:]:
        assert e == outs[0][0]
        assert np.array_equal(g, outs[0][1]) and np.array_equal(v, outs[0][2])


def test_nan_raises_value_error():
    s = make_system((3, 3, 3), seed=5)
    fpm = ForcePartMechanical(s)
    s.pos[4, 1] = np.nan
    with pytest.raises(ValueError):
        fpm.compute(np.zeros(s.pos.shape), np.zeros((3, 3)))
    # the device flag path of the asynchronous entry point reports it as well
    pos = torch.from_numpy(s.pos).cuda()
    en = torch.zeros(1, dtype=torch.float64, device="cuda")
    gp = torch.zeros((s.nnodes, 3), dtype=torch.float64, device="cuda")
    fpm.compute_device(pos, fpm.rvecs_device(), en, gp, None)
    with pytest.raises(ValueError):
        fpm.field.check()


def test_pressure_part():
    s = make_system((3, 3, 3), seed=6)
    p = 1e3 * units.bar
    fpm = ForcePartMechanical(s)
    fpp = ForcePartPressure(s, p)
    mmf = MicMecForceField(s, [fpm, fpp])
    g = np.zeros(s.pos.shape)
    v = np.zeros((3, 3))
    e = mmf.compute(g, v)
    g0 = np.zeros(s.pos.shape)
    v0 = np.zeros((3, 3))
    e0 = fpm.compute(g0, v0)
    V = s.domain.volume
    assert abs(e - (e0 + p * V)) <= 1e-14 * abs(e)
    assert rel(v, v0 + p * V * np.identity(3)) < 1e-14
    assert np.array_equal(g, g0)


def test_partial_face_contact_is_not_folded():
    """Two consecutive cells that share only one node (corner 3 of the first is corner 0 of the
    second, the other upper/lower corners differ) must not have their faces folded together:
    the device's gradient of every node equals the oracle's."""
    from micmec_amd.pes.mmff import DeviceField
    from micmec_amd.workloads import type_a
    from oracle import cforce
    print("partial face contact", flush=True)
    sn = np.array([[0, 1, 2, 3, 4, 5, 6, 7],
                   [3, 8, 9, 10, 11, 12, 13, 14]], dtype=np.int64)
    sc = -np.ones((15, 8), dtype=np.int64)
    for c in range(2):
        for k in range(8):
            sc[sn[c, k], k] = c
    offs = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1),
                     (0, 1, 1), (1, 1, 1)], float)
    a = 20.7 * units.angstrom
    rng = np.random.default_rng(3)
    pos = np.zeros((15, 3))
    pos[sn[0]] = offs * a
    pos[sn[1]] = pos[3] + offs * a + np.array([0.3, -0.2, 0.1]) * a
    pos += rng.normal(0, 0.05 * units.angstrom, pos.shape)
    t = type_a()
    params = {"type1/" + k: v for k, v in t.items()}
    types = np.array([1, 1])
    tab = cforce.ParamTable(params, types)
    e0, g0, v0 = cforce.compute_c(pos, np.zeros((0, 3)), types, sn, sc, np.zeros(2, np.uint8),
                                  params, "corner8")
    for model in ("corner8",):
        f = DeviceField(sn, sc, np.zeros(2, np.uint8), tab.cell_row, tab.nstates,
                        tab.h0.reshape(-1, tab.smax, 3, 3), tab.c0.reshape(-1, tab.smax, 3, 3, 3, 3),
                        tab.free, tab.temp, model=model)
        kw = dict(dtype=torch.float64, device=f.device)
        P = torch.tensor(pos, **kw)
        E = torch.zeros(1, **kw)
        G = torch.zeros((15, 3), **kw)
        V = torch.zeros((3, 3), **kw)
        f.compute(P, torch.zeros((3, 3), **kw), E, G, V)
        torch.cuda.synchronize()
        assert rel(E.item(), e0) < TOL
        assert rel(G.cpu().numpy(), g0) < TOL
        assert rel(V.cpu().numpy(), v0) < TOL
