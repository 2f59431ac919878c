"""The tiled cell pass (one workgroup per box of cells, node sums in LDS) against the untiled
kernels (one lane per cell, vertex gradients in HBM) and the CPU oracle: energy, node gradient
and virial of single systems (both models, multistate types, holes, non-periodic grids, grid
sizes that are no multiple of the tile, tile shapes down to 2 x 2 x 2) and of batches; a
tiled MD trajectory against the untiled one."""

import ctypes

import numpy as np
import pytest
import torch

from micmec_amd import _native, units
from micmec_amd.pes import ForcePartMechanical, MicMecForceField
from micmec_amd.pes.mmff import enable_tiles, grid_tables
from micmec_amd.topology import Grid
from micmec_amd.workloads import make_system
from oracle import mmff as omm

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b)))


def evaluate(system, model, tiles, shapes=None):
    fpm = ForcePartMechanical(system, model=model, tiles=tiles and shapes is None)
    if tiles and shapes is not None:
        pbc = np.asarray(system.domain.rvecs).shape[0] > 0
        fpm.field.tile_shape = enable_tiles(fpm.field._lib, fpm.field.handle,
                                            grid_tables([system], [pbc]), shapes)
    mmf = MicMecForceField(system, [fpm])
    gpos = np.zeros(system.pos.shape)
    vtens = np.zeros((3, 3))
    e = mmf.compute(gpos, vtens)
    e_only = mmf.compute()
    return fpm, e, gpos, vtens, e_only


def oracle_eval(system, model):
    wb = Grid(np.rint(np.asarray(system.grid, dtype=float)).astype(int),
              pbc=len(system.domain.rvecs) > 0).wrap_bits()
    return omm.compute_vectorised(system.pos, system.domain.rvecs, system.types,
                                  system.surrounding_nodes, system.surrounding_cells, wb,
                                  system.params, model)


def tiles_info(fpm):
    en, nt, lanes, ent = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int64()
    rc = _native.lib().mm_field_tiles_info(fpm.field.handle, ctypes.byref(en), ctypes.byref(nt),
                                           ctypes.byref(lanes), ctypes.byref(ent))
    _native.check(rc, "mm_field_tiles_info")
    return en.value, nt.value, lanes.value, ent.value


CASES = [
    ((16, 16, 16), 0.2, 0.05, 1, 0.02, True),
    ((9, 5, 11), 0.4, 0.1, 2, 0.03, True),
    ((8, 8, 8), 0.0, 0.0, 3, 0.01, True),
    ((7, 6, 5), 0.5, 0.2, 4, 0.02, False),
    ((2, 3, 2), 0.5, 0.0, 5, 0.02, True),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("model", ["corner8", "original"])
def test_tiled_matches_untiled_and_oracle(case, model):
    shape, p_b, p_hole, seed, strain, periodic = CASES[case]
    print("tiles", shape, p_b, p_hole, periodic, model, flush=True)
    s = make_system(shape, p_b=p_b, p_hole=p_hole, seed=seed, strain=strain,
                    noise=0.1 * units.angstrom, pbc=periodic)
    fpm, e, g, v, e1 = evaluate(s, model, True)
    assert fpm.field.tile_shape is not None
    en, nt, lanes, ent = tiles_info(fpm)
    assert en == 1 and nt > 0 and lanes <= 1024
    assert ent >= s.nnodes
    _, e0, g0, v0, e10 = evaluate(s, model, False)
    assert rel(e, e0) < 1e-13 and rel(e1, e10) < 1e-13 and rel(e1, e) < 1e-13
    assert rel(g, g0) < 1e-12
    assert rel(v, v0) < 1e-12
    eo, go, vo = oracle_eval(s, model)
    assert rel(e, eo) < 1e-10 and rel(g, go) < 1e-10 and rel(v, vo) < 1e-10


@pytest.mark.parametrize("shape", [(2, 2, 2), (4, 2, 8), (3, 5, 7)])
def test_tile_shapes(shape):
    s = make_system((10, 9, 7), p_b=0.3, p_hole=0.1, seed=6, strain=0.02,
                    noise=0.1 * units.angstrom)
    fpm, e, g, v, _ = evaluate(s, "corner8", True, shapes=(shape,))
    assert fpm.field.tile_shape == shape
    _, e0, g0, v0, _ = evaluate(s, "corner8", False)
    assert rel(e, e0) < 1e-13 and rel(g, g0) < 1e-12 and rel(v, v0) < 1e-12


def test_inconsistent_grid_keeps_the_untiled_kernels():
    s = make_system((4, 4, 4), p_b=0.3, seed=7, strain=0.02)
    fpm = ForcePartMechanical(s, tiles=False)
    bad = grid_tables([s], [True])
    bad[0][[0, 1]] = bad[0][[1, 0]]  # two cells swapped: corners no longer match the grid
    assert enable_tiles(fpm.field._lib, fpm.field.handle, bad) is None
    assert tiles_info(fpm)[0] == 0
    assert b"grid point" in _native.lib().mm_last_error() or \
        b"wrap" in _native.lib().mm_last_error()


def test_batched_tiled_matches_untiled():
    from micmec_amd.adjoint import DiffField

    systems = [make_system((n, n + 1, n), p_b=0.3, p_hole=0.05, seed=10 + n, strain=0.02,
                           noise=0.05 * units.angstrom) for n in (3, 5, 8, 9)]
    out = []
    for tiles in (True, False):
        df = DiffField(systems, tiles=tiles)
        assert (df.tile_shape is not None) == tiles
        pos, rv = df.positions()
        out.append([x.detach().cpu().numpy() for x in df.compute(pos, rv)])
    for a, b in zip(out[0], out[1]):
        assert rel(a, b) < 1e-12


def test_tiled_md_matches_untiled_md():
    from micmec_amd.sampling import MTKBarostat, NHCThermostat, VerletIntegrator

    runs = []
    for tiles in (True, False):
        s = make_system((6, 5, 7), p_b=0.3, p_hole=0.05, seed=8, strain=0.01)
        mmf = MicMecForceField(s, [ForcePartMechanical(s, tiles=tiles)])
        np.random.seed(8)
        hooks = [NHCThermostat(temp=300.0, timecon=100 * units.femtosecond),
                 MTKBarostat(mmf, temp=300.0, press=0.1 * units.gigapascal,
                             timecon=10000 * units.femtosecond)]
        v = VerletIntegrator(mmf, timestep=1.0 * units.femtosecond, hooks=hooks, temp0=300.0)
        v.run(50)
        runs.append((v.pos.copy(), v.econs, v.rvecs.copy()))
    (p1, e1, r1), (p0, e0, r0) = runs
    assert np.abs(p1 - p0).max() / units.angstrom < 1e-9
    assert abs(e1 - e0) <= 1e-10 * abs(e0)
    assert np.abs(r1 - r0).max() / units.angstrom < 1e-9
