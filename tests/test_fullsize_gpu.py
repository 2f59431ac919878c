"""Checks at the full 128^3 size through periodic replication, which fixes the answer without
a CPU evaluation of the large system: a random periodic 16^3 configuration tiled 8 times along
every axis has 512 times the energy and virial of the small one, every node the gradient of
its image in the small one (mmff.py:288-323: the energy is a sum over cells, each cell sees only
its corner positions), and under NVE (no global coupling, verlet.py:140-166) every node follows
its image's trajectory shifted by the lattice vector between them."""

import numpy as np
import pytest

from micmec_amd import units
from micmec_amd.pes import ForcePartMechanical, MicMecForceField
from micmec_amd.sampling import VerletIntegrator
from micmec_amd.system import System
from micmec_amd.topology import build_system, build_system_device
from micmec_amd.workloads import make_system, type_a, type_b

pytestmark = pytest.mark.gpu
ANG = units.angstrom


def replicated(shape, k, p_b, p_hole, seed):
    """(small System, large System tiled k times per axis, image index of every large node)."""
    small = make_system(shape, p_b=p_b, p_hole=p_hole, seed=seed, noise=0.05 * ANG)
    grid = np.rint(small.grid).astype(np.int64)
    data = {1: type_a(), 2: type_b()}
    data = {t: data[t] for t in np.unique(grid) if t != 0}
    lat_s = build_system(data, grid, pbc=True)["pos"]
    big = build_system_device(data, np.tile(grid, (k, k, k)), pbc=True)
    lat_b = np.asarray(big["pos"])
    spacing = np.diag(np.asarray(small.domain.rvecs)) / np.array(shape)
    cs, cb = np.rint(lat_s / spacing).astype(np.int64), np.rint(lat_b / spacing).astype(np.int64)
    assert np.allclose(cs * spacing, lat_s) and np.allclose(cb * spacing, lat_b)
    lut = -np.ones(shape, dtype=np.int64)
    lut[cs[:, 0] % shape[0], cs[:, 1] % shape[1], cs[:, 2] % shape[2]] = np.arange(len(cs))
    image = lut[cb[:, 0] % shape[0], cb[:, 1] % shape[1], cb[:, 2] % shape[2]]
    assert (image >= 0).all() and len(lat_b) == k ** 3 * len(lat_s)
    big["pos"] = lat_b + (small.pos - lat_s)[image]
    big["rvecs"] = k * np.asarray(small.domain.rvecs)
    return small, System.from_dict(big), image, lat_b - lat_s[image]


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("model", ["corner8", "original"])
def test_forces_at_128_cubed_by_replication(model):
    print("replication forces", model, flush=True)
    k = 8
    small, big, image, _ = replicated((16, 16, 16), k, 0.2, 0.05, 41)
    out = []
    for s in (small, big):
        g = np.zeros(s.pos.shape)
        v = np.zeros((3, 3))
        e = ForcePartMechanical(s, model=model).compute(g, v)
        out.append((e, g, v))
    (es, gs, vs), (eb, gb, vb) = out
    # the large system's coordinates reach ~5000 bohr, so the corner differences it forms round
    # ~1e-11 relative to strain-sized gradients: the force tolerance of the parity contract
    assert abs(eb - k ** 3 * es) <= 1e-10 * abs(k ** 3 * es)
    assert rel(gb, gs[image]) < 1e-10
    assert rel(vb, k ** 3 * vs) < 1e-10


def test_nve_at_128_cubed_by_replication():
    print("replication nve", flush=True)
    k = 8
    small, big, image, shift = replicated((16, 16, 16), k, 0.2, 0.05, 42)
    runs = []
    for s in (small, big):
        np.random.seed(42)
        verlet = VerletIntegrator(MicMecForceField(s, [ForcePartMechanical(s)]),
                                  timestep=1.0 * units.femtosecond, temp0=300.0)
        runs.append(verlet)
    vs, vb = runs
    vb.vel = np.asarray(vs.vel)[image]
    vs.run(20)
    vb.run(20)
    assert np.abs(np.asarray(vb.pos) - (np.asarray(vs.pos)[image] + shift)).max() / ANG < 1e-8
    assert rel(vb.vel, np.asarray(vs.vel)[image]) < 1e-8
    assert abs(vb.epot - k ** 3 * vs.epot) <= 1e-10 * abs(k ** 3 * vs.epot)
    assert abs(vb.ekin - k ** 3 * vs.ekin) <= 1e-10 * abs(k ** 3 * vs.ekin)
