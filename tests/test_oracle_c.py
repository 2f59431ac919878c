"""Pins of the C restatement of the force oracle (oracle/c/mm_oracle.c via oracle/cforce.py).

It is checked against the pinned numpy oracle, not against the kernels: the golden vectors
written by the literal per-cell loop (every case, both models), the vectorised numpy oracle on
random grids with holes, two types and open boundaries, and directly against the closed forms
(zero at rest, the affine-strain energy ncells det(h0) eps:C:eps/2).
"""

import os
import subprocess

import numpy as np
import pytest

from micmec_amd import units
from micmec_amd.topology import Grid
from micmec_amd.workloads import cubic_elasticity, make_system
from oracle import cforce
from oracle import mmff as omm

from golden.make_golden import load_system

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = np.load(os.path.join(HERE, "golden", "force_cases.npz"))


@pytest.fixture(scope="module", autouse=True)
def built():
    if not cforce.available():
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle", "c")], check=True,
                       stdout=subprocess.DEVNULL)


def _wb(s):
    return Grid(np.rint(np.asarray(s.grid, float)).astype(int),
                pbc=len(s.domain.rvecs) > 0).wrap_bits()


def _c(s, model, pos=None, rvecs=None):
    return cforce.compute_c(s.pos if pos is None else pos,
                            s.domain.rvecs if rvecs is None else rvecs, s.types,
                            s.surrounding_nodes, s.surrounding_cells, _wb(s), s.params, model)


def _close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("case", [str(c) for c in GOLDEN["cases"]])
@pytest.mark.parametrize("model", ["corner8", "original"])
def test_c_oracle_reproduces_golden(case, model):
    s = load_system(GOLDEN, case + "/")
    e, g, v = _c(s, model)
    key = case + "/" + model + ":"
    assert _close(e, GOLDEN[key + "energy"], 1e-12)
    assert _close(g, GOLDEN[key + "gpos"], 1e-12)
    # the virial is a sum of g (x) r over absolute positions (|r| ~ 100 bohr, cells near rest):
    # its rounding differs between summation orders by up to ~1e-11 of the result
    assert _close(v, GOLDEN[key + "vtens"], 1e-10)


@pytest.mark.parametrize("model", ["corner8", "original"])
@pytest.mark.parametrize("shape,pbc,seed", [((6, 5, 7), True, 1), ((5, 4, 3), False, 2),
                                            ((2, 2, 2), True, 3)])
def test_c_oracle_matches_vectorised(model, shape, pbc, seed):
    s = make_system(shape, p_b=0.3, p_hole=0.15, seed=seed, pbc=pbc)
    e, g, v = _c(s, model)
    e0, g0, v0 = omm.compute_vectorised(s.pos, s.domain.rvecs, s.types, s.surrounding_nodes,
                                        s.surrounding_cells, _wb(s), s.params, model)
    assert _close(e, e0, 1e-12) and _close(g, g0, 1e-12) and _close(v, v0, 1e-10)


@pytest.mark.parametrize("model", ["corner8", "original"])
def test_c_oracle_closed_forms(model):
    s = make_system((3, 4, 3), seed=0, noise=0.0)
    e, g, v = _c(s, model)
    assert abs(e) < 1e-25 and np.abs(g).max() < 1e-14 and np.abs(v).max() < 1e-10
    rng = np.random.default_rng(5)
    A = np.identity(3) + rng.uniform(-0.03, 0.03, (3, 3))
    e, g, v = _c(s, model, pos=s.pos @ A, rvecs=s.domain.rvecs @ A)
    eps = 0.5 * (A @ A.T - np.identity(3))
    C = cubic_elasticity(40.0, 20.0, 10.0)
    e_cell = (20.7 * units.angstrom) ** 3 * 0.5 * np.einsum("ij,ijkl,kl", eps, C, eps)
    assert abs(e - s.ncells * e_cell) < 1e-12 * e
    assert np.abs(g).max() < 1e-10 * np.abs(v).max() / (20.7 * units.angstrom)


def test_c_oracle_force_callable_matches():
    s = make_system((4, 3, 5), p_b=0.3, p_hole=0.1, seed=31)
    f = cforce.make_force(s, "corner8")
    e, g, v = f(s.pos, np.array(s.domain.rvecs))
    e0, g0, v0 = _c(s, "corner8")
    assert e == e0 and np.array_equal(g, g0) and np.array_equal(v, v0)
