"""Pins of the CPU oracle's force field (oracle/nanocell.py, oracle/mmff.py).

Each check is chosen so that a plausible restatement mistake (a dropped term, a sign, an index,
a transposed operand, a wrong prefactor) fails one of them:

* closed forms: zero at rest; under an affine map every corner matrix is h A and the energy is
  ncells det(h0) eps:C:eps/2 with eps = (A A^T - I)/2; the textbook uniaxial value C11 eps^2/2;
* central finite differences of the oracle's own energy for gpos and for the virial
  (vtens = dE/dD at D = I under pos -> pos D, rvecs -> rvecs D, micmec/pes/mmff_jax.py:249-328);
* invariants: translation, global rotation, sum of node gradients, virial symmetry;
* the multistate free energy against scipy's logsumexp of independently evaluated states;
* the literal per-cell loop against the vectorised form, and both against the golden vectors.
"""

import os

import numpy as np
import pytest
from scipy.special import logsumexp
from scipy.spatial.transform import Rotation

from micmec_amd import units
from micmec_amd.topology import Grid
from micmec_amd.workloads import cubic_elasticity, make_system
from oracle import mmff as omm
from oracle import nanocell as onc
from oracle.constants import boltzmann

from golden.make_golden import load_system

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = np.load(os.path.join(HERE, "golden", "force_cases.npz"))


def sysdict(s):
    return {"pos": s.pos, "rvecs": s.domain.rvecs, "grid": s.grid, "types": s.types,
            "surrounding_nodes": s.surrounding_nodes, "surrounding_cells": s.surrounding_cells,
            "params": s.params}


def vec_eval(s, model, pos=None, rvecs=None):
    wb = Grid(np.rint(np.asarray(s.grid, float)).astype(int),
              pbc=len(s.domain.rvecs) > 0).wrap_bits()
    return omm.compute_vectorised(s.pos if pos is None else pos,
                                  s.domain.rvecs if rvecs is None else rvecs, s.types,
                                  s.surrounding_nodes, s.surrounding_cells, wb, s.params, model)


def test_multiplicator_rows_are_positive_edges():
    # corner 7 = (1,1,1): rows r7-r6, r7-r5, r7-r4 (nanocell_utils.py:71-75)
    m = onc.MULTIPLICATOR
    assert list(m[7, 0]) == [0, 0, 0, 0, 0, 0, -1, 1]
    assert list(m[7, 1]) == [0, 0, 0, 0, 0, -1, 0, 1]
    assert list(m[7, 2]) == [0, 0, 0, 0, -1, 0, 0, 1]
    # corner 0: r1-r0, r2-r0, r3-r0
    assert list(m[0, 0]) == [-1, 1, 0, 0, 0, 0, 0, 0]
    assert (m.sum(axis=2) == 0).all() and (np.abs(m).sum(axis=2) == 2).all()


@pytest.mark.parametrize("model", ["corner8", "original"])
def test_rest_is_zero(model):
    s = make_system((3, 4, 3), seed=0, noise=0.0)
    e, g, v = vec_eval(s, model)
    assert abs(e) < 1e-25 and np.abs(g).max() < 1e-14 and np.abs(v).max() < 1e-10


@pytest.mark.parametrize("model", ["corner8", "original"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_affine_strain_closed_form(model, seed):
    s = make_system((3, 3, 4), seed=seed, noise=0.0)   # type A only, lattice = h0
    rng = np.random.default_rng(100 + seed)
    A = np.identity(3) + rng.uniform(-0.03, 0.03, (3, 3))
    e, g, v = vec_eval(s, model, pos=s.pos @ A, rvecs=s.domain.rvecs @ A)
    h0 = 20.7 * units.angstrom
    C = cubic_elasticity(40.0, 20.0, 10.0)
    eps = 0.5 * (A @ A.T - np.identity(3))
    e_cell = h0 ** 3 * 0.5 * np.einsum("ij,ijkl,kl", eps, C, eps)
    assert abs(e - s.ncells * e_cell) < 1e-12 * e
    # uniform strain: every node is in mechanical equilibrium
    assert np.abs(g).max() < 1e-10 * np.abs(v).max() / h0


def test_uniaxial_textbook_value():
    s = make_system((3, 3, 3), seed=0, noise=0.0)
    lam = 1.01
    A = np.diag([lam, 1.0, 1.0])
    for model in ("corner8", "original"):
        e, _, v = vec_eval(s, model, pos=s.pos @ A, rvecs=s.domain.rvecs @ A)
        c11 = 40.0 * units.gigapascal
        exx = 0.5 * (lam ** 2 - 1.0)
        V0 = (20.7 * units.angstrom) ** 3
        assert abs(e - 27 * V0 * 0.5 * c11 * exx ** 2) < 1e-12 * e
        # virial of a uniaxial stretch: only the xx and yy/zz (Poisson) diagonal terms
        assert abs(v[0, 1]) < 1e-12 * abs(v[0, 0]) and abs(v[1, 2]) < 1e-12 * abs(v[0, 0])


def _fd_gpos(s, model, idx, h=1e-4):
    out = []
    for (n, j) in idx:
        vals = []
        for step in (-2, -1, 1, 2):
            p = s.pos.copy()
            p[n, j] += step * h
            vals.append(vec_eval(s, model, pos=p)[0])
        out.append((vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h))
    return np.array(out)


def _major_symmetrised(s):
    """Copy of ``s`` whose elasticity tensors are replaced by (C + C^T_major) / 2."""
    s2 = load_system(GOLDEN, s._case + "/")
    for key in list(s2.params):
        if key.endswith("/elasticity"):
            C = np.asarray(s2.params[key], dtype=float)
            s2.params[key] = 0.5 * (C + C.transpose(0, 3, 4, 1, 2))
    return s2


def _case(name):
    s = load_system(GOLDEN, name + "/")
    s._case = name
    return s


@pytest.mark.parametrize("model", ["corner8", "original"])
@pytest.mark.parametrize("case", ["mix4_hollow", "reo2_nondiag_h0", "finite_3x3x2",
                                  "mix2_strained"])
def test_gradient_finite_differences(model, case):
    # The reference's gradient contracts with sigma = C:eps (nanocell.py:124), the exact
    # derivative of its energy only for a major-symmetric C; the fitted tensors of the reo/fcu
    # fixtures break that symmetry by ~1e-4, so the finite-difference pin uses (C + C^T)/2.
    s = _major_symmetrised(_case(case))
    _, g, _ = vec_eval(s, model)
    rng = np.random.default_rng(7)
    idx = [(int(rng.integers(s.nnodes)), int(rng.integers(3))) for _ in range(8)]
    fd = _fd_gpos(s, model, idx)
    an = np.array([g[n, j] for n, j in idx])
    assert np.abs(fd - an).max() < 1e-7 * np.abs(g).max()


@pytest.mark.parametrize("model", ["corner8", "original"])
def test_asymmetric_elasticity_quirk(model):
    """With a non-major-symmetric C the energy sees only (C + C^T)/2 while the gradient is
    linear in C itself: g(C) = g(C_sym) + g(C_anti) and E(C) = E(C_sym)."""
    s = _case("reo2_nondiag_h0")
    ssym = _major_symmetrised(s)
    santi = _major_symmetrised(s)
    for key in list(santi.params):
        if key.endswith("/elasticity"):
            C = np.asarray(s.params[key], dtype=float)
            santi.params[key] = 0.5 * (C - C.transpose(0, 3, 4, 1, 2))
    e, g, _ = vec_eval(s, model)
    es, gs, _ = vec_eval(ssym, model)
    ea, ga, _ = vec_eval(santi, model)
    assert abs(e - es) < 1e-13 * e and abs(ea) < 1e-13 * e
    assert np.abs(g - (gs + ga)).max() < 1e-13 * np.abs(g).max()
    assert np.abs(ga).max() > 1e-6 * np.abs(g).max()   # the quirk is real for this fixture


@pytest.mark.parametrize("model", ["corner8", "original"])
@pytest.mark.parametrize("case", ["mix4_hollow", "reo2_nondiag_h0", "mix2_strained"])
def test_virial_is_deformation_derivative(model, case):
    s = _major_symmetrised(_case(case))
    _, _, v = vec_eval(s, model)
    h = 1e-5
    fd = np.zeros((3, 3))
    for k in range(3):
        for l in range(3):
            vals = []
            for step in (-2, -1, 1, 2):
                D = np.identity(3)
                D[k, l] += step * h
                vals.append(vec_eval(s, model, pos=s.pos @ D, rvecs=s.domain.rvecs @ D)[0])
            fd[k, l] = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
    # vtens[k, l] = sum g_k r_l = dE/dD_lk
    assert np.abs(fd.T - v).max() < 1e-7 * np.abs(v).max()
    assert np.abs(v - v.T).max() < 1e-12 * np.abs(v).max()


@pytest.mark.parametrize("model", ["corner8", "original"])
def test_translation_and_rotation_invariance(model):
    s = load_system(GOLDEN, "mix4_hollow/")
    e, g, v = vec_eval(s, model)
    assert np.abs(g.sum(axis=0)).max() < 1e-13 * np.abs(g).max()
    e_t, g_t, _ = vec_eval(s, model, pos=s.pos + np.array([0.3, -1.7, 2.9]))
    assert abs(e_t - e) < 1e-12 * e and np.abs(g_t - g).max() < 1e-11 * np.abs(g).max()
    R = Rotation.from_euler("xyz", [0.3, -0.7, 1.1]).as_matrix()
    e_r, g_r, v_r = vec_eval(s, model, pos=s.pos @ R.T, rvecs=s.domain.rvecs @ R.T)
    assert abs(e_r - e) < 1e-12 * e
    assert np.abs(g_r - g @ R.T).max() < 1e-11 * np.abs(g).max()
    assert np.abs(v_r - R @ v @ R.T).max() < 1e-11 * np.abs(v).max()


def _two_state_params(f1, temp=300.0, same=False):
    h0 = 20.9 * units.angstrom * np.identity(3)
    C = cubic_elasticity(30.0, 18.0, 7.0)
    h1, C1 = (h0, C) if same else (0.97 * h0, 0.8 * C)
    return {"type1/cell": [h0, h1], "type1/elasticity": [C, C1],
            "type1/free_energy": [0.0, f1], "type1/effective_temp": temp}


@pytest.mark.parametrize("model", ["corner8", "original"])
def test_lse_equal_states_closed_form(model):
    s = make_system((3, 3, 3), seed=3, noise=0.1 * units.angstrom)
    s.params = dict(_two_state_params(0.0, same=True))
    d = sysdict(s)
    e2, g2, v2 = omm.compute_loop(d, model)
    s.params = {k: (v[:1] if isinstance(v, list) else v) for k, v in s.params.items()}
    s.params["type1/free_energy"] = [0.0]
    e1, g1, v1 = omm.compute_loop(sysdict(s), model)
    kt = boltzmann * 300.0
    assert abs(e2 - (e1 - s.ncells * kt * np.log(2.0))) < 1e-12 * abs(e1)
    assert np.abs(g2 - g1).max() < 1e-12 * np.abs(g1).max()


@pytest.mark.parametrize("model", ["corner8", "original"])
@pytest.mark.parametrize("f1", [0.0, 2e-3, -5e-3])
def test_lse_against_logsumexp(model, f1):
    s = make_system((2, 3, 2), seed=4, noise=0.15 * units.angstrom)
    s.params = dict(_two_state_params(f1))
    epot, gcell, _ = omm.deformation_loop(s.pos, s.domain.rvecs,
                                          __import__("oracle.topology", fromlist=["x"]).get_mic(
                                              np.rint(s.grid).astype(int)),
                                          s.types, s.surrounding_nodes, s.params, model)
    efun, gfun = onc.MODELS[model]
    kt = boltzmann * 300.0
    rv = s.domain.rvecs
    wb = Grid(np.rint(s.grid).astype(int)).wrap_bits()
    verts = omm.cell_vertices(s.pos, rv, np.asarray(s.surrounding_nodes), omm.corner_shifts(wb))
    for c in range(s.ncells):
        es = np.array([efun(verts[c], h, C) + f for h, C, f in
                       zip(s.params["type1/cell"], s.params["type1/elasticity"],
                           s.params["type1/free_energy"])])
        assert abs(epot[c] - (-kt * logsumexp(-es / kt))) < 1e-12 * max(abs(epot[c]), 1e-6)
        gs = [gfun(verts[c], h, C) for h, C in zip(s.params["type1/cell"],
                                                 s.params["type1/elasticity"])]
        p = np.exp(-(es - logsumexp(-es / kt) * -kt) / kt)
        assert abs(p.sum() - 1.0) < 1e-12
        gexp = p[0] * gs[0] + p[1] * gs[1]
        assert np.abs(gcell[c] - gexp).max() < 1e-12 * np.abs(gexp).max()


@pytest.mark.parametrize("case", [str(c) for c in GOLDEN["cases"]])
@pytest.mark.parametrize("model", ["corner8", "original"])
def test_oracle_reproduces_golden(case, model):
    s = load_system(GOLDEN, case + "/")
    e, g, v = omm.compute_loop(sysdict(s), model)
    assert e == GOLDEN[case + "/" + model + ":energy"]
    assert np.array_equal(g, GOLDEN[case + "/" + model + ":gpos"])
    assert np.array_equal(v, GOLDEN[case + "/" + model + ":vtens"])
    e2, g2, v2 = vec_eval(s, model)
    assert abs(e2 - e) < 1e-13 * abs(e)
    assert np.abs(g2 - g).max() < 1e-12 * np.abs(g).max()
    assert np.abs(v2 - v).max() < 1e-11 * np.abs(v).max()
