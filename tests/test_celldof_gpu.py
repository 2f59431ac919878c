"""Cell degrees of freedom on the GPU (micmec/sampling/dof.py:199-697): the gradient of both
parametrisations against finite differences of the energy, host and device paths against each
other, CG cell relaxations that end at a stationary point (zero node gradient and zero virial,
a physical condition independent of the parametrisation), and the elastic constants of a
homogeneous lattice from stress-strain differences (the recipe of simulations/stress_strain.py)
against the input C11, C12, C44."""

import numpy as np
import pytest
import torch

from micmec_amd import units
from micmec_amd.adjoint import voigt_from_tensor
from micmec_amd.pes import ForcePartMechanical, MicMecForceField
from micmec_amd.sampling import CartesianDOF, CGOptimizer, FullCellDOF, StrainCellDOF
from micmec_amd.workloads import cubic_elasticity, make_system

pytestmark = pytest.mark.gpu
A = units.angstrom


def field(s, model="corner8"):
    return MicMecForceField(s, [ForcePartMechanical(s, model=model)])


@pytest.mark.parametrize("cls", [StrainCellDOF, FullCellDOF])
def test_gradient_matches_finite_differences(cls):
    s = make_system((3, 3, 4), p_b=0.3, p_hole=0.1, seed=21, strain=0.03, noise=0.1 * A)
    dof = cls(field(s))
    # random moves of all variables: the trapezoidal check's error scales as eps^2 (about
    # 2e-7 relative at eps = 1e-6 on this system)
    np.random.seed(4)
    worst = dof.check_delta(eps=1e-6, ncheck=20)
    assert worst < 1e-5, worst
    # every cell variable by central differences
    x = dof.x0.copy()
    e, g = dof.fun(x, True)
    ncd = dof.ncelldof
    fd = np.zeros(ncd)
    for k in range(ncd):
        h = 1e-6
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        fd[k] = (dof.fun(xp) - dof.fun(xm)) / (2 * h)
    assert np.abs(fd - g[:ncd]).max() <= 1e-5 * np.abs(g[:ncd]).max()


@pytest.mark.parametrize("cls", [StrainCellDOF, FullCellDOF])
def test_device_path_matches_host_path(cls):
    s = make_system((4, 3, 3), p_b=0.4, p_hole=0.05, seed=22, strain=0.02, noise=0.05 * A)
    dof = cls(field(s))
    rng = np.random.default_rng(5)
    x = dof.x0 + rng.uniform(-1e-3, 1e-3, dof.x0.shape)
    e_h, g_h = dof.fun(x, True)
    xd = torch.tensor(x, dtype=torch.float64, device="cuda")
    e_d, g_d = dof.fun(xd, True)
    assert abs(e_d - e_h) <= 1e-12 * abs(e_h)
    g_d = g_d.cpu().numpy()
    assert np.abs(g_d - g_h).max() <= 1e-10 * np.abs(g_h).max()


def stationary(mmf, opt):
    g = np.zeros(mmf.system.pos.shape)
    v = np.zeros((3, 3))
    e = mmf.compute(g, v)
    return e, g, v


@pytest.mark.parametrize("cls", [StrainCellDOF, FullCellDOF])
def test_cg_cell_relaxation_reaches_zero_stress(cls):
    s = make_system((3, 4, 3), p_b=0.4, p_hole=0.1, seed=23, strain=0.02, noise=0.05 * A)
    mmf = field(s)
    e0, g0, v0 = stationary(mmf, None)
    dof = cls(mmf, gpos_rms=1e-10, dpos_rms=1e-8, grvecs_rms=1e-10, drvecs_rms=1e-8)
    opt = CGOptimizer(dof)
    opt.run(20000)
    assert dof.converged, (dof.conv_worst, dof.conv_val, opt.counter)
    e, g, v = stationary(mmf, opt)
    assert abs(e - opt.epot) <= 1e-10 * abs(e0)
    assert e < e0
    # a stationary point of the energy in positions and domain: no node force, no virial
    assert np.abs(g).max() <= 1e-7 * np.abs(g0).max()
    assert np.abs(v).max() <= 1e-7 * np.abs(v0).max()
    # with the domain fixed at the relaxed one, the positions are already relaxed
    cdof = CartesianDOF(mmf, gpos_rms=1e-10, dpos_rms=1e-8)
    copt = CGOptimizer(cdof)
    copt.run(50)
    assert copt.epot >= e - 1e-10 * abs(e0)


def test_strain_and_full_cell_relaxations_agree():
    def relax(cls):
        s = make_system((3, 3, 3), p_b=0.5, p_hole=0.1, seed=24, strain=0.02, noise=0.05 * A)
        mmf = field(s)
        dof = cls(mmf, gpos_rms=1e-10, dpos_rms=1e-8, grvecs_rms=1e-10, drvecs_rms=1e-8)
        opt = CGOptimizer(dof)
        opt.run(20000)
        assert dof.converged
        return opt.epot, s.domain.volume
    e_s, v_s = relax(StrainCellDOF)
    e_f, v_f = relax(FullCellDOF)
    assert abs(e_s - e_f) <= 1e-9 * max(abs(e_s), 1e-30) + 1e-14
    assert abs(v_s - v_f) <= 1e-8 * v_s


@pytest.mark.parametrize("model", ["corner8", "original"])
def test_elastic_constants_of_a_homogeneous_lattice(model):
    """Deform a perfect lattice of one type by F = sqrt(2 eps + I) (Green-Lagrange strain eps,
    simulations/stress_strain.py:71-76), read the stress vtens / V, and differentiate: the
    result is the Voigt matrix of the type (C11, C12, C44 = 40, 20, 10 GPa)."""
    s0 = make_system((3, 3, 3), p_b=0.0, seed=25, noise=0.0)
    pos0 = np.array(s0.pos)
    rv0 = np.array(s0.domain.rvecs)
    mmf = field(s0, model)
    pairs = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))
    delta = 1e-4
    C = np.zeros((6, 6))
    for J in range(6):
        stress = []
        for sign in (1.0, -1.0):
            e = np.zeros(6)
            e[J] = sign * delta
            eps = np.zeros((3, 3))
            for I, (i, j) in enumerate(pairs):
                eps[i, j] = eps[j, i] = e[I] if I < 3 else 0.5 * e[I]
            w, Q = np.linalg.eigh(2.0 * eps + np.identity(3))
            F = (Q * np.sqrt(w)) @ Q.T
            mmf.update_rvecs(rv0 @ F)
            mmf.update_pos(pos0 @ F)
            v = np.zeros((3, 3))
            mmf.compute(vtens=v)
            sig = v / s0.domain.volume
            stress.append(np.array([sig[i, j] for (i, j) in pairs]))
        C[:, J] = (stress[0] - stress[1]) / (2 * delta)
    Cv = voigt_from_tensor(cubic_elasticity(40.0, 20.0, 10.0))
    assert np.abs(C - Cv).max() <= 1e-6 * Cv.max(), (C / units.gigapascal).round(6)
