"""Cell degrees of freedom on the host (no GPU): the Jacobians of the cell parametrisations
(micmec/sampling/dof.py:487-697) against central differences of their own rvecs maps, the
freemask handling and the argument checks."""

import numpy as np
import pytest

from micmec_amd.sampling.dof import FullCellDOF, StrainCellDOF
from micmec_amd.workloads import make_system


class _TopologyOnly(object):
    """Stand-in force field that only carries a system (the Jacobians need nothing else)."""

    def __init__(self, system):
        self.system = system

    def update_rvecs(self, rvecs):
        self.system.domain.update_rvecs(np.asarray(rvecs, dtype=float))

    def update_pos(self, pos):
        self.system.pos[:] = pos


def _dof(cls, **kw):
    s = make_system((3, 3, 4), p_b=0.3, seed=5, strain=0.04)
    return cls(_TopologyOnly(s), **kw)


@pytest.mark.parametrize("cls", [StrainCellDOF, FullCellDOF])
def test_jacobian_matches_central_differences(cls):
    dof = _dof(cls)
    rng = np.random.default_rng(3)
    x = dof.domainvars0 + rng.uniform(-0.05, 0.05, dof.ncellvar)
    J = dof._get_celldofs_jacobian(x)
    h = 1e-6
    fd = np.zeros_like(J)
    for k in range(dof.ncellvar):
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        fd[:, k] = (dof._cellvars_to_rvecs(xp) - dof._cellvars_to_rvecs(xm)).ravel() / (2 * h)
    assert np.abs(J - fd).max() <= 1e-8 * np.abs(J).max()


@pytest.mark.parametrize("cls", [StrainCellDOF, FullCellDOF])
def test_initial_cellvars_reproduce_the_domain(cls):
    dof = _dof(cls)
    rv = dof._cellvars_to_rvecs(dof.domainvars0)
    assert np.allclose(rv, dof.mmf.system.domain.rvecs, rtol=1e-14, atol=0)
    # x = [celldofs, frac]; frac maps back to the positions
    frac = dof.x0[dof.ncelldof:].reshape(-1, 3)
    pos = frac @ dof.mmf.system.domain._get_rvecs(full=True)
    assert np.allclose(pos, dof.mmf.system.pos, rtol=0, atol=1e-9)


def test_strain_dof_has_no_rotation():
    dof = _dof(StrainCellDOF)
    A = dof._deform(np.array([1.01, 0.99, 1.02, 0.004, -0.006, 0.002]))
    assert np.array_equal(A, A.T)


def test_freemask():
    free = np.array([True, True, True, False, False, False])
    dof = _dof(StrainCellDOF, freemask=free)
    assert dof.ncelldof == 3 and len(dof.x0) == 3 + 3 * dof.mmf.system.nnodes
    assert dof._get_celldofs_jacobian(dof.x0[:3]).shape == (9, 3)
    with pytest.raises(TypeError):
        _dof(StrainCellDOF, freemask=np.array([1, 0, 1]))
    with pytest.raises(TypeError):
        _dof(StrainCellDOF, freemask=np.array([True, False]))
    frozen = _dof(FullCellDOF, do_frozen=True)
    assert len(frozen.x0) == 9


def test_requires_a_periodic_system():
    s = make_system((2, 2, 2), pbc=False, seed=1)
    with pytest.raises(ValueError):
        StrainCellDOF(_TopologyOnly(s))


class _OracleForceField(_TopologyOnly):
    """Stand-in force field evaluated by the oracle's vectorised restatement (host only)."""

    def __init__(self, system):
        from micmec_amd.pes.mmff import system_wrap_bits

        _TopologyOnly.__init__(self, system)
        self.wrap_bits = system_wrap_bits(system)

    def compute(self, gpos=None, vtens=None):
        from oracle.mmff import compute_vectorised

        s = self.system
        e, g, v = compute_vectorised(s.pos, s.domain.rvecs, s.types, s.surrounding_nodes,
                                     s.surrounding_cells, self.wrap_bits, s.params)
        if gpos is not None:
            gpos += g
        if vtens is not None:
            vtens += v
        return e


def test_cg_relaxes_cell_and_positions_to_the_same_minimum():
    # the first trial step 1/|g| is orders of magnitude too long along the cell degrees of
    # freedom (their curvature is the bulk stiffness times the volume); the line search has to
    # bracket back from it, and both parametrisations must reach the same stress-free minimum
    from micmec_amd.sampling.opt import ConjugateGradientSearch

    energies = []
    for cls in (StrainCellDOF, FullCellDOF):
        s = make_system((3, 4, 3), p_b=0.4, p_hole=0.1, seed=23, strain=0.02)
        dof = cls(_OracleForceField(s))
        cg = ConjugateGradientSearch(dof.x0.copy(), dof.fun)
        cg.initialize()
        f0 = cg.f
        assert cg.propagate(), cls.__name__
        assert cg.f < f0
        for _ in range(200):
            if not cg.propagate() or np.linalg.norm(cg.g) < 1e-7:
                break
        assert np.linalg.norm(cg.g) < 1e-6, (cls.__name__, np.linalg.norm(cg.g))
        f, g = dof.fun(cg.x, True)
        vtens = np.zeros((3, 3))
        _OracleForceField(s).compute(None, vtens)  # fun(x) left the state of x in the system
        assert np.abs(vtens).max() < 1e-6, (cls.__name__, vtens)
        energies.append(f)
    assert abs(energies[0] - energies[1]) <= 1e-10 * abs(energies[0])
