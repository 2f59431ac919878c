"""Geometry optimisation: the CG search (host, numpy) and CGOptimizer on the GPU.

molmod's minimiser (the reference's CG engine, micmec/sampling/opt.py:153-162) is not available,
so the search path is parity-unpinned; what is pinned is the outcome: monotone energy, the
reference's convergence criteria (dof.py:155-193), and the minimum energy of an independent
scipy optimiser on the CPU oracle.
"""

import numpy as np
import pytest
import scipy.optimize

from micmec_amd import units
from micmec_amd.sampling.opt import ConjugateGradientSearch


def test_cg_search_minimises_a_quadratic():
    rng = np.random.default_rng(0)
    A = rng.normal(size=(30, 30))
    A = A @ A.T + 30 * np.identity(30)
    b = rng.normal(size=30)
    fun = lambda x, g=True: (0.5 * x @ A @ x - b @ x, A @ x - b)
    cg = ConjugateGradientSearch(np.zeros(30), fun)
    cg.initialize()
    fs = [cg.f]
    for _ in range(60):
        assert cg.propagate()
        fs.append(cg.f)
        if np.linalg.norm(cg.g) < 1e-12:
            break
    assert all(b <= a + 1e-12 for a, b in zip(fs, fs[1:]))
    assert np.allclose(cg.x, np.linalg.solve(A, b), atol=1e-9)


@pytest.mark.parametrize("x0", [[-1.2, 1.0], [1.3, 0.7, 0.8, 1.9, 1.2]])
def test_cg_search_on_rosenbrock(x0):
    # the textbook starts; both lie in the basin of the global minimum at x = 1
    f = lambda x, g=True: (scipy.optimize.rosen(x), scipy.optimize.rosen_der(x))
    cg = ConjugateGradientSearch(np.array(x0), f)
    cg.initialize()
    for _ in range(2000):
        if not cg.propagate() or np.linalg.norm(cg.g) < 1e-10:
            break
    assert np.allclose(cg.x, 1.0, atol=1e-5)


def test_cg_search_line_search_recovers_from_an_overshooting_first_step():
    # a stiff direction (curvature 1e6) next to a soft one: the first trial step 1/|g| overshoots
    # the minimum along the steepest descent by orders of magnitude, and the line search has to
    # bracket back instead of giving up
    k = np.array([1e6, 1.0, 3.0])
    f = lambda x, g=True: (0.5 * np.dot(k * x, x), k * x)
    cg = ConjugateGradientSearch(np.array([1e-3, 2.0, -1.0]), f)
    cg.initialize()
    assert cg.propagate()
    assert cg.f < 0.5 * np.dot(k * np.array([1e-3, 2.0, -1.0]), [1e-3, 2.0, -1.0])
    for _ in range(50):
        if not cg.propagate() or np.linalg.norm(cg.g) < 1e-12:
            break
    assert np.linalg.norm(cg.x) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["corner8", "original"])
def test_cg_optimizer_reaches_oracle_minimum(model):
    from micmec_amd.pes import ForcePartMechanical, MicMecForceField
    from micmec_amd.sampling import CartesianDOF, CGOptimizer
    from micmec_amd.topology import Grid
    from micmec_amd.workloads import make_system
    from oracle import mmff as omm

    s = make_system((4, 4, 3), p_b=0.3, p_hole=0.1, seed=12, strain=0.02,
                    noise=0.05 * units.angstrom)
    pos0 = s.pos.copy()
    mmf = MicMecForceField(s, [ForcePartMechanical(s, model=model)])
    dof = CartesianDOF(mmf, gpos_rms=1e-8, dpos_rms=1e-6)
    opt = CGOptimizer(dof)
    energies = [opt.epot if hasattr(opt, "epot") else opt.minimizer.f]
    for _ in range(400):
        done = opt.propagate()
        energies.append(opt.epot)
        if done:
            break
    assert dof.converged, (dof.conv_worst, dof.conv_val)
    assert all(b <= a + 1e-15 * abs(a) for a, b in zip(energies, energies[1:]))
    # independent minimiser on the CPU oracle
    wb = Grid(np.rint(s.grid).astype(int)).wrap_bits()

    def f(x):
        e, g, _ = omm.compute_vectorised(x.reshape(-1, 3), s.domain.rvecs, s.types,
                                         s.surrounding_nodes, s.surrounding_cells, wb, s.params,
                                         model)
        return e, g.ravel()
    res = scipy.optimize.minimize(f, pos0.ravel(), jac=True, method="L-BFGS-B",
                                  options=dict(gtol=1e-12, ftol=1e-16, maxiter=5000))
    assert abs(opt.epot - res.fun) <= 1e-8 * abs(res.fun)
    # the state of the system is the optimised geometry
    e_sys = mmf.compute()
    assert abs(e_sys - opt.epot) <= 1e-12 * abs(opt.epot)


def test_cg_search_converges_through_energy_noise():
    """Near a minimum the energy is known only to the rounding of its strains (about 1e-13
    relative per cell): on a quadratic whose energy carries 1e-14 relative noise (a fixed
    function of x, as rounding is) the search must keep following the gradient to |g| < 1e-10
    instead of giving up when no trial lowers the energy beyond the noise."""
    import zlib
    rng = np.random.default_rng(0)
    n = 40
    A = rng.normal(size=(n, n))
    A = A @ A.T / n + np.diag(np.logspace(0, 4, n))
    b = rng.normal(size=n)

    def fun(x, g=True):
        f = 0.5 * x @ A @ x - b @ x + 10.0
        h = zlib.crc32(np.ascontiguousarray(x).tobytes()) / 2.0 ** 32 - 0.5
        return f * (1.0 + 2e-14 * h), A @ x - b

    cg = ConjugateGradientSearch(np.zeros(n), fun)
    cg.initialize()
    for _ in range(3000):
        assert cg.propagate(), np.linalg.norm(cg.g)
        if np.linalg.norm(cg.g) < 1e-10:
            break
    assert np.linalg.norm(cg.g) < 1e-10
    assert np.allclose(cg.x, np.linalg.solve(A, b), atol=1e-12)
