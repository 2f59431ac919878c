"""Fused conjugate-gradient kernels (mm_cg_trial, mm_cg_stats) and the device engine of
CGOptimizer over CartesianDOF (micmec/sampling/opt.py:147-180, dof.py:132-193).

The trial kernels are checked against the same steps done separately (the direction update
and the trial point with torch arithmetic, energy and gradient with mm_field_compute, the slope
with torch.dot); the statistics against torch reductions; the engine against the generic
vector path of the same optimiser, and its convergence measures against the reference's
formulas (dof.py:165-181) evaluated on the host copies of the accepted points."""

import ctypes

import numpy as np
import pytest
import torch

from micmec_amd import units
from micmec_amd.pes import ForcePartMechanical, ForcePartPressure, MicMecForceField
from micmec_amd.sampling import CartesianDOF, CGOptimizer
from micmec_amd.workloads import make_system

pytestmark = pytest.mark.gpu


def make(model, shape=(6, 5, 4), seed=3, press=None):
    s = make_system(shape, p_b=0.3, p_hole=0.1, seed=seed, strain=0.02,
                    noise=0.05 * units.angstrom)
    parts = [ForcePartMechanical(s, model=model)]
    if press is not None:
        parts.append(ForcePartPressure(s, press))
    return s, MicMecForceField(s, parts), parts[0]


def randn(n, seed, scale, device):
    gen = torch.Generator().manual_seed(seed)
    return (torch.randn(n, generator=gen, dtype=torch.float64) * scale).to(device)


@pytest.mark.parametrize("model", ["corner8", "original"])
@pytest.mark.parametrize("set_dir,beta", [(1, 0.0), (1, 0.37), (0, 0.0)])
def test_cg_trial_equals_the_separate_steps(model, set_dir, beta):
    print("cg_trial", model, set_dir, beta, flush=True)
    s, _, fpm = make(model)
    dev = fpm.device
    n = 3 * s.nnodes
    x = torch.as_tensor(np.ascontiguousarray(s.pos).ravel(), device=dev)
    g = randn(n, 1, 1e-3, dev)
    d = randn(n, 2, 1e-2, dev)
    rvecs = fpm.rvecs_device()
    alpha = 0.7
    d_ref = (beta * d - g if beta != 0.0 else -g) if set_dir else d.clone()
    d_run = d.clone()
    xa = torch.empty(n, dtype=torch.float64, device=dev)
    ga = torch.empty(n, dtype=torch.float64, device=dev)
    out = torch.zeros(7, dtype=torch.float64, device=dev)
    host = (ctypes.c_double * 7)()
    fpm.field.cg_trial(x, g, d_run, beta, set_dir, alpha, rvecs, xa, ga, out, host_out=host)
    # the same force kernels at the trial point the kernel wrote
    e_ref = torch.zeros(1, dtype=torch.float64, device=dev)
    g_ref = torch.zeros((s.nnodes, 3), dtype=torch.float64, device=dev)
    fpm.compute_device(xa.view(-1, 3), rvecs, e_ref, g_ref)
    torch.cuda.synchronize()
    got = out.tolist()
    assert list(host) == got
    # direction and trial point to rounding (the kernel may fuse the multiply-add)
    assert torch.allclose(d_run, d_ref, rtol=0, atol=1e-15 * float(d_ref.abs().max()))
    assert torch.allclose(xa, x + alpha * d_run, rtol=0, atol=1e-15 * float(x.abs().max()))
    # energy and gradient: the same kernels and summation order as mm_field_compute
    assert got[0] == float(e_ref[0])
    assert torch.equal(ga, g_ref.view(-1))
    # the reductions, against torch in another order
    gsq = (ga.view(-1, 3) ** 2).sum(dim=1)
    dsq = ((xa - x).view(-1, 3) ** 2).sum(dim=1)
    ref = [torch.dot(ga, d_run), gsq.sum(), torch.dot(ga, g), gsq.max(), dsq.sum(), dsq.max()]
    scale = [torch.dot(ga.abs(), d_run.abs()), gsq.sum(), torch.dot(ga.abs(), g.abs()),
             gsq.max(), dsq.sum(), dsq.max()]
    for q in range(6):
        assert abs(got[1 + q] - float(ref[q])) <= 1e-13 * float(scale[q]), q


def test_cg_stats_equal_torch_reductions():
    s, _, fpm = make("corner8", shape=(9, 7, 5))
    dev = fpm.device
    n = 3 * s.nnodes
    x, xo, g, go = (randn(n, k, sc, dev) for k, sc in ((1, 10.0), (2, 10.0), (3, 1e-3),
                                                     (4, 1e-3)))
    out = torch.zeros(5, dtype=torch.float64, device=dev)
    host = (ctypes.c_double * 5)()
    fpm.field.cg_stats(x, xo, g, go, out, host_out=host)
    torch.cuda.synchronize()
    assert list(host) == out.tolist()
    gsq = (g.view(-1, 3) ** 2).sum(dim=1)
    dsq = ((x - xo).view(-1, 3) ** 2).sum(dim=1)
    ref = [gsq.sum(), torch.dot(g, go), dsq.sum(), gsq.max(), dsq.max()]
    got = out.tolist()
    for q, r in enumerate(ref):
        r = float(r)
        tol = 1e-13 * (float(torch.dot(g.abs(), go.abs())) if q == 1 else abs(r))
        assert abs(got[q] - r) <= tol, (q, got[q], r)
    # with x_old = x and g_old = g: zero displacement, and g.g_old = |g|^2
    fpm.field.cg_stats(x, x, g, g, out)
    torch.cuda.synchronize()
    got = out.tolist()
    assert got[2] == 0.0 and got[4] == 0.0 and got[0] == got[1]


def run_cg(model, engine, iterations, press=None):
    s, mmf, _ = make(model, press=press)
    dof = CartesianDOF(mmf, gpos_rms=1e-30, dpos_rms=1e-30)
    if not engine:
        dof.cg_engine = lambda x0: None
    opt = CGOptimizer(dof)
    assert (opt.minimizer.engine.__class__.__name__ == "_DeviceCGEngine") == engine
    energies = [opt.epot]
    for _ in range(iterations):
        opt.propagate()
        energies.append(opt.epot)
    dof.sync_from_device()
    return np.array(energies), dof._pos.copy(), dof


@pytest.mark.parametrize("model,press", [("corner8", None), ("original", None),
                                         ("corner8", 0.3 * units.bar * 1e4)])
def test_engine_follows_the_generic_vector_path(model, press):
    print("engine vs generic", model, press, flush=True)
    e_eng, p_eng, _ = run_cg(model, True, 8, press)
    e_gen, p_gen, _ = run_cg(model, False, 8, press)
    assert np.all(np.diff(e_eng) <= 1e-15 * np.abs(e_eng[:-1]))
    assert np.allclose(e_eng, e_gen, rtol=1e-10, atol=0)
    assert np.abs(p_eng - p_gen).max() <= 1e-8 * units.angstrom


def test_engine_convergence_measures_follow_the_reference_formulas():
    s, mmf, _ = make("corner8")
    dof = CartesianDOF(mmf, gpos_rms=1e-30, dpos_rms=1e-30)
    opt = CGOptimizer(dof)
    assert dof.conv_worst == "first_step"
    last = dof._pos.copy()
    for _ in range(5):
        opt.propagate()
        dof.sync_from_device()
        pos, gpos = dof._pos.copy(), dof._gpos.copy()
        gsq = (gpos ** 2).sum(axis=1)
        dsq = ((pos - last) ** 2).sum(axis=1)
        assert np.isclose(dof.gpos_max, np.sqrt(gsq.max()), rtol=1e-14, atol=0)
        assert np.isclose(dof.gpos_rms, np.sqrt(gsq.mean()), rtol=1e-13, atol=0)
        assert np.isclose(dof.dpos_max, np.sqrt(dsq.max()), rtol=1e-14, atol=0)
        assert np.isclose(dof.dpos_rms, np.sqrt(dsq.mean()), rtol=1e-13, atol=0)
        last = pos
    # a second check without a step: zero displacement, the same gradient measures
    gmax, grms = dof.gpos_max, dof.gpos_rms
    dof.check_convergence()
    assert dof.dpos_max == 0.0 and dof.dpos_rms == 0.0
    assert np.isclose(dof.gpos_max, gmax, rtol=1e-15) and np.isclose(dof.gpos_rms, grms,
                                                                       rtol=1e-15)
    # the system holds the accepted point
    assert np.array_equal(np.asarray(s.pos), dof._pos)


def test_engine_restarts_along_the_steepest_descent():
    """A pending direction that is not a descent direction (slope -g.g + beta g.d_old >= 0)
    makes the device engine restart along -g, as the generic path does."""
    s, mmf, _ = make("corner8")
    dof = CartesianDOF(mmf, gpos_rms=1e-30, dpos_rms=1e-30)
    opt = CGOptimizer(dof)
    opt.propagate()
    eng = opt.minimizer.engine
    e0 = opt.epot
    eng.set_direction(1e6, abs(eng.gg))   # slope = -g.g + 1e6 |g.g| > 0
    assert eng.slope() > 0.0
    opt.propagate()
    assert opt.minimizer.restarts == 1
    assert opt.epot < e0
    # after the restart the search direction was -g of the accepted point before it
    torch.cuda.synchronize()


def test_engine_trial_with_a_nan_position_raises():
    """A not-a-number energy at a trial point raises ValueError, as the reference's
    ForcePart.compute does (micmec/pes/mmff.py:135-148)."""
    s, mmf, _ = make("corner8")
    dof = CartesianDOF(mmf)
    opt = CGOptimizer(dof)
    eng = opt.minimizer.engine
    eng.x[0] = float("nan")
    with pytest.raises(ValueError):
        eng.trial(1e-3)
