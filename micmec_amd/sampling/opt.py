"""Geometry optimisation (drop-in for micmec/sampling/opt.py:46-180).

``CGOptimizer`` keeps the reference API (``CGOptimizer(dof, state, hooks, counter0)``, ``run``,
``propagate`` returning convergence, ``OptScreenLog``). The reference delegates the search to
molmod's ``Minimizer(ConjugateGradient(), NewtonLineSearch())`` (opt.py:153-162), which is not
part of this build; this module implements its own Polak-Ribiere (PR+) conjugate gradient with
a Newton/secant line search on the directional derivative. When the degrees of freedom are
device-capable, x, gradient and search direction live on the GPU and every energy/gradient is
one launch of the HIP force kernels; only scalars return to the host.
"""

import time
from math import isfinite

import numpy as np
import torch

from ..log import log
from .iterative import (AttributeStateItem, DomainStateItem, EPotContribStateItem, Hook,
                        Iterative, PosStateItem, VolumeStateItem)

__all__ = ["OptScreenLog", "BaseOptimizer", "CGOptimizer", "ConjugateGradientSearch",
           "VectorEngine"]


class OptScreenLog(Hook):
    def __init__(self, start=0, step=1):
        Hook.__init__(self, start, step)
        self.time0 = None

    def expects_call(self, counter):
        return log.do_medium and Hook.expects_call(self, counter)

    def __call__(self, iterative):
        if self.time0 is None:
            self.time0 = time.time()
            log.hline()
            log("counter  Conv.val.  N           Worst     Energy   Walltime")
            log.hline()
        log("%7i % 10.3e %2i %15s % 14.8e %10.1f" % (
            iterative.counter, iterative.dof.conv_val, iterative.dof.conv_count,
            iterative.dof.conv_worst, iterative.epot, time.time() - self.time0))


class BaseOptimizer(Iterative):
    default_state = [
        AttributeStateItem("counter"),
        AttributeStateItem("epot"),
        PosStateItem(),
        VolumeStateItem(),
        DomainStateItem(),
        EPotContribStateItem(),
    ]
    log_name = "XXOPT"

    def __init__(self, dof, state=None, hooks=None, counter0=0):
        self.dof = dof
        Iterative.__init__(self, dof.mmf, state, hooks, counter0)

    def _add_default_hooks(self):
        if not any(isinstance(hook, OptScreenLog) for hook in self.hooks):
            self.hooks.append(OptScreenLog())

    def fun(self, x, do_gradient=False):
        if do_gradient:
            self.epot, gx = self.dof.fun(x, True)
            return self.epot, gx
        self.epot = self.dof.fun(x, False)
        return self.epot

    def initialize(self):
        self.dof.check_convergence()
        Iterative.initialize(self)

    def propagate(self):
        self.dof.check_convergence()
        Iterative.propagate(self)
        return self.dof.converged

    def _sync_for_hooks(self):
        # device-resident degrees of freedom: host positions and gradient for the hooks
        sync = getattr(self.dof, "sync_from_device", None)
        if sync is not None:
            sync()

    def finalize(self):
        self._sync_for_hooks()


def _dot(a, b):
    if isinstance(a, torch.Tensor):
        return float(torch.dot(a, b).item())
    return float(np.dot(a, b))


class VectorEngine(object):
    """The vectors of a conjugate-gradient search (position x, gradient g, direction d) as numpy
    arrays or torch tensors, each energy and gradient one call of ``fun(x, True)``.

    The search talks to its vectors only through these methods, so a device implementation
    can fuse them (see ``CartesianDOF.cg_engine``):

    ``start()`` evaluates f and g at x, sets d = -g and returns (f, g.g); ``slope()`` is g.d;
    ``restart()`` sets d = -g; ``trial(a)`` evaluates the point x + a d and returns (f, g'.d)
    there; ``keep()`` marks the last trial as the best of the line search; ``accept()`` moves
    to the kept point and returns (g.g, g.(g - g_old)) there; ``set_direction(beta, dp)`` sets
    d = -g + beta d, where dp is g.d with the old d.
    """

    def __init__(self, x0, fun):
        self.x = x0
        self.fun = fun
        self.f = None
        self.g = None
        self.d = None

    def start(self):
        self.f, self.g = self.fun(self.x, True)
        self.d = -self.g
        return self.f, _dot(self.g, self.g)

    def slope(self):
        return _dot(self.g, self.d)

    def restart(self):
        self.d = -self.g

    def trial(self, a):
        xa = self.x + a * self.d
        fa, ga = self.fun(xa, True)
        self._last = (xa, fa, ga)
        return fa, _dot(ga, self.d)

    def keep(self):
        self._best = self._last

    def accept(self):
        xa, fa, ga = self._best
        g_old = self.g
        self.x, self.f, self.g = xa, fa, ga
        return _dot(ga, ga), _dot(ga, ga - g_old)

    def set_direction(self, beta, dp):
        self.d = -self.g + beta * self.d


class ConjugateGradientSearch(object):
    """Polak-Ribiere (PR+) conjugate gradient with a secant line search on phi'(alpha).

    One ``propagate`` is one line search and one direction update, the unit of one iteration
    of the reference's minimiser. The vectors live in ``engine`` (a ``VectorEngine`` over
    ``fun`` unless another one is given).
    """

    def __init__(self, x0, fun, max_ls=20, curvature=1e-3, max_step=None, engine=None):
        self.engine = engine if engine is not None else VectorEngine(x0, fun)
        self.fun = fun
        self.max_ls = max_ls
        self.curvature = curvature
        self.max_step = max_step

    x = property(lambda self: self.engine.x)
    g = property(lambda self: self.engine.g)
    d = property(lambda self: self.engine.d)

    def initialize(self):
        self.f, self.gg = self.engine.start()
        gnorm = np.sqrt(self.gg)
        self.alpha = 1.0 / gnorm if gnorm > 0 else 1.0
        self.restarts = 0

    def propagate(self):
        E = self.engine
        dphi0 = E.slope()
        if not dphi0 < 0.0:
            # not a descent direction: restart along the steepest descent
            E.restart()
            dphi0 = E.slope()
            self.restarts += 1
            if dphi0 == 0.0:
                return True
        f0 = self.f
        scale = max(abs(f0), np.finfo(float).tiny)
        ftol = 8.0 * np.finfo(float).eps * scale
        # Energies within fnoise of phi(0) count as level with it: near a minimum the energy is
        # resolved only to the rounding of its strains (differences of numbers close to 1,
        # about 1e-13 relative per cell), so a trial there is judged by its slope; a level
        # trial is kept only when no trial decreases the energy.
        fnoise = 1e-12 * scale
        # bracket [lo, hi] of the minimum along d: lo has phi'(lo) < 0 and phi(lo) <= phi(0),
        # hi has phi'(hi) > 0 or no decrease; the next trial is the secant zero of phi' in the
        # bracket (exact on quadratics), kept a hundredth of the width away from its ends
        lo_a, lo_dp = 0.0, dphi0
        hi_a, hi_dp = None, None
        a = self.alpha
        best = near = None
        for _ in range(self.max_ls):
            fa, dp = E.trial(a)
            finite = isfinite(fa) and isfinite(dp)
            ok = finite and fa <= f0 + ftol
            level = finite and fa <= f0 + fnoise
            if ok and (best is None or abs(dp) < abs(best[2])):
                best = (a, fa, dp)
                E.keep()
            elif level and best is None and (near is None or abs(dp) < abs(near[2])):
                near = (a, fa, dp)
                E.keep()
            if level and abs(dp) <= self.curvature * abs(dphi0):
                break
            if level and dp < 0.0:
                prev_a, prev_dp = lo_a, lo_dp
                lo_a, lo_dp = a, dp
            else:
                hi_a, hi_dp = a, (dp if ok or (isfinite(fa) and isfinite(dp)) else None)
            if hi_a is None:
                # still descending: extrapolate the secant of phi' through the last two points
                # (beyond a when phi' rises), at most ten times as far
                a_new = 4.0 * a
                if dp > prev_dp:
                    a_new = min(a - dp * (a - prev_a) / (dp - prev_dp), 10.0 * a)
            else:
                if hi_dp is not None and hi_dp > 0.0:
                    a_new = lo_a - lo_dp * (hi_a - lo_a) / (hi_dp - lo_dp)
                else:
                    a_new = 0.5 * (lo_a + hi_a)
                width = hi_a - lo_a
                if not isfinite(a_new):
                    a_new = lo_a + 0.5 * width
                a_new = min(max(a_new, lo_a + 0.01 * width), hi_a - 0.01 * width)
            if self.max_step is not None:
                a_new = min(a_new, self.max_step)
            a = a_new
        if best is None:
            best = near
        if best is None:
            return False
        a, fa, dp = best
        gg_old = self.gg
        gg, num = E.accept()
        beta = max(0.0, num / gg_old) if gg_old > 0 else 0.0
        E.set_direction(beta, dp)
        self.f, self.gg, self.alpha = fa, gg, a
        return True


class CGOptimizer(BaseOptimizer):
    """Conjugate-gradient optimiser (opt.py:147-180)."""
    log_name = "CGOPT"

    def __init__(self, dof, state=None, hooks=None, counter0=0, device=True):
        x0 = dof.x0
        engine = None
        if device and getattr(dof, "device", None) is not None:
            x0 = torch.as_tensor(np.ascontiguousarray(x0), dtype=torch.float64,
                                 device=dof.device)
            make = getattr(dof, "cg_engine", None)
            engine = make(x0) if make is not None else None
        self.minimizer = ConjugateGradientSearch(x0, self.fun, engine=engine)
        BaseOptimizer.__init__(self, dof, state, hooks, counter0)

    def initialize(self):
        self.minimizer.initialize()
        self.epot = self.minimizer.f
        self.x = self.minimizer.x
        self.dof._accept(self.x, self.minimizer.g)
        BaseOptimizer.initialize(self)

    def propagate(self):
        success = self.minimizer.propagate()
        self.x = self.minimizer.x
        # make the accepted point the current state of the degrees of freedom
        self.epot = self.minimizer.f
        self.dof._accept(self.x, self.minimizer.g)
        if not success:
            log.warn("Line search failed in optimizer. Aborting optimization.")
            return True
        return BaseOptimizer.propagate(self)
