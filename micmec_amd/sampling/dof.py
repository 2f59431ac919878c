"""Degrees of freedom for geometry optimisation (micmec/sampling/dof.py:36-193).

``CartesianDOF`` keeps the reference constructor, ``fun(x, do_gradient)`` and
``check_convergence`` (gpos_rms / gpos_max / dpos_rms / dpos_max thresholds, dof.py:155-193).
When the force field is one ``ForcePartMechanical`` (plus optionally a ``ForcePartPressure``),
``fun`` also accepts device tensors and evaluates the energy and gradient with the HIP kernels
without host round trips; numpy input follows the reference path through ``mmf.compute``.
On the device path the convergence measures are reduced on the GPU, and the host positions and
gradient (and the system's positions) are refreshed at convergence, before host hooks and at
the end of a run (``sync_from_device``), not after every iteration.
"""

import ctypes

import numpy as np
import torch

from .. import _native

__all__ = ["DOF", "CartesianDOF", "BaseCellDOF", "FullCellDOF", "StrainCellDOF"]


def _pinned_like(t):
    """A pinned host tensor of t's shape and dtype (for DMA copies from the device)."""
    return torch.empty(tuple(t.shape), dtype=t.dtype, pin_memory=True)


def _to_host(src, dst, stage):
    """Copy the device tensor ``src`` into the numpy array ``dst`` through the pinned tensor
    ``stage`` (one DMA transfer, then a host copy)."""
    stage.copy_(src.detach())
    dst[...] = stage.numpy().reshape(dst.shape)


class _DeviceEnergy(object):
    """Energy and gradient of a device position tensor [N, 3]."""

    def __init__(self, mmf):
        from ..pes.mmff import ForcePartMechanical, ForcePartPressure
        parts = list(getattr(mmf, "parts", [mmf]))
        mech = [p for p in parts if isinstance(p, ForcePartMechanical)]
        press = [p for p in parts if isinstance(p, ForcePartPressure)]
        if len(mech) != 1 or len(press) > 1 or len(mech) + len(press) != len(parts):
            raise NotImplementedError("device evaluation needs one ForcePartMechanical")
        self.fpm = mech[0]
        self.press = press[0] if press else None
        self.device = self.fpm.device
        kw = dict(dtype=torch.float64, device=self.device)
        self.energy = torch.zeros(1, **kw)
        self.gpos = torch.zeros((mmf.system.nnodes, 3), **kw)
        self.vtens = torch.zeros((3, 3), **kw)
        self.rvecs = self.fpm.rvecs_device()
        self.calls = 0

    def __call__(self, pos, do_gradient=True, with_vtens=False):
        """Energy and gradient at device positions ``pos`` and the system's current domain
        (refresh ``self.rvecs`` with ``fpm.rvecs_device()`` after the domain changes); with
        ``with_vtens`` also the virial (host 3x3 array), pressure part included."""
        self.calls += 1
        self.fpm.compute_device(pos, self.rvecs, self.energy, self.gpos if do_gradient else None,
                                self.vtens if with_vtens else None)
        e = float(self.energy.item())
        if np.isnan(e):
            raise ValueError("The energy is not-a-number (``nan``).")
        volume = self.fpm.system.domain.volume
        if self.press is not None:
            e += self.press.pext * volume
        if not with_vtens:
            return e, self.gpos
        vt = self.vtens.cpu().numpy()
        if self.press is not None:
            vt = vt + self.press.pext * volume * np.identity(3)
        return e, self.gpos, vt


class _DeviceCGEngine(object):
    """Device vectors of a conjugate-gradient search over all node positions of a
    ``CartesianDOF`` (the engine protocol of ``sampling.opt.VectorEngine``).

    A trial point is one ``mm_cg_trial`` call: the pending direction update d <- beta d - g and
    the point x + a d in one kernel, the force kernels, the slope g'.d and the measures an
    accepted step needs (g'.g', g'.g, max |g'_n|^2, and the displacement from the current point,
    which is where the last convergence check left ``_last_pos``) reduced inside the node gather
    and the trial kernel, then one synchronising read of seven doubles. Accepting a trial
    launches nothing: it swaps buffers and keeps the trial's measures for the Polak-Ribiere
    coefficient and ``check_convergence``. Trial points alternate between two buffers, so no
    vector is copied.
    """

    def __init__(self, dof, x0):
        self.energy = dof._device
        self.field = self.energy.fpm.field
        kw = dict(dtype=torch.float64, device=x0.device)
        n = x0.numel()
        self.x = x0.detach().reshape(-1).clone()
        self.g = torch.zeros(n, **kw)
        self.d = torch.zeros(n, **kw)
        self._slots = [[torch.empty(n, **kw), torch.empty(n, **kw)] for _ in range(2)]
        self._out = torch.zeros(8, **kw)
        self._host = (ctypes.c_double * 8)()
        # raw arguments of the library calls, made once (the buffers never move)
        self._lib = _native.lib()
        self._check = _native.check
        self._handle = self.field.handle
        self._stream = ctypes.c_void_p(torch.cuda.current_stream(x0.device).cuda_stream)
        self._ptr = {id(t): ctypes.c_void_p(t.data_ptr())
                     for t in [self.x, self.g, self.d, self._out, self.energy.rvecs] +
                     [t for slot in self._slots for t in slot]}
        self._stats = [None, None]
        self._last = None
        self._best = None
        self._beta = 0.0
        self._set_dir = True  # d <- beta d - g happens in the next trial
        self._dp = 0.0
        self.gg = None
        self._measures = None
        self._checked = False

    def _energy(self, e):
        if np.isnan(e):
            raise ValueError("The energy is not-a-number (``nan``).")
        press = self.energy.press
        if press is not None:
            e += press.pext * self.energy.fpm.system.domain.volume
        return e

    def _measure(self, x, x_old, g, g_old):
        """(sum |g_n|^2, sum g_n . g_old_n, sum |dx_n|^2, max |g_n|^2, max |dx_n|^2)."""
        self.field.cg_stats(x, x_old, g, g_old, self._out, host_out=self._host)
        return tuple(self._host[:5])

    def start(self):
        en = self.energy
        en.calls += 1
        en.fpm.compute_device(self.x.view(-1, 3), en.rvecs, self._out[7:8], self.g.view(-1, 3))
        e = float(self._out[7].item())
        self.gg = self._measure(self.x, self.x, self.g, self.g)[0]
        self._beta, self._set_dir, self._dp = 0.0, True, 0.0
        self._measures, self._checked, self._best = None, False, None
        return self._energy(e), self.gg

    def slope(self):
        if self._set_dir:
            # g.(beta d - g) with g.d = dp, the slope of the accepted trial
            return -self.gg + self._beta * self._dp
        return float(torch.dot(self.g, self.d).item())

    def restart(self):
        self._beta, self._set_dir, self._dp = 0.0, True, 0.0

    def trial(self, a):
        k = 1 if self._best == 0 else 0
        xa, ga = self._slots[k]
        en = self.energy
        en.calls += 1
        P = self._ptr
        rc = self._lib.mm_cg_trial(self._handle, P[id(self.x)], P[id(self.g)], P[id(self.d)],
                                   self._beta, int(self._set_dir), a, P[id(en.rvecs)], P[id(xa)],
                                   P[id(ga)], P[id(self._out)], self._host, self._stream)
        if rc:
            self._check(rc, "mm_cg_trial")
        self._set_dir = False
        e, dp, gg, g_gold, gsq_max, dsq, dsq_max = self._host[:7]
        self._stats[k] = (gg, g_gold, gsq_max, dsq, dsq_max)
        self._last = k
        return self._energy(e), dp

    def keep(self):
        self._best = self._last

    def accept(self):
        k = self._best
        gg, g_gold, gsq_max, dsq, dsq_max = self._stats[k]
        xa, ga = self._slots[k]
        self._slots[k] = [self.x, self.g]
        self.x, self.g = xa, ga
        self._best = None
        self._measures = (gg, gsq_max, dsq, dsq_max)
        self.gg = gg
        return gg, gg - g_gold

    def set_direction(self, beta, dp):
        self._beta, self._set_dir, self._dp = beta, True, dp

    def take_measures(self):
        """(sum, max of |g_n|^2, sum, max of |x_n - x_last_n|^2) for a convergence check, or
        None for the first check after ``start``."""
        if not self._checked:
            self._checked = True
            self._measures = None
            return None
        m = self._measures
        if m is None:
            # no step since the last check: the displacement is zero
            gsq, _, _, gsq_max, _ = self._measure(self.x, self.x, self.g, self.g)
            m = (gsq, gsq_max, 0.0, 0.0)
        self._measures = None
        return m


class DOF(object):
    def __init__(self, mmf):
        self.mmf = mmf
        self.x0 = None
        self._init_initial()
        self._gx = np.zeros(self.ndof, float)

    def _init_initial(self):
        raise NotImplementedError

    ndof = property(lambda self: len(self.x0))

    def _update(self, x):
        raise NotImplementedError

    def reset(self):
        self._update(self.x0)

    def check_delta(self, x=None, eps=1e-4, zero=None, ncheck=100):
        """Compare the analytic gradient with finite differences along random displacements
        (the recipe of dof.py:62-69). Returns the largest relative discrepancy."""
        if x is None:
            x = self.x0
        dxs = np.random.uniform(-eps, eps, (ncheck, len(x)))
        if zero is not None:
            dxs[:, zero] = 0.0
        worst = 0.0
        for dx in dxs:
            f0, g0 = self.fun(x - 0.5 * dx, True)
            f1, g1 = self.fun(x + 0.5 * dx, True)
            exact = f1 - f0
            approx = 0.5 * np.dot(g0 + g1, dx)
            worst = max(worst, abs(exact - approx) / max(abs(exact), 1e-300))
        self._update(x)
        return worst


class CartesianDOF(DOF):
    """Cartesian coordinates of the nodes (dof.py:75-193); the domain is kept fixed."""

    def __init__(self, mmf, gpos_rms=1e-5, dpos_rms=1e-3, select=None):
        self.th_gpos_rms = gpos_rms
        self.th_dpos_rms = dpos_rms
        self.select = select
        DOF.__init__(self, mmf)
        self._last_pos = None
        self._last_pos_dev = None
        self._cg_engine = None
        try:
            self._device = _DeviceEnergy(mmf)
        except NotImplementedError:
            self._device = None

    def _init_initial(self):
        pos = np.asarray(self.mmf.system.pos)
        self.x0 = (pos.ravel() if self.select is None else pos[self.select].ravel()).copy()
        self._pos = pos.copy()
        self._dpos = np.zeros(pos.shape, float)
        self._gpos = np.zeros(pos.shape, float)

    @property
    def device(self):
        return None if self._device is None else self._device.device

    def _update(self, x):
        if isinstance(x, torch.Tensor):
            x = x.detach().cpu().numpy()
        if self.select is None:
            self._pos[:] = x.reshape(-1, 3)
        else:
            self._pos[self.select] = x.reshape(-1, 3)
        self.mmf.update_pos(self._pos[:])

    def cg_engine(self, x0):
        """Fused device vectors for ``ConjugateGradientSearch`` starting at the device tensor
        x0 (all nodes free, device force field), or None. While the engine drives the search,
        ``check_convergence`` takes its measures from the engine's accepted steps."""
        if self._device is None or self.select is not None or \
                not (isinstance(x0, torch.Tensor) and x0.is_cuda):
            return None
        self._cg_engine = _DeviceCGEngine(self, x0)
        self._pin_host_arrays()  # made before the run: page-locking 2 x 24 B per node is slow
        return self._cg_engine

    def fun(self, x, do_gradient=False):
        """Energy and, with do_gradient, its gradient with respect to x (dof.py:132-153)."""
        # an evaluation outside the engine: the engine no longer holds the current state
        self._cg_engine = None
        if isinstance(x, torch.Tensor) and x.is_cuda:
            if self._device is None:
                raise NotImplementedError("device evaluation needs one ForcePartMechanical")
            if self.select is None:
                pos = x.view(-1, 3)
            else:
                pos = torch.as_tensor(self._pos, device=x.device).clone()
                pos[torch.as_tensor(self.select, device=x.device)] = x.view(-1, 3)
            self._pos_dev = pos
            e, g = self._device(pos.contiguous(), do_gradient)
            if not do_gradient:
                return e
            self._gpos_dev = g
            gx = g.reshape(-1) if self.select is None else \
                g[torch.as_tensor(self.select, device=g.device)].reshape(-1)
            return e, gx.clone()
        self._update(x)
        if do_gradient:
            self._gpos[:] = 0.0
            v = self.mmf.compute(self._gpos)
            self._gx[:] = self._gpos.ravel() if self.select is None else \
                self._gpos[self.select].ravel()
            return v, self._gx.copy()
        return self.mmf.compute()

    def _accept(self, x, g):
        """Make the accepted point x (with gradient g) of an optimiser the current state."""
        if isinstance(x, torch.Tensor):
            if self.select is None:
                self._pos_dev = x.view(-1, 3)
                self._gpos_dev = g.view(-1, 3)
        else:
            self._update(x)
            if self.select is None:
                self._gpos[:] = np.asarray(g).reshape(-1, 3)
            else:
                self._gpos[self.select] = np.asarray(g).reshape(-1, 3)

    def _pin_host_arrays(self):
        """Back the host position and gradient arrays with pinned memory, so that a copy from
        the device is one DMA transfer straight into them."""
        if getattr(self, "_pos_pin", None) is None:
            self._pos_pin = torch.from_numpy(self._pos).pin_memory()
            self._gpos_pin = torch.from_numpy(self._gpos).pin_memory()
            self._pos = self._pos_pin.numpy()
            self._gpos = self._gpos_pin.numpy()

    def sync_from_device(self):
        """Copy the last device positions and gradient into the host arrays and the system."""
        pinned = getattr(self, "_pos_pin", None) is not None
        if getattr(self, "_pos_dev", None) is not None:
            if pinned and self._pos_dev.shape == self._pos_pin.shape:
                self._pos_pin.copy_(self._pos_dev.detach())
            else:
                self._pos[:] = self._pos_dev.detach().cpu().numpy()
            self.mmf.update_pos(self._pos[:])
        if getattr(self, "_gpos_dev", None) is not None:
            if pinned and self._gpos_dev.shape == self._gpos_pin.shape:
                self._gpos_pin.copy_(self._gpos_dev.detach())
            else:
                self._gpos[:] = self._gpos_dev.detach().cpu().numpy()

    def _device_measures(self):
        """gpos_max, gpos_rms, dpos_max, dpos_rms of the device state (one small copy)."""
        pos, g = self._pos_dev, self._gpos_dev
        if self.select is not None:
            sel = torch.as_tensor(self.select, device=pos.device)
            pos, g = pos[sel], g[sel]
        gsq = (g * g).sum(dim=1)
        first = self._last_pos_dev is None
        if first:
            dsq = torch.zeros_like(gsq)
        else:
            d = pos - self._last_pos_dev
            dsq = (d * d).sum(dim=1)
        m = torch.stack([gsq.max(), gsq.mean(), dsq.max(), dsq.mean()]).sqrt().cpu().numpy()
        self._last_pos_dev = pos.detach().clone()
        return first, m

    def check_convergence(self):
        """Convergence criteria of dof.py:155-193. On the device path the measures are reduced
        on the GPU; the host arrays are refreshed by ``sync_from_device`` when needed."""
        if self._cg_engine is not None:
            m = self._cg_engine.take_measures()
            if m is None:
                self.converged = False
                self.conv_val = 2
                self.conv_worst = "first_step"
                self.conv_count = -1
                return
            gsq_sum, gsq_max, dsq_sum, dsq_max = m
            nn = self._cg_engine.x.numel() // 3
            self.gpos_max, self.gpos_rms = np.sqrt(gsq_max), np.sqrt(gsq_sum / nn)
            self.dpos_max, self.dpos_rms = np.sqrt(dsq_max), np.sqrt(dsq_sum / nn)
            self._finish_convergence()
            if self.converged:
                self.sync_from_device()
            return
        if getattr(self, "_pos_dev", None) is not None and getattr(self, "_gpos_dev", None) \
                is not None:
            first, (self.gpos_max, self.gpos_rms, self.dpos_max, self.dpos_rms) = \
                self._device_measures()
            if first:
                self.converged = False
                self.conv_val = 2
                self.conv_worst = "first_step"
                self.conv_count = -1
                return
            self._finish_convergence()
            if self.converged:
                self.sync_from_device()
            return
        self.sync_from_device()
        if self._last_pos is None:
            self._last_pos = self._pos.copy()
            self.converged = False
            self.conv_val = 2
            self.conv_worst = "first_step"
            self.conv_count = -1
            return
        g = self._gpos if self.select is None else self._gpos[self.select]
        gpossq = (g ** 2).sum(axis=1)
        self.gpos_max = np.sqrt(gpossq.max())
        self.gpos_rms = np.sqrt(gpossq.mean())
        self._dpos[:] = self._pos
        self._dpos -= self._last_pos
        d = self._dpos if self.select is None else self._dpos[self.select]
        dpossq = (d ** 2).sum(axis=1)
        self.dpos_max = np.sqrt(dpossq.max())
        self.dpos_rms = np.sqrt(dpossq.mean())
        self._last_pos[:] = self._pos
        self._finish_convergence()

    def _finish_convergence(self):
        conv_vals = []
        if self.th_gpos_rms is not None:
            conv_vals.append((self.gpos_rms / self.th_gpos_rms, "gpos_rms"))
            conv_vals.append((self.gpos_max / (self.th_gpos_rms * 3), "gpos_max"))
        if self.th_dpos_rms is not None:
            conv_vals.append((self.dpos_rms / self.th_dpos_rms, "dpos_rms"))
            conv_vals.append((self.dpos_max / (self.th_dpos_rms * 3), "dpos_max"))
        if len(conv_vals) == 0:
            raise RuntimeError("At least one convergence criterion must be present.")
        self.conv_val, self.conv_worst = max(conv_vals)
        self.conv_count = sum(int(v >= 1) for v, n in conv_vals)
        self.converged = self.conv_count == 0


class BaseCellDOF(DOF):
    """Cell variables and fractional coordinates (dof.py:199-484).

    x = [celldofs, frac] (frac omitted with ``do_frozen``): ``celldofs`` is the ``freemask``
    selection of the subclass's cell variables, ``frac`` the fractional node coordinates in the
    full domain matrix. The gradient with respect to the cell variables is
    (gvecs . vtens) : d rvecs / d celldofs, with respect to frac gpos . rvecs^T; ``grvecs`` for the
    convergence test is gvecs . vtens projected on the span of the Jacobian. Device tensors x are
    evaluated with the HIP kernels: fractional coordinates, positions and gradients stay on the
    GPU, only the cell variables (at most 9 numbers) and the virial cross to the host.
    """

    def __init__(self, mmf, gpos_rms=1e-5, dpos_rms=1e-3, grvecs_rms=1e-5, drvecs_rms=1e-3,
                 do_frozen=False, freemask=None):
        if freemask is not None:
            if not (isinstance(freemask, np.ndarray) and issubclass(freemask.dtype.type, np.bool_)
                    and len(freemask.shape) == 1 and freemask.sum() > 0):
                raise TypeError("When given, freemask must be a vector of booleans.")
        self.th_gpos_rms = gpos_rms
        self.th_dpos_rms = dpos_rms
        self.th_grvecs_rms = grvecs_rms
        self.th_drvecs_rms = drvecs_rms
        self.do_frozen = do_frozen
        self.freemask = freemask
        DOF.__init__(self, mmf)
        self._last_pos = None
        self._last_rvecs = None
        self._last_pos_dev = None
        self._pos_dev = None
        self._gpos_dev = None
        self._eval_x = None
        try:
            self._device = _DeviceEnergy(mmf)
        except NotImplementedError:
            self._device = None

    @property
    def device(self):
        return None if self._device is None else self._device.device

    ncellvar = property(lambda self: len(self.domainvars0))

    @property
    def ncelldof(self):
        return len(self.domainvars0) if self.freemask is None else int(self.freemask.sum())

    def _reduce_cellvars(self, cellvars):
        return cellvars if self.freemask is None else cellvars[self.freemask]

    def _expand_celldofs(self, celldofs):
        if self.freemask is None:
            return celldofs
        cellvars = self.domainvars0.copy()
        cellvars[self.freemask] = celldofs
        return cellvars

    def _isfree(self, icellvar):
        return True if self.freemask is None else bool(self.freemask[icellvar])

    def _init_initial(self):
        self.domainvars0 = self._get_initial_cellvars()
        if self.freemask is not None and len(self.freemask) != self.ncellvar:
            raise TypeError("The length of the freemask vector (%i) does not match the number "
                            "of cellvars (%i)." % (len(self.freemask), len(self.domainvars0)))
        celldofs0 = self._reduce_cellvars(self.domainvars0)
        domain = self.mmf.system.domain
        frac = np.dot(np.asarray(self.mmf.system.pos), domain._get_gvecs(full=True).T)
        if self.do_frozen:
            self.x0 = celldofs0.copy()
            self._frac0 = frac
        else:
            self.x0 = np.concatenate([celldofs0, frac.ravel()])
        pos = np.asarray(self.mmf.system.pos)
        self._pos = pos.copy()
        self._dpos = np.zeros(pos.shape, float)
        self._gpos = np.zeros(pos.shape, float)
        self._rvecs = np.array(domain.rvecs, dtype=float)
        self._dcell = np.zeros(self._rvecs.shape, float)
        self._vtens = np.zeros((3, 3), float)
        self._grvecs = np.zeros(self._rvecs.shape, float)

    # -------------------------------------------------------------------------- evaluation
    def _set_cell(self, celldofs):
        """Domain vectors of the cell variables, stored in the force field (host)."""
        self._rvecs = self._cellvars_to_rvecs(self._expand_celldofs(celldofs))
        self.mmf.update_rvecs(self._rvecs[:])
        return self.mmf.system.domain._get_rvecs(full=True)

    def _cell_gradient(self, celldofs, vtens):
        """Cell part of the gradient, and grvecs projected on the Jacobian's span."""
        grvecs = np.dot(self.mmf.system.domain.gvecs, vtens)
        jacobian = self._get_celldofs_jacobian(celldofs)
        assert jacobian.shape == (grvecs.size, self.ncelldof)
        gcell = np.dot(grvecs.ravel(), jacobian)
        U, S, Vt = np.linalg.svd(jacobian, full_matrices=False)
        self._grvecs[:] = np.dot(U, np.dot(U.T, grvecs.ravel())).reshape(-1, 3)
        return gcell

    def _update(self, x):
        ncd = self.ncelldof
        rvecs_full = self._set_cell(x[:ncd])
        frac = self._frac0 if self.do_frozen else x[ncd:].reshape(-1, 3)
        self._pos[:] = np.dot(frac, rvecs_full)
        self.mmf.update_pos(self._pos[:])

    def fun(self, x, do_gradient=False):
        """Energy and, with do_gradient, its gradient with respect to x (dof.py:360-392)."""
        ncd = self.ncelldof
        if isinstance(x, torch.Tensor) and x.is_cuda:
            if self._device is None:
                raise NotImplementedError("device evaluation needs one ForcePartMechanical")
            celldofs = x[:ncd].detach().cpu().numpy()
            rvecs_full = self._set_cell(celldofs)
            self._device.rvecs = self._device.fpm.rvecs_device()
            R = torch.tensor(np.array(rvecs_full, dtype=float), device=x.device)
            frac = (torch.as_tensor(self._frac0, dtype=torch.float64, device=x.device)
                    if self.do_frozen else x[ncd:].view(-1, 3))
            pos = torch.matmul(frac, R).contiguous()
            self._pos_dev = pos
            self._eval_x = x
            if not do_gradient:
                return self._device(pos, False)[0]
            e, g, vtens = self._device(pos, True, with_vtens=True)
            self._gpos_dev = g.clone()
            self._vtens[:] = vtens
            gx = torch.empty_like(x)
            gx[:ncd] = torch.as_tensor(self._cell_gradient(celldofs, vtens), device=x.device)
            if not self.do_frozen:
                gx[ncd:] = torch.matmul(self._gpos_dev, R.t()).reshape(-1)
            return e, gx
        self._update(x)
        self._eval_x = x
        if do_gradient:
            self._gpos[:] = 0.0
            self._vtens[:] = 0.0
            v = self.mmf.compute(self._gpos, self._vtens)
            self._gx[:ncd] = self._cell_gradient(x[:ncd], self._vtens)
            if not self.do_frozen:
                self._gx[ncd:] = np.dot(self._gpos, self._rvecs.T).ravel()
            return v, self._gx.copy()
        return self.mmf.compute()

    def _accept(self, x, g):
        """Make the accepted point x of an optimiser the current state (re-evaluated unless it
        was the last point evaluated)."""
        if x is not self._eval_x:
            self.fun(x, True)

    def sync_from_device(self):
        if self._pos_dev is not None:
            stage = getattr(self, "_stage", None)
            if stage is None or stage.shape != self._pos_dev.shape:
                stage = self._stage = _pinned_like(self._pos_dev)
            _to_host(self._pos_dev, self._pos, stage)
            self.mmf.update_pos(self._pos[:])
            if self._gpos_dev is not None:
                _to_host(self._gpos_dev, self._gpos, stage)
        elif self._gpos_dev is not None:
            self._gpos[:] = self._gpos_dev.cpu().numpy()

    # -------------------------------------------------------------------------- convergence
    def check_convergence(self):
        """Convergence criteria of dof.py:394-457 (positions and gradients reduced on the GPU on
        the device path)."""
        device = self._pos_dev is not None and self._gpos_dev is not None
        if self._last_rvecs is None:
            self._last_rvecs = self._rvecs.copy()
            if device:
                self._last_pos_dev = self._pos_dev.detach().clone()
            else:
                self._last_pos = self._pos.copy()
            self.converged = False
            self.conv_val = 2
            self.conv_worst = "first_step"
            self.conv_count = -1
            return
        if device:
            g, d = self._gpos_dev, self._pos_dev - self._last_pos_dev
            gsq, dsq = (g * g).sum(dim=1), (d * d).sum(dim=1)
            m = torch.stack([gsq.max(), gsq.mean(), dsq.max(), dsq.mean(),
                             gsq.argmax().to(torch.float64)]).cpu().numpy()
            gmax, grms = np.sqrt(m[0]), np.sqrt(m[1])
            self.dpos_max, self.dpos_rms = np.sqrt(m[2]), np.sqrt(m[3])
            gind = int(m[4])
            self._last_pos_dev = self._pos_dev.detach().clone()
        else:
            gsq = (self._gpos ** 2).sum(axis=1)
            gmax, grms, gind = np.sqrt(gsq.max()), np.sqrt(gsq.mean()), int(gsq.argmax())
            self._dpos[:] = self._pos - self._last_pos
            dsq = (self._dpos ** 2).sum(axis=1)
            self.dpos_max, self.dpos_rms = np.sqrt(dsq.max()), np.sqrt(dsq.mean())
            self._last_pos[:] = self._pos
        if not self.do_frozen:
            self.gpos_max, self.gpos_rms, self.gpos_indmax = gmax, grms, gind
        grsq = (self._grvecs ** 2).sum(axis=1)
        self.grvecs_max = np.sqrt(grsq.max())
        self.grvecs_rms = np.sqrt(grsq.mean())
        self._dcell[:] = self._rvecs - self._last_rvecs
        dcsq = (self._dcell ** 2).sum(axis=1)
        self.drvecs_max = np.sqrt(dcsq.max())
        self.drvecs_rms = np.sqrt(dcsq.mean())
        conv_vals = []
        if not self.do_frozen and self.th_gpos_rms is not None:
            conv_vals.append((self.gpos_rms / self.th_gpos_rms, "gpos_rms"))
            conv_vals.append((self.gpos_max / (self.th_gpos_rms * 3),
                              "gpos_max(%i)" % self.gpos_indmax))
        if self.th_dpos_rms is not None:
            conv_vals.append((self.dpos_rms / self.th_dpos_rms, "dpos_rms"))
            conv_vals.append((self.dpos_max / (self.th_dpos_rms * 3), "dpos_max"))
        if self.th_grvecs_rms is not None:
            conv_vals.append((self.grvecs_rms / self.th_grvecs_rms, "grvecs_rms"))
            conv_vals.append((self.grvecs_max / (self.th_grvecs_rms * 3), "grvecs_max"))
        if self.th_drvecs_rms is not None:
            conv_vals.append((self.drvecs_rms / self.th_drvecs_rms, "drvecs_rms"))
            conv_vals.append((self.drvecs_max / (self.th_drvecs_rms * 3), "drvecs_max"))
        if len(conv_vals) == 0:
            raise RuntimeError("At least one convergence criterion must be present.")
        self.conv_val, self.conv_worst = max(conv_vals)
        self.conv_count = sum(int(v >= 1) for v, n in conv_vals)
        self.converged = self.conv_count == 0
        self._last_rvecs[:] = self._rvecs
        if self.converged and device:
            self.sync_from_device()

    # -------------------------------------------------------------------------- subclass API
    def _get_initial_cellvars(self):
        raise NotImplementedError

    def _cellvars_to_rvecs(self, cellvars):
        raise NotImplementedError

    def _get_celldofs_jacobian(self, x):
        raise NotImplementedError


class FullCellDOF(BaseCellDOF):
    """All components of the domain vectors, divided by s = V^(1/nvec) of the initial domain
    so that the cell variables are dimensionless (dof.py:487-514)."""

    def _get_initial_cellvars(self):
        domain = self.mmf.system.domain
        if domain.nvec == 0:
            raise ValueError("A cell optimization requires a system that is periodic.")
        self._rvecs_scale = domain.volume ** (1.0 / domain.nvec)
        return np.array(domain.rvecs, dtype=float).ravel() / self._rvecs_scale

    def _cellvars_to_rvecs(self, cellvars):
        return np.asarray(cellvars, dtype=float).reshape(-1, 3) * self._rvecs_scale

    def _get_celldofs_jacobian(self, x):
        result = np.identity(self.ncellvar) * self._rvecs_scale
        if self.freemask is not None:
            result = result[:, self.freemask]
        return result


class StrainCellDOF(BaseCellDOF):
    """A symmetric transformation A of the initial domain vectors, rotations eliminated
    (dof.py:517-697). Cell variables: 3D [A00, A11, A22, 2 A12, 2 A20, 2 A01],
    2D [A00, A11, 2 A01], 1D [A00]."""

    # (row, column) pairs of the symmetric matrix and their weight per cell variable
    _PAIRS = {3: [((0, 0),), ((1, 1),), ((2, 2),), ((1, 2), (2, 1)), ((2, 0), (0, 2)),
                  ((0, 1), (1, 0))],
              2: [((0, 0),), ((1, 1),), ((0, 1), (1, 0))],
              1: [((0, 0),)]}

    def _get_initial_cellvars(self):
        domain = self.mmf.system.domain
        if domain.nvec == 0:
            raise ValueError("A cell optimization requires a system that is periodic.")
        self.rvecs0 = np.array(domain.rvecs, dtype=float)
        nvec = domain.nvec
        return np.array([1.0] * nvec + [0.0] * (nvec * (nvec + 1) // 2 - nvec))

    def _deform(self, cellvars):
        nvec = len(self.rvecs0)
        A = np.zeros((nvec, nvec))
        for k, pairs in enumerate(self._PAIRS[nvec]):
            w = 1.0 if len(pairs) == 1 else 0.5
            for (i, j) in pairs:
                A[i, j] = w * cellvars[k]
        return A

    def _cellvars_to_rvecs(self, cellvars):
        nvec = len(self.rvecs0)
        return np.dot(self._deform(np.asarray(cellvars)[: nvec * (nvec + 1) // 2]), self.rvecs0)

    def _get_celldofs_jacobian(self, x):
        # d rvecs[i, :] / d cellvar k = sum over the (i, j) pairs of cellvar k of w * rvecs0[j]
        nvec = len(self.rvecs0)
        cols = []
        for k, pairs in enumerate(self._PAIRS[nvec]):
            if not self._isfree(k):
                continue
            w = 1.0 if len(pairs) == 1 else 0.5
            col = np.zeros((nvec, 3))
            for (i, j) in pairs:
                col[i] += w * self.rvecs0[j]
            cols.append(col.ravel())
        return np.array(cols).T
