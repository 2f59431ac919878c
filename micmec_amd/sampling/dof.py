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

import numpy as np
import torch

__all__ = ["DOF", "CartesianDOF"]


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
        self.rvecs = self.fpm.rvecs_device()
        self.calls = 0

    def __call__(self, pos, do_gradient=True):
        self.calls += 1
        self.fpm.compute_device(pos, self.rvecs, self.energy, self.gpos if do_gradient else None)
        e = float(self.energy.item())
        if np.isnan(e):
            raise ValueError("The energy is not-a-number (``nan``).")
        if self.press is not None:
            e += self.press.pext * self.fpm.system.domain.volume
        return e, self.gpos


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

    def fun(self, x, do_gradient=False):
        """Energy and, with do_gradient, its gradient with respect to x (dof.py:132-153)."""
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

    def sync_from_device(self):
        """Copy the last device positions and gradient into the host arrays and the system."""
        if getattr(self, "_pos_dev", None) is not None:
            self._pos[:] = self._pos_dev.cpu().numpy()
            self.mmf.update_pos(self._pos[:])
        if getattr(self, "_gpos_dev", None) is not None:
            self._gpos[:] = self._gpos_dev.cpu().numpy()

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
