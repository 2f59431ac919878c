"""Differentiation form: energy, forces and virial as differentiable functions of node
positions, domain vectors and the cell-type parameters.

``micmec_energy(field, pos, rvecs, params)`` returns per-system energies [nsys], the node
gradient gpos [N, 3] and the virials [nsys, 3, 3] of a :class:`DiffField` (one or many systems
in one device object) and is a ``torch.autograd.Function``: a loss built from any of the three
outputs back-propagates to ``pos``, ``rvecs`` and the parameter tensors

* ``cv``   [T, S, 6, 6]  symmetric Voigt elasticity matrix of every (type, state),
  u = e.Cv.e / 2 with the engineering strain e = (eps00, eps11, eps22, 2 eps12, 2 eps02,
  2 eps01); its gradient is returned symmetric and holds for symmetric perturbations (the 21
  independent entries). The forces use the stress Cv e as the reference does (sigma = C:eps,
  micmec/pes/nanocell.py:98-132), which is the gradient of the energy only for symmetric Cv;
* ``h0``   [T, S, 3, 3]  rest cell matrix (rows = cell vectors);
* ``free`` [T, S]        free energy of every state;
* ``teff`` [T]           effective temperature (kelvin) of the state combination.

The energy model is the reference's (micmec/pes/nanocell.py:40-132, nanocell_original.py:46-132,
micmec/pes/mmff.py:380-398); the reference itself differentiates only with respect to positions
(mmff.py:384-398) and, in mmff_jax.py:249-250, a deformation of the domain. The forward pass is
the plain force path (``mm_batch_compute``); the backward pass is one ``mm_adjoint_vjp`` call:
the tangent direction below, one launch of the hand-written adjoint kernel in dual numbers
(``mm_adjoint_compute``), which recomputes everything from the positions (no per-corner
intermediate is stored between the passes), and the position cotangent.

Backward, for cotangents (gE, gG, gV) of (E, gpos, vtens): with the tangent direction
u = (dpos, drvecs) = (gG + pos gV^T, rvecs gV^T) the kernel returns the values (gpos, dE/drvecs,
dE/dtheta) and their derivatives along u; then

    dL/dpos   = gE gpos + d(gpos)[u] + gpos gV
    dL/drvecs = gE dE/drvecs + d(dE/drvecs)[u] + (dE/drvecs) gV
    dL/dtheta = gE dE/dtheta + d(dE/dtheta)[u]

(the last terms of the first two lines come from vtens = sum_v g_v (x) r_v, whose derivative
along a deformation D is the derivative of E along (pos D, rvecs D)). The parameter entries of
the kernel (derivatives with respect to Cv, P = h0^-1, the weight w = det(h0)/8 or det(h0), the
free energies and kT) are mapped here to (cv, h0, free, teff).
"""

import ctypes
import weakref

import numpy as np
import torch
from torch.autograd.function import once_differentiable

from . import _native, units
from .batch import concat_systems

__all__ = ["DiffField", "CellParams", "micmec_energy", "voigt_from_tensor", "tensor_from_voigt",
           "voigt_pairs", "fit_workload", "FitProblem", "fit"]

_P = ctypes.c_void_p
_i32, _i64, _f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
_native.register_signature("mm_batch_field_create",
                           [ctypes.POINTER(_P), ctypes.c_int, ctypes.c_int, _i32, _i64, _i64, _P,
                            _P, _P, _P, _P, _i32, _i32, _P, _P, _P, _P, _P, _f64])
_native.register_signature("mm_batch_compute", [_P, _P, _P, _P, _P, _P, _P])
_native.register_signature("mm_field_set_params", [_P, _P, _P, _P, _P, _P, _f64, _P])
_native.register_signature("mm_adjoint_create", [ctypes.POINTER(_P), _P, ctypes.POINTER(_i64)])
_native.register_signature("mm_adjoint_compute", [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P])
_native.register_signature("mm_adjoint_vjp", [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P])
_native.register_signature("mm_adjoint_param_grads",
                           [_P, _P, _i64, ctypes.c_int, _P, _f64, _P, _P, _P, _P, _P])

NPS = 32       # parameter entries per (type, state) of the adjoint kernel
SYSQ = 19      # energy, virial (9), domain-vector gradient (9)
# Voigt index -> tensor index pair (micmec/analysis/tensor.py:188-197)
_VOIGT = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))


def voigt_pairs():
    """The 21 (I, J), I <= J, entries of the upper triangle, row by row."""
    return [(I, J) for I in range(6) for J in range(I, 6)]


def voigt_from_tensor(C):
    """6x6 Voigt matrix of a 3x3x3x3 elasticity tensor of full symmetry (leading axes kept)."""
    C = np.asarray(C, dtype=np.float64)
    out = np.zeros(C.shape[:-4] + (6, 6))
    for I, (i, j) in enumerate(_VOIGT):
        for J, (k, l) in enumerate(_VOIGT):
            out[..., I, J] = C[..., i, j, k, l]
    return out


def tensor_from_voigt(cv):
    """3x3x3x3 tensor with C[i,j,k,l] = cv[I(ij), J(kl)] (leading axes kept)."""
    cv = np.asarray(cv, dtype=np.float64)
    inv = {}
    for I, (i, j) in enumerate(_VOIGT):
        inv[(i, j)] = inv[(j, i)] = I
    out = np.zeros(cv.shape[:-2] + (3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    out[..., i, j, k, l] = cv[..., inv[(i, j)], inv[(k, l)]]
    return out


class CellParams(object):
    """Differentiable cell-type parameters (float64 tensors) and the fixed state counts."""

    def __init__(self, cv, h0, free, teff, nstates, type_ids=None):
        self.cv, self.h0, self.free, self.teff = cv, h0, free, teff
        self.nstates = np.asarray(nstates, dtype=np.int32)
        self.type_ids = list(type_ids) if type_ids is not None else list(range(len(nstates)))

    @classmethod
    def from_table(cls, table, device, requires_grad=False):
        ids, ns, h0, c0, free, temp = table
        kw = dict(dtype=torch.float64, device=device)
        t = [torch.tensor(voigt_from_tensor(c0), **kw), torch.tensor(h0, **kw),
             torch.tensor(free, **kw), torch.tensor(temp, **kw)]
        if requires_grad:
            for x in t:
                x.requires_grad_(True)
        return cls(*t, ns, ids)

    def tensors(self):
        return self.cv, self.h0, self.free, self.teff

    def detached(self):
        return CellParams(*[x.detach().clone() for x in self.tensors()], self.nstates,
                          self.type_ids)


class DiffField(object):
    """One or more systems in one batched device object, with the adjoint work buffers.

    Nodes of system k occupy rows ``node_start[k]:node_start[k] + nnodes_k`` of the position
    array; ``node_sys`` [N] gives each row's system (padding rows: the preceding system, and
    they carry no cell)."""

    def __init__(self, systems, model="corner8", device=None):
        if model not in _native.MODELS:
            raise ValueError("Unknown model %r; choose one of %s" % (model, list(_native.MODELS)))
        self.systems = list(systems)
        self.nsys = len(self.systems)
        if self.nsys < 1:
            raise ValueError("need at least one system")
        self.model = model
        self.device = torch.device(device if device is not None else "cuda")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        topo = concat_systems(self.systems, 1)
        self.table = topo["types"]
        ids, ns, h0, c0, free, temp = self.table
        self.nrows, self.max_states = len(ids), int(h0.shape[1])
        self.node_start = topo["node_start"]
        self.N = topo["nnodes"]
        self.ncells = topo["ncells"]
        self.lib = _native.lib()
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        handle = ctypes.c_void_p()
        rc = self.lib.mm_batch_field_create(
            ctypes.byref(handle), self.device.index, _native.MODELS[model], self.nsys,
            topo["ncells"], self.N, vp(topo["surrounding_nodes"]), vp(topo["surrounding_cells"]),
            vp(topo["wrap_bits"]), vp(topo["cell_row"]), vp(topo["cell_sys"]), self.nrows,
            self.max_states, vp(ns), vp(h0), vp(c0), vp(free), vp(temp),
            float(units.boltzmann))
        _native.check(rc, "mm_batch_field_create")
        self._field = handle
        adj = ctypes.c_void_p()
        npar = _i64(0)
        rc = self.lib.mm_adjoint_create(ctypes.byref(adj), self._field, ctypes.byref(npar))
        if rc != _native.MM_OK:
            self.lib.mm_field_destroy(self._field)
            self._field = None
        _native.check(rc, "mm_adjoint_create")
        self._adj = adj
        self.npar = int(npar.value)
        counts = np.diff(self.node_start)
        self.node_sys = torch.tensor(np.repeat(np.arange(self.nsys), counts), dtype=torch.long,
                                     device=self.device)
        self._node_sys32 = self.node_sys.to(torch.int32)

    # ------------------------------------------------------------------ inputs
    def positions(self):
        """Concatenated positions [N, 3] and domain vectors [nsys, 3, 3] of the systems."""
        pos = np.zeros((self.N, 3))
        rv = np.zeros((self.nsys, 3, 3))
        for k, s in enumerate(self.systems):
            a = self.node_start[k]
            pos[a:a + s.nnodes] = s.pos
            r = np.asarray(s.domain.rvecs)
            rv[k, : len(r)] = r
        kw = dict(dtype=torch.float64, device=self.device)
        return torch.tensor(pos, **kw), torch.tensor(rv, **kw)

    def initial_params(self, requires_grad=False):
        return CellParams.from_table(self.table, self.device, requires_grad)

    def set_params(self, params):
        """Make ``params`` (a CellParams) the object's parameter table: a kernel on the current
        stream builds it from the device tensors, so nothing synchronises. Uploading the same,
        unmodified tensors again (the backward pass after its forward) is skipped; a change is
        recognised by the tensors' version counters, which in-place operations on the tensors
        (or their detached views) advance but writes through ``.data`` do not: after such a
        write, set ``field._uploaded = None`` to force the upload. The record holds weak
        references only, so that it keeps no autograd history alive."""
        ts = params.tensors()
        key = (tuple(weakref.ref(x) for x in ts), tuple(x._version for x in ts),
               params.nstates.tobytes())
        last = getattr(self, "_uploaded", None)
        if last is not None and len(last[0]) == len(ts) and \
                all(r() is x for r, x in zip(last[0], ts)) and last[1:] == key[1:]:
            return
        kw = dict(dtype=torch.float64, device=self.device)
        cv, h0, free, teff = [x.detach().to(**kw).contiguous() for x in ts]
        ns = np.ascontiguousarray(params.nstates, dtype=np.int32)
        rc = self.lib.mm_field_set_params_device(
            self._field, ns.ctypes.data_as(ctypes.c_void_p), _native.ptr(h0), _native.ptr(cv),
            _native.ptr(free), _native.ptr(teff), float(units.boltzmann), _native.stream_ptr())
        _native.check(rc, "mm_field_set_params_device")
        self._uploaded = key

    # ------------------------------------------------------------------ kernels
    def compute(self, pos, rvecs):
        """Energies [nsys], gpos [N, 3], vtens [nsys, 3, 3] (plain force path)."""
        kw = dict(dtype=torch.float64, device=self.device)
        e = torch.empty(self.nsys, **kw)
        g = torch.empty((self.N, 3), **kw)
        v = torch.empty((self.nsys, 3, 3), **kw)
        rc = self.lib.mm_batch_compute(self._field, _native.ptr(pos), _native.ptr(rvecs),
                                       _native.ptr(e), _native.ptr(g), _native.ptr(v),
                                       _native.stream_ptr())
        _native.check(rc, "mm_batch_compute")
        return e, g, v

    def adjoint(self, pos, rvecs, dpos=None, drvecs=None, energy_weight=None):
        """One adjoint launch. Returns a dict of device tensors: ``energy`` [nsys],
        ``vtens`` [nsys, 3, 3], ``grvecs`` [nsys, 3, 3], ``gpos`` [N, 3], ``params``
        [nrows, npar] (raw kernel entries) and, with a direction, the tangents ``d*`` of each."""
        kw = dict(dtype=torch.float64, device=self.device)
        dual = dpos is not None
        if dual != (drvecs is not None):
            raise ValueError("dpos and drvecs go together")
        for t in (pos, rvecs, dpos, drvecs, energy_weight):
            if t is not None and (t.dtype != torch.float64 or not t.is_contiguous()
                                  or t.device != self.device):
                raise TypeError("adjoint inputs must be contiguous float64 tensors on %s"
                                % self.device)
        if pos.shape != (self.N, 3) or rvecs.numel() != 9 * self.nsys:
            raise ValueError("pos must be [%d, 3] and rvecs [%d, 3, 3]" % (self.N, self.nsys))
        if dual and (dpos.shape != pos.shape or drvecs.numel() != rvecs.numel()):
            raise ValueError("direction shapes differ from pos / rvecs")
        if energy_weight is not None and energy_weight.numel() != self.nsys:
            raise ValueError("energy_weight must have one entry per system")
        sys_out = torch.empty((self.nsys, 2 * SYSQ), **kw)
        gpos = torch.empty((self.N, 3), **kw)
        dgpos = torch.empty((self.N, 3), **kw) if dual else None
        params = torch.empty((self.nrows, self.npar, 2), **kw)
        rc = self.lib.mm_adjoint_compute(self._adj, _native.ptr(pos), _native.ptr(dpos),
                                         _native.ptr(rvecs), _native.ptr(drvecs),
                                         _native.ptr(energy_weight), _native.ptr(sys_out),
                                         _native.ptr(gpos), _native.ptr(dgpos),
                                         _native.ptr(params), _native.stream_ptr())
        _native.check(rc, "mm_adjoint_compute")
        out = dict(energy=sys_out[:, 0], vtens=sys_out[:, 1:10].reshape(-1, 3, 3),
                   grvecs=sys_out[:, 10:19].reshape(-1, 3, 3), gpos=gpos,
                   params=params[:, :, 0])
        if dual:
            out.update(denergy=sys_out[:, SYSQ], dvtens=sys_out[:, SYSQ + 1:SYSQ + 10]
                       .reshape(-1, 3, 3), dgrvecs=sys_out[:, SYSQ + 10:].reshape(-1, 3, 3),
                       dgpos=dgpos, dparams=params[:, :, 1])
        return out

    def vjp(self, pos, rvecs, gE, gG=None, gV=None):
        """Vector-Jacobian product of (energies [nsys], gpos [N, 3], vtens [nsys, 3, 3]) with the
        cotangents gE [nsys], gG [N, 3] and gV [nsys, 3, 3] (None: zero), in one library call
        (``mm_adjoint_vjp``). Returns a dict: ``g_pos`` [N, 3] (the full position cotangent),
        ``grvecs``/``dgrvecs`` [nsys, 3, 3] (dL/drvecs = gE grvecs + dgrvecs + grvecs gV) and
        ``params``/``dparams`` [nrows, npar] (raw entries of dL/dtheta = params + dparams)."""
        kw = dict(dtype=torch.float64, device=self.device)
        for t in (pos, rvecs, gE, gG, gV):
            if t is not None and (t.dtype != torch.float64 or not t.is_contiguous()
                                  or t.device != self.device):
                raise TypeError("vjp inputs must be contiguous float64 tensors on %s"
                                % self.device)
        if pos.shape != (self.N, 3) or rvecs.numel() != 9 * self.nsys or gE.numel() != self.nsys:
            raise ValueError("pos must be [%d, 3], rvecs [%d, 3, 3] and gE [%d]"
                             % (self.N, self.nsys, self.nsys))
        if (gG is not None and gG.shape != pos.shape) or \
                (gV is not None and gV.numel() != 9 * self.nsys):
            raise ValueError("gG must be [%d, 3] and gV [%d, 3, 3]" % (self.N, self.nsys))
        sys_out = torch.empty((self.nsys, 2 * SYSQ), **kw)
        g_pos = torch.empty((self.N, 3), **kw)
        params = torch.empty((self.nrows, self.npar, 2), **kw)
        rc = self.lib.mm_adjoint_vjp(self._adj, _native.ptr(pos), _native.ptr(rvecs),
                                     _native.ptr(self._node_sys32), _native.ptr(gE),
                                     _native.ptr(gG), _native.ptr(gV), _native.ptr(sys_out),
                                     _native.ptr(g_pos), _native.ptr(params),
                                     _native.stream_ptr())
        _native.check(rc, "mm_adjoint_vjp")
        return dict(g_pos=g_pos, grvecs=sys_out[:, 10:19].reshape(-1, 3, 3),
                    dgrvecs=sys_out[:, SYSQ + 10:].reshape(-1, 3, 3), params=params[:, :, 0],
                    dparams=params[:, :, 1], raw_params=params)

    def param_grads(self, entries, params, add_next=False):
        """Map raw kernel entries [nrows, npar] to gradients of (cv, h0, free, teff), in one
        kernel (``mm_adjoint_param_grads``: the symmetric Voigt gradient, dE/dh0 from dE/dP and
        dE/dw through the cofactor inverse of h0, the states a row lacks zeroed). With
        ``add_next`` the entries are the [nrows, npar, 2] value and tangent rows of an adjoint
        launch, summed on the way."""
        T, S = self.nrows, self.max_states
        kw = dict(dtype=torch.float64, device=self.device)
        # the state counts the kernel masks with are the field's: those of params
        self.set_params(params)
        ent = entries.contiguous()
        if ent.dtype != torch.float64 or ent.device != self.device or \
                ent.numel() != T * self.npar * (2 if add_next else 1):
            raise ValueError("entries must be float64 [%d, %d%s] on %s"
                             % (T, self.npar, ", 2" if add_next else "", self.device))
        h0 = params.h0.detach().to(**kw).contiguous()
        if h0.shape != (T, S, 3, 3):
            raise ValueError("h0 must be [%d, %d, 3, 3]" % (T, S))
        g_cv = torch.empty((T, S, 6, 6), **kw)
        g_h0 = torch.empty((T, S, 3, 3), **kw)
        g_free = torch.empty((T, S), **kw)
        g_teff = torch.empty((T,), **kw)
        rc = self.lib.mm_adjoint_param_grads(
            self._adj, _native.ptr(ent), 2 if add_next else 1, 1 if add_next else 0,
            _native.ptr(h0), float(units.boltzmann), _native.ptr(g_cv), _native.ptr(g_h0),
            _native.ptr(g_free), _native.ptr(g_teff), _native.stream_ptr())
        _native.check(rc, "mm_adjoint_param_grads")
        return g_cv, g_h0, g_free, g_teff

    def __del__(self):
        lib = getattr(self, "lib", None)
        if lib is None:
            return
        a = getattr(self, "_adj", None)
        if a is not None and a.value:
            lib.mm_adjoint_destroy(a)
            self._adj = None
        f = getattr(self, "_field", None)
        if f is not None and f.value:
            lib.mm_field_destroy(f)
            self._field = None


_native.register_signature("mm_adjoint_destroy", [_P])


class _MicMecEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, params, pos, rvecs, cv, h0, free, teff):
        field.set_params(params)
        pos_c = pos.detach().contiguous()
        rv_c = rvecs.detach().reshape(field.nsys, 3, 3).contiguous()
        e, g, v = field.compute(pos_c, rv_c)
        ctx.field = field
        ctx.params = params
        ctx.rv_shape = rvecs.shape
        ctx.save_for_backward(pos_c, rv_c)
        return e, g, v

    @staticmethod
    @once_differentiable
    def backward(ctx, gE, gG, gV):
        field, params = ctx.field, ctx.params
        pos, rv = ctx.saved_tensors
        kw = dict(dtype=torch.float64, device=pos.device)
        gE = torch.zeros(field.nsys, **kw) if gE is None else gE.contiguous()
        field.set_params(params)
        # one library call: direction (gG + pos gV^T, rvecs gV^T), the dual-number adjoint along
        # it with energy weights gE, and the position cotangent gE gpos + d(gpos) + gpos gV
        out = field.vjp(pos, rv, gE, None if gG is None else gG.contiguous(),
                        None if gV is None else gV.contiguous().reshape(-1, 3, 3))
        g_pos = out["g_pos"]
        grv = out["grvecs"]
        g_rv = gE[:, None, None] * grv + out["dgrvecs"]
        if gV is not None:
            gV = gV.contiguous().reshape(-1, 3, 3)
            g_rv = g_rv + (grv.unsqueeze(3) * gV.unsqueeze(1)).sum(2)
        g_cv, g_h0, g_free, g_teff = field.param_grads(out["raw_params"], params, add_next=True)
        return (None, None, g_pos, g_rv.reshape(ctx.rv_shape), g_cv, g_h0, g_free, g_teff)


def micmec_energy(field, pos, rvecs, params):
    """Differentiable (energies [nsys], gpos [N, 3], vtens [nsys, 3, 3]) of ``field`` at
    ``pos`` [N, 3] and ``rvecs`` [nsys, 3, 3] with parameters ``params`` (a CellParams)."""
    return _MicMecEnergy.apply(field, params, pos, rvecs, *params.tensors())


# ---------------------------------------------------------------------------- fit workload
class FitProblem(object):
    """Force-and-stress matching of cell-type parameters (the differentiation workload).

    Targets: forces -gpos and stresses -vtens / V of every configuration, computed with the
    target parameters. Loss: mean squared force error over nodes plus mean squared stress error
    over configurations, each divided by the squared scale of its targets."""

    def __init__(self, field, pos, rvecs, target_params):
        self.field, self.pos, self.rvecs = field, pos, rvecs
        with torch.no_grad():
            e, g, v = micmec_energy(field, pos, rvecs, target_params)
        self.volume = torch.tensor(np.abs(np.linalg.det(rvecs.cpu().numpy())),
                                   dtype=torch.float64, device=rvecs.device)
        self.f_target = -g
        self.s_target = -v / self.volume[:, None, None]
        self.f_scale = float((self.f_target ** 2).mean())
        self.s_scale = float((self.s_target ** 2).mean())
        # the same squared errors with the signs folded into the targets: mean((g - gt)^2)
        # and mean((v / V - vt / V)^2), each one fused kernel forward and backward
        self._g_target = g.contiguous()
        self._vv_target = (v / self.volume[:, None, None]).contiguous()
        self._inv_volume = (1.0 / self.volume)[:, None, None].contiguous()

    def loss(self, params):
        e, g, v = micmec_energy(self.field, self.pos, self.rvecs, params)
        lf = torch.nn.functional.mse_loss(g, self._g_target) / self.f_scale
        ls = torch.nn.functional.mse_loss(v * self._inv_volume, self._vv_target) / self.s_scale
        return lf + ls


def fit_workload(count=64, n=16, seed=3, c_noise=0.10, device=None, model="corner8"):
    """The differentiation workload: ``count`` configurations on an n^3 grid with types A and B
    (probability of B 0.2), node noise 0.1 Angstrom and a +-3% random symmetric strain
    (configuration i seeded with seed + i); targets from the elasticity matrices of every
    (type, state) perturbed by ``c_noise`` relative seeded noise (symmetric). Returns
    (FitProblem, initial CellParams with gradients on cv and h0, target CellParams)."""
    from .workloads import make_system
    systems = [make_system((n, n, n), p_b=0.2, seed=seed + i, noise=0.1 * units.angstrom,
                           strain=0.03) for i in range(count)]
    field = DiffField(systems, model=model, device=device)
    pos, rvecs = field.positions()
    target = field.initial_params()
    rng = np.random.default_rng(seed + 1000)
    noise = rng.uniform(-c_noise, c_noise, tuple(target.cv.shape))
    noise = 0.5 * (noise + np.swapaxes(noise, -1, -2))
    target.cv = target.cv * (1.0 + torch.tensor(noise, dtype=torch.float64,
                                                device=target.cv.device))
    problem = FitProblem(field, pos, rvecs, target)
    init = field.initial_params()
    init.cv.requires_grad_(True)
    init.h0.requires_grad_(True)
    return problem, init, target


def fit(problem, init, steps, lr_cv=1e-2, lr_h0=1e-4, callback=None, graph=True):
    """Adam on relative corrections of the elasticity matrices and rest cells:
    cv = cv_init * (1 + a) with a symmetric, h0 = h0_init * (1 + b). Returns the loss history
    and the final CellParams.

    graph: after three eager steps, one whole step (parameter table, forward force pass,
    loss, backward adjoint pass, Adam update) is captured in a CUDA graph and replayed for the
    remaining steps, so that the host launches one graph per step instead of some sixty
    kernels; the steps are the same kernels on the same buffers either way. A problem whose
    loss synchronises with the host or allocates outside torch cannot be captured: pass
    graph=False for it."""
    cv0, h00 = init.cv.detach(), init.h0.detach()
    free, teff = init.free.detach(), init.teff.detach()
    a = torch.zeros_like(cv0, requires_grad=True)
    b = torch.zeros_like(h00, requires_grad=True)
    groups = [{"params": [a], "lr": lr_cv}, {"params": [b], "lr": lr_h0}]
    graph = bool(graph) and a.is_cuda and steps > 3
    try:
        # one fused update kernel per group instead of a dozen foreach launches; capturable
        # keeps the step count on the device, as a graph replay needs
        opt = torch.optim.Adam(groups, fused=True, capturable=graph)
    except (RuntimeError, TypeError, ValueError):
        opt = torch.optim.Adam(groups, capturable=graph)

    def params_of(a, b):
        asym = 0.5 * (a + a.transpose(-1, -2))
        return CellParams(cv0 * (1.0 + asym), h00 * (1.0 + b), free, teff, init.nstates,
                          init.type_ids)

    def one_step():
        loss = problem.loss(params_of(a, b))
        loss.backward()
        opt.step()
        return loss

    losses = []

    def record(it, loss):
        # the history stays on the device (no synchronisation per step) unless a callback
        # wants the value now
        losses.append(loss.detach().clone())
        if callback is not None:
            callback(it, float(losses[-1]))

    if not graph:
        for it in range(steps):
            opt.zero_grad()
            record(it, one_step())
    else:
        # eager steps on a side stream first (the library's and torch's lazy set-up, the
        # optimizer state), as capture requires
        side = torch.cuda.Stream(device=a.device)
        side.wait_stream(torch.cuda.current_stream(a.device))
        with torch.cuda.stream(side):
            for it in range(3):
                opt.zero_grad(set_to_none=True)
                record(it, one_step())
        torch.cuda.current_stream(a.device).wait_stream(side)
        opt.zero_grad(set_to_none=True)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            static_loss = one_step()
        for it in range(3, steps):
            g.replay()
            record(it, static_loss)
    with torch.no_grad():
        final = params_of(a, b)
    losses = torch.stack(losses).cpu().tolist() if losses else []
    return losses, final
