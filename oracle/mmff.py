"""Force-field assembly on the CPU (test infrastructure only).

``deformation_loop`` restates micmec/pes/mmff.py:326-403 cell by cell; ``assemble`` restates
``_compute_gpos`` (mmff.py:303-318), ``_compute_vtens`` (320-323) and ``_compute_epot``
(299-301). ``parse_params`` restates the key parsing of mmff.py:218-231. ``compute_vectorised``
evaluates the same arithmetic batched over cells (the reference's derivative-table gradient,
chunked) so that grids of tens of thousands of cells stay practical.
"""

import numpy as np

from .constants import boltzmann
from .nanocell import MODELS, MULTIPLICATOR, CELL_DERIVS, CELL_DERIV_ORIGINAL
from .topology import get_mic


def parse_params(params):
    """(temps, free energies, cells, elasticities) keyed by type id (mmff.py:218-231)."""
    p1, p2, p3, p4 = {}, {}, {}, {}
    for key, item in params.items():
        t = int(key.split("/")[0][4:]) if key.startswith("type") and "/" in key else None
        if t is None:
            continue
        if "effective_temp" in key:
            p1[t] = item
        if "free_energy" in key:
            p2[t] = item
        if "cell" in key:
            p3[t] = item
        if "elasticity" in key:
            p4[t] = item
    return p1, p2, p3, p4


def deformation_loop(pos, rvecs, mic, types, surrounding_nodes, params, model="corner8"):
    """Per-cell energies [C], gradients [C, 8, 3] and vertices [C, 8, 3] (mmff.py:326-403)."""
    efun, gfun = MODELS[model]
    p1, p2, p3, p4 = parse_params(params)
    rv = np.zeros((3, 3))
    if len(rvecs):
        rv[: len(rvecs)] = rvecs

    def delta(i, j):
        d = pos[j] - pos[i]
        d += rv[0] * mic[i, j, 0] + rv[1] * mic[i, j, 1] + rv[2] * mic[i, j, 2]
        return d

    epots, gposs, vertss = [], [], []
    for c in range(len(types)):
        t = int(types[c])
        idx = surrounding_nodes[c]
        r0 = pos[idx[0]]
        verts = np.array([r0] + [r0 + delta(idx[0], idx[k]) for k in range(1, 8)])
        es, gs = [], []
        temp = p1[t]
        for f, h0, C0 in zip(p2[t], p3[t], p4[t]):
            es.append(efun(verts, h0, C0) + f)
            gs.append(gfun(verts, h0, C0))
        es = np.array(es)
        emin = np.min(es)
        w = np.exp(-(es - emin) / (boltzmann * temp))
        wn = w / np.sum(w)
        epots.append(emin - temp * boltzmann * np.log(np.sum(w)))
        g = np.zeros((8, 3))
        for wi, gi in zip(wn, gs):
            g += wi * gi
        gposs.append(g)
        vertss.append(verts)
    return np.array(epots), np.array(gposs).reshape(-1, 8, 3), np.array(vertss).reshape(-1, 8, 3)


def assemble(epot_cells, gpos_cells, verts_cells, surrounding_cells):
    """(energy, gpos [N, 3], vtens [3, 3]) from per-cell results (mmff.py:299-323)."""
    N = len(surrounding_cells)
    gpos = np.zeros((N, 3))
    for n in range(N):
        g = np.zeros(3)
        for k, c in enumerate(surrounding_cells[n]):
            if c < 0:
                continue
            g += gpos_cells[c][k]
        gpos[n] = g
    vtens = np.einsum("ijk,ijl->kl", gpos_cells, verts_cells)
    return np.sum(epot_cells), gpos, vtens


def compute_loop(system_like, model="corner8", mic=None):
    """Full literal evaluation for a small system; returns (energy, gpos, vtens)."""
    s = system_like
    rvecs = np.asarray(s["rvecs"], dtype=float)
    pbc = rvecs.shape[0] > 0
    if mic is None:
        mic = get_mic(np.rint(s["grid"]).astype(int), pbc)
    e, g, v = deformation_loop(np.asarray(s["pos"], dtype=float), rvecs, mic,
                               np.rint(s["types"]).astype(int),
                               np.asarray(s["surrounding_nodes"]), s["params"], model)
    return assemble(e, g, v, np.asarray(s["surrounding_cells"]))


# --------------------------------------------------------------------------------------------
# vectorised form of the same arithmetic


def cell_vertices(pos, rvecs, surrounding_nodes, shifts):
    """Vertices [C, 8, 3]: r_0 + (pos[n_k] - pos[n_0]) + shift_k . rvecs (mmff.py:347-371).

    ``shifts[C, 8, 3]`` holds the image multiplicity of each corner along each domain vector.
    """
    rv = np.zeros((3, 3))
    if len(rvecs):
        rv[: len(rvecs)] = rvecs
    p = pos[surrounding_nodes]                      # [C, 8, 3]
    d = p - p[:, :1]
    d = d + (rv[0][None, None] * shifts[..., 0:1] + rv[1][None, None] * shifts[..., 1:2]
             + rv[2][None, None] * shifts[..., 2:3])
    verts = p[:, :1] + d
    verts[:, 0] = p[:, 0]
    return verts


def corner_shifts(wrap_bits):
    """shifts[C, 8, 3] from per-cell wrap bits (bit d: the cell is the last layer along d)."""
    bits = np.asarray(wrap_bits, dtype=np.int64)
    offs = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
                     (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)])
    out = np.zeros((len(bits), 8, 3))
    for d in range(3):
        out[:, :, d] = ((bits[:, None] >> d) & 1) * offs[None, :, d]
    return out


def _state_energy_grad(verts, h0, C0, model):
    """Batched energy [B] and gradient [B, 8, 3] of one state for cells sharing (h0, C0)."""
    h0inv = np.linalg.inv(h0)
    det = np.linalg.det(h0)
    mats = np.einsum("adv,cvj->cadj", MULTIPLICATOR, verts)
    if model == "corner8":
        X = np.einsum("caji,kj->caik", mats, h0inv)
        strains = 0.5 * (np.einsum("caji,cajk->caik", X, X) - np.identity(3))
        dens = 0.5 * np.einsum("caij,ijkl,cakl->ca", strains, C0, strains)
        e = 0.125 * dens.sum(axis=1) * det
        stresses = np.einsum("ijkl,cakl->caij", C0, strains)
        g = np.zeros(verts.shape)
        for axis in range(3):
            m = np.einsum("caki,vakm,jm->cvaij", X, CELL_DERIVS[axis], h0inv)
            eps_d = 0.5 * (np.swapaxes(m, -1, -2) + m)
            g[:, :, axis] = 0.125 * np.einsum("cvaji,caij->cv", eps_d, stresses)
        return e, det * g
    mat = 0.125 * mats.sum(axis=1)
    X = np.einsum("cji,kj->cik", mat, h0inv)
    strain = 0.5 * (np.einsum("cji,cjk->cik", X, X) - np.identity(3))
    e = 0.5 * np.einsum("cij,ijkl,ckl->c", strain, C0, strain) * det
    stress = np.einsum("ijkl,ckl->cij", C0, strain)
    g = np.zeros(verts.shape)
    for axis in range(3):
        m = np.einsum("cki,vkm,jm->cvij", X, CELL_DERIV_ORIGINAL[axis], h0inv)
        eps_d = 0.5 * (np.swapaxes(m, -1, -2) + m)
        g[:, :, axis] = np.einsum("cvji,cij->cv", eps_d, stress)
    return e, det * g


def deformation_vectorised(pos, rvecs, types, surrounding_nodes, wrap_bits, params,
                           model="corner8", chunk=4096):
    """Vectorised ``deformation``: per-cell energies, gradients and vertices."""
    p1, p2, p3, p4 = parse_params(params)
    types = np.asarray(types).astype(int)
    sn = np.asarray(surrounding_nodes)
    shifts = corner_shifts(wrap_bits) if len(rvecs) else np.zeros((len(types), 8, 3))
    C = len(types)
    epot = np.zeros(C)
    gcell = np.zeros((C, 8, 3))
    verts_all = np.zeros((C, 8, 3))
    for start in range(0, C, chunk):
        sl = slice(start, min(C, start + chunk))
        verts = cell_vertices(pos, rvecs, sn[sl], shifts[sl])
        verts_all[sl] = verts
        tsl = types[sl]
        for t in np.unique(tsl):
            sel = np.nonzero(tsl == t)[0]
            es, gs = [], []
            for f, h0, C0 in zip(p2[t], p3[t], p4[t]):
                e, g = _state_energy_grad(verts[sel], np.asarray(h0), np.asarray(C0), model)
                es.append(e + f)
                gs.append(g)
            es = np.array(es)                       # [S, B]
            emin = es.min(axis=0)
            kt = boltzmann * p1[t]
            w = np.exp(-(es - emin) / kt)
            wsum = w.sum(axis=0)
            wn = w / wsum
            epot[start + sel] = emin - p1[t] * boltzmann * np.log(wsum)
            g = np.zeros((len(sel), 8, 3))
            for s in range(len(gs)):
                g += wn[s][:, None, None] * gs[s]
            gcell[start + sel] = g
    return epot, gcell, verts_all


def assemble_vectorised(epot_cells, gpos_cells, verts_cells, surrounding_cells):
    """Vectorised ``_compute_gpos``/``_compute_vtens``/``_compute_epot`` (k summed in order)."""
    sc = np.asarray(surrounding_cells)
    N = len(sc)
    gpos = np.zeros((N, 3))
    for k in range(8):
        c = sc[:, k]
        ok = c >= 0
        gpos[ok] += gpos_cells[c[ok], k]
    vtens = np.einsum("ijk,ijl->kl", gpos_cells, verts_cells)
    return np.sum(epot_cells), gpos, vtens


def compute_vectorised(pos, rvecs, types, surrounding_nodes, surrounding_cells, wrap_bits,
                       params, model="corner8"):
    """(energy, gpos, vtens) for any grid size."""
    e, g, v = deformation_vectorised(np.asarray(pos, dtype=float), np.asarray(rvecs, float),
                                     types, surrounding_nodes, wrap_bits, params, model)
    return assemble_vectorised(e, g, v, surrounding_cells)
