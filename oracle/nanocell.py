"""Per-cell nanocell energy and gradient, both models (test infrastructure only).

Restates micmec/pes/nanocell.py:40-132 (the eight-corner model, "corner8"),
micmec/pes/nanocell_original.py:38-132 (the averaged-matrix model, "original") and the
constant tables of micmec/pes/nanocell_utils.py:32-110. The gradient follows the reference's
route: derivatives of the corner matrices with respect to each vertex coordinate
(``cell_derivs``), contracted with the stress, not an adjoint.
"""

import numpy as np

from .topology import NEIGHBOR_CELLS, NEIGHBOR_NODES


def _multiplicator():
    """mult[a, d, v]: row d of corner matrix a is sum_v mult[a, d, v] r_v (nanocell_utils.py:32-75).

    Row d of corner a is the edge along axis d through corner a, from its lower to its upper end.
    """
    offs = [tuple(o) for o in NEIGHBOR_NODES]
    mult = np.zeros((8, 3, 8), dtype=int)
    for a, oa in enumerate(offs):
        for d in range(3):
            lo = list(oa)
            hi = list(oa)
            lo[d] = 0
            hi[d] = 1
            mult[a, d, offs.index(tuple(lo))] -= 1
            mult[a, d, offs.index(tuple(hi))] += 1
    return mult


def _cell_derivs():
    """derivs[axis][v][a] = (d M_a / d r_v[axis])^T (nanocell_utils.py:78-110)."""
    out = []
    for axis in range(3):
        per_v = []
        for nv in NEIGHBOR_CELLS:
            per_a = []
            for na in NEIGHBOR_CELLS:
                m = np.zeros((3, 3))
                deriv = np.array([1.0 if n == -1 else -1.0 for n in nv])
                dist_vec = np.abs(np.array(nv) - np.array(na))
                dist = dist_vec.sum()
                if dist == 0:
                    m[:, axis] = deriv
                elif dist == 1:
                    sel = dist_vec == 1
                    m[sel, axis] = deriv[sel]
                per_a.append(m.T)
            per_v.append(per_a)
        out.append(np.array(per_v))
    return out


MULTIPLICATOR = _multiplicator()
CELL_DERIVS = _cell_derivs()          # three arrays [8 (v), 8 (a), 3, 3]
CELL_DERIV_ORIGINAL = [0.25 * np.array([d[v, v] for v in range(8)]) for d in CELL_DERIVS]


def corner_matrices(verts):
    """M[a] (rows = the three +axis edges at corner a) from verts[8, 3]."""
    return np.einsum("adv,vj->adj", MULTIPLICATOR, verts)


def energy_corner8(verts, h0, C0):
    """Eight-corner energy (nanocell.py:40-77)."""
    mats = corner_matrices(verts)
    h0inv = np.linalg.inv(h0)
    X = np.einsum("aji,kj->aik", mats, h0inv)
    strains = 0.5 * (np.einsum("aji,ajk->aik", X, X) - np.identity(3)[None])
    dens = 0.5 * np.einsum("aij,ijkl,akl->a", strains, C0, strains)
    return 0.125 * dens.sum() * np.linalg.det(h0)


def grad_corner8(verts, h0, C0):
    """Eight-corner gradient [8, 3] (nanocell.py:80-132)."""
    h0inv = np.linalg.inv(h0)
    det = np.linalg.det(h0)
    mats = corner_matrices(verts)
    X = np.einsum("aji,kj->aik", mats, h0inv)
    strains = 0.5 * (np.einsum("aji,ajk->aik", X, X) - np.identity(3)[None])
    stresses = np.einsum("ijkl,akl->aij", C0, strains)
    g = np.zeros((8, 3))
    for axis in range(3):
        m = np.einsum("aki,vakm,jm->vaij", X, CELL_DERIVS[axis], h0inv)
        eps_d = 0.5 * (np.swapaxes(m, -1, -2) + m)
        g[:, axis] = 0.125 * np.einsum("vaji,aij->v", eps_d, stresses)
    return det * g


def energy_original(verts, h0, C0):
    """Averaged-matrix energy (nanocell_original.py:46-84)."""
    h0inv = np.linalg.inv(h0)
    mat = 0.125 * corner_matrices(verts).sum(axis=0)
    X = np.einsum("ji,kj->ik", mat, h0inv)
    strain = 0.5 * (X.T @ X - np.identity(3))
    return 0.5 * np.einsum("ij,ijkl,kl", strain, C0, strain) * np.linalg.det(h0)


def grad_original(verts, h0, C0):
    """Averaged-matrix gradient [8, 3] (nanocell_original.py:87-132)."""
    h0inv = np.linalg.inv(h0)
    det = np.linalg.det(h0)
    mat = 0.125 * corner_matrices(verts).sum(axis=0)
    X = np.einsum("ji,kj->ik", mat, h0inv)
    strain = 0.5 * (X.T @ X - np.identity(3))
    stress = np.einsum("ijkl,kl->ij", C0, strain)
    g = np.zeros((8, 3))
    for axis in range(3):
        m = np.einsum("ki,vkm,jm->vij", X, CELL_DERIV_ORIGINAL[axis], h0inv)
        eps_d = 0.5 * (np.swapaxes(m, -1, -2) + m)
        g[:, axis] = np.einsum("vji,ij->v", eps_d, stress)
    return det * g


MODELS = {
    "corner8": (energy_corner8, grad_corner8),
    "original": (energy_original, grad_original),
}
