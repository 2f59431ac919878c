"""Initial velocities, momentum cleaning and degree-of-freedom counts (host side, numpy).

Mirrors micmec/sampling/utils.py:28-501 with the same numpy operations and the same order of
draws from numpy's legacy global random stream, so seeded runs start from identical states.
"""

import numpy as np

from ..units import boltzmann

__all__ = ["get_random_vel", "remove_com_moment", "remove_angular_moment", "clean_momenta",
           "inertia_tensor", "angular_moment", "angular_velocity",
           "rigid_body_angular_velocities", "get_ndof_internal_md", "domain_symmetrize",
           "get_random_vel_press", "get_ndof_baro"]


def get_random_vel(temp0, scalevel0, masses, select=None):
    """Maxwell-Boltzmann velocities, optionally rescaled to exactly temp0 (utils.py:28-65)."""
    if select is not None:
        masses = masses[select]
    vel0 = np.random.normal(0, 1, (len(masses), 3)) * np.sqrt(
        boltzmann * temp0 / masses).reshape(-1, 1)
    if scalevel0 and temp0 > 0:
        temp = np.mean((vel0 ** 2 * masses.reshape(-1, 1))) / boltzmann
        vel0 *= np.sqrt(temp0 / temp)
    return vel0


def remove_com_moment(vel, masses):
    velcom = np.dot(masses, vel) / np.sum(masses)
    vel[:] -= velcom


def inertia_tensor(pos, masses):
    return np.identity(3) * (masses.reshape(-1, 1) * pos ** 2).sum() - np.dot(
        pos.T, masses.reshape(-1, 1) * pos)


def angular_moment(pos, vel, masses):
    lin = masses.reshape(-1, 1) * vel
    amom = np.zeros(3, float)
    amom[0] = np.sum(pos[:, 1] * lin[:, 2] - pos[:, 2] * lin[:, 1])
    amom[1] = np.sum(pos[:, 2] * lin[:, 0] - pos[:, 0] * lin[:, 2])
    amom[2] = np.sum(pos[:, 0] * lin[:, 1] - pos[:, 1] * lin[:, 0])
    return amom


def angular_velocity(amom, itens, epsilon=1e-10):
    evals, evecs = np.linalg.eigh(itens)
    mask = evals > epsilon
    evals = evals[mask]
    evecs = evecs[:, mask]
    return np.dot(evecs, np.dot(evecs.T, amom) / evals)


def rigid_body_angular_velocities(pos, avel):
    vel = np.zeros(pos.shape, float)
    vel[:, 0] = avel[1] * pos[:, 2] - avel[2] * pos[:, 1]
    vel[:, 1] = avel[2] * pos[:, 0] - avel[0] * pos[:, 2]
    vel[:, 2] = avel[0] * pos[:, 1] - avel[1] * pos[:, 0]
    return vel


def remove_angular_moment(pos, vel, masses):
    pos = pos.copy()
    pos -= np.dot(masses, pos) / np.sum(masses)
    itens = inertia_tensor(pos, masses)
    amom = angular_moment(pos, vel, masses)
    avel = angular_velocity(amom, itens)
    vel[:] -= rigid_body_angular_velocities(pos, avel)


def clean_momenta(pos, vel, masses, domain):
    """Remove the centre-of-mass momentum, and the angular momentum of isolated systems
    (utils.py:142-174)."""
    remove_com_moment(vel, masses)
    if domain.nvec == 0:
        remove_angular_moment(pos, vel, masses)
    elif domain.nvec == 1:
        raise NotImplementedError


def get_ndof_internal_md(nnodes, nper):
    if nper == 0:
        return 3 * nnodes - 6
    elif nper == 1:
        return 3 * nnodes - 4
    return 3 * nnodes - 3


def domain_symmetrize(mmf, vector_lst=None, tensor_lst=None):
    """Rotate the domain to a symmetric cell tensor and move the nodes along (utils.py:346-383)."""
    domain = np.asarray(mmf.system.domain.rvecs).copy()
    U, s, Vt = np.linalg.svd(domain)
    rot_mat = np.dot(Vt.T, U.T)
    domain = np.dot(domain, rot_mat)
    mmf.update_rvecs(domain)
    mmf.update_pos(np.dot(mmf.system.pos, rot_mat))
    new_vectors = [np.dot(v, rot_mat) for v in (vector_lst or [])]
    new_tensors = [np.dot(np.dot(rot_mat.T, t), rot_mat) for t in (tensor_lst or [])]
    return new_vectors, new_tensors


def get_random_vel_press(mass, temp):
    """Symmetric random barostat velocity tensor (utils.py:450-475)."""
    rand = np.random.normal(0, np.sqrt(mass * boltzmann * temp), (3, 3)) / mass
    vel_press = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            vel_press[i, j] = rand[i, j] if i >= j else rand[j, i]
            if i != j:
                vel_press[i, j] /= np.sqrt(2)
    return vel_press


def get_ndof_baro(dim, anisotropic, vol_constraint):
    ndof = 1
    if anisotropic:
        ndof = dim * (dim + 1) // 2
    if vol_constraint:
        ndof -= 1
    if ndof == 0:
        raise AssertionError("Isotropic barostat called with a volume constraint.")
    return ndof
