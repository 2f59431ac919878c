"""Host side of the differentiation form (no GPU): the Voigt convention of the parameter
interface and the mapping helpers."""

import numpy as np

from micmec_amd.adjoint import tensor_from_voigt, voigt_from_tensor, voigt_pairs
from micmec_amd.workloads import cubic_elasticity


def random_full_symmetry_tensor(rng):
    m = rng.normal(size=(6, 6))
    return tensor_from_voigt(m + m.T)


def test_voigt_round_trip():
    rng = np.random.default_rng(0)
    C = random_full_symmetry_tensor(rng)
    # the tensor has minor and major symmetry
    assert np.allclose(C, C.transpose(1, 0, 2, 3))
    assert np.allclose(C, C.transpose(2, 3, 0, 1))
    assert np.array_equal(tensor_from_voigt(voigt_from_tensor(C)), C)
    cub = voigt_from_tensor(cubic_elasticity(40.0, 20.0, 10.0))
    assert cub[0, 0] == cub[1, 1] == cub[2, 2] and cub[3, 3] == cub[4, 4] == cub[5, 5]
    assert cub[0, 3] == 0.0 and cub[3, 4] == 0.0


def test_voigt_energy_convention():
    """u = eps:C:eps / 2 equals e.Cv.e / 2 with the engineering strain
    e = (eps00, eps11, eps22, 2 eps12, 2 eps02, 2 eps01)."""
    rng = np.random.default_rng(1)
    for _ in range(5):
        C = random_full_symmetry_tensor(rng)
        a = rng.normal(size=(3, 3))
        eps = 0.5 * (a + a.T)
        u_t = 0.5 * np.einsum("ij,ijkl,kl->", eps, C, eps)
        e = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[1, 2], 2 * eps[0, 2],
                      2 * eps[0, 1]])
        u_v = 0.5 * e @ voigt_from_tensor(C) @ e
        assert abs(u_t - u_v) <= 1e-13 * max(1.0, abs(u_t))


def test_voigt_pairs():
    pairs = voigt_pairs()
    assert len(pairs) == 21 and len(set(pairs)) == 21
    assert pairs[0] == (0, 0) and pairs[5] == (0, 5) and pairs[6] == (1, 1) and pairs[-1] == (5, 5)
