"""Counter-based Gaussian noise of the device integrator, restated in numpy (test infrastructure
only: the tests and the CPU baseline use it as the checker, never the product path).

Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
SC'11): ten rounds of

    (c0, c1, c2, c3) <- (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2), hi(M0 c0) ^ c3 ^ k1, lo(M0 c0))

with the key bumped by the Weyl constants (W0, W1) between rounds. Pinned by
tests/test_langevin_cpu.py against numpy's independent ``numpy.random.Philox`` bit generator (the
same function of key and counter; numpy increments its counter before the first block) and the
published known-answer vector for a zero key and counter.

Uniforms u = ((w >> 12) + 1/2) 2^-52 lie in (0, 1) and are exact doubles; pairs become normals
by the Box-Muller transform. The reference draws its Langevin noise from numpy's global
Mersenne Twister (micmec/sampling/nvt.py:211-218, npt.py:489-505); a counter-based generator
is what lets every node draw its own noise on the device, so the oracle uses this one to check
the device trajectories draw for draw, while the distribution itself is checked statistically.
"""

import numpy as np

M0 = 0xD2E7470EE14C6C93
M1 = 0xCA5A826395121157
W0 = 0x9E3779B97F4A7C15
W1 = 0xBB67AE8584CAA73B
MASK64 = (1 << 64) - 1
MASK32 = np.uint64(0xFFFFFFFF)
S32 = np.uint64(32)

STREAM_THERMOSTAT = 1
STREAM_BAROSTAT = 2


def _mulhilo(m, b):
    """(hi, lo) of the 128-bit product of the constant m and the uint64 array b."""
    m_lo = np.uint64(m & 0xFFFFFFFF)
    m_hi = np.uint64(m >> 32)
    b_lo = b & MASK32
    b_hi = b >> S32
    ll = m_lo * b_lo
    lh = m_lo * b_hi
    hl = m_hi * b_lo
    hh = m_hi * b_hi
    mid = (ll >> S32) + (lh & MASK32) + (hl & MASK32)
    hi = hh + (lh >> S32) + (hl >> S32) + (mid >> S32)
    lo = np.uint64(m) * b
    return hi, lo


def philox4x64_10(ctr, key):
    """ctr: (n, 4) uint64, key: two ints. Returns (n, 4) uint64."""
    with np.errstate(over="ignore"):
        c = [np.array(ctr[:, i], dtype=np.uint64) for i in range(4)]
        k0, k1 = int(key[0]) & MASK64, int(key[1]) & MASK64
        for _ in range(10):
            hi0, lo0 = _mulhilo(M0, c[0])
            hi1, lo1 = _mulhilo(M1, c[2])
            c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
            k0 = (k0 + W0) & MASK64
            k1 = (k1 + W1) & MASK64
    return np.stack(c, axis=1)


def open_uniform(w):
    return ((w >> np.uint64(12)).astype(np.float64) + 0.5) * 2.0 ** -52


def normals4(ctr, key):
    """(n, 4) standard normals of the counters ctr (n, 4) under key."""
    u = open_uniform(philox4x64_10(ctr, key))
    z = np.empty(u.shape)
    for p in range(2):
        r = np.sqrt(-2.0 * np.log(u[:, 2 * p]))
        t = 6.283185307179586 * u[:, 2 * p + 1]
        z[:, 2 * p] = r * np.cos(t)
        z[:, 2 * p + 1] = r * np.sin(t)
    return z


def thermostat_noise(seed, counter, post, nnodes):
    """(nnodes, 3) normals of the Langevin thermostat at step ``counter`` (pre: post=0)."""
    ctr = np.zeros((nnodes, 4), dtype=np.uint64)
    ctr[:, 0] = np.arange(nnodes, dtype=np.uint64)
    ctr[:, 1] = np.uint64(2 * counter + post)
    return normals4(ctr, (seed, STREAM_THERMOSTAT))[:, :3]


def barostat_noise(seed, counter, post, sub, system=0):
    """The 3x3 draw of LangevinBarostat.getR (npt.py:489-505) whose lower triangle is used:
    entries (0,0), (1,0), (1,1), (2,0), (2,1), (2,2) take the normals in that order."""
    ctr = np.zeros((2, 4), dtype=np.uint64)
    ctr[:, 0] = np.uint64(system)
    ctr[:, 1] = np.uint64(2 * counter + post)
    ctr[:, 2] = np.uint64(sub)
    ctr[:, 3] = np.arange(2, dtype=np.uint64)
    z = normals4(ctr, (seed, STREAM_BAROSTAT)).ravel()
    rand = np.zeros((3, 3))
    k = 0
    for i in range(3):
        for j in range(i + 1):
            rand[i, j] = z[k]
            k += 1
    return rand
