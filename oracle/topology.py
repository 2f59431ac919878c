"""Brute-force loop restatement of the reference topology (test infrastructure only).

Follows micmec/utils.py:113-218 (node and cell enumeration, surrounding tables, masses) and
micmec/pes/mmff.py:258-286 (the pairwise image-shift table ``get_mic``) with plain Python
loops. Meant for grids of a few cells per axis.
"""

import numpy as np

NEIGHBOR_CELLS = [(0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1),
                  (-1, -1, 0), (-1, 0, -1), (0, -1, -1), (-1, -1, -1)]
NEIGHBOR_NODES = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
                  (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]


def enumerate_grid(grid, pbc=(True, True, True)):
    """Return (nodes, cells, types, node dims) following utils.py:113-161."""
    nx, ny, nz = grid.shape
    nnx, nny, nnz = nx + 1 - pbc[0], ny + 1 - pbc[1], nz + 1 - pbc[2]
    nodes = []
    for kk in range(nnx):
        for ll in range(nny):
            for mm in range(nnz):
                for off in NEIGHBOR_CELLS:
                    a, b, c = kk + off[0], ll + off[1], mm + off[2]
                    if (not pbc[0] and (a < 0 or a >= nx)) or \
                       (not pbc[1] and (b < 0 or b >= ny)) or \
                       (not pbc[2] and (c < 0 or c >= nz)):
                        continue
                    if grid[a % nx, b % ny, c % nz] != 0:
                        nodes.append((kk, ll, mm))
                        break
    cells, types = [], []
    for a in range(nx):
        for b in range(ny):
            for c in range(nz):
                if grid[a, b, c] != 0:
                    cells.append((a, b, c))
                    types.append(int(grid[a, b, c]))
    return nodes, cells, types, (nnx, nny, nnz)


def surrounding_tables(grid, type_masses, pbc=(True, True, True)):
    """surrounding_nodes, surrounding_cells and masses by the double loop of utils.py:201-218."""
    nodes, cells, types, nn = enumerate_grid(grid, pbc)
    sn = -np.ones((len(cells), 8), dtype=int)
    sc = -np.ones((len(nodes), 8), dtype=int)
    masses = np.zeros(len(nodes))
    for ci, (a, b, c) in enumerate(cells):
        t = int(grid[a, b, c])
        for k, off in enumerate(NEIGHBOR_NODES):
            target = ((a + off[0]) % nn[0], (b + off[1]) % nn[1], (c + off[2]) % nn[2])
            for ni, node in enumerate(nodes):
                if node == target:
                    sn[ci, k] = ni
                    sc[ni, k] = ci
                    masses[ni] += 0.125 * type_masses[t]
                    break
    return sn, sc, masses, nodes, cells, types


def get_mic(grid, pbc=True):
    """Pairwise image-shift table mic[N, N, 3] (mmff.py:258-286)."""
    p = (pbc, pbc, pbc)
    nodes, _, _, nn = enumerate_grid(grid, p)
    N = len(nodes)
    mic = np.zeros((N, N, 3))
    if not pbc:
        return mic
    boundary = [i for i, (k, l, m) in enumerate(nodes)
                if k in (0, nn[0] - 1) or l in (0, nn[1] - 1) or m in (0, nn[2] - 1)]
    bset = set(boundary)
    arr = np.array(nodes)
    sign = lambda x: 2 * (x > 0) - 1
    for j in range(N):
        for i in range(j):
            if i not in bset or j not in bset:
                continue
            for d in range(3):
                dij = arr[j, d] - arr[i, d]
                dmax = nn[d] - 1
                if abs(dij) == dmax:
                    if dmax != 1:
                        mic[i, j, d] = -sign(dij)
                        mic[j, i, d] = sign(dij)
                    else:
                        mic[i, j, d] = float(dij == -1)
                        mic[j, i, d] = float(dij == 1)
    return mic
