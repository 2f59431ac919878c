"""Grid topology: node and cell enumeration, neighbour tables and system construction.

Vectorised equivalent of ``Grid``, ``build_system`` and ``build_type`` (micmec/utils.py:55-329).
The semantics are those of the reference:

* cells are the non-empty grid sites (grid value != 0), enumerated x-major, z fastest
  (utils.py:149-161);
* nodes are the lattice points that touch at least one non-empty cell, enumerated in the same
  order (utils.py:113-136); under periodic boundaries there are ``n`` node planes per axis,
  otherwise ``n + 1``;
* ``surrounding_nodes[c, k]`` is the node at corner ``k`` of cell ``c`` and
  ``surrounding_cells[n, k]`` the cell in which node ``n`` is corner ``k`` (-1 if that cell is
  empty or absent) (utils.py:198-218); the corner order is ``NEIGHBOR_NODES``;
* node masses are the sum of one eighth of the type mass of each adjacent cell, accumulated in
  cell order (utils.py:217);
* lattice positions use the diagonal of the first equilibrium cell matrix of the most common
  type (utils.py:220-233) and the domain vectors span the periodic axes (utils.py:235-245).

The reference builds these tables with an O(cells x nodes) Python search; here every step is a
numpy array operation, so grids of 128^3 cells build in seconds.
"""

import numpy as np

from . import units

__all__ = [
    "NEIGHBOR_CELLS", "NEIGHBOR_NODES", "Grid", "build_system", "build_system_device",
    "build_type", "periodic_wrap_bits",
]

# Offsets of the eight cells around a node; slot k holds the cell in which the node is corner k
# (utils.py:32-41).
NEIGHBOR_CELLS = np.array(
    [(0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1),
     (-1, -1, 0), (-1, 0, -1), (0, -1, -1), (-1, -1, -1)], dtype=np.int64)
# Offsets of the eight corner nodes of a cell (utils.py:43-52).
NEIGHBOR_NODES = -NEIGHBOR_CELLS


def _pbc_list(pbc):
    if pbc is None or pbc is True:
        return [True, True, True]
    if pbc is False:
        return [False, False, False]
    return [bool(p) for p in pbc]


class Grid(object):
    """Node and cell enumeration of a type grid (utils.py:55-161).

    Parameters
    ----------
    grid : array of int, shape (nx, ny, nz)
        Cell types; 0 marks an empty cell.
    pbc : bool or list of 3 bool, optional
        Periodicity per axis (default: periodic in all three).
    """

    def __init__(self, grid, pbc=None):
        self.grid = np.asarray(grid)
        self.pbc = _pbc_list(pbc)
        self.nx, self.ny, self.nz = self.grid.shape
        self.nx_nodes = self.nx + 1 - self.pbc[0]
        self.ny_nodes = self.ny + 1 - self.pbc[1]
        self.nz_nodes = self.nz + 1 - self.pbc[2]
        self._built = False

    @property
    def shape_nodes(self):
        return (self.nx_nodes, self.ny_nodes, self.nz_nodes)

    def _build(self):
        if self._built:
            return
        g = np.rint(self.grid).astype(np.int64)
        nonempty = g != 0
        nn = self.shape_nodes
        # A lattice point is a node if any of its eight neighbouring cells is non-empty.
        kk, ll, mm = np.meshgrid(np.arange(nn[0]), np.arange(nn[1]), np.arange(nn[2]),
                                 indexing="ij")
        valid = np.zeros(nn, dtype=bool)
        dims = (self.nx, self.ny, self.nz)
        for off in NEIGHBOR_CELLS:
            idx = []
            ok = np.ones(nn, dtype=bool)
            for d, (base, o) in enumerate(zip((kk, ll, mm), off)):
                c = base + o
                if self.pbc[d]:
                    c = np.mod(c, dims[d])
                else:
                    ok &= (c >= 0) & (c < dims[d])
                    c = np.clip(c, 0, dims[d] - 1)
                idx.append(c)
            valid |= ok & nonempty[idx[0], idx[1], idx[2]]
        self.node_mask = valid
        self.node_ijk = np.argwhere(valid)              # lexicographic: x-major, z fastest
        self.node_index = -np.ones(nn, dtype=np.int64)  # lattice point -> node index
        self.node_index[valid] = np.arange(len(self.node_ijk))
        self.cell_ijk = np.argwhere(nonempty)
        self.cell_index = -np.ones(g.shape, dtype=np.int64)
        self.cell_index[nonempty] = np.arange(len(self.cell_ijk))
        self.cell_types = g[nonempty]
        self._built = True

    @property
    def nodes(self):
        self._build()
        return [tuple(int(v) for v in r) for r in self.node_ijk]

    @property
    def nnodes(self):
        self._build()
        return len(self.node_ijk)

    @property
    def cells(self):
        self._build()
        return [tuple(int(v) for v in r) for r in self.cell_ijk]

    @property
    def ncells(self):
        self._build()
        return len(self.cell_ijk)

    @property
    def types(self):
        self._build()
        return self.cell_types.copy()

    @property
    def boundary_nodes(self):
        """Nodes on a face of the node lattice (utils.py:138-147)."""
        self._build()
        ijk = self.node_ijk
        nn = self.shape_nodes
        b = np.zeros(len(ijk), dtype=bool)
        for d in range(3):
            b |= (ijk[:, d] == 0) | (ijk[:, d] == nn[d] - 1)
        return np.nonzero(b)[0]

    def surrounding(self):
        """Return ``(surrounding_nodes [C,8], surrounding_cells [N,8])`` as int64 arrays."""
        self._build()
        nn = self.shape_nodes
        C = len(self.cell_ijk)
        sn = np.empty((C, 8), dtype=np.int64)
        for k, off in enumerate(NEIGHBOR_NODES):
            p = [np.mod(self.cell_ijk[:, d] + off[d], nn[d]) for d in range(3)]
            sn[:, k] = self.node_index[p[0], p[1], p[2]]
        if C and (sn < 0).any():
            raise RuntimeError("corner node missing from the node list")
        sc = -np.ones((len(self.node_ijk), 8), dtype=np.int64)
        cells = np.arange(C)
        for k in range(8):
            sc[sn[:, k], k] = cells
        return sn, sc

    def wrap_bits(self):
        """Per cell, bit d set when the cell's +d corners wrap through the periodic boundary."""
        self._build()
        return periodic_wrap_bits(self.cell_ijk, (self.nx, self.ny, self.nz), self.pbc)


def periodic_wrap_bits(cell_ijk, dims, pbc):
    """Image-shift bits of each cell, replacing the N x N table of ``get_mic`` (mmff.py:258-286).

    Corner k of cell (i, j, l) is shifted by domain vector d exactly when the cell index along d
    plus the corner offset along d equals the grid size, i.e. when the cell sits in the last
    layer of a periodic axis. This equals the reference's pairwise table for every axis with at
    least two cells.
    """
    cell_ijk = np.asarray(cell_ijk, dtype=np.int64).reshape(-1, 3)
    bits = np.zeros(len(cell_ijk), dtype=np.uint8)
    for d in range(3):
        if pbc[d]:
            bits |= ((cell_ijk[:, d] == dims[d] - 1).astype(np.uint8) << d)
    return bits


def build_system(data, grid, pbc=None):
    """Build the CHK-ready dict of a micromechanical system (utils.py:164-263).

    ``data`` maps integer type ids to type dicts from :func:`build_type`; ``grid`` holds the type
    of every cell (0 = empty).
    """
    grid = np.asarray(grid)
    gs = Grid(grid, pbc=pbc)
    sn, sc = gs.surrounding()
    N = gs.nnodes
    types = gs.types
    # Masses: 1/8 of the type mass of each adjacent cell, summed in cell order.
    type_ids = sorted(int(t) for t in np.unique(types)) if len(types) else []
    mass_of = np.zeros(max(type_ids + [0]) + 1)
    for t in type_ids:
        mass_of[t] = data[t]["mass"]
    masses = np.zeros(N)
    contrib = np.repeat(0.125 * mass_of[types], 8)
    np.add.at(masses, sn.ravel(), contrib)
    # Lattice positions from the majority type's first equilibrium cell.
    dx, dy, dz = _majority_spacing(data, grid)
    ijk = gs.node_ijk
    pos = np.zeros((N, 3))
    pos += np.stack([ijk[:, 0] * dx, ijk[:, 1] * dy, ijk[:, 2] * dz], axis=1)
    rvecs = []
    if gs.pbc[0]:
        rvecs.append(np.array([gs.nx * dx, 0.0, 0.0]))
    if gs.pbc[1]:
        rvecs.append(np.array([0.0, gs.ny * dy, 0.0]))
    if gs.pbc[2]:
        rvecs.append(np.array([0.0, 0.0, gs.nz * dz]))
    rvecs = np.array(rvecs) if rvecs else np.zeros((0, 3))
    output = {
        "pos": pos,
        "rvecs": rvecs,
        "types": types,
        "grid": grid,
        "pbc": gs.pbc,
        "surrounding_cells": sc,
        "surrounding_nodes": sn,
        "boundary_nodes": gs.boundary_nodes,
        "masses": masses,
    }
    for type_idx, type_data in data.items():
        for key, value in type_data.items():
            output["type%i/%s" % (type_idx, key)] = value
    return output


def _majority_spacing(data, grid):
    counts = np.bincount(np.rint(grid).astype(np.int64).ravel())
    maj = int(np.argmax(counts[1:]) + 1)
    return np.diag(np.asarray(data[maj]["cell"][0], dtype=np.float64))


def build_system_device(data, grid, pbc=None, device=None, return_extra=False):
    """``build_system`` with the tables built on the GPU (``mm_topology_build``): the same dict,
    equal entry for entry (indices, masses summed in the same order, positions), for grids
    where a host pass is slow. ``return_extra`` adds ``wrap_bits`` (the force kernels' image
    shifts) and ``cell_point`` (linear grid index of every cell)."""
    import ctypes

    import torch

    from . import _native

    grid = np.asarray(grid)
    g = np.ascontiguousarray(np.rint(grid).astype(np.int32))
    if g.ndim != 3:
        raise ValueError("grid must be three-dimensional")
    pbc = _pbc_list(pbc)
    ntypes = int(g.max()) if g.size else 0
    type_mass = np.zeros(ntypes + 1)
    for t in np.unique(g):
        if t != 0:
            type_mass[int(t)] = data[int(t)]["mass"]
    spacing = np.ascontiguousarray(_majority_spacing(data, g), dtype=np.float64)
    dev = torch.device(device if device is not None else "cuda")
    dev_index = dev.index if dev.index is not None else torch.cuda.current_device()
    lib = _native.lib()
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    handle = ctypes.c_void_p()
    pbc_bits = sum(1 << d for d in range(3) if pbc[d])
    rc = lib.mm_topology_build(ctypes.byref(handle), dev_index, vp(g), *[int(x) for x in g.shape],
                               pbc_bits, vp(type_mass), ntypes, vp(spacing),
                               _native.stream_ptr(torch.cuda.current_stream(dev_index)))
    _native.check(rc, "mm_topology_build")
    try:
        C, N, B = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _native.check(lib.mm_topology_sizes(handle, ctypes.byref(C), ctypes.byref(N),
                                            ctypes.byref(B)), "mm_topology_sizes")
        C, N, B = C.value, N.value, B.value
        sn = np.empty((C, 8), dtype=np.int64)
        sc = np.empty((N, 8), dtype=np.int64)
        bn = np.empty(B, dtype=np.int64)
        masses = np.empty(N)
        pos = np.empty((N, 3))
        types = np.empty(C, dtype=np.int32)
        wrap = np.empty(C, dtype=np.uint8)
        cpt = np.empty(C, dtype=np.int32)
        _native.check(lib.mm_topology_copy(handle, vp(sn), vp(sc), vp(bn), vp(masses), vp(pos),
                                           vp(types), vp(wrap), vp(cpt),
                                           _native.stream_ptr(torch.cuda.current_stream(dev_index))),
                      "mm_topology_copy")
    finally:
        lib.mm_topology_destroy(handle)
    dx, dy, dz = spacing
    nx, ny, nz = g.shape
    rvecs = []
    if pbc[0]:
        rvecs.append(np.array([nx * dx, 0.0, 0.0]))
    if pbc[1]:
        rvecs.append(np.array([0.0, ny * dy, 0.0]))
    if pbc[2]:
        rvecs.append(np.array([0.0, 0.0, nz * dz]))
    rvecs = np.array(rvecs) if rvecs else np.zeros((0, 3))
    output = {
        "pos": pos,
        "rvecs": rvecs,
        "types": types.astype(np.int64),
        "grid": grid,
        "pbc": pbc,
        "surrounding_cells": sc,
        "surrounding_nodes": sn,
        "boundary_nodes": bn,
        "masses": masses,
    }
    for type_idx, type_data in data.items():
        for key, value in type_data.items():
            output["type%i/%s" % (type_idx, key)] = value
    if return_extra:
        output["wrap_bits"] = wrap
        output["cell_point"] = cpt
    return output


def build_type(material, mass, cell0, elasticity0, free_energy=None, effective_temp=None,
               topology=None):
    """Describe one cell type (utils.py:266-329); lists hold one entry per metastable state."""
    if not isinstance(cell0, list):
        cell0 = [cell0]
    if not isinstance(elasticity0, list):
        elasticity0 = [elasticity0]
    if topology is None:
        topology = "UNKNOWN"
    if effective_temp is None:
        effective_temp = 300 * units.kelvin
    if free_energy is None:
        free_energy = [0.0] * len(cell0)
    if not isinstance(free_energy, list):
        free_energy = [free_energy]
    if len(cell0) != len(elasticity0) or len(cell0) != len(free_energy):
        raise ValueError("cell0, elasticity0 and free_energy need one entry per state")
    return {
        "material": material,
        "topology": topology,
        "mass": mass,
        "cell": cell0,
        "elasticity": elasticity0,
        "free_energy": free_energy,
        "effective_temp": effective_temp,
    }
