// Topology of a grid of cells built on the device (micmec/utils.py:55-263: Grid and
// build_system), for grids too large for a host pass or built many times.
//
// The tables are those of the reference, in its orders:
//   * cells: the non-empty grid points (type != 0), x-major, z fastest (utils.py:149-161);
//   * nodes: the lattice points next to at least one non-empty cell, same order
//     (utils.py:113-136); node planes per axis: n (periodic) or n + 1 (not periodic);
//   * surrounding_nodes[c][k]: the node at corner k of cell c (corner offsets of
//     utils.py:43-52); surrounding_cells[n][k]: the cell in which node n is corner k, -1 if
//     that grid point is empty or outside the grid (utils.py:198-218);
//   * masses: one eighth of the type mass of every (cell, corner) that lands on the node, added
//     in increasing (cell, corner) order (utils.py:203-218);
//   * positions: node lattice index times the cell spacing (utils.py:220-233);
//   * boundary_nodes: the nodes on a face of the node lattice (utils.py:138-147);
//   * wrap bits: bit d when the cell is in the last layer of periodic axis d (the image shifts of
//     mmff.py:258-286 as the force kernels apply them).
// Indices come from exclusive prefix sums over the node and cell masks (integer block scans, so
// every index is exact and independent of scheduling).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/micmec_amd.h"
#include "mm_internal.h"

namespace mm {

constexpr int kScanBlock = 1024;

// corner offsets of a cell (node k of the cell = cell + kCornerOff[k])
__constant__ int kCornerOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                     {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};

struct TopoGeom {
  int nx, ny, nz;     // cells per axis
  int mx, my, mz;     // node lattice points per axis
  int px, py, pz;     // periodic flags
};

__device__ __forceinline__ int64_t cell_lin(const TopoGeom& g, int i, int j, int k) {
  return ((int64_t)i * g.ny + j) * g.nz + k;
}
__device__ __forceinline__ int64_t node_lin(const TopoGeom& g, int i, int j, int k) {
  return ((int64_t)i * g.my + j) * g.mz + k;
}

// Grid point of the cell in which lattice point (i, j, k) is corner c, or -1 if there is none.
__device__ __forceinline__ int64_t corner_cell_point(const TopoGeom& g, int i, int j, int k, int c) {
  int a = i - kCornerOff[c][0], b = j - kCornerOff[c][1], d = k - kCornerOff[c][2];
  if (g.px) a = (a % g.nx + g.nx) % g.nx; else if (a < 0 || a >= g.nx) return -1;
  if (g.py) b = (b % g.ny + g.ny) % g.ny; else if (b < 0 || b >= g.ny) return -1;
  if (g.pz) d = (d % g.nz + g.nz) % g.nz; else if (d < 0 || d >= g.nz) return -1;
  return cell_lin(g, a, b, d);
}

__global__ void topo_masks(const int32_t* __restrict__ grid, TopoGeom g, int64_t ncell_pts,
                           int64_t nnode_pts, int32_t* __restrict__ cell_flag,
                           int32_t* __restrict__ node_flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ncell_pts) cell_flag[t] = grid[t] != 0 ? 1 : 0;
  if (t < nnode_pts) {
    const int i = (int)(t / ((int64_t)g.my * g.mz));
    const int j = (int)((t / g.mz) % g.my);
    const int k = (int)(t % g.mz);
    int any = 0;
    for (int c = 0; c < 8; ++c) {
      const int64_t p = corner_cell_point(g, i, j, k, c);
      if (p >= 0 && grid[p] != 0) any = 1;
    }
    node_flag[t] = any;
  }
}

// Exclusive scan of flags within blocks of kScanBlock; block totals to sums.
__global__ void __launch_bounds__(kScanBlock)
scan_blocks(const int32_t* __restrict__ flags, int64_t n, int32_t* __restrict__ out,
            int32_t* __restrict__ sums) {
  __shared__ int32_t wsum[kScanBlock / 64];
  const int64_t t = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  const int v = t < n ? flags[t] : 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) {
      const int s = wsum[w];
      wsum[w] = run;
      run += s;
    }
    sums[blockIdx.x] = run;
  }
  __syncthreads();
  if (t < n) out[t] = wsum[wave] + x - v;
}

// Exclusive scan of the block totals in place (one block, chunks of kScanBlock); total to *total.
__global__ void __launch_bounds__(kScanBlock)
scan_sums(int32_t* __restrict__ sums, int nb, int32_t* __restrict__ total) {
  __shared__ int32_t wsum[kScanBlock / 64];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int base = 0; base < nb; base += kScanBlock) {
    const int t = base + threadIdx.x;
    const int v = t < nb ? sums[t] : 0;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      int run = 0;
      for (int w = 0; w < kScanBlock / 64; ++w) {
        const int s = wsum[w];
        wsum[w] = run;
        run += s;
      }
    }
    __syncthreads();
    const int c0 = carry;
    if (t < nb) sums[t] = c0 + wsum[wave] + x - v;
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1) carry = c0 + wsum[wave] + x;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// index[t] = flag ? prefix : -1
__global__ void scan_finish(const int32_t* __restrict__ flags, const int32_t* __restrict__ sums,
                            int64_t n, int32_t* __restrict__ index) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  index[t] = flags[t] ? index[t] + sums[t / kScanBlock] : -1;
}

__global__ void topo_cells(const int32_t* __restrict__ grid, TopoGeom g, int64_t ncell_pts,
                           const int32_t* __restrict__ cell_index,
                           const int32_t* __restrict__ node_index, int64_t* __restrict__ sn,
                           int32_t* __restrict__ types, uint8_t* __restrict__ wrap,
                           int32_t* __restrict__ cell_point) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell_pts) return;
  const int32_t c = cell_index[t];
  if (c < 0) return;
  const int i = (int)(t / ((int64_t)g.ny * g.nz));
  const int j = (int)((t / g.nz) % g.ny);
  const int k = (int)(t % g.nz);
  for (int q = 0; q < 8; ++q) {
    const int a = (i + kCornerOff[q][0]) % g.mx;
    const int b = (j + kCornerOff[q][1]) % g.my;
    const int d = (k + kCornerOff[q][2]) % g.mz;
    sn[8 * (int64_t)c + q] = node_index[node_lin(g, a, b, d)];
  }
  types[c] = grid[t];
  wrap[c] = (uint8_t)(((g.px && i == g.nx - 1) ? 1 : 0) | ((g.py && j == g.ny - 1) ? 2 : 0) |
                      ((g.pz && k == g.nz - 1) ? 4 : 0));
  cell_point[c] = (int32_t)t;
}

__global__ void topo_nodes(const int32_t* __restrict__ grid, TopoGeom g, int64_t nnode_pts,
                           const int32_t* __restrict__ cell_index,
                           const int32_t* __restrict__ node_index,
                           const double* __restrict__ type_mass, int ntypes, double sx,
                           double sy, double sz, int64_t* __restrict__ sc,
                           double* __restrict__ masses, double* __restrict__ pos,
                           int32_t* __restrict__ bflag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode_pts) return;
  const int32_t n = node_index[t];
  if (n < 0) return;
  const int i = (int)(t / ((int64_t)g.my * g.mz));
  const int j = (int)((t / g.mz) % g.my);
  const int k = (int)(t % g.mz);
  int32_t cells[8];
  for (int q = 0; q < 8; ++q) {
    const int64_t p = corner_cell_point(g, i, j, k, q);
    const int32_t c = p >= 0 ? cell_index[p] : -1;
    cells[q] = c;
    sc[8 * (int64_t)n + q] = c;
  }
  // mass: 1/8 of the type mass of every (cell, corner) landing here, in (cell, corner) order
  int64_t key[8];
  for (int q = 0; q < 8; ++q) key[q] = cells[q] >= 0 ? (int64_t)cells[q] * 8 + q : INT64_MAX;
  for (int a = 1; a < 8; ++a) {
    const int64_t v = key[a];
    int b = a - 1;
    while (b >= 0 && key[b] > v) {
      key[b + 1] = key[b];
      --b;
    }
    key[b + 1] = v;
  }
  double m = 0.0;
  for (int q = 0; q < 8; ++q) {
    if (key[q] == INT64_MAX) break;
    const int64_t p = corner_cell_point(g, i, j, k, (int)(key[q] & 7));
    const int ty = grid[p];
    const double tm = (ty >= 0 && ty <= ntypes) ? type_mass[ty] : 0.0;
    m += 0.125 * tm;
  }
  masses[n] = m;
  pos[3 * (int64_t)n + 0] = (double)i * sx;
  pos[3 * (int64_t)n + 1] = (double)j * sy;
  pos[3 * (int64_t)n + 2] = (double)k * sz;
  bflag[n] = (i == 0 || i == g.mx - 1 || j == 0 || j == g.my - 1 || k == 0 || k == g.mz - 1) ? 1 : 0;
}

__global__ void topo_boundary(const int32_t* __restrict__ bflag, const int32_t* __restrict__ bidx,
                              int64_t nnodes, int64_t* __restrict__ boundary) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < nnodes && bflag[n]) boundary[bidx[n]] = n;
}

}  // namespace mm

using namespace mm;

// Device-resident topology tables (include/micmec_amd.h).
struct mm_topology {
  int device = 0;
  int64_t C = 0, N = 0, B = 0;
  int64_t* sn = nullptr;        // [C][8]
  int64_t* sc = nullptr;        // [N][8]
  int64_t* boundary = nullptr;  // [B]
  double* masses = nullptr;     // [N]
  double* pos = nullptr;        // [N][3]
  int32_t* types = nullptr;     // [C]
  uint8_t* wrap = nullptr;      // [C]
  int32_t* cell_point = nullptr;  // [C] linear grid index of every cell
};

namespace {

struct Guard {
  int prev = -1;
  explicit Guard(int d) {
    (void)hipGetDevice(&prev);
    if (prev != d) (void)hipSetDevice(d);
  }
  ~Guard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// exclusive scan of flags[n] into index[n] (-1 where the flag is 0); returns the count
hipError_t scan_index(const int32_t* flags, int64_t n, int32_t* index, int32_t* sums,
                      int32_t* total_dev, int32_t* total_host, hipStream_t s) {
  const int nb = (int)((n + kScanBlock - 1) / kScanBlock);
  if (nb == 0) {
    *total_host = 0;
    return hipSuccess;
  }
  hipLaunchKernelGGL(scan_blocks, dim3(nb), dim3(kScanBlock), 0, s, flags, n, index, sums);
  hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanBlock), 0, s, sums, nb, total_dev);
  hipLaunchKernelGGL(scan_finish, dim3(nb), dim3(kScanBlock), 0, s, flags, sums, n, index);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(total_host, total_dev, sizeof(int32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

void free_topology(mm_topology* t) {
  if (!t) return;
  (void)hipFree(t->sn);
  (void)hipFree(t->sc);
  (void)hipFree(t->boundary);
  (void)hipFree(t->masses);
  (void)hipFree(t->pos);
  (void)hipFree(t->types);
  (void)hipFree(t->wrap);
  (void)hipFree(t->cell_point);
  delete t;
}

}  // namespace

extern "C" {

int mm_topology_build(mm_topology** out, int device, const int32_t* grid, int32_t nx,
                      int32_t ny, int32_t nz, int32_t pbc_bits, const double* type_mass,
                      int32_t ntypes, const double* spacing, void* stream) {
  if (!out) return set_error(MM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!grid || !type_mass || !spacing) return set_error(MM_ERR_ARG, "missing argument");
  if (nx < 1 || ny < 1 || nz < 1) return set_error(MM_ERR_ARG, "grid %d x %d x %d", nx, ny, nz);
  if (ntypes < 0 || ntypes > (1 << 20)) return set_error(MM_ERR_ARG, "bad type count %d", ntypes);
  TopoGeom g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.px = (pbc_bits & 1) ? 1 : 0;
  g.py = (pbc_bits & 2) ? 1 : 0;
  g.pz = (pbc_bits & 4) ? 1 : 0;
  g.mx = nx + 1 - g.px;
  g.my = ny + 1 - g.py;
  g.mz = nz + 1 - g.pz;
  const int64_t ncp = (int64_t)nx * ny * nz, nnp = (int64_t)g.mx * g.my * g.mz;
  if (nnp >= (1LL << 31) - 1) return set_error(MM_ERR_ARG, "grid too large for 32-bit indices");
  for (int64_t t = 0; t < ncp; ++t)
    if (grid[t] < 0 || grid[t] > ntypes)
      return set_error(MM_ERR_ARG, "grid point %lld has type %d (0..%d allowed)", (long long)t,
                       grid[t], ntypes);
  Guard guard(device);
  hipStream_t s = (hipStream_t)stream;
  int32_t *grid_d = nullptr, *cflag = nullptr, *nflag = nullptr, *cidx = nullptr, *nidx = nullptr;
  int32_t *sums = nullptr, *total = nullptr, *bflag = nullptr, *bidx = nullptr;
  double* tmass = nullptr;
  const int64_t nmax = ncp > nnp ? ncp : nnp;
  const int nbmax = (int)((nmax + kScanBlock - 1) / kScanBlock) + 1;
  mm_topology* t = new mm_topology();
  t->device = device;
  int32_t C32 = 0, N32 = 0, B32 = 0;
  hipError_t e = hipMalloc((void**)&grid_d, sizeof(int32_t) * ncp);
  if (e == hipSuccess) e = hipMalloc((void**)&cflag, sizeof(int32_t) * ncp);
  if (e == hipSuccess) e = hipMalloc((void**)&cidx, sizeof(int32_t) * ncp);
  if (e == hipSuccess) e = hipMalloc((void**)&nflag, sizeof(int32_t) * nnp);
  if (e == hipSuccess) e = hipMalloc((void**)&nidx, sizeof(int32_t) * nnp);
  if (e == hipSuccess) e = hipMalloc((void**)&sums, sizeof(int32_t) * nbmax);
  if (e == hipSuccess) e = hipMalloc((void**)&total, sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&tmass, sizeof(double) * (ntypes + 1));
  if (e == hipSuccess)
    e = hipMemcpyAsync(grid_d, grid, sizeof(int32_t) * ncp, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(tmass, type_mass, sizeof(double) * (ntypes + 1), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    const int64_t nt = nmax;
    hipLaunchKernelGGL(topo_masks, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, grid_d, g,
                       ncp, nnp, cflag, nflag);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = scan_index(cflag, ncp, cidx, sums, total, &C32, s);
  if (e == hipSuccess) e = scan_index(nflag, nnp, nidx, sums, total, &N32, s);
  if (e == hipSuccess) {
    t->C = C32;
    t->N = N32;
    e = hipMalloc((void**)&t->sn, sizeof(int64_t) * 8 * (t->C > 0 ? t->C : 1));
    if (e == hipSuccess) e = hipMalloc((void**)&t->sc, sizeof(int64_t) * 8 * (t->N > 0 ? t->N : 1));
    if (e == hipSuccess) e = hipMalloc((void**)&t->masses, sizeof(double) * (t->N > 0 ? t->N : 1));
    if (e == hipSuccess) e = hipMalloc((void**)&t->pos, sizeof(double) * 3 * (t->N > 0 ? t->N : 1));
    if (e == hipSuccess) e = hipMalloc((void**)&t->types, sizeof(int32_t) * (t->C > 0 ? t->C : 1));
    if (e == hipSuccess) e = hipMalloc((void**)&t->wrap, (t->C > 0 ? t->C : 1));
    if (e == hipSuccess) e = hipMalloc((void**)&t->cell_point, sizeof(int32_t) * (t->C > 0 ? t->C : 1));
    if (e == hipSuccess) e = hipMalloc((void**)&bflag, sizeof(int32_t) * (t->N > 0 ? t->N : 1));
    if (e == hipSuccess) e = hipMalloc((void**)&bidx, sizeof(int32_t) * (t->N > 0 ? t->N : 1));
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(topo_cells, dim3((unsigned)((ncp + 255) / 256)), dim3(256), 0, s, grid_d, g,
                       ncp, cidx, nidx, t->sn, t->types, t->wrap, t->cell_point);
    hipLaunchKernelGGL(topo_nodes, dim3((unsigned)((nnp + 255) / 256)), dim3(256), 0, s, grid_d, g,
                       nnp, cidx, nidx, tmass, ntypes, spacing[0], spacing[1], spacing[2], t->sc,
                       t->masses, t->pos, bflag);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = scan_index(bflag, t->N, bidx, sums, total, &B32, s);
  if (e == hipSuccess) {
    t->B = B32;
    e = hipMalloc((void**)&t->boundary, sizeof(int64_t) * (t->B > 0 ? t->B : 1));
  }
  if (e == hipSuccess && t->N > 0) {
    hipLaunchKernelGGL(topo_boundary, dim3((unsigned)((t->N + 255) / 256)), dim3(256), 0, s, bflag,
                       bidx, t->N, t->boundary);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(grid_d);
  (void)hipFree(cflag);
  (void)hipFree(cidx);
  (void)hipFree(nflag);
  (void)hipFree(nidx);
  (void)hipFree(sums);
  (void)hipFree(total);
  (void)hipFree(tmass);
  (void)hipFree(bflag);
  (void)hipFree(bidx);
  if (e != hipSuccess) {
    free_topology(t);
    return hip_error(e, "mm_topology_build");
  }
  *out = t;
  return MM_OK;
}

int mm_topology_sizes(const mm_topology* t, int64_t* ncells, int64_t* nnodes, int64_t* nboundary) {
  if (!t) return set_error(MM_ERR_ARG, "topology is NULL");
  if (ncells) *ncells = t->C;
  if (nnodes) *nnodes = t->N;
  if (nboundary) *nboundary = t->B;
  return MM_OK;
}

int mm_topology_copy(const mm_topology* t, int64_t* surrounding_nodes, int64_t* surrounding_cells,
                     int64_t* boundary_nodes, double* masses, double* pos, int32_t* types,
                     uint8_t* wrap_bits, int32_t* cell_point, void* stream) {
  if (!t) return set_error(MM_ERR_ARG, "topology is NULL");
  Guard guard(t->device);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSuccess;
  auto cp = [&](void* dst, const void* src, size_t bytes) {
    if (e == hipSuccess && dst && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s);
  };
  cp(surrounding_nodes, t->sn, sizeof(int64_t) * 8 * t->C);
  cp(surrounding_cells, t->sc, sizeof(int64_t) * 8 * t->N);
  cp(boundary_nodes, t->boundary, sizeof(int64_t) * t->B);
  cp(masses, t->masses, sizeof(double) * t->N);
  cp(pos, t->pos, sizeof(double) * 3 * t->N);
  cp(types, t->types, sizeof(int32_t) * t->C);
  cp(wrap_bits, t->wrap, t->C);
  cp(cell_point, t->cell_point, sizeof(int32_t) * t->C);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_error(e, "mm_topology_copy");
  return MM_OK;
}

void mm_topology_destroy(mm_topology* t) {
  if (!t) return;
  Guard guard(t->device);
  free_topology(t);
}

}  // extern "C"
