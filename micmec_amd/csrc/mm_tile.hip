// Tiled cell pass: the force kernels of mm_force.hip reorganised around spatial tiles.
//
// Reference path: ForcePartMechanical._internal_compute (micmec/pes/mmff.py:288-323) with
// deformation (mmff.py:326-403), _compute_gpos (303-318) and _compute_vtens (320-323).
//
// Every workgroup owns one box of tx * ty * tz cells of one system and the (tx+1)(ty+1)(tz+1)
// node slots of the box:
//   1. the positions of the slots' nodes are staged in LDS, unwrapped through the periodic
//      boundary (a slot on the upper face of a periodic axis holds its node + that domain
//      vector, which is the image shift the reference applies to the cell's upper corners);
//   2. one lane per cell, lanes sorted by parameter row so that every wave holds one cell type
//      (parameters in scalar registers); multistate types first evaluate the energies of all
//      states and their Boltzmann weights p_s (mmff.py:386-398);
//   3. corner by corner (corner8: the eight corner matrices; original: the averaged matrix),
//      the lane builds the matrix from the LDS slots, evaluates dU/dM (weighted by p_s over
//      the states) and adds the corner's contribution to each of the four vertices it touches
//      to the LDS slot sums, one vertex role per barrier-separated sub-phase. For a fixed role
//      the target vertex is the same corner of every cell, so no two lanes of a sub-phase
//      touch one slot: the sums need no atomics and their order is fixed;
//   4. the slot sums go to HBM ([tile][slot][3]), the cell energies and the tile's virial
//      (sum_slots g_slot (x) (r_slot - r_ref), equal to the cells' sum_v g_v (x) (r_v - r_0)
//      since every cell's vertex gradients sum to zero) to the tile's partials;
//   5. a node pass sums, per node, its slots over the tiles that hold it (a CSR list in
//      increasing tile order; one slot for a node inside a tile, 2, 4 or 8 on tile faces).
// Compared with one lane per cell writing 24 vertex gradients to HBM and a node pass reading
// them back, the HBM traffic of a force evaluation drops from about 1 GB to 0.3 GB at 128^3,
// and without the 24 vertex-gradient accumulators per lane the kernel fits in far fewer
// registers, so more waves per SIMD hide the FP64 latency.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/micmec_amd.h"
#include "mm_device.h"
#include "mm_internal.h"
#include "mm_kernels.h"

namespace mm {

struct TileArgs {
  const double* pos;
  const double* rvecs;
  int64_t rvecs_stride;
  const int4* meta;
  const int4* geom;
  const int32_t* node;
  const int32_t* lane;
  const int32_t* wrow;
  const double* tab;
  const double* kt;
  const int32_t* nstates;
  int max_states;
  int tx, ty, tz, S, L, maxw, ntiles;
  double* nodeacc;
  double* partials;
};

constexpr int kTileThreadsMax = 1024;

__device__ __forceinline__ int slot_off(int v, int sx, int sy) {
  return kOff[v][0] * sx + kOff[v][1] * sy + kOff[v][2];
}

// Corner matrix of corner a (rows: the edges through a along x, y, z) from the LDS slots.
__device__ __forceinline__ void tile_corner_matrix(const double* __restrict__ px,
                                                   const double* __restrict__ py,
                                                   const double* __restrict__ pz, int base,
                                                   int sx, int sy, int a, double M[9]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int hi = base + slot_off(kHi[a][d], sx, sy);
    const int lo = base + slot_off(kLo[a][d], sx, sy);
    M[3 * d + 0] = px[hi] - px[lo];
    M[3 * d + 1] = py[hi] - py[lo];
    M[3 * d + 2] = pz[hi] - pz[lo];
  }
}

// Averaged matrix H = (1/8) sum_a M_a of the original model (nanocell_original.py:38-132).
__device__ __forceinline__ void tile_mean_matrix(const double* __restrict__ px,
                                                 const double* __restrict__ py,
                                                 const double* __restrict__ pz, int base,
                                                 int sx, int sy, double H[9]) {
  double r[8][3];
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int s = base + slot_off(v, sx, sy);
    r[v][0] = px[s];
    r[v][1] = py[s];
    r[v][2] = pz[s];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < 8; ++a) acc += r[kHi[a][d]][j] - r[kLo[a][d]][j];
      H[3 * d + j] = 0.125 * acc;
    }
  }
}

// Energy of one state (without its free energy), no gradient.
template <int kModel>
__device__ __forceinline__ double tile_state_energy(const double* __restrict__ tab,
                                                    const double* __restrict__ px,
                                                    const double* __restrict__ py,
                                                    const double* __restrict__ pz, int base,
                                                    int sx, int sy) {
  double G[9];
  if (kModel == 0) {
    double usum = 0.0;
    // rolled: unrolled corners are interleaved by the scheduler and no longer fit the registers
#pragma unroll 1
    for (int a = 0; a < 8; ++a) {
      double M[9];
      tile_corner_matrix(px, py, pz, base, sx, sy, a, M);
      usum += corner_energy<false>(tab, M, G, kCvwOffset);
    }
    return usum;
  } else {
    double H[9];
    tile_mean_matrix(px, py, pz, base, sx, sy, H);
    return corner_energy<false>(tab, H, G, kCvwOffset);
  }
}

// dU/dM of the lane's cell for matrix M: single-state cells return their own derivative
// (and add the energy to usum); multistate cells the p_s-weighted sum over the states.
__device__ __forceinline__ void tile_matrix_gradient(const double* __restrict__ tab, int ns,
                                                     const double* __restrict__ psh, int L,
                                                     const double M[9], double G[9],
                                                     double& usum) {
  if (ns == 1) {
    usum += corner_energy<true>(tab, M, G, kCvwOffset);
  } else {
#pragma unroll
    for (int q = 0; q < 9; ++q) G[q] = 0.0;
    for (int s = 0; s < ns; ++s) {
      double Gs[9];
      (void)corner_energy<true>(tab + s * kTabStride, M, Gs, kCvwOffset);
      const double p = psh[s * L];
#pragma unroll
      for (int q = 0; q < 9; ++q) G[q] += p * Gs[q];
    }
  }
}

__device__ __forceinline__ void slot_add(double* __restrict__ gx, double* __restrict__ gy,
                                         double* __restrict__ gz, int s, const double c[3]) {
  gx[s] += c[0];
  gy[s] += c[1];
  gz[s] += c[2];
}

template <int kModel, bool kGrad>
__global__ void __launch_bounds__(kTileThreadsMax) tile_kernel(TileArgs a) {
  extern __shared__ double lds[];
  const int S = a.S, L = a.L;
  double* px = lds;
  double* py = px + S;
  double* pz = py + S;
  double* gx = pz + S;
  double* gy = gx + S;
  double* gz = gy + S;
  double* psh = gz + S;                  // [max_states][L] state energies, then weights
  double* red = psh + a.max_states * L;  // [10][waves]
  const int tid = threadIdx.x;
  const int nthreads = blockDim.x;
  const int t = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const int4 tm = a.meta[t];
  const int4 sg = a.geom[tm.w];
  const double* rvb = a.rvecs + (int64_t)tm.w * a.rvecs_stride;
  const int sy = a.tz + 1, sx = (a.ty + 1) * (a.tz + 1);

  // 1. slot positions (unwrapped) and zeroed sums; ref_slot = first slot with a node
  __shared__ int ref_slot;
  if (tid == 0) ref_slot = S - 1;
  __syncthreads();
  for (int s = tid; s < S; s += nthreads) {
    const int n = a.node[(int64_t)t * S + s];
    if (n >= 0) atomicMin(&ref_slot, s);
    const int i = s / sx, rem = s - i * sx, j = rem / sy, k = rem - j * sy;
    double p[3] = {0.0, 0.0, 0.0};
    if (n >= 0) {
      const bool wx = sg.w && tm.x + i == sg.x;
      const bool wy = sg.w && tm.y + j == sg.y;
      const bool wz = sg.w && tm.z + k == sg.z;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double sh = ((wx ? rvb[0 + q] : 0.0) + (wy ? rvb[3 + q] : 0.0)) + (wz ? rvb[6 + q] : 0.0);
        p[q] = a.pos[3 * (int64_t)n + q] + sh;
      }
    }
    px[s] = p[0];
    py[s] = p[1];
    pz[s] = p[2];
    gx[s] = 0.0;
    gy[s] = 0.0;
    gz[s] = 0.0;
  }
  __syncthreads();

  // 2. the lane's cell and its type (uniform per wave)
  const int wave = tid >> 6;
  const int row = __builtin_amdgcn_readfirstlane(a.wrow[(int64_t)t * a.maxw + wave]);
  int c = a.lane[(int64_t)t * L + tid];
  const bool act = row >= 0 && c >= 0;
  if (c < 0) c = 0;
  const int ci = c / (a.ty * a.tz), cr = c - ci * (a.ty * a.tz);
  const int cj = cr / a.tz, ck = cr - cj * a.tz;
  const int base = ci * sx + cj * sy + ck;
  const int ns = row >= 0 ? a.nstates[row] : 1;
  const double* tab = a.tab + (int64_t)(row >= 0 ? row : 0) * a.max_states * kTabStride;
  double* pl = psh + tid;
  double acc[10];
#pragma unroll
  for (int q = 0; q < 10; ++q) acc[q] = 0.0;
  double ecell = 0.0;
  if (row >= 0 && ns > 1) {
    // energies of all states, E = E_min - kT ln sum exp(-(E_s - E_min)/kT)
    const double kt = a.kt[row];
    double emin = 0.0;
    for (int s = 0; s < ns; ++s) {
      const double* ts = tab + s * kTabStride;
      const double es = ts[10] + tile_state_energy<kModel>(ts, px, py, pz, base, sx, sy);
      pl[s * L] = es;
      emin = (s == 0 || es < emin) ? es : emin;
    }
    double zsum = 0.0;
    for (int s = 0; s < ns; ++s) zsum += exp(-(pl[s * L] - emin) / kt);
    ecell = emin - kt * log(zsum);
    if (kGrad)
      for (int s = 0; s < ns; ++s) pl[s * L] = exp(-(pl[s * L] - emin) / kt) / zsum;
  }

  if (!kGrad) {
    if (row >= 0 && ns == 1) ecell = tab[10] + tile_state_energy<kModel>(tab, px, py, pz, base, sx, sy);
  } else {
    // 3. matrix gradients, accumulated into the slot sums one vertex role at a time
    double usum = 0.0;
    if (kModel == 0) {
#pragma unroll 1
      for (int ca = 0; ca < 8; ++ca) {
        double M[9], G[9];
        tile_corner_matrix(px, py, pz, base, sx, sy, ca, M);
        if (row >= 0) {
          tile_matrix_gradient(tab, ns, pl, L, M, G, usum);
        } else {
#pragma unroll
          for (int q = 0; q < 9; ++q) G[q] = 0.0;
        }
        // role 0: the corner's own vertex (+G_d where it is the upper end of edge d, -G_d
        // where the lower); roles 1..3: the other end of edge d (the opposite sign)
        {
          double self[3];
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            double sq = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const double g = G[3 * d + q];
              sq += (kHi[ca][d] == ca) ? g : -g;
            }
            self[q] = sq;
          }
          __syncthreads();
          if (act) slot_add(gx, gy, gz, base + slot_off(ca, sx, sy), self);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const bool is_hi = kHi[ca][d] == ca;
          const int v = is_hi ? kLo[ca][d] : kHi[ca][d];
          double far[3];
#pragma unroll
          for (int q = 0; q < 3; ++q) far[q] = is_hi ? -G[3 * d + q] : G[3 * d + q];
          __syncthreads();
          if (act) slot_add(gx, gy, gz, base + slot_off(v, sx, sy), far);
        }
      }
    } else {
      double H[9], G[9];
      tile_mean_matrix(px, py, pz, base, sx, sy, H);
      if (row >= 0) {
        tile_matrix_gradient(tab, ns, pl, L, H, G, usum);
      } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) G[q] = 0.0;
      }
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        double cv[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          double sq = 0.0;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double tq = 0.25 * G[3 * d + q];
            sq += kOff[v][d] ? tq : -tq;
          }
          cv[q] = sq;
        }
        __syncthreads();
        if (act) slot_add(gx, gy, gz, base + slot_off(v, sx, sy), cv);
      }
    }
    if (row >= 0 && ns == 1) ecell = tab[10] + usum;
  }
  acc[0] = act ? ecell : 0.0;

  // 4. slot sums to HBM; the tile's virial sum_cells sum_v g_v (x) (r_v - r_0) equals
  // sum_slots g_slot (x) (r_slot - r_ref) because the vertex gradients of every cell sum to
  // zero; r_ref is a slot of the tile, so the products stay at the tile's scale.
  __syncthreads();
  if (kGrad) {
    double* out = a.nodeacc + (int64_t)t * S * 3;
    for (int q = tid; q < 3 * S; q += nthreads) {
      const int s = q / 3, j = q - 3 * s;
      out[q] = j == 0 ? gx[s] : (j == 1 ? gy[s] : gz[s]);
    }
    const int rs = ref_slot;
    const double rx = px[rs], ry = py[rs], rz = pz[rs];
    for (int s = tid; s < S; s += nthreads) {
      const double g[3] = {gx[s], gy[s], gz[s]};
      const double d[3] = {px[s] - rx, py[s] - ry, pz[s] - rz};
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) acc[1 + 3 * k + l] += g[k] * d[l];
    }
  }
  wave_reduce<10>(acc);
  const int nw = nthreads >> 6;
  if ((tid & 63) == 0) {
#pragma unroll
    for (int q = 0; q < 10; ++q) red[q * nw + wave] = acc[q];
  }
  __syncthreads();
  if (tid < 10) {
    double s = red[tid * nw];
    for (int w = 1; w < nw; ++w) s += red[tid * nw + w];
    a.partials[(int64_t)tid * a.ntiles + t] = s;
  }
}

// gpos[n] (+)= sum of the node's slots over its tiles, in list order.
__global__ void __launch_bounds__(256)
tile_gather_kernel(const int32_t* __restrict__ hptr, const int32_t* __restrict__ hent,
                   const double* __restrict__ nodeacc, int64_t N, double* __restrict__ gpos,
                   int accumulate, int* __restrict__ nanflag) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  double g[3] = {0.0, 0.0, 0.0};
  const int e1 = hptr[n + 1];
  for (int e = hptr[n]; e < e1; ++e) {
    const int64_t idx = 3 * (int64_t)hent[e];
#pragma unroll
    for (int j = 0; j < 3; ++j) g[j] += nodeacc[idx + j];
  }
  bool bad = false;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    bad |= (g[j] != g[j]);
    gpos[3 * n + j] = accumulate ? gpos[3 * n + j] + g[j] : g[j];
  }
  if (bad) atomicOr(nanflag, 2);
}

size_t tile_lds_bytes(const mm_tiles& tl, int max_states) {
  const int L = 64 * tl.maxw;
  return sizeof(double) * ((size_t)6 * tl.S + (size_t)max_states * L + (size_t)10 * tl.maxw);
}

hipError_t launch_tile_cells(const mm_field* f, const double* pos, const double* rvecs,
                             int64_t rvecs_stride, bool grad, hipStream_t stream) {
  const mm_tiles& tl = f->tiles;
  if (tl.ntiles == 0) return hipSuccess;
  TileArgs a;
  a.pos = pos;
  a.rvecs = rvecs;
  a.rvecs_stride = rvecs_stride;
  a.meta = tl.meta;
  a.geom = tl.geom;
  a.node = tl.node;
  a.lane = tl.lane;
  a.wrow = tl.wrow;
  a.tab = f->tab;
  a.kt = f->kt;
  a.nstates = f->nstates;
  a.max_states = f->max_states;
  a.tx = tl.tx;
  a.ty = tl.ty;
  a.tz = tl.tz;
  a.S = tl.S;
  a.L = 64 * tl.maxw;
  a.maxw = tl.maxw;
  a.ntiles = tl.ntiles;
  a.nodeacc = tl.nodeacc;
  a.partials = tl.partials;
  const size_t lds = tile_lds_bytes(tl, f->max_states);
  const dim3 grid(tl.ntiles), block(64 * tl.maxw);
#define MM_TILE(MODEL, GRAD) \
  hipLaunchKernelGGL((tile_kernel<MODEL, GRAD>), grid, block, lds, stream, a)
  if (f->model == 0) {
    if (grad) MM_TILE(0, true); else MM_TILE(0, false);
  } else {
    if (grad) MM_TILE(1, true); else MM_TILE(1, false);
  }
#undef MM_TILE
  return hipGetLastError();
}

hipError_t launch_tile_gather(const mm_field* f, double* gpos, int accumulate, hipStream_t stream) {
  if (!gpos || f->N <= 0) return hipSuccess;
  const int gb = (int)((f->N + 255) / 256);
  hipLaunchKernelGGL(tile_gather_kernel, dim3(gb), dim3(256), 0, stream, f->tiles.hptr,
                     f->tiles.hent, f->tiles.nodeacc, f->N, gpos, accumulate, f->nanflag);
  return hipGetLastError();
}

void free_tiles(mm_tiles& t) {
  (void)hipFree(t.meta);
  (void)hipFree(t.geom);
  (void)hipFree(t.node);
  (void)hipFree(t.lane);
  (void)hipFree(t.wrow);
  (void)hipFree(t.nodeacc);
  (void)hipFree(t.partials);
  (void)hipFree(t.sys_ptr);
  (void)hipFree(t.sys_idx);
  (void)hipFree(t.hptr);
  (void)hipFree(t.hent);
  t = mm_tiles();
}

}  // namespace mm

using namespace mm;

namespace {

// host copy of the corner offsets kOff
const int kOffHost[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                            {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    (void)hipGetDevice(&prev);
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

template <typename T>
hipError_t upload(T** dst, const std::vector<T>& src) {
  const size_t n = src.empty() ? 1 : src.size();
  hipError_t e = hipMalloc((void**)dst, n * sizeof(T));
  if (e == hipSuccess && !src.empty())
    e = hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

}  // namespace

extern "C" {

int mm_field_set_grid(mm_field* f, const int32_t* cell_ijk, const int32_t* sys_dims,
                      const int32_t* sys_pbc, int32_t tx, int32_t ty, int32_t tz) {
  if (!f) return set_error(MM_ERR_ARG, "field is NULL");
  if (f->C > 0 && (!cell_ijk || !sys_dims || !sys_pbc))
    return set_error(MM_ERR_ARG, "missing grid tables");
  if (tx < 1 || ty < 1 || tz < 1 || tx > 16 || ty > 16 || tz > 16)
    return set_error(MM_ERR_ARG, "tile %d x %d x %d out of range", tx, ty, tz);
  const int nsys = f->nsys;
  const int64_t C = f->C, W = f->W, N = f->N;
  for (int s = 0; s < nsys; ++s)
    for (int d = 0; d < 3; ++d)
      if (sys_dims[3 * s + d] < 1 || sys_dims[3 * s + d] > (1 << 20))
        return set_error(MM_ERR_ARG, "system %d: %d cells along axis %d", s, sys_dims[3 * s + d], d);
  DeviceGuard guard(f->device);
  // per-cell corner nodes, wrap bits, parameter row and system, from the work order
  std::vector<int32_t> cnode((size_t)8 * (W > 0 ? W : 1)), cwrap(W > 0 ? W : 1), bsys(W / kCellBlock + 1);
  hipError_t e = hipSuccess;
  if (W > 0) {
    e = hipMemcpy(cnode.data(), f->cnode, sizeof(int32_t) * 8 * W, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(cwrap.data(), f->cwrap, sizeof(int32_t) * W, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = hipMemcpy(bsys.data(), f->blk_sys, sizeof(int32_t) * (W / kCellBlock), hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess) return hip_error(e, "mm_field_set_grid");
  // tiles of every system, numbered system by system, x-major
  std::vector<int64_t> tile_base(nsys + 1, 0);
  std::vector<int> ntx(nsys), nty(nsys), ntz(nsys);
  for (int s = 0; s < nsys; ++s) {
    ntx[s] = (sys_dims[3 * s] + tx - 1) / tx;
    nty[s] = (sys_dims[3 * s + 1] + ty - 1) / ty;
    ntz[s] = (sys_dims[3 * s + 2] + tz - 1) / tz;
    tile_base[s + 1] = tile_base[s] + (int64_t)ntx[s] * nty[s] * ntz[s];
  }
  const int64_t nt_all = tile_base[nsys];
  if (nt_all > (1LL << 28)) return set_error(MM_ERR_ARG, "too many tiles");
  const int S = (tx + 1) * (ty + 1) * (tz + 1), sy = tz + 1, sx = (ty + 1) * (tz + 1);
  const int cells_per_tile = tx * ty * tz;
  std::vector<int64_t> cell_tile(C);
  std::vector<int32_t> cell_local(C), cell_row(C);
  std::vector<int64_t> tile_count(nt_all, 0);
  for (int64_t c = 0; c < C; ++c) {
    const int64_t w = f->slot_of_cell[c];
    const int sys = bsys[w / kCellBlock];
    const int* dims = sys_dims + 3 * sys;
    int p[3];
    for (int d = 0; d < 3; ++d) {
      p[d] = cell_ijk[3 * c + d];
      if (p[d] < 0 || p[d] >= dims[d])
        return set_error(MM_ERR_ARG, "cell %lld lies outside its grid", (long long)c);
    }
    int expect = 0;
    for (int d = 0; d < 3; ++d)
      if (sys_pbc[sys] && p[d] == dims[d] - 1) expect |= 1 << d;
    if (cwrap[w] != expect)
      return set_error(MM_ERR_ARG, "cell %lld: wrap bits %d do not follow the grid (%d)",
                       (long long)c, cwrap[w], expect);
    const int64_t t = tile_base[sys] +
                      ((int64_t)(p[0] / tx) * nty[sys] + p[1] / ty) * ntz[sys] + p[2] / tz;
    cell_tile[c] = t;
    cell_local[c] = ((p[0] % tx) * ty + p[1] % ty) * tz + p[2] % tz;
    cell_row[c] = f->blk_row_host[w / kCellBlock];
    ++tile_count[t];
  }
  // non-empty tiles get consecutive indices
  std::vector<int64_t> tile_new(nt_all, -1);
  std::vector<int4> meta;
  std::vector<int32_t> sys_ptr(nsys + 1, 0);
  for (int s = 0; s < nsys; ++s) {
    for (int64_t t = tile_base[s]; t < tile_base[s + 1]; ++t) {
      if (tile_count[t] == 0) continue;
      const int64_t r = t - tile_base[s];
      const int ix = (int)(r / ((int64_t)nty[s] * ntz[s]));
      const int iy = (int)((r / ntz[s]) % nty[s]);
      const int iz = (int)(r % ntz[s]);
      tile_new[t] = (int64_t)meta.size();
      meta.push_back(make_int4(ix * tx, iy * ty, iz * tz, s));
    }
    sys_ptr[s + 1] = (int32_t)meta.size();
  }
  const int64_t nt = (int64_t)meta.size();
  // cells of every tile, grouped by parameter row (multistate rows first, the work order's
  // row order), each group padded to whole waves
  std::vector<std::vector<int64_t>> tcells(nt);
  for (int64_t c = 0; c < C; ++c) tcells[tile_new[cell_tile[c]]].push_back(c);
  std::vector<int> row_rank(f->nrows, 0);
  {
    int r = 0;
    std::vector<int> seen(f->nrows, 0);
    for (size_t b = 0; b < f->blk_row_host.size(); ++b) {
      const int row = f->blk_row_host[b];
      if (!seen[row]) {
        seen[row] = 1;
        row_rank[row] = r++;
      }
    }
    for (int row = 0; row < f->nrows; ++row)
      if (!seen[row]) row_rank[row] = r++;
  }
  int maxw = 1;
  for (int64_t t = 0; t < nt; ++t) {
    auto& v = tcells[t];
    std::stable_sort(v.begin(), v.end(), [&](int64_t x, int64_t y) {
      return row_rank[cell_row[x]] < row_rank[cell_row[y]];
    });
    int waves = 0;
    for (size_t i = 0; i < v.size();) {
      size_t j = i;
      while (j < v.size() && cell_row[v[j]] == cell_row[v[i]]) ++j;
      waves += (int)((j - i + 63) / 64);
      i = j;
    }
    maxw = std::max(maxw, waves);
  }
  const int L = 64 * maxw;
  mm_tiles tl;
  tl.tx = tx;
  tl.ty = ty;
  tl.tz = tz;
  tl.S = S;
  tl.maxw = maxw;
  tl.ntiles = (int)nt;
  if (L > kTileThreadsMax)
    return set_error(MM_ERR_ARG, "tiles of %d x %d x %d need %d lanes (at most %d)", tx, ty, tz,
                     L, kTileThreadsMax);
  if (tile_lds_bytes(tl, f->max_states) > 65536 - 64)
    return set_error(MM_ERR_ARG, "tiles of %d x %d x %d need %zu bytes of LDS", tx, ty, tz,
                     tile_lds_bytes(tl, f->max_states));
  std::vector<int32_t> lane((size_t)nt * L, -1), wrow((size_t)nt * maxw, -1),
      node((size_t)nt * S, -1);
  for (int64_t t = 0; t < nt; ++t) {
    const auto& v = tcells[t];
    int wv = 0;
    for (size_t i = 0; i < v.size();) {
      size_t j = i;
      while (j < v.size() && cell_row[v[j]] == cell_row[v[i]]) ++j;
      for (size_t q = i; q < j; ++q) {
        const size_t l = (size_t)wv * 64 + (q - i);
        lane[(size_t)t * L + l] = cell_local[v[q]];
      }
      const int nwv = (int)((j - i + 63) / 64);
      for (int q = 0; q < nwv; ++q) wrow[(size_t)t * maxw + wv + q] = cell_row[v[i]];
      wv += nwv;
      i = j;
    }
    // slots: node at every corner of every cell; a slot is shared by the cells around it
    std::vector<char> used(cells_per_tile, 0);
    for (int64_t c : v) {
      const int lc = cell_local[c];
      if (used[lc]) return set_error(MM_ERR_ARG, "two cells at one grid point (cell %lld)", (long long)c);
      used[lc] = 1;
      const int ci = lc / (ty * tz), cj = (lc / tz) % ty, ck = lc % tz;
      const int base = ci * sx + cj * sy + ck;
      const int64_t w = f->slot_of_cell[c];
      for (int k = 0; k < 8; ++k) {
        const int s = base + kOffHost[k][0] * sx + kOffHost[k][1] * sy + kOffHost[k][2];
        const int32_t n = cnode[(size_t)w * 8 + k];
        int32_t& slot = node[(size_t)t * S + s];
        if (slot < 0) {
          slot = n;
        } else if (slot != n) {
          return set_error(MM_ERR_ARG, "cell %lld: corner %d is node %d, its grid point holds %d",
                           (long long)c, k, n, slot);
        }
      }
    }
  }
  // CSR: for every node, its slots in increasing tile (and slot) order
  std::vector<int32_t> hptr(N + 1, 0);
  for (size_t q = 0; q < node.size(); ++q)
    if (node[q] >= 0) ++hptr[node[q] + 1];
  // (nodes of no cell, such as the padding rows of a batch, get no entry and a zero gradient)
  for (int64_t n = 0; n < N; ++n) hptr[n + 1] += hptr[n];
  const int64_t E = hptr[N];
  if (E > (1LL << 31) - 1 || (int64_t)nt * S > (1LL << 31) - 1)
    return set_error(MM_ERR_ARG, "too many node slots");
  std::vector<int32_t> hent(E), fill(hptr.begin(), hptr.end() - 1);
  for (size_t q = 0; q < node.size(); ++q)
    if (node[q] >= 0) hent[fill[node[q]]++] = (int32_t)q;
  std::vector<int4> geom(nsys);
  for (int s = 0; s < nsys; ++s)
    geom[s] = make_int4(sys_dims[3 * s], sys_dims[3 * s + 1], sys_dims[3 * s + 2],
                        sys_pbc[s] ? 1 : 0);
  std::vector<int32_t> sys_idx(nt);
  for (int64_t t = 0; t < nt; ++t) sys_idx[t] = (int32_t)t;
  tl.E = E;
  if (e == hipSuccess) e = upload(&tl.meta, meta);
  if (e == hipSuccess) e = upload(&tl.geom, geom);
  if (e == hipSuccess) e = upload(&tl.node, node);
  if (e == hipSuccess) e = upload(&tl.lane, lane);
  if (e == hipSuccess) e = upload(&tl.wrow, wrow);
  if (e == hipSuccess) e = upload(&tl.sys_ptr, sys_ptr);
  if (e == hipSuccess) e = upload(&tl.sys_idx, sys_idx);
  if (e == hipSuccess) e = upload(&tl.hptr, hptr);
  if (e == hipSuccess) e = upload(&tl.hent, hent);
  if (e == hipSuccess)
    e = hipMalloc((void**)&tl.nodeacc, sizeof(double) * 3 * (size_t)(nt > 0 ? nt : 1) * S);
  if (e == hipSuccess)
    e = hipMalloc((void**)&tl.partials, sizeof(double) * kPartialStride * (size_t)(nt > 0 ? nt : 1));
  if (e != hipSuccess) {
    free_tiles(tl);
    return hip_error(e, "mm_field_set_grid");
  }
  free_tiles(f->tiles);
  tl.enabled = 1;
  f->tiles = tl;
  return MM_OK;
}

int mm_field_tiles_info(const mm_field* f, int32_t* enabled, int64_t* ntiles, int32_t* lanes,
                        int64_t* entries) {
  if (!f) return set_error(MM_ERR_ARG, "field is NULL");
  if (enabled) *enabled = f->tiles.enabled;
  if (ntiles) *ntiles = f->tiles.ntiles;
  if (lanes) *lanes = 64 * f->tiles.maxw;
  if (entries) *entries = f->tiles.E;
  return MM_OK;
}

}  // extern "C"
