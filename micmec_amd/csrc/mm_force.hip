// Force-field kernels: per-cell energies and vertex gradients, node gather, reductions.
//
// Reference path: ForcePartMechanical._internal_compute (micmec/pes/mmff.py:288-323) and
// deformation (mmff.py:326-403). Layout in HBM (all FP64 unless noted):
//   pos      [N][3]              node positions (caller's buffer)
//   rvecs    [3][3]              domain vectors, rows (device copy)
//   cnode    [W][8] int32        node at corner k of work slot w (-1 = padding slot)
//   cwrap    [W]    int32        wrap bits of the cell in slot w
//   blk_row  [W/256] int32       parameter-table row shared by all cells of a block
//   ncell    [8][N] int32        node n -> record (slot << 2 | face class) of its corner k, -1
//                                none or folded into the +z neighbour's record
//   gcell    [4][3][S]           vertex gradients by face class: the lower face (corners 0, 1,
//                                2, 4) of slot w at w, with the upper face of the -z neighbour
//                                in slot w - 1 added when the two share a wave; the other upper
//                                faces (corners 3, 5, 6, 7) at compact slots W + i (S = W + their
//                                number)
//   cbase    [W/64] int32        first compact slot of every wave
//   partials [16][W/256]         per-block energy (0) and virial (1..9), SoA
//   cgpart   [8][(N+255)/256]    per-workgroup sums of the line-search kernels (cg_*)
// Cells are grouped by type into 256-slot segments, so the type, its number of states and its
// parameter row are uniform per block and the parameters are read with scalar loads.
//
// One lane per cell: the lane gathers the cell's eight nodes, evaluates the corner matrices of
// every state and accumulates the eight vertex gradients in registers (for multistate cells
// each state's; their weighted sum over the states is kept in LDS).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mm_device.h"
#include "mm_kernels.h"

namespace mm {

// Deterministic block reduction (the reduce-scatter butterfly per wave, 17 exchanges for up to
// 16 values, then the waves in order); the block's totals are written SoA to out[q * stride].
template <int NV, int BS>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double* smem, double* out,
                                                   int64_t stride) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  {
    const double t = wave_reduce_scatter<NV>(v, lane);
    const int q = scatter_index(lane);
    if ((lane & 3) == 0 && q < NV) smem[wave * NV + q] = t;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = smem[threadIdx.x];
#pragma unroll
    for (int w = 1; w < BS / 64; ++w) s += smem[w * NV + threadIdx.x];
    out[threadIdx.x * stride] = s;
  }
}

// Gather the eight vertices of a cell relative to corner 0: r_0 = 0, r_k = (pos[n_k] -
// pos[n_0]) + sum_d shift_kd rvecs[d] with shift_kd = 1 when the cell is in the last layer
// along a periodic axis d and corner k lies at +d (mmff.py:347-371 with get_mic 258-286). The
// reference adds pos[n_0] back before taking edge differences; the energy depends on the edges
// only and the virial on r_v - r_0, and leaving the origin out avoids rounding the vertices at
// the magnitude of the absolute coordinates.
// All 24 loads are issued before any arithmetic, and the shifts are formed with factors 0 or 1
// (exact products) rather than data-dependent selects, so that no load ends up behind a branch
// and a full wait; the sums are those of the selects, in the same order.
// With kXmat every position is first taken through the pending transforms: xm2 (when not
// null), then xm (vecmat, the products the scaling pass stores, in its order).
template <bool kXmat = false>
__device__ __forceinline__ void gather_vertices(const double* __restrict__ pos,
                                                const double* __restrict__ rv, const int n[8],
                                                int wrap, double r[8][3],
                                                const double* __restrict__ xm = nullptr,
                                                const double* __restrict__ xm2 = nullptr) {
  double p[8][3];
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) p[k][j] = pos[3 * (int64_t)n[k] + j];
  if (kXmat && xm2) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double t[3];
      vecmat(p[k], xm2, t);
#pragma unroll
      for (int j = 0; j < 3; ++j) p[k][j] = t[j];
    }
  }
  if (kXmat) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double t[3];
      vecmat(p[k], xm, t);
#pragma unroll
      for (int j = 0; j < 3; ++j) p[k][j] = t[j];
    }
  }
  const double fx = (double)(wrap & 1), fy = (double)((wrap >> 1) & 1),
               fz = (double)((wrap >> 2) & 1);
  double sx[3], sy[3], sz[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    sx[j] = fx * rv[0 + j];
    sy[j] = fy * rv[3 + j];
    sz[j] = fz * rv[6 + j];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) r[0][j] = 0.0;
#pragma unroll
  for (int k = 1; k < 8; ++k) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double sh = ((kOff[k][0] ? sx[j] : 0.0) + (kOff[k][1] ? sy[j] : 0.0)) +
                        (kOff[k][2] ? sz[j] : 0.0);
      r[k][j] = (p[k][j] - p[0][j]) + sh;
    }
  }
}

// Corners of the lower (z = 0) and upper (z = 1) face of a cell, in face-class order: upper
// corner kUpperCorner[q] of a cell is lower corner kLowerCorner[q] of its +z neighbour.
__device__ constexpr int kLowerCorner[4] = {0, 1, 2, 4};
__device__ constexpr int kUpperCorner[4] = {3, 5, 6, 7};

// One lane per cell. Energies of all states of the block's cell type, combined by the
// log-sum-exp of mmff.py:386-398; the vertex gradient of the combined energy goes to gcell and
// the cell's energy and virial sum_v g_v (x) (r_v - r_0) enter the block partials. A block holds
// one cell type; single-state and multistate blocks run in separate instances (kMulti), each
// launched over its own range of blocks (blk0 + blockIdx.x) when the work order allows, so that
// the single-state instance does not carry the registers of the state combination.
// kVir: the virial as the sum of the corners' B^T S B (state_energy), which the vertex
// gradients give as sum_v g_v (x) r_v; with kGrad false the pass then forms no gradient at all
// (energy and virial only). The passes with gradients and the vertex virial, and those with the
// corner virial and no gradients, take the node positions through the pending transforms xmat
// (and xmat2 first, when its flag is set) when they are given: a barostat call whose position
// scaling is stored by a later node pass.
#define MM_CELL_PARAMS                                                                        \
  const double *__restrict__ pos, const double *__restrict__ rvecs, int64_t rvecs_stride,     \
      const int32_t *__restrict__ blk_sys, const int32_t *__restrict__ cnode,                 \
      const int32_t *__restrict__ cwrap, const int32_t *__restrict__ blk_row,                 \
      const double *__restrict__ tab_g, const double *__restrict__ kt_g,                      \
      const int32_t *__restrict__ nstates_g, int32_t max_states, int64_t W,                   \
      const int32_t *__restrict__ cbase, int64_t S, double *__restrict__ gcell,               \
      double *__restrict__ partials, int blk0, int nblocks, const double *__restrict__ xmat,  \
      const double *__restrict__ xmat2, const double *__restrict__ xmat2_on
#define MM_CELL_ARGS                                                                          \
  pos, rvecs, rvecs_stride, blk_sys, cnode, cwrap, blk_row, tab_g, kt_g, nstates_g,            \
      max_states, W, cbase, S, gcell, partials, blk0, nblocks, xmat, xmat2, xmat2_on

// The work of cell block bid (see cell_kernel).
template <int kModel, bool kGrad, bool kMulti, bool kVir>
__device__ __forceinline__ void cell_body(int bid, MM_CELL_PARAMS) {
  __shared__ double red[10 * (kCellBlock / 64)];
  __shared__ double es_sh[(kMulti && !kGrad && !kVir) ? kMaxStates * kCellBlock : 1];
  // running weighted gradient sum of the one-pass multistate gradient: [24][kCellBlock]
  __shared__ double gs_sh[(kMulti && kGrad) ? 24 * kCellBlock : 1];
  const int trow = blk_row[bid];
  // a block of the other kind (full-range launches only) leaves its partials to the other
  // instance
  if (kMulti != (nstates_g[trow] > 1)) return;
  const int ns = nstates_g[trow];
  const double kt = kt_g[trow];
  const double* tab = tab_g + (int64_t)trow * max_states * kTabStride;
  const double* rvb = rvecs + (blk_sys ? (int64_t)blk_sys[bid] * rvecs_stride : 0);
  double rv[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) rv[q] = rvb[q];
  constexpr bool kCanX = kVir != kGrad;
  const int64_t soff = blk_sys ? (int64_t)blk_sys[bid] * rvecs_stride : 0;
  const double* xm = (kCanX && xmat) ? xmat + soff : nullptr;
  const double* xm2 =
      (kCanX && xmat2 && xmat2_on[soff] != 0.0) ? xmat2 + soff : nullptr;

  const int64_t w = (int64_t)bid * kCellBlock + threadIdx.x;
  int n[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) n[k] = cnode[w * 8 + k];
  // sums start at -0.0, the identity of addition ((-0.0) + x == x for every x), so that the
  // first term of each needs no add
  double acc[10];
#pragma unroll
  for (int q = 0; q < 10; ++q) acc[q] = -0.0;
  double gacc[8][3];
#pragma unroll
  for (int v = 0; v < 8; ++v)
#pragma unroll
    for (int j = 0; j < 3; ++j) gacc[v][j] = -0.0;
  double vir[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) vir[q] = -0.0;
  const bool valid = n[0] >= 0;
  const int wrap = valid ? cwrap[w] : 0;
  if (valid) {
    double r[8][3];
    if (kCanX && xm)
      gather_vertices<true>(pos, rv, n, wrap, r, xm, xm2);
    else
      gather_vertices(pos, rv, n, wrap, r);
    double ecell;
    if (!kMulti) {
      ecell = tab[10] + state_energy<kModel, kGrad, true, kVir>(tab, r, gacc, 1.0, vir);
    } else if (!kGrad && !kVir) {
      // energies of all states, E = E_min - kT ln sum exp(-(E_s - E_min)/kT) (mmff.py:386-395);
      // the state energies are kept in LDS
      double emin = 0.0;
      for (int s = 0; s < ns; ++s) {
        const double* ts = tab + s * kTabStride;
        const double es = ts[10] + state_energy<kModel, false>(ts, r, gacc);
        es_sh[s * kCellBlock + threadIdx.x] = es;
        emin = (s == 0 || es < emin) ? es : emin;
      }
      double zsum = 0.0;
      for (int s = 0; s < ns; ++s) zsum += exp(-(es_sh[s * kCellBlock + threadIdx.x] - emin) / kt);
      ecell = emin - kt * log(zsum);
    } else {
      // one pass over the states: energy and gradient of each, and the gradient
      // sum_s p_s dE_s/dr with p_s = exp(-(E_s - E_min)/kT) / Z (mmff.py:386-398) accumulated
      // in LDS with weights relative to the lowest energy so far, rescaled when it moves
      // (kVir: the state virials combine the same way, in registers)
      double* gsum = gs_sh + threadIdx.x;
      double emin = 0.0, zsum = 0.0, vsum[6];
      for (int s = 0; s < ns; ++s) {
        const double* ts = tab + s * kTabStride;
        double gs[8][3], vs[6];
#pragma unroll
        for (int v = 0; v < 8; ++v)
#pragma unroll
          for (int j = 0; j < 3; ++j) gs[v][j] = -0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) vs[q] = -0.0;
        const double es = ts[10] + state_energy<kModel, kGrad, true, kVir>(ts, r, gs, 1.0, vs);
        if (s == 0) {
          emin = es;
          zsum = 1.0;
          if (kGrad) {
#pragma unroll
            for (int q = 0; q < 24; ++q) gsum[q * kCellBlock] = gs[q / 3][q % 3];
          }
#pragma unroll
          for (int q = 0; q < 6; ++q) vsum[q] = vs[q];
        } else if (es < emin) {
          const double c = exp(-(emin - es) / kt);
          zsum = zsum * c + 1.0;
          if (kGrad) {
#pragma unroll
            for (int q = 0; q < 24; ++q)
              gsum[q * kCellBlock] = gsum[q * kCellBlock] * c + gs[q / 3][q % 3];
          }
#pragma unroll
          for (int q = 0; q < 6; ++q) vsum[q] = vsum[q] * c + vs[q];
          emin = es;
        } else {
          const double c = exp(-(es - emin) / kt);
          zsum += c;
          if (kGrad) {
#pragma unroll
            for (int q = 0; q < 24; ++q) gsum[q * kCellBlock] += c * gs[q / 3][q % 3];
          }
#pragma unroll
          for (int q = 0; q < 6; ++q) vsum[q] += c * vs[q];
        }
      }
      ecell = emin - kt * log(zsum);
      const double inv = 1.0 / zsum;
      if (kGrad) {
#pragma unroll
        for (int q = 0; q < 24; ++q) gacc[q / 3][q % 3] = gsum[q * kCellBlock] * inv;
      }
      if (kVir) {
#pragma unroll
        for (int q = 0; q < 6; ++q) vir[q] = vsum[q] * inv;
      }
    }
    acc[0] = ecell;
    if (kVir) {
      // vtens[k][l] from the symmetric corner sums (00, 11, 22, 12, 02, 01)
      constexpr int kSym[9] = {0, 5, 4, 5, 1, 3, 4, 3, 2};
#pragma unroll
      for (int q = 0; q < 9; ++q) acc[1 + q] = vir[kSym[q]];
    } else if (kGrad) {
      // virial: vtens[k][l] = sum_v g_v[k] (r_v - r_0)[l]
#pragma unroll
      for (int v = 1; v < 8; ++v) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l) acc[1 + 3 * k + l] += gacc[v][k] * r[v][l];
      }
    }
  }
  if (kGrad) {
    // The upper face (corners 3, 5, 6, 7) of the slot below is added to this slot's lower face
    // (corners 0, 1, 2, 4) when that slot holds the -z neighbour (wrap bit 3); the lower faces
    // go to their class planes, the upper faces that were not folded (wrap bit 4 clear) to the
    // wave's compact slots, in lane order. The move from lane - 1 is a DPP wave shift (lane 0,
    // which never folds, keeps its own value).
    const bool fold = (wrap & 8) != 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double up = gacc[kUpperCorner[q]][j];
        const double t = dpp_f64<0x138, 0xF>(up, up);  // wave_shr:1
        const double lo = gacc[kLowerCorner[q]][j];
        gacc[kLowerCorner[q]][j] = fold ? lo + t : lo;
      }
    }
    if (valid) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          __builtin_nontemporal_store(gacc[kLowerCorner[q]][j], &gcell[(int64_t)(3 * q + j) * S + w]);
    }
    const bool open_top = valid && !(wrap & 16);
    const uint64_t tops = __ballot(open_top);
    if (open_top) {
      const int lane = threadIdx.x & 63;
      const int64_t ci = W + cbase[w >> 6] + __popcll(tops & ((1ull << lane) - 1ull));
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          __builtin_nontemporal_store(gacc[kUpperCorner[q]][j], &gcell[(int64_t)(3 * q + j) * S + ci]);
    }
  }
  block_reduce_store<10, kCellBlock>(acc, red, partials + bid, nblocks);
}

template <int kModel, bool kGrad, bool kMulti, bool kVir>
__global__ void __launch_bounds__(kCellBlock, (kVir && !kGrad && !kMulti) ? 3 : 2)
cell_kernel(MM_CELL_PARAMS) {
  cell_body<kModel, kGrad, kMulti, kVir>(blk0 + xcd_block((int)blockIdx.x, (int)gridDim.x),
                                         MM_CELL_ARGS);
}

// Both kinds of block in one launch, each block running the body of its own kind: for a pass
// that the device holds at once, where the two instances would otherwise run one after the
// other, each taking about one block's time.
template <int kModel, bool kGrad, bool kVir>
__global__ void __launch_bounds__(kCellBlock, 2) cell_kernel_both(MM_CELL_PARAMS) {
  const int bid = blk0 + xcd_block((int)blockIdx.x, (int)gridDim.x);
  if (nstates_g[blk_row[bid]] > 1)
    cell_body<kModel, kGrad, true, kVir>(bid, MM_CELL_ARGS);
  else
    cell_body<kModel, kGrad, false, kVir>(bid, MM_CELL_ARGS);
}
#undef MM_CELL_ARGS
#undef MM_CELL_PARAMS

// NaN-propagating maximum (the result is NaN when either operand is).
__device__ __forceinline__ double nan_max(double a, double b) {
  return (a > b || a != a) ? a : b;
}

// Block reduction of NS sums (fixed order) followed by NM NaN-propagating maxima; thread q
// writes total q to part[q * stride] (256-thread blocks).
template <int NS, int NM>
__device__ __forceinline__ void block_sum_max_store(double (&v)[NS + NM], double* part,
                                                    int64_t stride) {
  constexpr int NV = NS + NM;
  __shared__ double red[4 * NV];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const double o = __shfl_down(v[q], off, 64);
      v[q] = q < NS ? v[q] + o : nan_max(v[q], o);
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) red[wave * NV + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int q = threadIdx.x;
    double t = red[q];
    for (int w = 1; w < 4; ++w) t = q < NS ? t + red[w * NV + q] : nan_max(t, red[w * NV + q]);
    part[q * stride] = t;
  }
}

// gpos[n] (+)= sum_{k=0..7} gcell[k][:][w(n,k)] in fixed k order (mmff.py:303-318). A workgroup
// owns 256 consecutive nodes; their 768 gpos values (and, with DOT, dir and gold values) move
// between HBM and LDS in contiguous runs, and each thread takes its node's three from LDS. With
// DOT (a line-search trial), also the workgroup partials of sum_n gpos_n . dir_n,
// sum_n |gpos_n|^2, sum_n gpos_n . gold_n and max_n |gpos_n|^2 into dpart rows 0-3 (row stride =
// workgroups).
template <bool DOT, bool kAllLoads>
__global__ void __launch_bounds__(256)
gather_kernel(const int32_t* __restrict__ ncell, const double* __restrict__ gcell, int64_t N,
              int64_t RS, int rshift, double* __restrict__ gpos, int accumulate, int* __restrict__ nanflag,
              const double* __restrict__ dir, const double* __restrict__ gold,
              double* __restrict__ dpart) {
  __shared__ double sg[768];
  __shared__ double sd[DOT ? 768 : 1], so[DOT ? 768 : 1];
  const int t = threadIdx.x;
  // node blocks dealt to the XCDs in contiguous ranges: neighbouring blocks share record lines
  const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const int64_t n0 = (int64_t)blk * 256, n = n0 + t;
  const int64_t e0 = 3 * n0;
  const int64_t ne = 3 * (N < n0 + 256 ? N - n0 : (int64_t)256);  // values of this workgroup
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int li = r * 256 + t;
    if (li < ne) {
      if (accumulate) sg[li] = gpos[e0 + li];
      if (DOT) {
        sd[li] = dir[e0 + li];
        so[li] = gold[e0 + li];
      }
    }
  }
  double g[3] = {0.0, 0.0, 0.0};
  if (n < N) sum_node_records<kAllLoads>(ncell, gcell, N, RS, rshift, n, g);
  __syncthreads();
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (n < N) {
    bool bad = false;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      bad |= (g[j] != g[j]);
      const double v = accumulate ? sg[3 * t + j] + g[j] : g[j];
      sg[3 * t + j] = v;
      if (DOT) {
        acc[0] += v * sd[3 * t + j];
        acc[1] += v * v;
        acc[2] += v * so[3 * t + j];
      }
    }
    acc[3] = acc[1];
    if (bad) atomicOr(nanflag, 2);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int li = r * 256 + t;
    if (li < ne) gpos[e0 + li] = sg[li];
  }
  if (DOT) block_sum_max_store<3, 1>(acc, dpart + blk, gridDim.x);
}

// One block per item: item i reduces src[0..n) (sum in a fixed order, or a NaN-propagating
// maximum) into dst[0] (added to it when acc). A NaN total sets the item's bit in nanflag.
__global__ void __launch_bounds__(1024)
reduce_items_kernel(ReduceList list, int* __restrict__ nanflag) {
  constexpr int kUnroll = 8;
  __shared__ double red[16];
  const ReduceItem it = list.item[blockIdx.x];
  const bool is_max = it.op == kReduceMax;
  const double zero = is_max ? -HUGE_VAL : 0.0;
  double acc = zero;
  for (int64_t base = threadIdx.x; base < it.n; base += 1024 * kUnroll) {
    double v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + (int64_t)u * 1024;
      v[u] = i < it.n ? it.src[i] : zero;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) acc = is_max ? nan_max(acc, v[u]) : acc + v[u];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(acc, off, 64);
    acc = is_max ? nan_max(acc, o) : acc + o;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int w = 1; w < 16; ++w) t = is_max ? nan_max(t, red[w]) : t + red[w];
    if (it.acc) t += it.dst[0];
    it.dst[0] = t;
    if (t != t && it.nanbit) atomicOr(nanflag, it.nanbit);
  }
}

hipError_t launch_cells(const ForceArgs& a, hipStream_t stream) {
  const int nb = (int)(a.W / kCellBlock);
  const bool grad = a.gpos != nullptr || a.vtens != nullptr || a.force_grad;
  if (nb == 0) return hipSuccess;
  // Multistate blocks [0, split), single-state blocks [split, nb) when the work order is so
  // partitioned (split >= 0); otherwise both instances cover every block.
  const bool part = a.split >= 0;
  const int m0 = 0, mn = part ? a.split : nb;
  const int s0 = part ? a.split : 0, sn = part ? nb - a.split : nb;
  // The second instance runs on the side stream, concurrently with the first, when the pass
  // has more cell blocks than the device holds at once (256 CUs x 2 blocks of 256 lanes at two
  // waves a SIMD). A smaller pass takes the time of one block per instance, less than the fork
  // and join cost: it runs both kinds of block in one launch (cell_kernel_both). NVT-NHC / NPT
  // step at 32^3 (129 blocks), fork -> back to back on one stream -> one launch: 51.0 / 120.7
  // -> 47.8 / 109 -> 37.4 / 81.1 us (the last two with the gathers' all-loads form), at 48^3
  // (about 430 blocks) 59.2 / 138.7 -> 58.8 / 127.5 us, -> 53.8 / 117.5 us in one launch; at
  // 64^3 (about 1030 blocks) the fork is the faster (NVT 74.0 vs 76.4 us back to back), at
  // 96^3 too (NPT 329 vs 343 us). The results are the same bits in every form.
  const bool fork = a.side != nullptr && a.ev_fork != nullptr && a.ev_join != nullptr &&
                    mn > 0 && sn > 0 && nb > kForkMinBlocks;
  hipStream_t ms = fork ? a.side : stream;
  hipError_t e = hipSuccess;
  if (fork) {
    e = hipEventRecord(a.ev_fork, stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(a.side, a.ev_fork, 0);
    if (e != hipSuccess) return e;
  }
#define MM_LAUNCH_CELL(MODEL, GRAD, MULTI, S, B0, NBL)                                       \
  hipLaunchKernelGGL((cell_kernel<MODEL, GRAD, MULTI, VIR>), dim3(NBL), dim3(kCellBlock), 0, S, \
                     a.pos, a.rvecs, a.rvecs_stride, a.blk_sys, a.cnode, a.cwrap, a.blk_row, \
                     a.tab, a.kt, a.nstates, a.max_states, a.W, a.cbase, a.rstride, a.gcell,  \
                     a.partials, B0, nb, a.xmat, a.xmat2, a.xmat2_on)
  // corner_virial: the virial from the corners; without records the pass then forms no
  // gradient at all
  const bool vir = a.corner_virial != 0;
  const bool kgrad = grad && !(vir && !a.store_records);
  // pending position transforms: the gradient pass with the vertex virial or the corner-virial
  // pass without gradients, xmat with every use, xmat2 with its flag
  if ((a.xmat2 != nullptr) != (a.xmat2_on != nullptr)) return hipErrorInvalidValue;
  if (a.xmat2 != nullptr && a.xmat == nullptr) return hipErrorInvalidValue;
  if (a.xmat != nullptr && kgrad == vir) return hipErrorInvalidValue;
  // The instance that takes longer is launched first, on the caller's stream: it then starts
  // right behind the preceding kernel instead of after the side stream's. That is the one
  // with more state evaluations (a multistate block counted as two), and the multistate one
  // without gradients and with the corner virial, where the single-state instance runs three
  // waves a SIMD and the multistate one two.
  const bool multi_lead = (vir && !kgrad) || 2 * (int64_t)mn > (int64_t)sn;
  // a pass that the device holds at once, with blocks of both kinds: one launch
  const bool both = mn > 0 && sn > 0 && nb <= kForkMinBlocks;
#define MM_LAUNCH_BOTH(MODEL, GRAD)                                                          \
  hipLaunchKernelGGL((cell_kernel_both<MODEL, GRAD, VIR>), dim3(nb), dim3(kCellBlock), 0,      \
                     stream, a.pos, a.rvecs, a.rvecs_stride, a.blk_sys, a.cnode, a.cwrap,      \
                     a.blk_row, a.tab, a.kt, a.nstates, a.max_states, a.W, a.cbase, a.rstride, \
                     a.gcell, a.partials, 0, nb, a.xmat, a.xmat2, a.xmat2_on)
#define MM_LAUNCH_PAIR(MODEL, GRAD)                                                          \
  do {                                                                                       \
    if (both) {                                                                              \
      MM_LAUNCH_BOTH(MODEL, GRAD);                                                           \
    } else if (multi_lead) {                                                                 \
      if (mn > 0) MM_LAUNCH_CELL(MODEL, GRAD, true, stream, m0, mn);                        \
      if (sn > 0) MM_LAUNCH_CELL(MODEL, GRAD, false, ms, s0, sn);                           \
    } else {                                                                                 \
      if (sn > 0) MM_LAUNCH_CELL(MODEL, GRAD, false, stream, s0, sn);                       \
      if (mn > 0) MM_LAUNCH_CELL(MODEL, GRAD, true, ms, m0, mn);                            \
    }                                                                                        \
  } while (0)
#define MM_LAUNCH_MODEL(MODEL)                                                                \
  do {                                                                                        \
    if (kgrad) MM_LAUNCH_PAIR(MODEL, true); else MM_LAUNCH_PAIR(MODEL, false);               \
  } while (0)
  if (vir) {
    constexpr bool VIR = true;
    if (a.model == 0) MM_LAUNCH_MODEL(0); else MM_LAUNCH_MODEL(1);
  } else {
    constexpr bool VIR = false;
    if (a.model == 0) MM_LAUNCH_MODEL(0); else MM_LAUNCH_MODEL(1);
  }
#undef MM_LAUNCH_MODEL
#undef MM_LAUNCH_PAIR
#undef MM_LAUNCH_BOTH
#undef MM_LAUNCH_CELL
  e = hipGetLastError();
  if (e == hipSuccess && fork) {
    e = hipEventRecord(a.ev_join, a.side);
    if (e == hipSuccess) e = hipStreamWaitEvent(stream, a.ev_join, 0);
  }
  return e;
}

hipError_t launch_gather(const ForceArgs& a, hipStream_t stream) {
  if (a.gpos == nullptr || a.N <= 0 || a.W <= 0) return hipSuccess;
  const int gb = (int)((a.N + 255) / 256);
  if (gb <= kGatherAllLoadsMaxBlocks)
    hipLaunchKernelGGL((gather_kernel<false, true>), dim3(gb), dim3(256), 0, stream, a.ncell,
                       a.gcell, a.N, a.rstride, a.rshift, a.gpos, a.accumulate, a.nanflag,
                       nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL((gather_kernel<false, false>), dim3(gb), dim3(256), 0, stream, a.ncell,
                       a.gcell, a.N, a.rstride, a.rshift, a.gpos, a.accumulate, a.nanflag,
                       nullptr, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_gather_dot(const ForceArgs& a, const double* dir, const double* gold,
                             double* dpart, hipStream_t stream) {
  if (a.gpos == nullptr || a.N <= 0 || a.W <= 0) return hipSuccess;
  const int gb = (int)((a.N + 255) / 256);
  if (gb <= kGatherAllLoadsMaxBlocks)
    hipLaunchKernelGGL((gather_kernel<true, true>), dim3(gb), dim3(256), 0, stream, a.ncell,
                       a.gcell, a.N, a.rstride, a.rshift, a.gpos, a.accumulate, a.nanflag, dir,
                       gold, dpart);
  else
    hipLaunchKernelGGL((gather_kernel<true, false>), dim3(gb), dim3(256), 0, stream, a.ncell,
                       a.gcell, a.N, a.rstride, a.rshift, a.gpos, a.accumulate, a.nanflag, dir,
                       gold, dpart);
  return hipGetLastError();
}

hipError_t launch_reduce(const ReduceList& list, int* nanflag, hipStream_t stream) {
  if (list.count <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_items_kernel, dim3(list.count), dim3(1024), 0, stream, list,
                     nanflag);
  return hipGetLastError();
}

void add_force_totals(ReduceList& list, const ForceArgs& a, int nb) {
  if (a.energy) list.add(a.partials, nb, kReduceSum, a.energy, a.accumulate, 1);
  if (a.vtens)
    for (int q = 0; q < 9; ++q)
      list.add(a.partials + (int64_t)(1 + q) * nb, nb, kReduceSum, a.vtens + q, a.accumulate, 4);
}

hipError_t launch_force(const ForceArgs& a, hipStream_t stream) {
  const int nb = (int)(a.W / kCellBlock);
  hipError_t err = launch_cells(a, stream);
  if (err != hipSuccess) return err;
  err = launch_gather(a, stream);
  if (err != hipSuccess) return err;
  ReduceList list;
  add_force_totals(list, a, nb);
  return launch_reduce(list, a.nanflag, stream);
}

// Trial point of a line search over 256 nodes a workgroup: optionally d <- beta d - g first,
// then xa <- x + alpha d, in contiguous runs of the workgroup's 768 values; the squared
// displacements go through LDS to the node threads for the workgroup partials of
// sum_n |xa_n - x_n|^2 and max_n |xa_n - x_n|^2 (part rows 0-1, row stride = workgroups).
__global__ void __launch_bounds__(256)
cg_trial_kernel(const double* __restrict__ x, const double* __restrict__ g,
                double* __restrict__ d, double beta, int set_dir, double alpha,
                double* __restrict__ xa, int64_t N, double* __restrict__ part) {
  __shared__ double dx2[768];
  const int t = threadIdx.x;
  const int64_t n0 = (int64_t)blockIdx.x * 256;
  const int64_t e0 = 3 * n0;
  const int64_t ne = 3 * (N < n0 + 256 ? N - n0 : (int64_t)256);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int li = r * 256 + t;
    double v = 0.0;
    if (li < ne) {
      const int64_t i = e0 + li;
      double di;
      if (set_dir) {
        di = beta != 0.0 ? beta * d[i] - g[i] : -g[i];  // beta = 0: d is not read
        d[i] = di;
      } else {
        di = d[i];
      }
      const double xi = x[i];
      const double xn = xi + alpha * di;
      xa[i] = xn;
      const double dx = xn - xi;
      v = dx * dx;
    }
    dx2[li] = v;
  }
  __syncthreads();
  double acc[2];
  acc[0] = (dx2[3 * t] + dx2[3 * t + 1]) + dx2[3 * t + 2];
  acc[1] = acc[0];
  block_sum_max_store<1, 1>(acc, part + blockIdx.x, gridDim.x);
}

hipError_t launch_cg_trial(const double* x, const double* g, double* d, double beta, int set_dir,
                           double alpha, double* xa, int64_t N, double* part,
                           hipStream_t stream) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(cg_trial_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, stream,
                     x, g, d, beta, set_dir, alpha, xa, N, part);
  return hipGetLastError();
}

// Per-block partials over the nodes n of |g_n|^2, g_n . go_n, |x_n - xo_n|^2 (sums, rows 0-2)
// and of |g_n|^2, |x_n - xo_n|^2 (maxima, rows 3-4); part is [5][nb].
__global__ void __launch_bounds__(256)
cg_stats_kernel(const double* __restrict__ x, const double* __restrict__ xo,
                const double* __restrict__ g, const double* __restrict__ go, int64_t N,
                double* __restrict__ part) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (n < N) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double gj = g[3 * n + j], dj = x[3 * n + j] - xo[3 * n + j];
      v[0] += gj * gj;
      v[1] += gj * go[3 * n + j];
      v[2] += dj * dj;
    }
    v[3] = v[0];
    v[4] = v[2];
  }
  block_sum_max_store<3, 2>(v, part + blockIdx.x, gridDim.x);
}

hipError_t launch_cg_stats(const double* x, const double* xo, const double* g, const double* go,
                           int64_t N, double* part, double* out, int* nanflag,
                           hipStream_t stream) {
  if (N <= 0) return hipSuccess;
  const int nb = (int)((N + 255) / 256);
  hipLaunchKernelGGL(cg_stats_kernel, dim3(nb), dim3(256), 0, stream, x, xo, g, go, N, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  ReduceList list;
  for (int q = 0; q < 5; ++q)
    list.add(part + (int64_t)q * nb, nb, q < 3 ? kReduceSum : kReduceMax, out + q, 0, 0);
  return launch_reduce(list, nanflag, stream);
}

// Parameter table entries (row t, state s) from device arrays, one thread per entry, with the
// arithmetic of the host builder (Voigt input), so both give the same bits; entries of states
// beyond nstates[t] are zero. kt[t] = boltzmann * temp_eff[t].
__global__ void __launch_bounds__(64)
build_table_kernel(int corner8, int nrows, int max_states, const int32_t* __restrict__ nstates,
                   const double* __restrict__ h0, const double* __restrict__ cv,
                   const double* __restrict__ free_energy, const double* __restrict__ temp_eff,
                   double boltzmann, double* __restrict__ tab, double* __restrict__ kt) {
#pragma clang fp contract(off)
  const int ts = (int)blockIdx.x * 64 + threadIdx.x;
  if (ts >= nrows * max_states) return;
  const int t = ts / max_states, s = ts % max_states;
  double* e = tab + (int64_t)ts * kTabStride;
  if (s == 0) kt[t] = boltzmann * temp_eff[t];
  if (s >= nstates[t]) {
    for (int q = 0; q < kTabStride; ++q) e[q] = 0.0;
    return;
  }
  double inv[9];
  const double det = inverse3_cof(h0 + (int64_t)ts * 9, inv);
#pragma unroll
  for (int q = 0; q < 9; ++q) e[q] = inv[q];
  const double w = corner8 ? 0.125 * det : det;
  e[9] = w;
  e[10] = free_energy[ts];
  e[11] = 0.0;
#pragma unroll
  for (int q = 0; q < 36; ++q) {
    const double c = cv[(int64_t)ts * 36 + q];
    e[kCvOffset + q] = c;
    e[kCvwOffset + q] = w * c;
  }
}

hipError_t launch_build_table(int corner8, int nrows, int max_states, const int32_t* nstates,
                              const double* h0, const double* cv, const double* free_energy,
                              const double* temp_eff, double boltzmann, double* tab, double* kt,
                              hipStream_t stream) {
  const int n = nrows * max_states;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(build_table_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, corner8,
                     nrows, max_states, nstates, h0, cv, free_energy, temp_eff, boltzmann, tab,
                     kt);
  return hipGetLastError();
}

// Per-system totals of a batch: one block per system sums its own cell blocks in list order.
__global__ void __launch_bounds__(256)
batch_finalize_kernel(const double* __restrict__ partials, int nb,
                      const int32_t* __restrict__ sys_ptr, const int32_t* __restrict__ sys_idx,
                      double* __restrict__ energies, double* __restrict__ vtens,
                      int* __restrict__ nanflag) {
  __shared__ double red[10 * 4];
  const int s = blockIdx.x;
  const int b0 = sys_ptr[s], b1 = sys_ptr[s + 1];
  double acc[10];
#pragma unroll
  for (int q = 0; q < 10; ++q) acc[q] = 0.0;
  for (int i = b0 + threadIdx.x; i < b1; i += 256) {
    const int b = sys_idx[i];
#pragma unroll
    for (int q = 0; q < 10; ++q) acc[q] += partials[(int64_t)q * nb + b];
  }
  wave_reduce<10>(acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 10; ++q) red[wave * 10 + q] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int bad = 0;
    for (int q = 0; q < 10; ++q) {
      const double t = ((red[q] + red[10 + q]) + red[20 + q]) + red[30 + q];
      if (t != t) bad |= q == 0 ? 1 : 4;
      if (q == 0 && energies) energies[s] = t;
      if (q > 0 && vtens) vtens[9 * (int64_t)s + q - 1] = t;
    }
    if (bad) atomicOr(nanflag, bad);
  }
}

hipError_t launch_batch_finalize(const double* partials, int nb, const int32_t* sys_ptr,
                                 const int32_t* sys_idx, int nsys, double* energies,
                                 double* vtens, int* nanflag, hipStream_t stream) {
  if (nsys <= 0) return hipSuccess;
  hipLaunchKernelGGL(batch_finalize_kernel, dim3(nsys), dim3(256), 0, stream, partials, nb,
                     sys_ptr, sys_idx, energies, vtens, nanflag);
  return hipGetLastError();
}

}  // namespace mm
