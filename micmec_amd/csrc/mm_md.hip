// Device MD engine: velocity Verlet with Nose-Hoover chain or Langevin thermostat and MTK or
// Langevin barostat (FP64).
//
// Reference sequence (one step, micmec/sampling/verlet.py:140-166):
//   pre hooks  -> v += a dt/2 -> x += v dt -> force -> v += a dt/2 -> ekin -> post hooks
// With a thermostat and a barostat the hooks run as TBCombination (micmec/sampling/npt.py:99-148):
//   pre  = MTK.baro (npt.py:653-736) then NHC chain (nvt.py:410-451);
//   post = NHC chain then MTK.baro. Each MTK.baro call evaluates the force once more.
// Node-wise work runs in wide kernels; every scalar update (chain, barostat velocity, 3x3
// symmetric exponentials, energies, conserved-quantity tracking) runs in a small kernel on the
// scalar block S between the node kernels (its blocks reduce the partial sums, the last one
// runs the phase), so a whole step can be captured in a HIP graph. Velocity/position
// transforms produced by a scalar phase (thermostat factors, barostat matrices) are stored in
// S and applied by the next pass that reads the array; kinetic sums after a pending barostat
// matrix follow from the sums before it. The Langevin thermostat's noise is drawn inside the
// node passes that already touch the velocities (pre: the kick-drift pass, with a barostat
// the gather that is the kick-drift; post: the gather-kick pass), from a counter-based
// generator keyed by node and hook call, so it costs no extra pass over the nodes.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/micmec_amd.h"
#include "mm_device.h"
#include "mm_internal.h"
#include "mm_kernels.h"
#include "mm_rng.h"

namespace mm {

constexpr int kNodeBlock = 256;   // nodes per block (system boundaries are block aligned)
// What a node gather does with the face-class records (gather_kick_kernel).
constexpr int kGatherNone = 0, kGatherKeep = 1, kGatherStore = 2;
constexpr int kNodeThreads = 64;  // one wave per block, kNodeBlock / kNodeThreads nodes a lane
// node sums: [0] ekin [1..9] sum m v(x)v [10] sum g^2 [11] sum d^2 [12] nan
// [13] ekin before the Langevin post update [14] ekin after the Langevin pre update
constexpr int kNodePartial = 16;
constexpr int kNodeSums = 15;
constexpr int kScalarThreads = 256;
constexpr int kScalarSlices = 32;  // most blocks a scalar phase spreads its reduction over
constexpr int kSliceQ = 10 + kNodeSums;  // force and node sums per slice

enum NodeMode {
  kKickDrift = 0,
  kScalePos = 1,
  kVV = 3,
  kApplyV = 4,
  kRmsd = 5,
  kKickDriftLT = 6,  // kick-drift preceded by the Langevin thermostat's pre update
};

enum Phase {
  kPhInitForce = 0,      // epot (+ vtens) after an initial force evaluation
  kPhInitProps = 1,      // node sums -> ekin, svv; properties without advancing time
  kPhNveEnd = 2,         // force + node sums after the second kick; properties
  kPhNvtPre = 3,         // NHC pre
  kPhNvtEnd = 4,         // force + node sums; NHC post; properties
  kPhBaroPre = 5,        // TBC pre, barostat: v_g update, position/domain and velocity transforms
  kPhBaroEndPre = 7,     // force; kinetic sums after the pending velocity matrix; v_g; NHC pre
  kPhMainEndNpt = 8,     // force + node sums; NHC post; barostat transforms of the post call
  kPhBaroEndPost = 9,    // force (+ gather NaN flag); kinetic sums after the matrix; v_g; econs
  kPhRmsd = 10,          // rmsd_gpos / rmsd_delta
  kPhApplied = 11,       // pending transforms were applied
  kPhBaroEndPostPre = 12,  // kPhBaroEndPost, then the next step's kPhBaroPre (steps in a row)
  kPhNvtEndPre = 13,       // kPhNvtEnd, then the next step's kPhNvtPre (steps in a row)
};

// ----------------------------------------------------------------------------- node kernels

// Block totals of a node pass (one wave per block), written SoA: out[q * gridDim.x + blockIdx.x].
template <int NV>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double* part) {
  if (NV >= 4) {
    const int lane = threadIdx.x & 63;
    const double t = wave_reduce_scatter<NV>(v, lane);
    const int q = scatter_index(lane);
    if ((lane & 3) == 0 && q < NV) part[(int64_t)q * gridDim.x + blockIdx.x] = t;
    return;
  }
  wave_reduce<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) part[(int64_t)q * gridDim.x + blockIdx.x] = v[q];
  }
}

// The same for blocks of kNodeBlock threads (waves summed in order through LDS), into the
// partial slot of logical block blk.
template <int NV>
__device__ __forceinline__ void block_sum_store_wide(double (&v)[NV], double* part, int blk) {
  __shared__ double red[NV * (kNodeBlock / 64)];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  {
    const double t = wave_reduce_scatter<NV>(v, lane);
    const int q = scatter_index(lane);
    if ((lane & 3) == 0 && q < NV) red[wave * NV + q] = t;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = red[threadIdx.x];
    for (int w = 1; w < kNodeBlock / 64; ++w) s += red[w * NV + threadIdx.x];
    part[(int64_t)threadIdx.x * gridDim.x + blk] = s;
  }
}

// Kinetic-energy and m v (x) v sums of one node, in the reference's operation order:
// 0.5 * (v**2 * m) (verlet.py:169) and (v.T * m) . v (verlet.py:187-188, npt.py:661-663).
template <int NA>
__device__ __forceinline__ void node_vv(const double v[3], double m, double (&acc)[NA]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int j = 0; j < 3; ++j) acc[0] += 0.5 * ((v[j] * v[j]) * m);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[1 + 3 * i + j] += (v[i] * m) * v[j];
}

// LangevinThermostat.thermo (nvt.py:211-218): v = c1 v + sqrt((1 - c1^2) T k_B / m) xi with
// xi the standard normals of node n at hook call 2 counter + post.
__device__ __forceinline__ void langevin_thermo(const double* S, int64_t n, int post, double m,
                                                double v[3]) {
#pragma clang fp contract(off)
  const uint64_t ctr[4] = {(uint64_t)n, 2ull * (uint64_t)S[MM_S_COUNTER] + (uint64_t)post, 0ull,
                           0ull};
  double z[4];
  normals4(ctr, seed_from(S + MM_S_LT_SEED), kStreamThermostat, z);
  const double c1 = S[MM_S_LT_C1];
  const double c2 = sqrt(S[MM_S_LT_A] / m);
#pragma unroll
  for (int j = 0; j < 3; ++j) v[j] = c1 * v[j] + c2 * z[j];
}

__device__ __forceinline__ double node_ekin(const double v[3], double m) {
#pragma clang fp contract(off)
  double e = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) e += 0.5 * ((v[j] * v[j]) * m);
  return e;
}

// Pending velocity transforms, applied in the order the reference performs them:
// factor A (thermostat post of the previous half), matrix M (barostat), matrix M2 (the next
// barostat call, when M still waits), factor B (thermostat pre). A null pointer is a transform
// that is off.
__device__ __forceinline__ void apply_vtransform(const double* fa, const double* M,
                                                 const double* M2, const double* fb,
                                                 double v[3]) {
#pragma clang fp contract(off)
  if (fa) {
    const double f = *fa;
#pragma unroll
    for (int j = 0; j < 3; ++j) v[j] = v[j] * f;
  }
  if (M) {
    double t[3];
    vecmat(v, M, t);
#pragma unroll
    for (int j = 0; j < 3; ++j) v[j] = t[j];
  }
  if (M2) {
    double t[3];
    vecmat(v, M2, t);
#pragma unroll
    for (int j = 0; j < 3; ++j) v[j] = t[j];
  }
  if (fb) {
    const double f = *fb;
#pragma unroll
    for (int j = 0; j < 3; ++j) v[j] = v[j] * f;
  }
}

// The transforms as the scalar block holds them, read where they are applied.
__device__ __forceinline__ void apply_vtransform(const double* __restrict__ S, double v[3]) {
  apply_vtransform(S[MM_S_VSCALE_A_ON] != 0.0 ? S + MM_S_VSCALE_A : nullptr,
                   S[MM_S_VMAT_ON] != 0.0 ? S + MM_S_VMAT : nullptr,
                   S[MM_S_VMAT2_ON] != 0.0 ? S + MM_S_VMAT2 : nullptr,
                   S[MM_S_VSCALE_B_ON] != 0.0 ? S + MM_S_VSCALE_B : nullptr, v);
}

// The transforms read from the scalar block once, for a pass over several nodes a lane.
struct VTransform {
  bool a_on, m_on, m2_on, b_on;
  double fa, fb, m[9], m2[9];
};

__device__ __forceinline__ VTransform load_vtransform(const double* __restrict__ S) {
  VTransform T;
  T.a_on = S[MM_S_VSCALE_A_ON] != 0.0;
  T.m_on = S[MM_S_VMAT_ON] != 0.0;
  T.m2_on = S[MM_S_VMAT2_ON] != 0.0;
  T.b_on = S[MM_S_VSCALE_B_ON] != 0.0;
  T.fa = T.a_on ? S[MM_S_VSCALE_A] : 1.0;
  T.fb = T.b_on ? S[MM_S_VSCALE_B] : 1.0;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    T.m[q] = T.m_on ? S[MM_S_VMAT + q] : 0.0;
    T.m2[q] = T.m2_on ? S[MM_S_VMAT2 + q] : 0.0;
  }
  return T;
}

__device__ __forceinline__ void apply_vtransform(const VTransform& T, double v[3]) {
  apply_vtransform(T.a_on ? &T.fa : nullptr, T.m_on ? T.m : nullptr, T.m2_on ? T.m2 : nullptr,
                   T.b_on ? &T.fb : nullptr, v);
}

template <int kMode>
__global__ void __launch_bounds__(kNodeThreads)
node_kernel(double* __restrict__ x, double* __restrict__ v, const double* __restrict__ g,
            const double* __restrict__ m, const double* __restrict__ xold,
            const double* __restrict__ S0, const int32_t* __restrict__ blk_sys, int64_t N,
            double* __restrict__ partials) {
#pragma clang fp contract(off)
  const double* S = S0 + (blk_sys ? (int64_t)blk_sys[blockIdx.x] * MM_S_COUNT : 0);
  double acc[13];
#pragma unroll
  for (int q = 0; q < 13; ++q) acc[q] = 0.0;
  double lt[1] = {0.0};
  constexpr int kIt = kNodeBlock / kNodeThreads;
  const int64_t n0 = (int64_t)blockIdx.x * kNodeBlock + threadIdx.x;
  if (kMode == kKickDrift || kMode == kKickDriftLT) {
    // every load of the lane's nodes first, then the arithmetic and the stores, so that the
    // loads are in flight together (a pass over a small system is then one memory round trip
    // instead of two a node)
    const double dt = S[MM_S_TIMESTEP];
    const VTransform T = load_vtransform(S);
    double mass[kIt], vv[kIt][3], gg[kIt][3], xx[kIt][3];
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int64_t n = n0 + it * kNodeThreads;
      mass[it] = 1.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) vv[it][j] = gg[it][j] = xx[it][j] = 0.0;
      if (n < N) {
        mass[it] = m[n];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          vv[it][j] = v[3 * n + j];
          gg[it][j] = g[3 * n + j];
          xx[it][j] = x[3 * n + j];
        }
      }
    }
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int64_t n = n0 + it * kNodeThreads;
      if (n >= N) break;
      apply_vtransform(T, vv[it]);
      if (kMode == kKickDriftLT) {
        langevin_thermo(S, n, 0, mass[it], vv[it]);
        lt[0] += node_ekin(vv[it], mass[it]);
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double a = -gg[it][j] / mass[it];
        vv[it][j] = vv[it][j] + (0.5 * a) * dt;
        v[3 * n + j] = vv[it][j];
        x[3 * n + j] = xx[it][j] + dt * vv[it][j];
      }
    }
  }
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int64_t n = n0 + it * kNodeThreads;
    if (n >= N) break;
    if (kMode == kKickDrift || kMode == kKickDriftLT) {
      // done above
    } else if (kMode == kScalePos) {
      double p[3], t[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) p[j] = x[3 * n + j];
      // a pending earlier transform first (the previous post call's, of steps in a row)
      if (S[MM_S_XMAT2_ON] != 0.0) {
        vecmat(p, S + MM_S_XMAT2, t);
#pragma unroll
        for (int j = 0; j < 3; ++j) p[j] = t[j];
      }
      vecmat(p, S + MM_S_XMAT, t);
#pragma unroll
      for (int j = 0; j < 3; ++j) x[3 * n + j] = t[j];
    } else if (kMode == kVV || kMode == kApplyV) {
      double vv[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) vv[j] = v[3 * n + j];
      if (kMode != kVV) {
        apply_vtransform(S, vv);
#pragma unroll
        for (int j = 0; j < 3; ++j) v[3 * n + j] = vv[j];
      }
      if (kMode != kApplyV) node_vv(vv, m[n], acc);
    } else if (kMode == kRmsd) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double gj = g[3 * n + j];
        const double d = x[3 * n + j] - xold[3 * n + j];
        acc[10] += gj * gj;
        acc[11] += d * d;
      }
    }
  }
  if (kMode == kVV || kMode == kRmsd) {
    block_sum_store<13>(acc, partials);
  }
  if (kMode == kKickDriftLT) block_sum_store<1>(lt, partials + (int64_t)14 * gridDim.x);
}

// gpos[n] = sum of the node's face-class records (mmff.py:303-318; the cell pass has already
// added the faces shared along z), fused with the second half kick
// v += a dt / 2 (verlet.py:152-153) and the node sums of the new velocities; with kLT also
// the Langevin thermostat's post update (the node sums then hold the ekin before it in [13]).
// With kScaleV instead (a barostat call, npt.py:704-735), the pending velocity transforms are
// applied (the barostat's exp(-(v_g + tr(v_g)/ndof) dt / 2), known before its force) in the
// same pass; the phase after it takes the kinetic sums from those before the matrix, so this
// pass leaves only the NaN flag of the gradient.
// gmode: kGatherStore sums the records and stores gpos; kGatherKeep sums them for the kick only
// (the step overwrites gpos before anything reads it); kGatherNone (kScaleV only) reads no
// records: the force evaluation fed only the energy and the virial.
// With kDrift instead, the first half kick and the drift of the step (verlet.py:141-146) with
// the barostat's pre-call gradient, which nothing else reads: the pending velocity transforms
// (and the Langevin thermostat's pre update, its ekin after the update in partial row 14),
// then v += a dt / 2 and x += dt v, as the node pass kKickDrift does from a stored gpos, with
// x first taken through the barostat call's pending position scaling (the previous post
// call's when it still waits, then the pre call's), which its cell pass applied on the fly.
// One node per thread (the gathers need every wave in flight that they can get); the node
// blocks are dealt to the XCDs in contiguous ranges (xcd_block), and the partial sums are
// stored per logical block.
template <bool kKick, bool kLT, bool kScaleV, bool kDrift, bool kAllLoads>
__global__ void __launch_bounds__(kNodeBlock)
gather_kick_kernel(const int32_t* __restrict__ ncell, const double* __restrict__ gcell,
                   int64_t N, int64_t RS, double* __restrict__ gpos, double* __restrict__ v,
                   const double* __restrict__ m, const double* __restrict__ S0,
                   const int32_t* __restrict__ blk_sys, double* __restrict__ partials,
                   int gmode, double* __restrict__ x = nullptr) {
  const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const double* S = S0 + (blk_sys ? (int64_t)blk_sys[blk] * MM_S_COUNT : 0);
  constexpr int NA = kLT ? 14 : 13;
  double acc[NA];
#pragma unroll
  for (int q = 0; q < NA; ++q) acc[q] = 0.0;
  {
    const int64_t n = (int64_t)blk * kNodeBlock + threadIdx.x;
    if (n < N) {
    // the node's own values: with kAllLoads loaded before the records, so that they arrive
    // with them
    double mass = 1.0, v0[3] = {0.0, 0.0, 0.0}, x0[3] = {0.0, 0.0, 0.0};
    auto load_node = [&]() {
      if (kKick || kDrift) mass = m[n];
      if (kKick || kDrift || kScaleV) {
#pragma unroll
        for (int j = 0; j < 3; ++j) v0[j] = v[3 * n + j];
      }
      if (kDrift) {
#pragma unroll
        for (int j = 0; j < 3; ++j) x0[j] = x[3 * n + j];
      }
    };
    if (kAllLoads) load_node();
    double g[3] = {0.0, 0.0, 0.0};
    if (gmode != kGatherNone) {
      sum_node_records<kAllLoads>(ncell, gcell, N, RS, 2, n, g);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        if (gmode == kGatherStore) gpos[3 * n + j] = g[j];
        if (g[j] != g[j]) acc[12] = 1.0;
      }
    }
    if (!kAllLoads) load_node();
    if (kKick) {
#pragma clang fp contract(off)
      const double dt = S[MM_S_TIMESTEP];
      double vv[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double a = -g[j] / mass;
        vv[j] = v0[j] + (0.5 * a) * dt;
      }
      if (kLT) {
        acc[NA - 1] += node_ekin(vv, mass);
        langevin_thermo(S, n, 1, mass, vv);
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) v[3 * n + j] = vv[j];
      node_vv(vv, mass, acc);
    }
    if (kScaleV) {
      double vv[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) vv[j] = v0[j];
      apply_vtransform(S, vv);
#pragma unroll
      for (int j = 0; j < 3; ++j) v[3 * n + j] = vv[j];
    }
    if (kDrift) {
#pragma clang fp contract(off)
      const double dt = S[MM_S_TIMESTEP];
      double vv[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) vv[j] = v0[j];
      apply_vtransform(S, vv);
      if (kLT) {
        langevin_thermo(S, n, 0, mass, vv);
        acc[0] += node_ekin(vv, mass);
      }
      // the positions as the cell pass read them: the pending scaling(s) of the barostat call
      double p[3], t[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) p[j] = x0[j];
      if (S[MM_S_XMAT2_ON] != 0.0) {
        vecmat(p, S + MM_S_XMAT2, t);
#pragma unroll
        for (int j = 0; j < 3; ++j) p[j] = t[j];
      }
      vecmat(p, S + MM_S_XMAT, t);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double a = -g[j] / mass;
        vv[j] = vv[j] + (0.5 * a) * dt;
        v[3 * n + j] = vv[j];
        x[3 * n + j] = t[j] + dt * vv[j];
      }
    }
    }
  }
  if (kDrift) {
    if (kLT) {
      double lt[1] = {acc[0]};
      block_sum_store_wide<1>(lt, partials + (int64_t)14 * gridDim.x, blk);
    }
  } else if (kKick) {
    block_sum_store_wide<NA>(acc, partials, blk);
  } else {
    // only the NaN flag: any lane's flag, per block
    const unsigned long long bad = __ballot(acc[12] != 0.0);
    __shared__ int any_bad;
    if (threadIdx.x == 0) any_bad = 0;
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && bad) any_bad = 1;
    __syncthreads();
    if (threadIdx.x < 13)
      partials[(int64_t)threadIdx.x * gridDim.x + blk] =
          (threadIdx.x == 12 && any_bad) ? 1.0 : 0.0;
  }
}

// ----------------------------------------------------------------------------- scalar math

__device__ void matmul3(const double* A, const double* B, double* C) {
#pragma clang fp contract(off)
  double T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      T[3 * i + j] = (A[3 * i + 0] * B[0 + j] + A[3 * i + 1] * B[3 + j]) + A[3 * i + 2] * B[6 + j];
  for (int i = 0; i < 9; ++i) C[i] = T[i];
}

// 1 / k for the Horner steps of expm_sym_half (correctly rounded, as the division gives).
__device__ constexpr double kInvK[15] = {0.0,       1.0,        1.0 / 2,  1.0 / 3,  1.0 / 4,
                                         1.0 / 5,   1.0 / 6,    1.0 / 7,  1.0 / 8,  1.0 / 9,
                                         1.0 / 10,  1.0 / 11,   1.0 / 12, 1.0 / 13, 1.0 / 14};

// R = exp(sign * A * dt / 2) for symmetric A: the matrix the reference forms from
// eigh(A) = (w, Q) as Q diag(exp(sign w dt / 2)) Q^T (npt.py:686-690, 709-726), here by scaling
// and squaring. X = sign * A * dt / 2 is halved s times until r = ||X||_inf <= 1/2, exp(X) is
// the Taylor polynomial in Horner form to the smallest degree m <= 14 whose first omitted term
// bound r^(m+1) / (m+1)! is below 1e-17 (the remainder is at most that times e^r against
// ||exp(X)|| >= e^-r: below 3e-17 relative, the bound of the full degree 14 at r = 1/2; a
// barostat velocity times dt / 2 is of order 1e-5, where m = 3), and the result is squared s
// times. Multiply-adds only: no square roots or divisions on the serial path.
__device__ void expm_sym_half(const double* A, double sign, double dt, double* R) {
  const double h = (sign * dt) / 2.0;
  double X[9];
  double nrm = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double r = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      X[3 * i + j] = h * A[3 * i + j];
      r += fabs(X[3 * i + j]);
    }
    nrm = fmax(nrm, r);
  }
  int s = 0;
  double scale = 1.0;
  while (nrm * scale > 0.5 && s < 64) {
    scale *= 0.5;
    ++s;
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) X[q] *= scale;
  const double r = nrm * scale;
  int m = 1;
  double bound = (r * r) / 2.0;
  while (m < 14 && bound > 1e-17) {
    ++m;
    bound *= r / (double)(m + 1);
  }
  double P[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) P[q] = (q % 4 == 0) ? 1.0 : 0.0;
  for (int k = m; k >= 1; --k) {
    // P <- I + X P / k
    const double inv = kInvK[k];
    double T[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        T[3 * i + j] = X[3 * i + 0] * P[0 + j] + X[3 * i + 1] * P[3 + j] + X[3 * i + 2] * P[6 + j];
#pragma unroll
    for (int q = 0; q < 9; ++q) P[q] = T[q] * inv + ((q % 4 == 0) ? 1.0 : 0.0);
  }
  for (int i = 0; i < s; ++i) matmul3(P, P, P);
#pragma unroll
  for (int q = 0; q < 9; ++q) R[q] = P[q];
}

// |det| of the domain vectors, as micmec/pes/domain.c:42-48.
__device__ double volume3(const double* r) {
#pragma clang fp contract(off)
  return fabs(r[0] * (r[4] * r[8] - r[5] * r[7]) + r[1] * (r[5] * r[6] - r[3] * r[8]) +
              r[2] * (r[3] * r[7] - r[4] * r[6]));
}

// Where a Nose-Hoover chain keeps its state in the scalar block: the particle thermostat's
// chain, or the MTK barostat's own chain (baro_thermo), which acts on the barostat alone.
struct ChainSlots {
  int len, temp, ndof, mass, pos, vel;
};
constexpr ChainSlots kThermoChain{MM_S_CHAIN_LEN, MM_S_THERMO_TEMP, MM_S_CHAIN_NDOF,
                                  MM_S_CHAIN_MASS, MM_S_CHAIN_POS, MM_S_CHAIN_VEL};
constexpr ChainSlots kBaroChain{MM_S_BCHAIN_LEN, MM_S_BCHAIN_TEMP, MM_S_BCHAIN_NDOF,
                                MM_S_BCHAIN_MASS, MM_S_BCHAIN_POS, MM_S_BCHAIN_VEL};

// NHChain.__call__ (nvt.py:410-451): returns the velocity factor; updates ekin in place.
__device__ double nhc_chain(double* S, double* ekin_io, bool has_g1, double g1,
                            const ChainSlots& c = kThermoChain) {
#pragma clang fp contract(off)
  const int L = (int)S[c.len];
  const double dt = S[MM_S_TIMESTEP];
  const double temp = S[c.temp];
  const double kb = S[MM_S_BOLTZMANN];
  const double ndof = S[c.ndof];
  double* mass = S + c.mass;
  double* pos = S + c.pos;
  double* vel = S + c.vel;
  auto bead = [&](int k, double ekin) {
    double g;
    if (k == 0) {
      g = 2.0 * ekin - (ndof * temp) * kb;
      if (has_g1) g += g1;
    } else {
      g = mass[k - 1] * (vel[k - 1] * vel[k - 1]) - temp * kb;
    }
    g /= mass[k];
    if (k == L - 1) {
      vel[k] += (g * dt) / 4.0;
    } else {
      // the same factor before and after the force term (vel[k + 1] does not change between)
      const double e = exp(((-vel[k + 1]) * dt) / 8.0);
      vel[k] *= e;
      vel[k] += (g * dt) / 4.0;
      vel[k] *= e;
    }
  };
  double ekin = *ekin_io;
  for (int k = L - 1; k >= 0; --k) bead(k, ekin);
  for (int k = 0; k < L; ++k) pos[k] += (vel[k] * dt) / 2.0;
  const double factor = exp(((-vel[0]) * dt) / 2.0);
  ekin *= factor * factor;
  for (int k = 0; k < L; ++k) bead(k, ekin);
  *ekin_io = ekin;
  return factor;
}

// NHChain.get_econs_correction (nvt.py:453-458)
__device__ double nhc_econs(const double* S, const ChainSlots& c = kThermoChain) {
#pragma clang fp contract(off)
  const int L = (int)S[c.len];
  const double kt = S[MM_S_BOLTZMANN] * S[c.temp];
  double kin = 0.0, rest = 0.0;
  for (int k = 0; k < L; ++k) kin += (S[c.vel + k] * S[c.vel + k]) * S[c.mass + k];
  for (int k = 1; k < L; ++k) rest += S[c.pos + k];
  return 0.5 * kin + kt * (S[c.ndof] * S[c.pos] + rest);
}

__device__ __forceinline__ bool baro_chain_on(const double* S) {
  return S[MM_S_BCHAIN_ON] != 0.0;
}

// MTKBarostat._compute_ekin_baro (npt.py:748-757): W tr(v_g^T v_g) / 2, or W v_g^2 / 2 in the
// isotropic form.
__device__ double baro_ekin(const double* S) {
#pragma clang fp contract(off)
  if (S[MM_S_BARO_ANISO] == 0.0)
    return (0.5 * S[MM_S_MASS_PRESS]) * (S[MM_S_VEL_PRESS] * S[MM_S_VEL_PRESS]);
  double tr = 0.0;
  for (int i = 0; i < 9; ++i) tr += S[MM_S_VEL_PRESS + i] * S[MM_S_VEL_PRESS + i];
  return (0.5 * S[MM_S_MASS_PRESS]) * tr;
}

// MTKBarostat.add_press_cont (npt.py:738-746); with its own chain the barostat is decoupled
// from the particle thermostat and contributes 0.
__device__ double add_press_cont(const double* S) {
#pragma clang fp contract(off)
  if (baro_chain_on(S)) return 0.0;
  return 2.0 * baro_ekin(S) - S[MM_S_BARO_NDOF] * (S[MM_S_BARO_TEMP] * S[MM_S_BOLTZMANN]);
}

// The barostat chain's half step on the barostat kinetic energy (npt.py:625-629, 634-638): the
// chain advances, the velocity factor it returns is discarded (v_g is damped inside
// update_baro_vel through chainvel0).
__device__ void baro_chain_step(double* S) {
  double ekin = baro_ekin(S);
  (void)nhc_chain(S, &ekin, true, 0.0, kBaroChain);
}

__device__ __forceinline__ bool baro_langevin(const double* S) {
  return S[MM_S_BARO] == (double)MM_BARO_LANGEVIN;
}
__device__ __forceinline__ bool baro_aniso(const double* S) {
  return S[MM_S_BARO_ANISO] != 0.0;
}
__device__ __forceinline__ bool thermo_nhc(const double* S) {
  return S[MM_S_THERMO] == (double)MM_THERMO_NHC;
}

// LangevinBarostat.getR (npt.py:489-505): a symmetric matrix whose lower triangle is a draw of
// N(0, 1) x RSCALE, here the normals of system sys at hook call 2 counter + post, update sub.
__device__ void langevin_random_force(const double* S, int sys, int post, int sub, double R[9]) {
#pragma clang fp contract(off)
  const uint64_t call = 2ull * (uint64_t)S[MM_S_COUNTER] + (uint64_t)post;
  const uint64_t key = seed_from(S + MM_S_LB_SEED);
  double z[8];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const uint64_t ctr[4] = {(uint64_t)sys, call, (uint64_t)sub, (uint64_t)b};
    normals4(ctr, key, kStreamBarostat, z + 4 * b);
  }
  const double sc = S[MM_S_LB_RSCALE];
  int k = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      const double r = z[k++] * sc;
      R[3 * i + j] = r;
      R[3 * j + i] = r;
    }
  }
}

// update_baro_vel of MTKBarostat.baro (npt.py:654-681) and of LangevinBarostat.baro
// (npt.py:394-432: friction exp(-dt / 8 tau) on both sides, random force R, optional volume
// constraint and isotropic form). sub is 0 for the first and 1 for the second update of a call.
__device__ __forceinline__ void update_baro_vel(double* S, int sys, int post, int sub) {
#pragma clang fp contract(off)
  const double dt = S[MM_S_TIMESTEP];
  double* vp = S + MM_S_VEL_PRESS;
  const bool lb = baro_langevin(S);
  const bool aniso = baro_aniso(S);
  const int nv = aniso ? 9 : 1;
  const bool has = S[MM_S_HAS_CHAINVEL0] != 0.0;
  const double cv0 = S[MM_S_CHAINVEL0];
  const double fric = lb ? exp((-dt) / (8.0 * S[MM_S_LB_TIMECON])) : 1.0;
  // the chain's damping factor, applied before and after the force term
  const double fchain = has ? exp(((-dt) * cv0) / 8.0) : 1.0;
  if (lb)
    for (int i = 0; i < nv; ++i) vp[i] *= fric;
  if (has) {
    for (int i = 0; i < nv; ++i) vp[i] *= fchain;
  }
  double P[9], G[9];
  for (int i = 0; i < 9; ++i) P[i] = S[MM_S_SVV + i] - S[MM_S_VTENS + i];
  const double diag = (2.0 * S[MM_S_EKIN]) / S[MM_S_NDOF] - S[MM_S_BARO_PRESS] * S[MM_S_VOLUME];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      double p = 0.5 * (P[3 * j + i] + P[3 * i + j]);
      if (i == j) p = p + diag;
      G[3 * i + j] = p / S[MM_S_MASS_PRESS];
    }
  }
  if (!lb) {
    // MTK (npt.py:671-676): the isotropic form takes the trace, the volume constraint removes
    // it (the two together are refused at construction, utils.py:498-500)
    if (!aniso) {
      vp[0] += ((((G[0] + G[4]) + G[8])) * dt) / 4.0;
    } else {
      if (S[MM_S_BARO_VOLCON] != 0.0) {
        const double tg = ((G[0] + G[4]) + G[8]) / S[MM_S_NVEC];
        for (int i = 0; i < 9; i += 4) G[i] -= tg;
      }
      for (int i = 0; i < 9; ++i) vp[i] += (G[i] * dt) / 4.0;
    }
  } else {
    const double W = S[MM_S_MASS_PRESS];
    double R[9];
    langevin_random_force(S, sys, post, sub, R);
    if (S[MM_S_BARO_VOLCON] != 0.0) {
      const double dim = S[MM_S_NVEC];
      const double tg = ((G[0] + G[4]) + G[8]) / dim;
      const double tr = ((R[0] + R[4]) + R[8]) / dim;
      for (int i = 0; i < 9; i += 4) {
        G[i] -= tg;
        R[i] -= tr;
      }
    }
    if (aniso) {
      for (int i = 0; i < 9; ++i) vp[i] += ((G[i] - R[i] / W) * dt) / 4.0;
    } else {
      const double g = (G[0] + G[4]) + G[8];
      vp[0] += ((g - R[0] / W) * dt) / 4.0;
    }
    for (int i = 0; i < nv; ++i) vp[i] *= fric;
  }
  if (has) {
    for (int i = 0; i < nv; ++i) vp[i] *= fchain;
  }
}

// First half of the barostat call: v_g update, then positions and domain scaled by
// exp(v_g dt / 2) (npt.py:683-702, 434-452). The position transform is left in XMAT.
__device__ __forceinline__ void baro_part1(double* S, int sys, int post) {
#pragma clang fp contract(off)
  update_baro_vel(S, sys, post, 0);
  double R[9];
  if (baro_aniso(S)) {
    expm_sym_half(S + MM_S_VEL_PRESS, 1.0, S[MM_S_TIMESTEP], R);
  } else {
    const double c = exp((S[MM_S_VEL_PRESS] * S[MM_S_TIMESTEP]) / 2.0);
    for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? c : 0.0;
  }
  for (int i = 0; i < 9; ++i) S[MM_S_XMAT + i] = R[i];
  matmul3(S + MM_S_RVECS, R, S + MM_S_RVECS);
  S[MM_S_VOLUME] = volume3(S + MM_S_RVECS);
}

// Velocity transform of the barostat call: exp(-(v_g + tr(v_g)/ndof) dt / 2) (npt.py:709-726;
// without the trace term under a volume constraint, and exp(-(1 + 3/ndof) v_g dt / 2) in the
// isotropic form, npt.py:466-484), pending after any pending velocity factor and matrix (into
// the second matrix slot when the previous call's matrix still waits).
__device__ __forceinline__ void baro_part2(double* S) {
#pragma clang fp contract(off)
  const double* vp = S + MM_S_VEL_PRESS;
  const bool second = S[MM_S_VMAT_ON] != 0.0;
  double* dst = S + (second ? MM_S_VMAT2 : MM_S_VMAT);
  if (baro_aniso(S)) {
    double A[9];
    for (int i = 0; i < 9; ++i) A[i] = vp[i];
    if (S[MM_S_BARO_VOLCON] == 0.0) {
      const double tr = (vp[0] + vp[4]) + vp[8];
      const double sh = tr / S[MM_S_NDOF];
      A[0] += sh;
      A[4] += sh;
      A[8] += sh;
    }
    expm_sym_half(A, -1.0, S[MM_S_TIMESTEP], dst);
  } else {
    const double f = exp(((-((1.0 + 3.0 / S[MM_S_NDOF]) * vp[0])) * S[MM_S_TIMESTEP]) / 2.0);
    for (int i = 0; i < 9; ++i) dst[i] = (i % 4 == 0) ? f : 0.0;
  }
  S[second ? MM_S_VMAT2_ON : MM_S_VMAT_ON] = 1.0;
}

// LangevinBarostat.pre/post bookkeeping (npt.py:374-392): energies at the start of the call.
__device__ void baro_call_begin(double* S) {
  if (baro_langevin(S)) {
    S[MM_S_BARO_EPOT0] = S[MM_S_EPOT];
    S[MM_S_BARO_EKIN0] = S[MM_S_EKIN];
  }
}

// ... and the econs correction at its end: += epot0 - epot1 + ekin0 - ekin1.
__device__ void baro_call_end(double* S) {
#pragma clang fp contract(off)
  if (baro_langevin(S))
    S[MM_S_ECONS_BARO] += ((S[MM_S_BARO_EPOT0] - S[MM_S_EPOT]) + S[MM_S_BARO_EKIN0]) - S[MM_S_EKIN];
}

// LangevinThermostat.pre/post bookkeeping (nvt.py:197-209) for the two updates of a step made
// in the node passes: econs += ekin0 - ekin1 for each.
__device__ void langevin_thermo_econs(double* S, double ekin_before_pre, const double (&nd)[kNodeSums]) {
#pragma clang fp contract(off)
  S[MM_S_ECONS_THERMO] += ekin_before_pre - nd[14];
  S[MM_S_ECONS_THERMO] += nd[13] - S[MM_S_EKIN];
}

__device__ void clear_vtransforms(double* S) {
  S[MM_S_VSCALE_A_ON] = 0.0;
  S[MM_S_VMAT_ON] = 0.0;
  S[MM_S_VSCALE_B_ON] = 0.0;
  // the second matrix slot is used only by steps in a row; clearing its entries as well keeps
  // the scalar block the same as after steps taken one per call
  S[MM_S_VMAT2_ON] = 0.0;
  for (int i = 0; i < 9; ++i) S[MM_S_VMAT2 + i] = 0.0;
}

__device__ void set_vscale(double* S, int slot, double f) {
  S[slot] = f;
  S[slot + 1] = 1.0;
}

// Hook econs corrections as summed by compute_properties (verlet.py:176-181, npt.py:135-148).
__device__ double econs_correction(const double* S) {
#pragma clang fp contract(off)
  const bool th = S[MM_S_THERMO] != 0.0, ba = S[MM_S_BARO] != 0.0;
  if (th && ba) {
    const double corr = S[MM_S_ECONS_THERMO] + S[MM_S_ECONS_BARO];
    if (!thermo_nhc(S) || baro_chain_on(S)) return corr;
    // particle NHC thermostat also thermostats the barostat (npt.py:139-148)
    const double kt = S[MM_S_BOLTZMANN] * S[MM_S_THERMO_TEMP];
    return corr + S[MM_S_BARO_NDOF] * kt * S[MM_S_CHAIN_POS];
  }
  if (th) return S[MM_S_ECONS_THERMO];
  if (ba) return S[MM_S_ECONS_BARO];
  return 0.0;
}

// VerletIntegrator.compute_properties (verlet.py:171-190) minus the rmsd terms.
__device__ void properties(double* S, bool advance) {
#pragma clang fp contract(off)
  if (advance) {
    S[MM_S_TIME] += S[MM_S_TIMESTEP];
    S[MM_S_COUNTER] += 1.0;
  }
  const double ekin = S[MM_S_EKIN];
  S[MM_S_TEMP] = (ekin / S[MM_S_NDOF]) * (2.0 / S[MM_S_BOLTZMANN]);
  S[MM_S_ETOT] = ekin + S[MM_S_EPOT];
  S[MM_S_ECONS_CORR] = econs_correction(S);
  const double econs = S[MM_S_ETOT] + S[MM_S_ECONS_CORR];
  S[MM_S_ECONS] = econs;
  double* ce = S + MM_S_CE;  // ConsErrTracker (verlet.py:275-307)
  if (ce[0] == 0.0) {
    ce[1] = ekin;
    ce[3] = econs;
  } else {
    const double t1 = ekin - ce[1];
    ce[1] += t1 / (ce[0] + 1.0);
    ce[2] += t1 * (ekin - ce[1]);
    const double t2 = econs - ce[3];
    ce[3] += t2 / (ce[0] + 1.0);
    ce[4] += t2 * (econs - ce[3]);
  }
  ce[0] += 1.0;
  S[MM_S_CONS_ERR] = ce[0] > 1.0 ? sqrt(ce[4] / ce[2]) : 0.0;
  if (S[MM_S_NVEC] > 0.0) {
    const double V = S[MM_S_VOLUME];
    for (int i = 0; i < 9; ++i) S[MM_S_PTENS + i] = (S[MM_S_SVV + i] - S[MM_S_VTENS + i]) / V;
    S[MM_S_PRESS] = ((S[MM_S_PTENS + 0] + S[MM_S_PTENS + 4]) + S[MM_S_PTENS + 8]) / 3.0;
  }
}

// Energy and virial from the force partials, plus the external-pressure part (P V, P V I).
__device__ void take_force(double* S, const double (&f)[10], bool with_vtens) {
#pragma clang fp contract(off)
  S[MM_S_EPOT_MECH] = f[0];
  const double P = S[MM_S_EXT_PRESS];
  const double pv = P * S[MM_S_VOLUME];
  S[MM_S_EPOT] = P != 0.0 ? f[0] + S[MM_S_VOLUME] * P : f[0];
  if (with_vtens) {
    for (int q = 0; q < 9; ++q) {
      S[MM_S_VTENS + q] = f[1 + q];
      if (f[1 + q] != f[1 + q]) S[MM_S_NAN] = 1.0;
    }
    if (P != 0.0) {
      S[MM_S_VTENS + 0] += pv;
      S[MM_S_VTENS + 4] += pv;
      S[MM_S_VTENS + 8] += pv;
    }
  }
  if (f[0] != f[0]) S[MM_S_NAN] = 1.0;
}

__device__ void take_nodes(double* S, const double (&nd)[kNodeSums]) {
  S[MM_S_EKIN] = nd[0];
  for (int q = 0; q < 9; ++q) S[MM_S_SVV + q] = nd[1 + q];
  if (nd[12] != 0.0) S[MM_S_NAN] = 1.0;
}

// The node sums after the last pending barostat matrix M, from those before it: the barostat
// call changes every velocity to v M (row vectors, npt.py:709-726), so sum m (v M) (x) (v M) =
// M^T (sum m v (x) v) M and the kinetic energy is half its trace. Only the NaN flag comes from
// the node pass (the gather of the same force), when there is one.
__device__ void take_nodes_after_vmat(double* S, const double (&nd)[kNodeSums]) {
#pragma clang fp contract(off)
  const bool second = S[MM_S_VMAT2_ON] != 0.0;
  if (second || S[MM_S_VMAT_ON] != 0.0) {
    const double* M = S + (second ? MM_S_VMAT2 : MM_S_VMAT);
    const double* V = S + MM_S_SVV;
    double T[9], R[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        T[3 * i + j] = (V[3 * i + 0] * M[0 + j] + V[3 * i + 1] * M[3 + j]) + V[3 * i + 2] * M[6 + j];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        R[3 * i + j] = (M[0 + i] * T[0 + j] + M[3 + i] * T[3 + j]) + M[6 + i] * T[6 + j];
    for (int q = 0; q < 9; ++q) S[MM_S_SVV + q] = R[q];
    S[MM_S_EKIN] = 0.5 * ((R[0] + R[4]) + R[8]);
  }
  if (nd[12] != 0.0) S[MM_S_NAN] = 1.0;
}

// The force partials fpart[q * nbf + b] over the blocks b = fidx[ff + i] (or ff + i), i < fc,
// and the node partials npart[q * nbn + nf + i], i < nc, summed in a fixed order: each thread
// over its strided share in index order, then the wave by the reduce-scatter butterfly, then
// the waves in order. The loads of both kinds are issued before any sum, and every thread of
// the block receives the totals.
__device__ void reduce_force_nodes(const double* __restrict__ fpart, int nbf,
                                   const int32_t* __restrict__ fidx, int ff, int fc,
                                   const double* __restrict__ npart, int nbn, int nf, int nc,
                                   double (&f)[10], double (&nd)[kNodeSums]) {
  constexpr int kWaves = kScalarThreads / 64;
  constexpr int NV = 10 + kNodeSums;
  __shared__ double red[NV * kWaves];
  __shared__ double tot[NV];
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (int i = threadIdx.x; i < fc; i += kScalarThreads) {
    const int b = fidx ? fidx[ff + i] : ff + i;
#pragma unroll
    for (int q = 0; q < 10; ++q) acc[q] += fpart[(int64_t)q * nbf + b];
  }
  for (int i = threadIdx.x; i < nc; i += kScalarThreads) {
#pragma unroll
    for (int q = 0; q < kNodeSums; ++q) acc[10 + q] += npart[(int64_t)q * nbn + nf + i];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  {
    double lo[16], hi[NV - 16];
#pragma unroll
    for (int q = 0; q < 16; ++q) lo[q] = acc[q];
#pragma unroll
    for (int q = 16; q < NV; ++q) hi[q - 16] = acc[q];
    const double tl = wave_reduce_scatter<16>(lo, lane);
    const double th = wave_reduce_scatter<NV - 16>(hi, lane);
    const int q = scatter_index(lane);
    if ((lane & 3) == 0) {
      red[wave * NV + q] = tl;
      if (q < NV - 16) red[wave * NV + 16 + q] = th;
    }
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double t = red[threadIdx.x];
    for (int w = 1; w < kWaves; ++w) t += red[w * NV + threadIdx.x];
    tot[threadIdx.x] = t;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 10; ++q) f[q] = tot[q];
#pragma unroll
  for (int q = 0; q < kNodeSums; ++q) nd[q] = tot[10 + q];
  __syncthreads();
}

__device__ __forceinline__ bool phase_uses_force(int phase) {
  return phase == kPhInitForce || phase == kPhNveEnd || phase == kPhNvtEnd ||
         phase == kPhBaroEndPre || phase == kPhMainEndNpt || phase == kPhBaroEndPost ||
         phase == kPhBaroEndPostPre || phase == kPhNvtEndPre;
}
__device__ __forceinline__ bool phase_uses_nodes(int phase) {
  return phase == kPhInitProps || phase == kPhNveEnd || phase == kPhNvtEnd ||
         phase == kPhMainEndNpt || phase == kPhBaroEndPost || phase == kPhNvtEndPre ||
         phase == kPhRmsd;
}

// The scalar update of one phase on the scalar block S of system sys. chained: the phase closes
// a step that another step of the same call follows (kPhBaroEndPost: its velocity transforms
// stay pending for the next kick-drift pass).
__device__ void run_phase(int phase, double* S, const double (&f)[10],
                          const double (&nd)[kNodeSums], int with_vtens, int sys,
                          bool chained = false) {
  const bool nhc = thermo_nhc(S);
  const bool lt = S[MM_S_THERMO] == (double)MM_THERMO_LANGEVIN;
  // the barostat's pressure contribution enters the chain only for MTK (npt.py:109-113)
  const bool g1_on = S[MM_S_BARO] == (double)MM_BARO_MTK;
  switch (phase) {
    case kPhInitForce:
      take_force(S, f, with_vtens != 0);
      break;
    case kPhInitProps:
      take_nodes(S, nd);
      properties(S, false);
      break;
    case kPhNveEnd:
      clear_vtransforms(S);
      take_force(S, f, true);
      take_nodes(S, nd);
      properties(S, true);
      break;
    case kPhNvtPre: {
      double ekin = S[MM_S_EKIN];
      const double fac = nhc_chain(S, &ekin, false, 0.0);
      S[MM_S_EKIN] = ekin;
      set_vscale(S, MM_S_VSCALE_B, fac);
      break;
    }
    case kPhNvtEnd: {
      clear_vtransforms(S);
      take_force(S, f, true);
      if (lt) {
        const double ekin0 = S[MM_S_EKIN];
        take_nodes(S, nd);
        langevin_thermo_econs(S, ekin0, nd);
        properties(S, true);
        break;
      }
      take_nodes(S, nd);
      double ekin = S[MM_S_EKIN];
      const double fac = nhc_chain(S, &ekin, false, 0.0);
      set_vscale(S, MM_S_VSCALE_A, fac);
      for (int q = 0; q < 9; ++q) S[MM_S_SVV + q] *= fac * fac;
      S[MM_S_EKIN] = ekin;
      S[MM_S_ECONS_THERMO] = nhc_econs(S);
      properties(S, true);
      break;
    }
    case kPhBaroPre:
      // TBC.pre: chainvel0 = chain.vel[0] before the barostat (npt.py:101-107); the
      // barostat's own chain overrules it (npt.py:616-619)
      S[MM_S_HAS_CHAINVEL0] = nhc ? 1.0 : 0.0;
      S[MM_S_CHAINVEL0] = nhc ? S[MM_S_CHAIN_VEL] : 0.0;
      if (baro_chain_on(S)) {
        S[MM_S_HAS_CHAINVEL0] = 1.0;
        S[MM_S_CHAINVEL0] = S[MM_S_BCHAIN_VEL];
      }
      // (a pending thermostat factor is already in m v (x) v: the phases that set one scale
      // the sums by its square)
      baro_call_begin(S);
      baro_part1(S, sys, 0);
      // the velocity transform of this call does not depend on its force; it stays pending
      // for the kick-drift pass, which applies it before the chain's factor
      baro_part2(S);
      break;
    case kPhBaroEndPre: {
      // the barostat's velocity transform is still pending (see kPhBaroPre): the node sums
      // after it follow from those before it. The gather of this force comes after the phase
      // (it is the kick-drift pass), so a NaN gradient shows here through the virial.
      take_force(S, f, true);
      take_nodes_after_vmat(S, nd);
      update_baro_vel(S, sys, 0, 1);
      baro_call_end(S);
      if (baro_chain_on(S)) baro_chain_step(S);
      // a Langevin thermostat's pre update follows in the kick-drift pass
      if (nhc) {
        const double g1 = g1_on ? add_press_cont(S) : 0.0;
        double ekin = S[MM_S_EKIN];
        const double fac = nhc_chain(S, &ekin, g1_on, g1);
        S[MM_S_EKIN] = ekin;
        set_vscale(S, MM_S_VSCALE_B, fac);
      }
      break;
    }
    case kPhMainEndNpt: {
      // gather_drift has stored the positions through a pending earlier position transform; its
      // slot is cleared, entries included, as after steps taken one per call
      S[MM_S_XMAT2_ON] = 0.0;
      for (int q = 0; q < 9; ++q) S[MM_S_XMAT2 + q] = 0.0;
      clear_vtransforms(S);
      take_force(S, f, true);
      const double ekin0 = S[MM_S_EKIN];
      take_nodes(S, nd);
      if (nhc) {
        const double g1 = g1_on ? add_press_cont(S) : 0.0;
        double ekin = S[MM_S_EKIN];
        const double fac = nhc_chain(S, &ekin, g1_on, g1);
        S[MM_S_EKIN] = ekin;
        set_vscale(S, MM_S_VSCALE_A, fac);
        for (int q = 0; q < 9; ++q) S[MM_S_SVV + q] *= fac * fac;
        S[MM_S_ECONS_THERMO] = nhc_econs(S);
      } else if (lt) {
        langevin_thermo_econs(S, ekin0, nd);
      }
      // TBC.post: chainvel0 after the thermostat (npt.py:128-133); the barostat's own chain
      // steps first and overrules it (npt.py:631-639)
      S[MM_S_HAS_CHAINVEL0] = nhc ? 1.0 : 0.0;
      S[MM_S_CHAINVEL0] = nhc ? S[MM_S_CHAIN_VEL] : 0.0;
      if (baro_chain_on(S)) {
        baro_chain_step(S);
        S[MM_S_HAS_CHAINVEL0] = 1.0;
        S[MM_S_CHAINVEL0] = S[MM_S_BCHAIN_VEL];
      }
      baro_call_begin(S);
      baro_part1(S, sys, 1);
      baro_part2(S);
      break;
    }
    case kPhBaroEndPost:
      // the node sums after the call's velocity transform follow from those before it (as in
      // kPhBaroEndPre), whether the gather applied it (a call's last step) or it stays
      // pending (chained)
      take_force(S, f, true);
      take_nodes_after_vmat(S, nd);
      if (!chained) clear_vtransforms(S);
      // chained, the positions still wait for this call's transform (its force evaluation
      // applied it on the fly): it moves to the earlier slot, before the next call's
      if (chained) {
        for (int q = 0; q < 9; ++q) S[MM_S_XMAT2 + q] = S[MM_S_XMAT + q];
        S[MM_S_XMAT2_ON] = 1.0;
      }
      update_baro_vel(S, sys, 1, 1);
      if (baro_langevin(S)) {
        baro_call_end(S);
      } else {
        // MTKBarostat.post (npt.py:640-651)
        double e = baro_ekin(S);
        if (S[MM_S_BARO_VOLCON] == 0.0) e += S[MM_S_BARO_PRESS] * S[MM_S_VOLUME];
        if (baro_chain_on(S)) e += nhc_econs(S, kBaroChain);
        S[MM_S_ECONS_BARO] = e;
      }
      properties(S, true);
      break;
    case kPhRmsd: {
      const double denom = 3.0 * S[MM_S_NNODES];
      S[MM_S_RMSD_GPOS] = sqrt(nd[10] / denom);
      S[MM_S_RMSD_DELTA] = sqrt(nd[11] / denom);
      break;
    }
    case kPhApplied:
      clear_vtransforms(S);
      break;
  }
}

// Grid (slices, systems). Block (g, sys) reduces slice g of that system's force and node
// partials; the last block of the system to finish (counted with an atomic counter, which it
// resets) sums the slices in slice order and runs the phase on a copy of the system's scalar
// block in LDS, written back at the end.
__global__ void __launch_bounds__(kScalarThreads)
scalar_kernel(int phase, double* __restrict__ S0, const double* __restrict__ fpart, int nbf,
              const int32_t* __restrict__ fsys_ptr, const int32_t* __restrict__ fsys_idx,
              const double* __restrict__ npart, int nbn, const int32_t* __restrict__ nsys_ptr,
              int with_vtens, double* __restrict__ slices, unsigned int* __restrict__ counters) {
  static_assert(MM_S_COUNT <= kScalarThreads, "one scalar-block entry per thread");
  __shared__ double Sl[MM_S_COUNT];
  __shared__ double tot[kSliceQ];
  __shared__ int is_last;
  const int sys = blockIdx.y;
  const int G = gridDim.x, gi = blockIdx.x;
  double* Sg = S0 + (int64_t)sys * MM_S_COUNT;
  // the scalar block is loaded up front, so that its latency hides behind the reduction
  const double s_own = threadIdx.x < MM_S_COUNT ? Sg[threadIdx.x] : 0.0;
  double f[10], nd[kNodeSums];
  {
    int ff = 0, fc = 0, nf = 0, nc = 0;
    if (phase_uses_force(phase)) {
      const int b0 = fsys_ptr[sys], cnt = fsys_ptr[sys + 1] - b0;
      ff = b0 + (int)((int64_t)cnt * gi / G);
      fc = b0 + (int)((int64_t)cnt * (gi + 1) / G) - ff;
    }
    if (phase_uses_nodes(phase)) {
      const int b0 = nsys_ptr[sys], cnt = nsys_ptr[sys + 1] - b0;
      nf = b0 + (int)((int64_t)cnt * gi / G);
      nc = b0 + (int)((int64_t)cnt * (gi + 1) / G) - nf;
    }
    reduce_force_nodes(fpart, nbf, fsys_idx, ff, fc, npart, nbn, nf, nc, f, nd);
  }
  if (G > 1) {
    if (threadIdx.x == 0) {
      double* sl = slices + ((int64_t)sys * G + gi) * kSliceQ;
      for (int q = 0; q < 10; ++q) sl[q] = f[q];
      for (int q = 0; q < kNodeSums; ++q) sl[10 + q] = nd[q];
      __threadfence();
      const unsigned int prev = atomicAdd(&counters[sys], 1u);
      is_last = prev == (unsigned int)(G - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    // all slices into LDS in one parallel pass, then each total summed in slice order
    __shared__ double sl_sh[kScalarSlices * kSliceQ];
    const double* sl = slices + (int64_t)sys * G * kSliceQ;
    for (int i = threadIdx.x; i < G * kSliceQ; i += kScalarThreads) sl_sh[i] = sl[i];
    __syncthreads();
    if (threadIdx.x < kSliceQ) {
      double t = 0.0;
      for (int k = 0; k < G; ++k) t += sl_sh[k * kSliceQ + threadIdx.x];
      tot[threadIdx.x] = t;
    }
    if (threadIdx.x == 0) counters[sys] = 0u;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 10; ++q) f[q] = tot[q];
#pragma unroll
    for (int q = 0; q < kNodeSums; ++q) nd[q] = tot[10 + q];
  }
  if (threadIdx.x < MM_S_COUNT) Sl[threadIdx.x] = s_own;
  __syncthreads();
  if (threadIdx.x == 0) {
    // the end of one step and the start of the next, on the same copy of the block
    if (phase == kPhBaroEndPostPre) {
      run_phase(kPhBaroEndPost, Sl, f, nd, with_vtens, sys, true);
      run_phase(kPhBaroPre, Sl, f, nd, with_vtens, sys);
    } else if (phase == kPhNvtEndPre) {
      run_phase(kPhNvtEnd, Sl, f, nd, with_vtens, sys);
      run_phase(kPhNvtPre, Sl, f, nd, with_vtens, sys);
    } else {
      run_phase(phase, Sl, f, nd, with_vtens, sys);
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < MM_S_COUNT; q += kScalarThreads) Sg[q] = Sl[q];
}

}  // namespace mm

using namespace mm;

struct mm_md {
  mm_field* field = nullptr;
  double* pos = nullptr;
  double* vel = nullptr;
  double* gpos = nullptr;
  const double* masses = nullptr;
  double* posold = nullptr;
  double* S = nullptr;
  double* npart = nullptr;
  int thermo = 0, baro = 0;
  int64_t N = 0;
  int nbn = 0;
  int32_t nsys = 1;
  int32_t* node_blk_sys = nullptr;   // [nbn] system of every node block
  int32_t* sys_nblk = nullptr;       // [nsys + 1] node-block range of every system
  int slices = 1;                    // blocks per system of a reducing scalar phase
  double* slice_buf = nullptr;       // [nsys][slices][kSliceQ]
  unsigned int* counters = nullptr;  // [nsys] last-block counters (kept at zero between calls)
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // One step; runs whose step opens with a scalar phase replay, for several steps in a call,
  // that phase, then the lean step without it (which opens the next step) and then the full
  // step without it.
  hipGraph_t graph = nullptr, graph_lean = nullptr, graph_last = nullptr;
  hipGraphExec_t exec = nullptr, exec_lean = nullptr, exec_last = nullptr;
  int graph_split = 0;                // the field's block split the graphs were captured with
};

namespace {

struct Dev {
  int prev = -1;
  explicit Dev(int d) {
    (void)hipGetDevice(&prev);
    if (prev != d) (void)hipSetDevice(d);
  }
  ~Dev() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

ForceArgs force_args(mm_md* md) {
  mm_field* f = md->field;
  ForceArgs a;
  a.pos = md->pos;
  a.rvecs = md->S + MM_S_RVECS;
  a.rvecs_stride = MM_S_COUNT;
  a.blk_sys = f->blk_sys;
  a.cnode = f->cnode;
  a.cwrap = f->cwrap;
  a.blk_row = f->blk_row;
  a.tab = f->tab;
  a.kt = f->kt;
  a.nstates = f->nstates;
  a.max_states = f->max_states;
  a.W = f->W;
  a.N = f->N;
  a.ncell = f->nrec;
  a.gcell = f->gcell;
  a.rstride = f->S;
  a.rshift = 2;
  a.cbase = f->cbase;
  a.partials = f->partials;
  a.energy = nullptr;
  a.gpos = nullptr;
  a.vtens = nullptr;
  a.accumulate = 0;
  a.model = f->model;
  a.nanflag = f->nanflag;
  a.force_grad = 1;
  schedule_args(f, a);
  return a;
}

hipError_t node(mm_md* md, int mode, hipStream_t s) {
  const dim3 grid(md->nbn), block(kNodeThreads);
#define MM_NODE(M)                                                                          \
  hipLaunchKernelGGL((node_kernel<M>), grid, block, 0, s, md->pos, md->vel, md->gpos,        \
                     md->masses, md->posold, md->S, md->node_blk_sys, md->N, md->npart)
  switch (mode) {
    case kKickDrift: MM_NODE(kKickDrift); break;
    case kScalePos: MM_NODE(kScalePos); break;
    case kVV: MM_NODE(kVV); break;
    case kApplyV: MM_NODE(kApplyV); break;
    case kRmsd: MM_NODE(kRmsd); break;
    case kKickDriftLT: MM_NODE(kKickDriftLT); break;
  }
#undef MM_NODE
  return hipGetLastError();
}

bool reduces(int phase) {
  return phase == kPhInitForce || phase == kPhNveEnd || phase == kPhNvtEnd ||
         phase == kPhMainEndNpt || phase == kPhInitProps || phase == kPhBaroEndPre ||
         phase == kPhBaroEndPost || phase == kPhBaroEndPostPre || phase == kPhNvtEndPre ||
         phase == kPhRmsd;
}

hipError_t scalar(mm_md* md, int phase, int with_vtens, hipStream_t s) {
  const int nbf = (int)(md->field->W / kCellBlock);
  const int G = reduces(phase) ? md->slices : 1;
  // one system owns every cell block, in order: its block list is the identity
  const int32_t* cidx = md->nsys == 1 ? nullptr : md->field->sys_cblk_idx;
  hipLaunchKernelGGL(scalar_kernel, dim3(G, md->nsys), dim3(kScalarThreads), 0, s, phase, md->S,
                     md->field->partials, nbf, md->field->sys_cblk_ptr, cidx, md->npart, md->nbn,
                     md->sys_nblk, with_vtens, md->slice_buf, md->counters);
  return hipGetLastError();
}

// force evaluation: cells, then gather (+ second half kick when kick, + the Langevin
// thermostat's post update when lt; + the pending velocity transforms when scale_v). gmode
// (gather_kick_kernel): kGatherNone leaves out the records altogether, so the cell pass skips
// them too, and without scale_v there is then no node pass at all (energy and virial only).
hipError_t force(mm_md* md, bool kick, bool lt, hipStream_t s, bool scale_v = false,
                 int gmode = kGatherStore, bool corner_virial = false,
                 const double* xmat = nullptr) {
  ForceArgs a = force_args(md);
  a.corner_virial = corner_virial;
  a.xmat = xmat;
  if (gmode == kGatherNone) {
    if (kick) return hipErrorInvalidValue;
    a.store_records = 0;
  }
  hipError_t e = launch_cells(a, s);
  if (e != hipSuccess) return e;
  // with no records, no transform to apply and no kick, the evaluation has no node pass
  if (md->N > 0 && !(gmode == kGatherNone && !scale_v && !kick)) {
    const bool all = md->nbn <= kGatherAllLoadsMaxBlocks;
#define MM_GATHER1(K, T, V, A)                                                               \
  hipLaunchKernelGGL((gather_kick_kernel<K, T, V, false, A>), dim3(md->nbn), dim3(kNodeBlock), \
                     0, s, a.ncell, a.gcell, md->N, a.rstride, md->gpos, md->vel, md->masses,  \
                     md->S, md->node_blk_sys, md->npart, gmode)
#define MM_GATHER(K, T, V)          \
  do {                              \
    if (all)                        \
      MM_GATHER1(K, T, V, true);    \
    else                            \
      MM_GATHER1(K, T, V, false);   \
  } while (0)
    if (kick && lt)
      MM_GATHER(true, true, false);
    else if (kick)
      MM_GATHER(true, false, false);
    else if (scale_v)
      MM_GATHER(false, false, true);
    else
      MM_GATHER(false, false, false);
#undef MM_GATHER
#undef MM_GATHER1
    e = hipGetLastError();
  }
  return e;
}

#define MM_TRY(x)                         \
  do {                                    \
    hipError_t _e = (x);                  \
    if (_e != hipSuccess) return _e;      \
  } while (0)

// The first half kick and the drift from the records of the last cell pass (the barostat's
// pre call, whose gradient nothing else reads), after that call's scalar phase has set the
// pending velocity transforms; with lt also the Langevin thermostat's pre update.
hipError_t gather_drift(mm_md* md, bool lt, hipStream_t s) {
  if (md->N == 0) return hipSuccess;
  mm_field* f = md->field;
  const bool all = md->nbn <= kGatherAllLoadsMaxBlocks;
#define MM_DRIFT(T, A)                                                                       \
  hipLaunchKernelGGL((gather_kick_kernel<false, T, false, true, A>), dim3(md->nbn),          \
                     dim3(kNodeBlock), 0, s, f->nrec, f->gcell, md->N, f->S, md->gpos,       \
                     md->vel, md->masses, md->S, md->node_blk_sys, md->npart, kGatherKeep,   \
                     md->pos)
  if (lt && all)
    MM_DRIFT(true, true);
  else if (lt)
    MM_DRIFT(true, false);
  else if (all)
    MM_DRIFT(false, true);
  else
    MM_DRIFT(false, false);
#undef MM_DRIFT
  return hipGetLastError();
}

// The barostat call MTKBarostat.baro after its first scalar phase: positions scaled, force,
// then the second barostat-velocity update. The pre call (pre_call) scales no positions in a
// pass of its own: its cell pass reads them through the pending transforms, and it leaves its
// records for gather_drift, which stores the scaled positions and applies the velocity
// transform with the kick. The post call's velocity transform is applied by its gather
// (scale_v) or stays pending; with defer_x (a post call without records) its position scaling
// also waits (the cell pass applies it on the fly, the next gather_drift stores it).
hipError_t baro_call(mm_md* md, int end_phase, hipStream_t s, bool scale_v,
                     int gmode = kGatherStore, bool corner_virial = false,
                     bool pre_call = false, bool defer_x = false) {
  if (!defer_x && !pre_call) MM_TRY(node(md, kScalePos, s));
  if (pre_call) {
    // the pre call's scaling (after a pending post-call one) is applied by its cell pass as it
    // reads the positions and stored by gather_drift
    ForceArgs a = force_args(md);
    a.xmat = md->S + MM_S_XMAT;
    a.xmat2 = md->S + MM_S_XMAT2;
    a.xmat2_on = md->S + MM_S_XMAT2_ON;
    MM_TRY(launch_cells(a, s));
  } else {
    MM_TRY(force(md, false, false, s, scale_v, gmode, corner_virial,
                 defer_x ? md->S + MM_S_XMAT : nullptr));
  }
  MM_TRY(scalar(md, end_phase, 1, s));
  return hipSuccess;
}

// Ensembles whose step opens with a scalar phase (the barostat's or the chain's pre update).
bool opens_with_phase(const mm_md* md) {
  return md->baro != 0 || md->thermo == MM_THERMO_NHC;
}

// One step. lean: a step that another step of the same call follows, so that nothing reads
// the node gradients or the scalar block in between. With a barostat, the gradient of its main
// force evaluation is then used only by the kick in the gather, and that of the barostat's post
// call by nothing (the next step's pre call evaluates the force again before its kick,
// verlet.py:141-143 with npt.py:616-620), so those evaluations keep their gradients out of
// HBM; and a lean step ends with the next step's opening phase (kPhBaroEndPostPre,
// kPhNvtEndPre), so that step is enqueued with pre = false. Positions, velocities, energies,
// virials and every scalar are those of full steps, bit for bit.
hipError_t enqueue_step(mm_md* md, hipStream_t s, bool lean = false, bool pre = true) {
  const bool lt = md->thermo == MM_THERMO_LANGEVIN;
  if (md->baro) {
    if (pre) MM_TRY(scalar(md, kPhBaroPre, 1, s));
    // the pre call's gradient drives only the first half kick: its gather is the kick-drift
    MM_TRY(baro_call(md, kPhBaroEndPre, s, false, kGatherKeep, false, true));
    MM_TRY(gather_drift(md, lt, s));
    MM_TRY(force(md, true, lt, s, false, lean ? kGatherKeep : kGatherStore));
    MM_TRY(scalar(md, kPhMainEndNpt, 1, s));
    // the post call's velocity transform stays pending in a lean step (its node sums follow
    // from the ones before it), and so does its position scaling (the cell pass applies it
    // as it reads the positions, the next step's scaling pass stores it), so its force
    // evaluation has no node pass at all; its virial comes from the corners (B^T S B, no
    // vertex gradient needed), in the last step too, so that lean and full steps agree to
    // the bit
    MM_TRY(baro_call(md, lean ? kPhBaroEndPostPre : kPhBaroEndPost, s, !lean,
                     lean ? kGatherNone : kGatherStore, true, false, lean));
  } else if (md->thermo == MM_THERMO_NHC) {
    if (pre) MM_TRY(scalar(md, kPhNvtPre, 1, s));
    MM_TRY(node(md, kKickDrift, s));
    MM_TRY(force(md, true, false, s));
    MM_TRY(scalar(md, lean ? kPhNvtEndPre : kPhNvtEnd, 1, s));
  } else if (lt) {
    MM_TRY(node(md, kKickDriftLT, s));
    MM_TRY(force(md, true, true, s));
    MM_TRY(scalar(md, kPhNvtEnd, 1, s));
  } else {
    MM_TRY(node(md, kKickDrift, s));
    MM_TRY(force(md, true, false, s));
    MM_TRY(scalar(md, kPhNveEnd, 1, s));
  }
  return hipSuccess;
}

int begin(mm_md* md, void* stream) {
  hipError_t e = hipEventRecord(md->ev_in, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(md->stream, md->ev_in, 0);
  return e == hipSuccess ? MM_OK : hip_error(e, "stream ordering");
}

int end(mm_md* md, void* stream) {
  hipError_t e = hipEventRecord(md->ev_out, md->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent((hipStream_t)stream, md->ev_out, 0);
  return e == hipSuccess ? MM_OK : hip_error(e, "stream ordering");
}

}  // namespace

namespace {

// Capture one step into an executable graph (replayed once per step).
hipError_t capture_step(mm_md* md, bool lean, bool pre, hipGraph_t* graph,
                        hipGraphExec_t* exec) {
  hipError_t e = hipStreamBeginCapture(md->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  const hipError_t e2 = enqueue_step(md, md->stream, lean, pre);
  e = hipStreamEndCapture(md->stream, graph);
  if (e2 != hipSuccess) e = e2;
  if (e == hipSuccess) e = hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0);
  return e;
}

void drop_graph(mm_md* md) {
  if (md->exec) (void)hipGraphExecDestroy(md->exec);
  if (md->graph) (void)hipGraphDestroy(md->graph);
  if (md->exec_lean) (void)hipGraphExecDestroy(md->exec_lean);
  if (md->graph_lean) (void)hipGraphDestroy(md->graph_lean);
  if (md->exec_last) (void)hipGraphExecDestroy(md->exec_last);
  if (md->graph_last) (void)hipGraphDestroy(md->graph_last);
  md->exec = md->exec_lean = md->exec_last = nullptr;
  md->graph = md->graph_lean = md->graph_last = nullptr;
}

}  // namespace

extern "C" {

int mm_md_create_batch(mm_md** out, mm_field* field, int32_t nsys, const int64_t* sys_node_start,
                       double* pos, double* vel, double* gpos, const double* masses,
                       double* posold, double* scalars, int thermo, int baro) {
  if (!out || !field || !scalars || !sys_node_start) return set_error(MM_ERR_ARG, "missing argument");
  if (nsys != field->nsys)
    return set_error(MM_ERR_ARG, "field holds %d systems, not %d", field->nsys, nsys);
  if (field->N > 0 && (!pos || !vel || !gpos || !masses || !posold))
    return set_error(MM_ERR_ARG, "missing node buffer");
  if (thermo != MM_THERMO_NONE && thermo != MM_THERMO_NHC && thermo != MM_THERMO_LANGEVIN)
    return set_error(MM_ERR_ARG, "unknown thermostat %d", thermo);
  if (baro != MM_BARO_NONE && baro != MM_BARO_MTK && baro != MM_BARO_LANGEVIN)
    return set_error(MM_ERR_ARG, "unknown barostat %d", baro);
  if (sys_node_start[0] != 0 || sys_node_start[nsys] != field->N)
    return set_error(MM_ERR_ARG, "node ranges must cover [0, %lld)", (long long)field->N);
  const int nbn_total = (int)((field->N + kNodeBlock - 1) / kNodeBlock);
  std::vector<int32_t> node_blk_sys(nbn_total > 0 ? nbn_total : 1, 0), sys_nblk(nsys + 1, 0);
  for (int s = 0; s < nsys; ++s) {
    const int64_t a = sys_node_start[s], b = sys_node_start[s + 1];
    if (b < a || (nsys > 1 && a % kNodeBlock != 0))
      return set_error(MM_ERR_ARG, "system %d: node range [%lld, %lld) invalid", s,
                       (long long)a, (long long)b);
    const int blo = (int)(a / kNodeBlock), bhi = (int)((b + kNodeBlock - 1) / kNodeBlock);
    sys_nblk[s] = blo;
    sys_nblk[s + 1] = bhi;
    for (int k = blo; k < bhi; ++k) node_blk_sys[k] = s;
  }
  Dev dev(field->device);
  mm_md* md = new mm_md();
  md->field = field;
  md->pos = pos;
  md->vel = vel;
  md->gpos = gpos;
  md->masses = masses;
  md->posold = posold;
  md->S = scalars;
  md->thermo = thermo;
  md->baro = baro;
  md->N = field->N;
  md->nbn = nbn_total > 0 ? nbn_total : 1;
  md->nsys = nsys;
  // slices of a reducing scalar phase: about 256 partial blocks each, at most kScalarSlices
  std::vector<int32_t> cptr(nsys + 1, 0);
  hipError_t e = hipMemcpy(cptr.data(), field->sys_cblk_ptr, sizeof(int32_t) * (nsys + 1),
                           hipMemcpyDeviceToHost);
  int most = 0;
  for (int s = 0; s < nsys; ++s) {
    most = std::max(most, cptr[s + 1] - cptr[s]);
    most = std::max(most, sys_nblk[s + 1] - sys_nblk[s]);
  }
  md->slices = std::min(kScalarSlices, std::max(1, (most + 255) / 256));
  if (e == hipSuccess) e = hipMalloc((void**)&md->npart, sizeof(double) * kNodePartial * md->nbn);
  if (e == hipSuccess) e = hipMemset(md->npart, 0, sizeof(double) * kNodePartial * md->nbn);
  if (e == hipSuccess)
    e = hipMalloc((void**)&md->slice_buf, sizeof(double) * kSliceQ * md->slices * nsys);
  if (e == hipSuccess) e = hipMalloc((void**)&md->counters, sizeof(unsigned int) * nsys);
  if (e == hipSuccess) e = hipMemset(md->counters, 0, sizeof(unsigned int) * nsys);
  if (e == hipSuccess) e = hipMalloc((void**)&md->node_blk_sys, sizeof(int32_t) * md->nbn);
  if (e == hipSuccess) e = hipMalloc((void**)&md->sys_nblk, sizeof(int32_t) * (nsys + 1));
  if (e == hipSuccess)
    e = hipMemcpy(md->node_blk_sys, node_blk_sys.data(), sizeof(int32_t) * md->nbn,
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(md->sys_nblk, sys_nblk.data(), sizeof(int32_t) * (nsys + 1),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&md->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&md->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&md->ev_out, hipEventDisableTiming);
  if (e != hipSuccess) {
    mm_md_destroy(md);
    return hip_error(e, "mm_md_create");
  }
  *out = md;
  return MM_OK;
}

int mm_md_create(mm_md** out, mm_field* field, double* pos, double* vel, double* gpos,
                 const double* masses, double* posold, double* scalars, int thermo, int baro) {
  if (!field) return set_error(MM_ERR_ARG, "missing argument");
  if (field->nsys != 1) return set_error(MM_ERR_ARG, "batched field: use mm_md_create_batch");
  const int64_t range[2] = {0, field->N};
  return mm_md_create_batch(out, field, 1, range, pos, vel, gpos, masses, posold, scalars,
                            thermo, baro);
}

int mm_md_init_force(mm_md* md, int with_vtens, void* stream) {
  if (!md) return set_error(MM_ERR_ARG, "md is NULL");
  Dev dev(md->field->device);
  int rc = begin(md, stream);
  if (rc) return rc;
  hipError_t e = force(md, false, false, md->stream);
  if (e == hipSuccess) e = scalar(md, kPhInitForce, with_vtens, md->stream);
  if (e != hipSuccess) return hip_error(e, "mm_md_init_force");
  return end(md, stream);
}

int mm_md_init_properties(mm_md* md, void* stream) {
  if (!md) return set_error(MM_ERR_ARG, "md is NULL");
  Dev dev(md->field->device);
  int rc = begin(md, stream);
  if (rc) return rc;
  hipError_t e = node(md, kVV, md->stream);
  if (e == hipSuccess) e = scalar(md, kPhInitProps, 1, md->stream);
  if (e != hipSuccess) return hip_error(e, "mm_md_init_properties");
  return end(md, stream);
}

int mm_md_step(mm_md* md, int32_t nsteps, int use_graph, void* stream) {
  if (!md) return set_error(MM_ERR_ARG, "md is NULL");
  if (nsteps < 0) return set_error(MM_ERR_ARG, "negative step count");
  if (nsteps == 0) return MM_OK;
  Dev dev(md->field->device);
  int rc = begin(md, stream);
  if (rc) return rc;
  hipError_t e = hipSuccess;
  if (use_graph) {
    // the launch ranges of the cell pass follow the field's block split: a new split (new
    // state counts through mm_field_set_params) needs a new graph
    if (md->graph_split != md->field->split) drop_graph(md);
    md->graph_split = md->field->split;
    // every step but the last is followed by another before the call returns
    const bool lean = opens_with_phase(md) && nsteps > 1;
    if (!lean && !md->exec) e = capture_step(md, false, true, &md->graph, &md->exec);
    if (lean && !md->exec_lean) e = capture_step(md, true, false, &md->graph_lean, &md->exec_lean);
    if (lean && e == hipSuccess && !md->exec_last)
      e = capture_step(md, false, false, &md->graph_last, &md->exec_last);
    if (e != hipSuccess) return hip_error(e, "capturing the MD step");
    if (lean) {
      e = scalar(md, md->baro ? kPhBaroPre : kPhNvtPre, 1, md->stream);
      for (int i = 0; i + 1 < nsteps && e == hipSuccess; ++i)
        e = hipGraphLaunch(md->exec_lean, md->stream);
      if (e == hipSuccess) e = hipGraphLaunch(md->exec_last, md->stream);
    } else {
      for (int i = 0; i < nsteps && e == hipSuccess; ++i) e = hipGraphLaunch(md->exec, md->stream);
    }
  } else {
    const bool lean = opens_with_phase(md) && nsteps > 1;
    for (int i = 0; i < nsteps && e == hipSuccess; ++i)
      e = enqueue_step(md, md->stream, lean && i + 1 < nsteps, !lean || i == 0);
  }
  if (e != hipSuccess) return hip_error(e, "mm_md_step");
  return end(md, stream);
}

int mm_md_sync(mm_md* md, void* stream) {
  if (!md) return set_error(MM_ERR_ARG, "md is NULL");
  Dev dev(md->field->device);
  int rc = begin(md, stream);
  if (rc) return rc;
  hipError_t e = node(md, kApplyV, md->stream);
  if (e == hipSuccess) e = scalar(md, kPhApplied, 0, md->stream);
  if (e != hipSuccess) return hip_error(e, "mm_md_sync");
  return end(md, stream);
}

int mm_md_snapshot(mm_md* md, void* stream) {
  if (!md) return set_error(MM_ERR_ARG, "md is NULL");
  Dev dev(md->field->device);
  int rc = begin(md, stream);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(md->posold, md->pos, sizeof(double) * 3 * md->N,
                                hipMemcpyDeviceToDevice, md->stream);
  if (e != hipSuccess) return hip_error(e, "mm_md_snapshot");
  return end(md, stream);
}

int mm_md_rmsd(mm_md* md, void* stream) {
  if (!md) return set_error(MM_ERR_ARG, "md is NULL");
  Dev dev(md->field->device);
  int rc = begin(md, stream);
  if (rc) return rc;
  hipError_t e = node(md, kRmsd, md->stream);
  if (e == hipSuccess) e = scalar(md, kPhRmsd, 0, md->stream);
  if (e != hipSuccess) return hip_error(e, "mm_md_rmsd");
  return end(md, stream);
}

void mm_md_destroy(mm_md* md) {
  if (!md) return;
  drop_graph(md);
  if (md->stream) {
    (void)hipStreamSynchronize(md->stream);
    (void)hipStreamDestroy(md->stream);
  }
  if (md->ev_in) (void)hipEventDestroy(md->ev_in);
  if (md->ev_out) (void)hipEventDestroy(md->ev_out);
  if (md->npart) (void)hipFree(md->npart);
  if (md->slice_buf) (void)hipFree(md->slice_buf);
  if (md->counters) (void)hipFree(md->counters);
  if (md->node_blk_sys) (void)hipFree(md->node_blk_sys);
  if (md->sys_nblk) (void)hipFree(md->sys_nblk);
  delete md;
}

}  // extern "C"
