// Differentiation form: energy, node gradient, virial, domain-vector gradient and the gradient
// with respect to every cell-type parameter, optionally in dual numbers.
//
// Energy model: micmec/pes/nanocell.py:40-132 (eight corner matrices) and
// nanocell_original.py:46-132 (averaged matrix), combined over metastable states by the
// free-energy log-sum-exp of micmec/pes/mmff.py:380-398. The reference differentiates only with
// respect to positions (mmff.py:384-398; mmff_jax.py:249-250 also with respect to a deformation);
// the parameter derivatives have no counterpart there.
//
// In double precision the kernel returns first derivatives of the energy. With dual numbers
// (value, tangent) and a tangent direction (dpos, drvecs) the same pass also returns the
// directional derivative of each first derivative: the tangent of gpos is a Hessian-vector
// product and the tangent of dE/dtheta is (d^2E / dtheta dr) . dr, the quantity the parameter
// gradient of a loss on forces and stresses needs (forward-over-reverse).
//
// Parameter entries per (table row, state), kNps = 32 of them:
//   [0..20]  dE/dCv, Cv the 6x6 Voigt matrix in u = e.Cv.e / 2: 1/2 sum e_I e_J for I <= J
//            (upper triangle, row by row); the full 6x6 gradient is that entry at (I,J) and (J,I)
//   [21..29] dE/dP, P = h0^-1 (row-major)
//   [30]     dE/dw, w = det(h0)/8 (eight-corner model) or det(h0) (averaged model)
//   [31]     dE/d(free energy) = p_s, the state's Boltzmann weight
// followed by one entry per row, dE/d(kT). The caller maps (Cv, P, w, kT) to (Cv, h0, T_eff).
//
// Lanes: one per (cell, corner), 8 cells per wave, 32 cells per block; lane 8c + b owns the
// corner at offset (b & 1, b >> 1 & 1, b >> 2) of cell c. Blocks lie inside one 256-slot segment
// of the work order, so the type row and the system are uniform per block.
//
// HBM layout of the outputs:
//   part   [2 * kSysQ][W / 32]     per block: energy, virial (9), domain-vector gradient (9);
//                                  rows kSysQ.. hold the tangents
//   ppart  [2 * npar][W / 32]      per block: row 2i value, 2i + 1 tangent of parameter entry i
//   gcell, gcell2 [8][3][W]        vertex gradients and their tangents (gathered into gpos)
// All sums run in a fixed order; there are no floating-point atomics.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/micmec_amd.h"
#include "mm_device.h"
#include "mm_internal.h"
#include "mm_kernels.h"

namespace mm {

struct Dual {
  double v, d;
};

__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) {
  return {a.v * b.v, a.v * b.d + a.d * b.v};
}
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual dexp(Dual a) {
  const double e = exp(a.v);
  return {e, e * a.d};
}
__device__ __forceinline__ Dual dlog(Dual a) { return {log(a.v), a.d / a.v}; }
__device__ __forceinline__ double dexp(double a) { return exp(a); }
__device__ __forceinline__ double dlog(double a) { return log(a); }
__device__ __forceinline__ double val(double a) { return a; }
__device__ __forceinline__ double val(Dual a) { return a.v; }
__device__ __forceinline__ double tan_(double) { return 0.0; }
__device__ __forceinline__ double tan_(Dual a) { return a.d; }
template <typename T> __device__ __forceinline__ T make(double v, double d);
template <> __device__ __forceinline__ double make<double>(double v, double) { return v; }
template <> __device__ __forceinline__ Dual make<Dual>(double v, double d) { return {v, d}; }
template <typename T> struct IsDual { static constexpr bool value = false; };
template <> struct IsDual<Dual> { static constexpr bool value = true; };

// Lane exchanges inside groups of eight lanes, as DPP moves (VALU, no LDS round trip): lane l
// receives the value of lane l ^ mask. Masks 1 and 2 are quad permutations ([1,0,3,2] and
// [2,3,0,1]); mask 4 swaps the two 4-lane banks of each half row, as a row shift left by 4
// written to banks 0 and 2 and a row shift right by 4 written to banks 1 and 3 (dpp_f64: a
// double moves as its two 32-bit halves).
template <int kMask>
__device__ __forceinline__ double xs8(double v) {
  static_assert(kMask == 1 || kMask == 2 || kMask == 4, "exchanges inside 8 lanes");
  if (kMask == 1) return dpp_f64<0xB1, 0xF>(v, v);
  if (kMask == 2) return dpp_f64<0x4E, 0xF>(v, v);
  const double t = dpp_f64<0x104, 0x5>(v, v);
  return dpp_f64<0x114, 0xA>(t, v);
}
__device__ __forceinline__ double xs(double v, int mask) {
  return mask == 1 ? xs8<1>(v) : mask == 2 ? xs8<2>(v) : xs8<4>(v);
}
__device__ __forceinline__ Dual xs(Dual v, int mask) { return {xs(v.v, mask), xs(v.d, mask)}; }
__device__ __forceinline__ double bc(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ Dual bc(Dual v, int src) {
  return {__shfl(v.v, src, 64), __shfl(v.d, src, 64)};
}
// Sum over the 8 lanes of a cell (xor butterfly: every lane gets the same bits).
template <typename T>
__device__ __forceinline__ T sum8(T v) {
  v = v + xs(v, 1);
  v = v + xs(v, 2);
  v = v + xs(v, 4);
  return v;
}

// Sums over the 8 lanes of a cell of eight values c[0..7] at once: lane b returns the sum of
// c[b]. Each level keeps half of the lane's values and adds the partner's copy of them, over
// the same pairs, quads and halves as sum8, so every sum equals sum8 of that value bit for bit
// (addition commutes), with 7 exchanges instead of 24.
template <typename T>
__device__ __forceinline__ T rs8(const T (&c)[8], int b) {
  const bool b0 = (b & 1) != 0, b1 = (b & 2) != 0, b2 = (b & 4) != 0;
  T a[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const T keep = b0 ? c[2 * k + 1] : c[2 * k];
    const T send = b0 ? c[2 * k] : c[2 * k + 1];
    a[k] = keep + xs(send, 1);  // value 2k + b0
  }
  T a2[2];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const T keep = b1 ? a[2 * m + 1] : a[2 * m];
    const T send = b1 ? a[2 * m] : a[2 * m + 1];
    a2[m] = keep + xs(send, 2);  // value 4m + 2 b1 + b0
  }
  const T keep = b2 ? a2[1] : a2[0];
  const T send = b2 ? a2[0] : a2[1];
  return keep + xs(send, 4);
}

// Upper triangle of the 6x6 Voigt matrix, row by row: entry q is (kUpperI[q], kUpperJ[q]).
__device__ constexpr int kUpperI[21] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5};
__device__ constexpr int kUpperJ[21] = {0, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 2, 3, 4, 5, 3, 4, 5, 4, 5, 5};

constexpr int kAdjCells = 32;  // cells per block
constexpr int kAdjPerSeg = kCellBlock / kAdjCells;  // adjoint blocks per cell-pass block
static_assert(kCellBlock % kAdjCells == 0, "cell-pass blocks hold whole adjoint blocks");
constexpr int kNps = 32;       // parameter entries per (row, state)
constexpr int kSysQ = 19;      // energy, virial (9), domain-vector gradient (9)
// corner (x + 2y + 4z) -> corner index of the reference order (000,100,010,001,110,101,011,111)
__device__ constexpr int kRefCorner[8] = {0, 1, 2, 4, 3, 5, 6, 7};

// The lead lane of each cell stores the value (and tangent) of one per-cell quantity in the
// cell's staging row; padding cells store zero.
template <typename T>
__device__ __forceinline__ void stage(double* row, int q, T x, bool lead, bool valid) {
  if (lead) {
    row[q] = valid ? val(x) : 0.0;
    if (IsDual<T>::value) row[32 + q] = valid ? tan_(x) : 0.0;
  }
}

// Sum the staged rows of the block's 32 cells: value entry q goes to row vbase + vstep * q of
// out, tangent entry q to row tbase + tstep * q (rows of length stride). Thread t sums cells
// 8p..8p+7 of one (entry, half), p = t mod 4, and the four parts are added by two shuffles
// (a fixed order). Entry `extra` (< 0: none), if staged, goes to row xbase / xbase + 1.
template <typename T>
__device__ __forceinline__ void flush(const double* vsh, int nq, double* out, int64_t vbase,
                                      int64_t vstep, int64_t tbase, int64_t tstep,
                                      int64_t stride, int extra = -1, int64_t xbase = 0) {
  __syncthreads();
  const int part = threadIdx.x & 3;
  const int col = threadIdx.x >> 2;        // 0..63: half * 32 + q
  const int q = col & 31, half = col >> 5;
  double s = 0.0;
  // without dual numbers the tangent rows are written as zeros
  if ((q < nq || q == extra) && (half == 0 || IsDual<T>::value)) {
#pragma unroll
    for (int c = 0; c < 8; ++c) s += vsh[(part * 8 + c) * 64 + col];
  }
  s += xs8<1>(s);
  s += xs8<2>(s);
  if (part == 0 && (q < nq || q == extra)) {
    int64_t r;
    if (q == extra) r = xbase + half;
    else r = half == 0 ? vbase + vstep * q : tbase + tstep * q;
    out[r * stride + blockIdx.x] = s;
  }
  __syncthreads();
}

template <typename T, int kModel>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
adjoint_kernel(const double* __restrict__ pos, const double* __restrict__ dpos,
               const double* __restrict__ rvecs, const double* __restrict__ drvecs,
               const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ cnode,
               const int32_t* __restrict__ cwrap, const int32_t* __restrict__ blk_row,
               const double* __restrict__ tab_g, const double* __restrict__ kt_g,
               const int32_t* __restrict__ nstates_g, int32_t max_states, int64_t W,
               double* __restrict__ gcell, double* __restrict__ gcell2,
               double* __restrict__ part, double* __restrict__ ppart) {
  __shared__ double es_sh[kMaxStates * kAdjCells * 2];
  __shared__ double vsh[kAdjCells * 64];
  const int64_t nblk = gridDim.x;
  const int seg = blockIdx.x / kAdjPerSeg;
  const int trow = blk_row[seg];
  const int sys = blk_sys[seg];
  const int ns = nstates_g[trow];
  const double kt = kt_g[trow];
  const double* tab = tab_g + (int64_t)trow * max_states * kTabStride;
  const int b = threadIdx.x & 7;
  const int kref = kRefCorner[b];
  const int lane = threadIdx.x & 63;
  const int lead_lane = lane & ~7;
  const bool lead = b == 0;
  const int cl = threadIdx.x >> 3;
  double* vrow = vsh + cl * 64;
  const int64_t w = (int64_t)blockIdx.x * kAdjCells + cl;
  const int n = cnode[w * 8 + kref];
  const bool valid = cnode[w * 8] >= 0;
  const int wrap = valid ? cwrap[w] : 0;
  const double* rv = rvecs + 9 * (int64_t)sys;
  const double* drv = drvecs ? drvecs + 9 * (int64_t)sys : nullptr;

  // vertex of this lane: r_0 = p_0, r_k = p_0 + ((p_k - p_0) + shift_k) (mmff.py:347-371).
  // Every load is issued unconditionally (a padding slot reads node 0 and discards it) and the
  // shifts use factors 0 or 1 (exact products), so that no load sits behind a branch.
  const int64_t nn = valid ? n : 0;
  double pv[3], dv[3] = {0.0, 0.0, 0.0}, rvv[9], drvv[9];
#pragma unroll
  for (int j = 0; j < 3; ++j) pv[j] = pos[3 * nn + j];
  if (dpos) {
#pragma unroll
    for (int j = 0; j < 3; ++j) dv[j] = dpos[3 * nn + j];
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    rvv[q] = rv[q];
    drvv[q] = drv ? drv[q] : 0.0;
  }
  T p[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) p[j] = make<T>(valid ? pv[j] : 0.0, valid ? dv[j] : 0.0);
  T r[3], r0[3];
  {
    const int moved = b & wrap;
    const double fx = (double)(moved & 1), fy = (double)((moved >> 1) & 1),
                 fz = (double)((moved >> 2) & 1);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      r0[j] = bc(p[j], lead_lane);
      const T s0 = make<T>(fx * rvv[0 + j], fx * drvv[0 + j]);
      const T s1 = make<T>(fy * rvv[3 + j], fy * drvv[3 + j]);
      const T s2 = make<T>(fz * rvv[6 + j], fz * drvv[6 + j]);
      const T dvec = (p[j] - r0[j]) + ((s0 + s1) + s2);
      r[j] = b ? r0[j] + dvec : p[j];
    }
  }
  // X: the lane's corner matrix (rows = +x, +y, +z edges through the corner), or for the
  // averaged model H = (1/8) sum of the eight corner matrices, identical on the cell's lanes
  T X[9];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int bit = (b >> d) & 1;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const T other = xs(r[j], 1 << d);
      X[3 * d + j] = bit ? r[j] - other : other - r[j];
    }
  }
  if (kModel == 1) {
#pragma unroll
    for (int i = 0; i < 9; ++i) X[i] = 0.125 * sum8(X[i]);
  }

  // pass 1 (multistate rows): E_s = w * sum_a u_a + f_s for every state (mmff.py:380-384),
  // which the weights of pass 2 need first; a single-state cell's energy comes out of pass 2
  double emin = 0.0;
  for (int s = 0; s < (ns > 1 ? ns : 0); ++s) {
    const double* ts = tab + s * kTabStride;
    const double* P = ts;
    const double* Cv = ts + kCvOffset;
    T B[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        B[3 * i + j] = (P[3 * i + 0] * X[0 + j] + P[3 * i + 1] * X[3 + j]) + P[3 * i + 2] * X[6 + j];
    T e[6];
    e[0] = 0.5 * (((B[0] * B[0] + B[1] * B[1]) + B[2] * B[2]) - 1.0);
    e[1] = 0.5 * (((B[3] * B[3] + B[4] * B[4]) + B[5] * B[5]) - 1.0);
    e[2] = 0.5 * (((B[6] * B[6] + B[7] * B[7]) + B[8] * B[8]) - 1.0);
    e[3] = (B[3] * B[6] + B[4] * B[7]) + B[5] * B[8];
    e[4] = (B[0] * B[6] + B[1] * B[7]) + B[2] * B[8];
    e[5] = (B[0] * B[3] + B[1] * B[4]) + B[2] * B[5];
    T u = make<T>(0.0, 0.0);
#pragma unroll
    for (int I = 0; I < 6; ++I) {
      const T sI = ((((Cv[6 * I + 0] * e[0] + Cv[6 * I + 1] * e[1]) + Cv[6 * I + 2] * e[2]) +
                     Cv[6 * I + 3] * e[3]) + Cv[6 * I + 4] * e[4]) + Cv[6 * I + 5] * e[5];
      u = u + e[I] * sI;
    }
    u = 0.5 * u;
    const T usum = kModel == 0 ? sum8(u) : u;
    const T es = ts[9] * usum + ts[10];
    if (lead) {
      es_sh[(s * kAdjCells + cl) * 2 + 0] = val(es);
      es_sh[(s * kAdjCells + cl) * 2 + 1] = tan_(es);
    }
    emin = (s == 0 || val(es) < emin) ? val(es) : emin;
  }
  if (ns > 1) __syncthreads();  // uniform over the block (one row per block)
  // E = E_min - kT ln Z, Z = sum_s exp(-(E_s - E_min)/kT) (mmff.py:386-395); one state: E = E_0
  T ecell = make<T>(0.0, 0.0), zsum = make<T>(1.0, 0.0), dkt = make<T>(0.0, 0.0);
  if (ns > 1) {
    zsum = make<T>(0.0, 0.0);
    for (int s = 0; s < ns; ++s) {
      const T es = make<T>(es_sh[(s * kAdjCells + cl) * 2], es_sh[(s * kAdjCells + cl) * 2 + 1]);
      zsum = zsum + dexp(-(es - emin) / kt);
    }
    const T lnz = dlog(zsum);
    ecell = emin - kt * lnz;
    // dE/dkT = -ln Z - sum_s p_s (E_s - E_min) / kT
    dkt = -lnz;
  }

  // pass 2: gradients, weighted by p_s (mmff.py:396-398)
  T gacc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) gacc[i] = make<T>(0.0, 0.0);
  for (int s = 0; s < ns; ++s) {
    const double* ts = tab + s * kTabStride;
    const double* P = ts;
    const double wt = ts[9];
    const double* Cv = ts + kCvOffset;
    T ps = make<T>(1.0, 0.0);
    if (ns > 1) {
      const T es = make<T>(es_sh[(s * kAdjCells + cl) * 2], es_sh[(s * kAdjCells + cl) * 2 + 1]);
      const T x = es - emin;
      ps = dexp(-x / kt) / zsum;
      dkt = dkt - (ps * x) / kt;
    }
    T B[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        B[3 * i + j] = (P[3 * i + 0] * X[0 + j] + P[3 * i + 1] * X[3 + j]) + P[3 * i + 2] * X[6 + j];
    T e[6];
    e[0] = 0.5 * (((B[0] * B[0] + B[1] * B[1]) + B[2] * B[2]) - 1.0);
    e[1] = 0.5 * (((B[3] * B[3] + B[4] * B[4]) + B[5] * B[5]) - 1.0);
    e[2] = 0.5 * (((B[6] * B[6] + B[7] * B[7]) + B[8] * B[8]) - 1.0);
    e[3] = (B[3] * B[6] + B[4] * B[7]) + B[5] * B[8];
    e[4] = (B[0] * B[6] + B[1] * B[7]) + B[2] * B[8];
    e[5] = (B[0] * B[3] + B[1] * B[4]) + B[2] * B[5];
    T sv[6];
    T u = make<T>(0.0, 0.0);
#pragma unroll
    for (int I = 0; I < 6; ++I) {
      sv[I] = ((((Cv[6 * I + 0] * e[0] + Cv[6 * I + 1] * e[1]) + Cv[6 * I + 2] * e[2]) +
                Cv[6 * I + 3] * e[3]) + Cv[6 * I + 4] * e[4]) + Cv[6 * I + 5] * e[5];
      u = u + e[I] * sv[I];
    }
    u = 0.5 * u;
    if (ns == 1) ecell = wt * (kModel == 0 ? sum8(u) : u) + ts[10];
    // dU/dB = S B with S the symmetric stress of the Voigt vector Cv e; dU/dX = P^T (S B)
    const T S[9] = {sv[0], sv[5], sv[4], sv[5], sv[1], sv[3], sv[4], sv[3], sv[2]};
    T Tm[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        Tm[3 * i + j] = (S[3 * i + 0] * B[0 + j] + S[3 * i + 1] * B[3 + j]) + S[3 * i + 2] * B[6 + j];
    const T sw = wt * ps;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const T G = (P[0 + i] * Tm[0 + j] + P[3 + i] * Tm[3 + j]) + P[6 + i] * Tm[6 + j];
        gacc[3 * i + j] = gacc[3 * i + j] + sw * G;
      }
    // parameter entries of this state, summed over the cell's corners: entry q < 21 is
    // 1/2 e_I e_J (upper triangle), 21 <= q < 30 is (dU/dB X^T)_ik = (dU/dP)_ik, 30 is u
    auto entry = [&](int q) -> T {
      if (q < 21) return 0.5 * (e[kUpperI[q]] * e[kUpperJ[q]]);
      if (q < 30) {
        const int i = (q - 21) / 3, k = (q - 21) % 3;
        return (Tm[3 * i + 0] * X[3 * k + 0] + Tm[3 * i + 1] * X[3 * k + 1]) +
               Tm[3 * i + 2] * X[3 * k + 2];
      }
      return q == 30 ? u : make<T>(0.0, 0.0);
    };
    if (kModel == 0) {
      // eight entries at a time, summed over the cell's lanes by a reduce-scatter: lane b ends
      // with entry 8g + b and stages it (entry 31, the weight p_s, is not a sum)
#pragma unroll
      for (int g = 0; g < kNps / 8; ++g) {
        T c[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) c[i] = entry(8 * g + i);
        const T sum = rs8(c, b);
        const int q = 8 * g + b;
        const T v = q == 31 ? ps : (q == 30 ? ps : sw) * sum;
        vrow[q] = valid ? val(v) : 0.0;
        if (IsDual<T>::value) vrow[32 + q] = valid ? tan_(v) : 0.0;
      }
    } else {
      // the averaged model's entries are the same on all eight lanes of a cell
#pragma unroll
      for (int q = 0; q < 30; ++q) stage(vrow, q, sw * entry(q), lead, valid);
      stage(vrow, 30, ps * u, lead, valid);
      stage(vrow, 31, ps, lead, valid);
    }
    flush<T>(vsh, kNps, ppart, (int64_t)2 * s * kNps, 2, (int64_t)2 * s * kNps + 1, 2, nblk);
  }

  // vertex gradient: corner b is the upper end of the d edges of corners b and b ^ (1 << d)
  // when bit d of b is set, the lower end otherwise (averaged model: a quarter of the row)
  T g[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) g[j] = make<T>(0.0, 0.0);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int bit = (b >> d) & 1;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const T row = kModel == 0 ? gacc[3 * d + j] + xs(gacc[3 * d + j], 1 << d)
                                : 0.25 * gacc[3 * d + j];
      g[j] = bit ? g[j] + row : g[j] - row;
    }
  }
  if (valid) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      __builtin_nontemporal_store(val(g[j]), &gcell[(int64_t)(3 * kref + j) * W + w]);
      if (IsDual<T>::value)
        __builtin_nontemporal_store(tan_(g[j]), &gcell2[(int64_t)(3 * kref + j) * W + w]);
    }
  }
  // energy, virial sum_v g_v (x) (r_v - r_0) and the domain-vector gradient: corner b moves
  // with domain vector d when the cell wraps along d and the corner lies at +d
  // (entries 1..18 summed over the cell's lanes eight at a time, as the parameter entries;
  // the energy and dE/dkT, the same on every lane of the cell, ride along as entries 0 and
  // kSysQ, the latter into its parameter rows)
  auto sys_entry = [&](int q) -> T {
    if (q >= 1 && q <= 9) return g[(q - 1) / 3] * (r[(q - 1) % 3] - r0[(q - 1) % 3]);
    if (q >= 10 && q <= 18) {
      const int d = (q - 10) / 3;
      const double moves = (double)(((b & wrap) >> d) & 1);
      return moves * g[(q - 10) % 3];
    }
    return make<T>(0.0, 0.0);
  };
#pragma unroll
  for (int gq = 0; gq < 3; ++gq) {
    T c[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) c[i] = sys_entry(8 * gq + i);
    const T sum = rs8(c, b);
    const int q = 8 * gq + b;
    const T v = q == 0 ? ecell : q == kSysQ ? dkt : sum;
    if (q <= kSysQ) {
      vrow[q] = valid ? val(v) : 0.0;
      if (IsDual<T>::value) vrow[32 + q] = valid ? tan_(v) : 0.0;
    }
  }
  __syncthreads();
  {
    // the system rows go to `part`, the kT entry to `ppart`: two writers of one staged block
    const int pq = threadIdx.x & 3;
    const int col = threadIdx.x >> 2;
    const int q = col & 31, half = col >> 5;
    double sum = 0.0;
    if (q <= kSysQ && (half == 0 || IsDual<T>::value)) {
#pragma unroll
      for (int c = 0; c < 8; ++c) sum += vsh[(pq * 8 + c) * 64 + col];
    }
    sum += xs8<1>(sum);
    sum += xs8<2>(sum);
    if (pq == 0 && q < kSysQ)
      part[(int64_t)(half == 0 ? q : kSysQ + q) * nblk + blockIdx.x] = sum;
    if (pq == 0 && q == kSysQ)
      ppart[((int64_t)2 * max_states * kNps + half) * nblk + blockIdx.x] = sum;
  }
}

// Row sums of per-block partials over lists of segments (8 blocks each): one block per
// (list, row of the partial array). Rows q with q % weight_every == 0 are scaled per segment by
// weight[system of the segment] when weight is given.
__global__ void __launch_bounds__(256)
segment_sum_kernel(const double* __restrict__ part, int64_t nblk, int nrows_part,
                   const int32_t* __restrict__ list_ptr, const int32_t* __restrict__ list_idx,
                   const int32_t* __restrict__ blk_sys, const double* __restrict__ weight,
                   int weight_every, double* __restrict__ out) {
  __shared__ double red[4];
  const int li = blockIdx.x / nrows_part;
  const int q = blockIdx.x % nrows_part;
  const bool weighted = weight != nullptr && (q % weight_every) == 0;
  const int i0 = list_ptr[li], i1 = list_ptr[li + 1];
  double acc = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    const int seg = list_idx[i];
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kAdjPerSeg; ++k) s += part[(int64_t)q * nblk + (int64_t)seg * kAdjPerSeg + k];
    acc += weighted ? weight[blk_sys[seg]] * s : s;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Tangent direction of a vector-Jacobian product from the cotangents (gE, gG, gV) of
// (E, gpos, vtens): dpos_n = gG_n + pos_n gV_s^T (s = node_sys[n]), and for the threads past the
// nodes drvecs_s = rvecs_s gV_s^T. gG or gV may be null (zero).
__global__ void __launch_bounds__(256)
vjp_direction_kernel(const double* __restrict__ pos, const double* __restrict__ rvecs,
                     const int32_t* __restrict__ node_sys, const double* __restrict__ gG,
                     const double* __restrict__ gV, int64_t N, int nsys,
                     double* __restrict__ dpos, double* __restrict__ drv) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < N) {
    const double* V = gV ? gV + 9 * (int64_t)node_sys[t] : nullptr;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      double d = gG ? gG[3 * t + l] : 0.0;
      if (V)
        d += (pos[3 * t] * V[3 * l] + pos[3 * t + 1] * V[3 * l + 1]) + pos[3 * t + 2] * V[3 * l + 2];
      dpos[3 * t + l] = d;
    }
  } else if (t < N + 9 * (int64_t)nsys) {
    const int64_t q = t - N;
    const int64_t sy = q / 9;
    const int dd = (int)(q % 9) / 3, l = (int)(q % 3);
    const double* R = rvecs + 9 * sy;
    double d = 0.0;
    if (gV) {
      const double* V = gV + 9 * sy;
      d = (R[3 * dd] * V[3 * l] + R[3 * dd + 1] * V[3 * l + 1]) + R[3 * dd + 2] * V[3 * l + 2];
    }
    drv[q] = d;
  }
}

// Position cotangent of the product: g_n = gE_s gpos_n + dgpos_n + gpos_n gV_s (gV may be null:
// zero).
__global__ void __launch_bounds__(256)
vjp_position_kernel(const double* __restrict__ gpos, const double* __restrict__ dgpos,
                    const int32_t* __restrict__ node_sys, const double* __restrict__ gE,
                    const double* __restrict__ gV, int64_t N, double* __restrict__ out) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  const int sy = node_sys[n];
  const double e = gE[sy];
  const double* V = gV ? gV + 9 * (int64_t)sy : nullptr;
  const double g0 = gpos[3 * n], g1 = gpos[3 * n + 1], g2 = gpos[3 * n + 2];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    double v = e * gpos[3 * n + l] + dgpos[3 * n + l];
    if (V) v += (g0 * V[l] + g1 * V[3 + l]) + g2 * V[6 + l];
    out[3 * n + l] = v;
  }
}

}  // namespace mm

using namespace mm;

// Gradients of (Cv, h0, free energy, T_eff) from the parameter entries of one (row, state) per
// thread: the 21 upper-triangle entries give the symmetric Voigt gradient, the inverse P = h0^-1
// by cofactors (rows r1 x r2, r2 x r0, r0 x r1 of cof; P^T = cof / det), dE/dh0 =
// -P^T (dE/dP) P^T + (dE/dw) w P^T with w = det / 8 (corner8) or det; states beyond the row's
// count get zeros; thread (row, 0) also writes boltzmann dE/d(kT).
__global__ void __launch_bounds__(256)
param_grads_kernel(const double* __restrict__ entries, int64_t stride, int add_next,
                   int64_t npar, int nrows, int max_states, const int32_t* __restrict__ nstates,
                   const double* __restrict__ h0, int corner8, double boltzmann,
                   double* __restrict__ g_cv, double* __restrict__ g_h0,
                   double* __restrict__ g_free, double* __restrict__ g_teff) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows * max_states) return;
  const int row = t / max_states, s = t % max_states;
  const double* er = entries + (int64_t)row * npar * stride;
  auto ent = [&](int64_t q) {
    return add_next ? er[q * stride] + er[q * stride + 1] : er[q * stride];
  };
  const bool on = s < nstates[row];
  const int64_t q0 = (int64_t)s * kNps;
  double* gc = g_cv + (int64_t)t * 36;
  int q = 0;
  for (int I = 0; I < 6; ++I) {
    for (int J = I; J < 6; ++J, ++q) {
      const double v = on ? ent(q0 + q) : 0.0;
      gc[6 * I + J] = v;
      gc[6 * J + I] = v;
    }
  }
  const double* h = h0 + (int64_t)t * 9;
  double cof[9];
  // cof row 0 = r1 x r2, row 1 = r2 x r0, row 2 = r0 x r1
  for (int a = 0; a < 3; ++a) {
    const double* u = h + 3 * ((a + 1) % 3);
    const double* v = h + 3 * ((a + 2) % 3);
    cof[3 * a + 0] = u[1] * v[2] - u[2] * v[1];
    cof[3 * a + 1] = u[2] * v[0] - u[0] * v[2];
    cof[3 * a + 2] = u[0] * v[1] - u[1] * v[0];
  }
  const double det = (h[0] * cof[0] + h[1] * cof[1]) + h[2] * cof[2];
  double pt[9], gp[9], m1[9];
  for (int k = 0; k < 9; ++k) {
    pt[k] = cof[k] / det;
    gp[k] = ent(q0 + 21 + k);
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      m1[3 * i + j] = (pt[3 * i] * gp[j] + pt[3 * i + 1] * gp[3 + j]) + pt[3 * i + 2] * gp[6 + j];
  const double w = corner8 ? det / 8.0 : det;
  const double gww = ent(q0 + 30) * w;
  double* gh = g_h0 + (int64_t)t * 9;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double m2 =
          (m1[3 * i] * pt[j] + m1[3 * i + 1] * pt[3 + j]) + m1[3 * i + 2] * pt[6 + j];
      gh[3 * i + j] = on ? -m2 + gww * pt[3 * i + j] : 0.0;
    }
  g_free[t] = on ? ent(q0 + 31) : 0.0;
  if (s == 0) g_teff[row] = ent((int64_t)max_states * kNps) * boltzmann;
}

struct mm_adjoint {
  mm_field* field = nullptr;
  int64_t npar = 0;            // parameter entries per row: max_states * 32 + 1
  int64_t nblk = 0;            // W / 32
  double* part = nullptr;      // [2 kSysQ][nblk]
  double* ppart = nullptr;     // [2 npar][nblk]
  double* gcell2 = nullptr;    // [8][3][W]
  int32_t* row_ptr = nullptr;  // [nrows + 1] CSR of the segments of every table row
  int32_t* row_idx = nullptr;  // [W / kCellBlock]
  double* vjp = nullptr;       // [3][N][3] + [nsys][9]: dpos, gpos, dgpos, drvecs of mm_adjoint_vjp
};

extern "C" {

int mm_adjoint_create(mm_adjoint** out, mm_field* field, int64_t* npar) {
  if (!out || !field) return set_error(MM_ERR_ARG, "missing argument");
  *out = nullptr;
  const int64_t nseg = field->W / kCellBlock;
  std::vector<int32_t> blk_row(nseg > 0 ? nseg : 1, 0);
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(field->device);
  hipError_t e = hipSuccess;
  if (nseg > 0)
    e = hipMemcpy(blk_row.data(), field->blk_row, sizeof(int32_t) * nseg, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    (void)hipSetDevice(cur);
    return hip_error(e, "mm_adjoint_create");
  }
  std::vector<int32_t> row_ptr(field->nrows + 1, 0), row_idx;
  for (int t = 0; t < field->nrows; ++t) {
    for (int64_t s = 0; s < nseg; ++s)
      if (blk_row[s] == t) row_idx.push_back((int32_t)s);
    row_ptr[t + 1] = (int32_t)row_idx.size();
  }
  mm_adjoint* a = new mm_adjoint();
  a->field = field;
  a->npar = (int64_t)field->max_states * kNps + 1;
  a->nblk = field->W / kAdjCells;
  const size_t nb = (size_t)(a->nblk > 0 ? a->nblk : 1);
  e = hipMalloc((void**)&a->part, sizeof(double) * 2 * kSysQ * nb);
  if (e == hipSuccess) e = hipMalloc((void**)&a->ppart, sizeof(double) * 2 * a->npar * nb);
  if (e == hipSuccess)
    e = hipMalloc((void**)&a->gcell2, sizeof(double) * 24 * (field->W > 0 ? field->W : 1));
  if (e == hipSuccess)
    e = hipMalloc((void**)&a->vjp, sizeof(double) * (9 * (size_t)field->N + 9 * field->nsys));
  if (e == hipSuccess) e = hipMalloc((void**)&a->row_ptr, sizeof(int32_t) * row_ptr.size());
  if (e == hipSuccess)
    e = hipMalloc((void**)&a->row_idx, sizeof(int32_t) * (row_idx.empty() ? 1 : row_idx.size()));
  if (e == hipSuccess)
    e = hipMemcpy(a->row_ptr, row_ptr.data(), sizeof(int32_t) * row_ptr.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess && !row_idx.empty())
    e = hipMemcpy(a->row_idx, row_idx.data(), sizeof(int32_t) * row_idx.size(),
                  hipMemcpyHostToDevice);
  // rows of states that a type does not have are never written: they stay zero
  if (e == hipSuccess) e = hipMemset(a->ppart, 0, sizeof(double) * 2 * a->npar * nb);
  (void)hipSetDevice(cur);
  if (e != hipSuccess) {
    mm_adjoint_destroy(a);
    return hip_error(e, "mm_adjoint_create");
  }
  if (npar) *npar = a->npar;
  *out = a;
  return MM_OK;
}

int mm_adjoint_compute(mm_adjoint* a, const double* pos, const double* dpos, const double* rvecs,
                       const double* drvecs, const double* energy_weight, double* sys_out,
                       double* gpos, double* dgpos, double* params, void* stream) {
  if (!a) return set_error(MM_ERR_ARG, "adjoint is NULL");
  mm_field* f = a->field;
  if ((f->N > 0 && !pos) || !rvecs || !sys_out || !params)
    return set_error(MM_ERR_ARG, "missing argument");
  if ((dpos == nullptr) != (drvecs == nullptr))
    return set_error(MM_ERR_ARG, "dpos and drvecs go together");
  if (dgpos && !dpos) return set_error(MM_ERR_ARG, "dgpos needs a tangent direction");
  const bool dual = dpos != nullptr;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(f->device);
  hipStream_t s = (hipStream_t)stream;
  const int64_t nblk = a->nblk;
  hipError_t e = hipSuccess;
  if (nblk > 0) {
#define MM_ADJ(T, MODEL)                                                                        \
  hipLaunchKernelGGL((adjoint_kernel<T, MODEL>), dim3((unsigned)nblk), dim3(256), 0, s, pos,    \
                     dpos, rvecs, drvecs, f->blk_sys, f->cnode, f->cwrap, f->blk_row, f->tab,  \
                     f->kt, f->nstates, f->max_states, f->W, f->gcell, a->gcell2, a->part,     \
                     a->ppart)
    if (dual) {
      if (f->model == 0) MM_ADJ(Dual, 0); else MM_ADJ(Dual, 1);
    } else {
      if (f->model == 0) MM_ADJ(double, 0); else MM_ADJ(double, 1);
    }
#undef MM_ADJ
    e = hipGetLastError();
    // per system: energy, virial, domain-vector gradient (value rows, then tangent rows)
    if (e == hipSuccess) {
      hipLaunchKernelGGL(segment_sum_kernel, dim3((unsigned)(f->nsys * 2 * kSysQ)), dim3(256), 0,
                         s, a->part, nblk, 2 * kSysQ, f->sys_cblk_ptr, f->sys_cblk_idx,
                         f->blk_sys, (const double*)nullptr, 1, sys_out);
      e = hipGetLastError();
    }
    // per table row: parameter entries (value rows scaled by the system's energy weight)
    if (e == hipSuccess) {
      hipLaunchKernelGGL(segment_sum_kernel, dim3((unsigned)(f->nrows * 2 * a->npar)), dim3(256),
                         0, s, a->ppart, nblk, (int)(2 * a->npar), a->row_ptr, a->row_idx,
                         f->blk_sys, energy_weight, 2, params);
      e = hipGetLastError();
    }
    if (e == hipSuccess && f->N > 0 && gpos) {
      ForceArgs g{};
      g.ncell = f->ncell;
      g.gcell = f->gcell;
      g.rstride = f->W;
      g.rshift = 3;
      g.N = f->N;
      g.W = f->W;
      g.gpos = gpos;
      g.accumulate = 0;
      g.nanflag = f->nanflag;
      e = launch_gather(g, s);
      if (e == hipSuccess && dgpos) {
        g.gcell = a->gcell2;
        g.gpos = dgpos;
        e = launch_gather(g, s);
      }
    }
  } else {
    e = hipMemsetAsync(sys_out, 0, sizeof(double) * 2 * kSysQ * f->nsys, s);
    if (e == hipSuccess)
      e = hipMemsetAsync(params, 0, sizeof(double) * 2 * a->npar * f->nrows, s);
  }
  (void)hipSetDevice(cur);
  if (e != hipSuccess) return hip_error(e, "adjoint kernels");
  return MM_OK;
}

int mm_adjoint_vjp(mm_adjoint* a, const double* pos, const double* rvecs,
                   const int32_t* node_sys, const double* gE, const double* gG, const double* gV,
                   double* sys_out, double* g_pos, double* params, void* stream) {
  if (!a) return set_error(MM_ERR_ARG, "adjoint is NULL");
  mm_field* f = a->field;
  if ((f->N > 0 && (!pos || !node_sys || !g_pos)) || !rvecs || !gE || !sys_out || !params)
    return set_error(MM_ERR_ARG, "missing argument (gE is required: zeros for none)");
  double* dpos = a->vjp;
  double* gtmp = dpos + 3 * f->N;
  double* dgtmp = gtmp + 3 * f->N;
  double* drv = dgtmp + 3 * f->N;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(f->device);
  hipStream_t s = (hipStream_t)stream;
  const int64_t nt = f->N + 9 * (int64_t)f->nsys;
  hipLaunchKernelGGL(vjp_direction_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s,
                     pos, rvecs, node_sys, gG, gV, f->N, f->nsys, dpos, drv);
  hipError_t e = hipGetLastError();
  (void)hipSetDevice(cur);
  if (e != hipSuccess) return hip_error(e, "vector-Jacobian direction kernel");
  const int rc =
      mm_adjoint_compute(a, pos, dpos, rvecs, drv, gE, sys_out, gtmp, dgtmp, params, stream);
  if (rc != MM_OK) return rc;
  if (f->N > 0) {
    (void)hipSetDevice(f->device);
    hipLaunchKernelGGL(vjp_position_kernel, dim3((unsigned)((f->N + 255) / 256)), dim3(256), 0, s,
                       gtmp, dgtmp, node_sys, gE, gV, f->N, g_pos);
    e = hipGetLastError();
    (void)hipSetDevice(cur);
    if (e != hipSuccess) return hip_error(e, "vector-Jacobian position kernel");
  }
  return MM_OK;
}

int mm_adjoint_param_grads(mm_adjoint* a, const double* entries, int64_t stride, int add_next,
                           const double* h0, double boltzmann, double* g_cv, double* g_h0,
                           double* g_free, double* g_teff, void* stream) {
  if (!a) return set_error(MM_ERR_ARG, "adjoint is NULL");
  if (!entries || !h0 || !g_cv || !g_h0 || !g_free || !g_teff)
    return set_error(MM_ERR_ARG, "missing argument");
  if (stride < 1 || (add_next && stride < 2))
    return set_error(MM_ERR_ARG, "stride %lld with add_next %d", (long long)stride, add_next);
  mm_field* f = a->field;
  const int n = f->nrows * f->max_states;
  if (n == 0) return MM_OK;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(f->device);
  hipLaunchKernelGGL(param_grads_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, entries, stride, add_next, a->npar, f->nrows,
                     f->max_states, f->nstates, h0, f->model == MM_MODEL_CORNER8 ? 1 : 0,
                     boltzmann, g_cv, g_h0, g_free, g_teff);
  const hipError_t e = hipGetLastError();
  (void)hipSetDevice(cur);
  if (e != hipSuccess) return hip_error(e, "parameter gradient kernel");
  return MM_OK;
}

void mm_adjoint_destroy(mm_adjoint* a) {
  if (!a) return;
  if (a->part) (void)hipFree(a->part);
  if (a->ppart) (void)hipFree(a->ppart);
  if (a->gcell2) (void)hipFree(a->gcell2);
  if (a->row_ptr) (void)hipFree(a->row_ptr);
  if (a->row_idx) (void)hipFree(a->row_idx);
  if (a->vjp) (void)hipFree(a->vjp);
  delete a;
}

}  // extern "C"
