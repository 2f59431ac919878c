// Device-side building blocks of the nanocell force field (CDNA4 / gfx950, FP64).
//
// Geometry conventions follow micmec/pes/nanocell.py:40-132 and micmec/utils.py:43-52:
//   corner k of a cell sits at lattice offset OFF[k] = (000,100,010,001,110,101,011,111);
//   the corner matrix M_a has as rows the three edges through corner a, each oriented along
//   +x, +y, +z (micmec/pes/nanocell_utils.py:32-75);
//   B_a = h0^-1 M_a, eps_a = (B_a B_a^T - I)/2 (Green-Lagrange), u_a = eps_a : C0 : eps_a / 2.
// The gradient is the adjoint of that chain: dE/dB = S B with S the symmetric stress built from
// the Voigt vector C e, and dE/dM = h0^-T dE/dB; it equals the derivative-table route of
// nanocell.py:98-132.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mm {

// Doubles per (type, state) entry of the parameter table.
//   [0..8]   P = h0^-1 (row-major)
//   [9]      w  = det(h0) / 8 (corner8) or det(h0) (original)
//   [10]     free energy of the state
//   [11]     unused
//   [12..47] Voigt matrix Cv[6][6] (row-major), averaged over index orderings
//   [48..83] w * Cv, the weighted matrix of the force kernels
constexpr int kTabStride = 84;
constexpr int kCvOffset = 12;
constexpr int kCvwOffset = 48;

// 3x3 inverse (row-major) by cofactors; returns the determinant. Shared by the host and the
// device table builders, without contraction, so both round the same way.
__host__ __device__ inline double inverse3_cof(const double* h, double* inv) {
#pragma clang fp contract(off)
  const double a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7],
               l = h[8];
  const double c00 = e * l - f * k, c01 = f * g - d * l, c02 = d * k - e * g;
  const double det = a * c00 + b * c01 + c * c02;
  inv[0] = c00 / det;
  inv[1] = (c * k - b * l) / det;
  inv[2] = (b * f - c * e) / det;
  inv[3] = c01 / det;
  inv[4] = (a * l - c * g) / det;
  inv[5] = (c * d - a * f) / det;
  inv[6] = c02 / det;
  inv[7] = (b * g - a * k) / det;
  inv[8] = (a * e - b * d) / det;
  return det;
}

// Edge endpoints: corner a, axis d -> lower / upper corner of the edge along d through a.
__device__ constexpr int kLo[8][3] = {
    {0, 0, 0}, {0, 1, 1}, {2, 0, 2}, {3, 3, 0}, {2, 1, 4}, {3, 5, 1}, {6, 3, 2}, {6, 5, 4}};
__device__ constexpr int kHi[8][3] = {
    {1, 2, 3}, {1, 4, 5}, {4, 2, 6}, {5, 6, 3}, {4, 4, 7}, {5, 7, 5}, {7, 6, 6}, {7, 7, 7}};
// Corner offsets along x, y, z.
__device__ constexpr int kOff[8][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
// The twelve edges: four along x, y and z each; edge kEdge[a][d] is the edge along d through
// corner a, from corner kEdgeLo to corner kEdgeHi (so kEdgeLo/Hi[kEdge[a][d]] = kLo/kHi[a][d]).
__device__ constexpr int kEdgeLo[12] = {0, 2, 3, 6, 0, 1, 3, 5, 0, 1, 2, 4};
__device__ constexpr int kEdgeHi[12] = {1, 4, 5, 7, 2, 4, 6, 7, 3, 5, 6, 7};
__device__ constexpr int kEdge[8][3] = {{0, 4, 8},  {0, 5, 9},  {1, 4, 10}, {2, 6, 8},
                                        {1, 5, 11}, {2, 7, 9},  {3, 6, 10}, {3, 7, 11}};

// Strain energy density of one corner (or of the averaged matrix) and, if kGrad, the
// derivative with respect to the corner matrix M, both for one (type, state) entry `tab`.
// M is given row-major (rows = edges). Returns u = e.Cv.e / 2; writes G = dU/dM (unscaled).
// kVir: also adds the corner's virial G^T M = (S B)^T B = B^T S B (P M = B; symmetric) to
// V[6] = (00, 11, 22, 12, 02, 01) — the sum over its edges of dU/d(edge) (x) edge, which is what
// the vertex gradients give as sum_v g_v (x) r_v — without forming G when kGrad is false.
template <bool kGrad, bool kVir = false>
__device__ __forceinline__ double corner_energy(const double* __restrict__ tab, const double M[9],
                                                double G[9], int cv_offset = kCvOffset,
                                                double* V = nullptr) {
  const double* P = tab;
  double B[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      B[3 * i + j] = P[3 * i + 0] * M[0 + j] + P[3 * i + 1] * M[3 + j] + P[3 * i + 2] * M[6 + j];
    }
  }
  // Q = B B^T; engineering strain e = (eps00, eps11, eps22, 2eps12, 2eps02, 2eps01)
  double e[6];
  e[0] = 0.5 * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2] - 1.0);
  e[1] = 0.5 * (B[3] * B[3] + B[4] * B[4] + B[5] * B[5] - 1.0);
  e[2] = 0.5 * (B[6] * B[6] + B[7] * B[7] + B[8] * B[8] - 1.0);
  e[3] = B[3] * B[6] + B[4] * B[7] + B[5] * B[8];
  e[4] = B[0] * B[6] + B[1] * B[7] + B[2] * B[8];
  e[5] = B[0] * B[3] + B[1] * B[4] + B[2] * B[5];
  const double* Cv = tab + cv_offset;
  double s[6];
#pragma unroll
  for (int I = 0; I < 6; ++I) {
    s[I] = Cv[6 * I + 0] * e[0] + Cv[6 * I + 1] * e[1] + Cv[6 * I + 2] * e[2] +
           Cv[6 * I + 3] * e[3] + Cv[6 * I + 4] * e[4] + Cv[6 * I + 5] * e[5];
  }
  const double u = 0.5 * (e[0] * s[0] + e[1] * s[1] + e[2] * s[2] + e[3] * s[3] + e[4] * s[4] +
                          e[5] * s[5]);
  if (kGrad || kVir) {
    // S (symmetric): S00=s0 S11=s1 S22=s2 S12=s3 S02=s4 S01=s5 ; T = S B ; G = P^T T
    const double S[9] = {s[0], s[5], s[4], s[5], s[1], s[3], s[4], s[3], s[2]};
    double T[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        T[3 * i + j] = S[3 * i + 0] * B[0 + j] + S[3 * i + 1] * B[3 + j] + S[3 * i + 2] * B[6 + j];
      }
    }
    if (kGrad) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          G[3 * i + j] = P[0 + i] * T[0 + j] + P[3 + i] * T[3 + j] + P[6 + i] * T[6 + j];
        }
      }
    }
    if (kVir) {
      // (T^T B)_kl = sum_i T_ik B_il
      constexpr int kK[6] = {0, 1, 2, 1, 0, 0}, kL[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
      for (int q = 0; q < 6; ++q)
        V[q] += T[0 + kK[q]] * B[0 + kL[q]] + T[3 + kK[q]] * B[3 + kL[q]] + T[6 + kK[q]] * B[6 + kL[q]];
    }
  }
  return u;
}

// Energy of one metastable state of a cell (without the free-energy offset) and, if kGrad,
// the gradient with respect to the eight vertices, accumulated (scaled by gscale) into g[8][3].
// The weight w enters through the weighted Voigt matrix w Cv, so the corner terms come out
// already multiplied by w. kUnit: gscale is 1 (single-state cells) and is not applied.
// kModel: 0 = corner8 (nanocell.py), 1 = original (nanocell_original.py).
// kVir: also adds the state's virial sum_v g_v (x) r_v, as the sum of its corners' B^T S B
// (see corner_energy; for the averaged matrix H the vertex rows are +-1/4 of dU/dH and
// sum_v +-r_v along d is 4 H_d, so the same form), to V[6] — unscaled by gscale.
template <int kModel, bool kGrad, bool kUnit = false, bool kVir = false>
__device__ __forceinline__ double state_energy(const double* __restrict__ tab, const double r[8][3],
                                               double g[8][3], double gscale = 1.0,
                                               double* V = nullptr) {
  if (kModel == 0) {
    // the twelve edges once (each is a row of two corner matrices)
    double E[12][3];
#pragma unroll
    for (int e = 0; e < 12; ++e)
#pragma unroll
      for (int j = 0; j < 3; ++j) E[e][j] = r[kEdgeHi[e]][j] - r[kEdgeLo[e]][j];
    double usum = -0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      double M[9], G[9];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int j = 0; j < 3; ++j) M[3 * d + j] = E[kEdge[a][d]][j];
      }
      usum += corner_energy<kGrad, kVir>(tab, M, G, kCvwOffset, V);
      if (kGrad) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const double v = kUnit ? G[3 * d + j] : gscale * G[3 * d + j];
            g[kHi[a][d]][j] += v;
            g[kLo[a][d]][j] -= v;
          }
        }
      }
    }
    return usum;
  } else {
    // H = (1/8) sum_a M_a: row d is a quarter of the sum of the four edges along d.
    double H[9], G[9];
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
    const double u = corner_energy<kGrad, kVir>(tab, H, G, kCvwOffset, V);
    if (kGrad) {
      const double q = kUnit ? 0.25 : 0.25 * gscale;
#pragma unroll
      for (int v = 0; v < 8; ++v) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const double t = q * G[3 * d + j];
            g[v][j] += kOff[v][d] ? t : -t;
          }
        }
      }
    }
    return u;
  }
}

template <int NV>
__device__ __forceinline__ void wave_reduce(double (&v)[NV]) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] += __shfl_down(v[q], off, 64);
  }
}

// Row vector times 3x3 matrix, without contractions: the position and velocity transforms of
// the barostat (npt.py:700-702, 725-726) give the same bits in every kernel that applies them.
__device__ __forceinline__ void vecmat(const double v[3], const double* M, double out[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int j = 0; j < 3; ++j) out[j] = (v[0] * M[j] + v[1] * M[3 + j]) + v[2] * M[6 + j];
}

// g = the sum of node n's records in k order (mmff.py:303-318): record e = ncell[k * N + n]
// (none when e < 0) holds the three components gcell[(3 * (e & (2^rshift - 1)) + j) * RS +
// (e >> rshift)]; g starts at +0.0 and takes the records present in k order.
// kAllLoads: the eight indices are loaded first and then all 24 components, a none reading those
// of record 0 (always allocated) and left out of the sum, so that every load of the lane is in
// flight at once (two memory round trips instead of up to sixteen), at the cost of about 50
// more VGPRs; the same bits as skipping the nones. Without it each record is loaded behind
// the test of its index, which keeps the registers few: the form for passes with enough
// waves to hide the latency.
template <bool kAllLoads>
__device__ __forceinline__ void sum_node_records(const int32_t* __restrict__ ncell,
                                                 const double* __restrict__ gcell, int64_t N,
                                                 int64_t RS, int rshift, int64_t n,
                                                 double g[3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j) g[j] = 0.0;
  if (!kAllLoads) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int32_t e = __builtin_nontemporal_load(&ncell[k * N + n]);
      if (e >= 0) {
        const int64_t w = e >> rshift;
        const int p = e & ((1 << rshift) - 1);
#pragma unroll
        for (int j = 0; j < 3; ++j)
          g[j] += __builtin_nontemporal_load(&gcell[(int64_t)(3 * p + j) * RS + w]);
      }
    }
    return;
  }
  int32_t e[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) e[k] = __builtin_nontemporal_load(&ncell[k * N + n]);
  double r[8][3];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int32_t c = e[k] >= 0 ? e[k] : 0;
    const int64_t w = c >> rshift;
    const int p = c & ((1 << rshift) - 1);
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r[k][j] = __builtin_nontemporal_load(&gcell[(int64_t)(3 * p + j) * RS + w]);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) g[j] = e[k] >= 0 ? g[j] + r[k][j] : g[j];
}

// Workgroups are dealt to the 8 XCDs of the device in turn (block b runs on XCD b mod 8), and
// every XCD has its own L2. Neighbouring blocks share cache lines (the cell segments share
// nodes, the node blocks' record gathers share the lines at their ends), so each XCD is given
// one contiguous range of the n blocks: block b -> (b mod 8) * (n / 8) + min(b mod 8, n mod 8)
// + b / 8. The map is a bijection of [0, n) for any n.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int q = n / kXcds, r = n % kXcds, x = b % kXcds, i = b / kXcds;
  return x * q + (x < r ? x : r) + i;
}

// 64-bit DPP move (two 32-bit moves): the lanes that kBanks enables take v from the lane kCtrl
// names (row_shl:N reads lane + N, row_shr:N lane - N, quad_perm within four lanes), the others
// keep old. A register-to-register exchange, no LDS round trip.
template <int kCtrl, int kBanks>
__device__ __forceinline__ double dpp_f64(double old, double v) {
  const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, kCtrl, 0xF, kBanks, false);
  const int hi =
      __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), kCtrl, 0xF, kBanks, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}

// Row swap of two registers (v_permlane32_swap / v_permlane16_swap): with kDist 32 lanes 32-63
// of x trade places with lanes 0-31 of y, with kDist 16 the odd 16-lane rows of x with the even
// rows of y. Afterwards a lane whose bit kDist is clear holds its own x in x and its partner's
// (lane ^ kDist) x in y; a lane whose bit is set holds its own y in y and its partner's y in x.
template <int kDist>
__device__ __forceinline__ void row_swap(double& x, double& y) {
  static_assert(kDist == 16 || kDist == 32, "row swaps across 16 or 32 lanes");
  const unsigned long long xb = (unsigned long long)__double_as_longlong(x);
  const unsigned long long yb = (unsigned long long)__double_as_longlong(y);
  const unsigned xl = (unsigned)xb, xh = (unsigned)(xb >> 32);
  const unsigned yl = (unsigned)yb, yh = (unsigned)(yb >> 32);
  unsigned xl2, xh2, yl2, yh2;
  if (kDist == 32) {
    const auto lo = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
    xl2 = lo[0], yl2 = lo[1], xh2 = hi[0], yh2 = hi[1];
  } else {
    const auto lo = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
    xl2 = lo[0], yl2 = lo[1], xh2 = hi[0], yh2 = hi[1];
  }
  x = __longlong_as_double((long long)(((unsigned long long)xh2 << 32) | xl2));
  y = __longlong_as_double((long long)(((unsigned long long)yh2 << 32) | yl2));
}

// Wave sums of up to 16 values by a reduce-scatter butterfly: returns in lane l the wave total
// of value q(l) = 8 b5 + 4 b4 + 2 b3 + b2 (bk = bit k of l); values q >= NV are zero. At each
// level a lane keeps the upper half of its values when its bit is set, the lower half
// otherwise, and adds its partner's copy of them: across 32 and 16 lanes by row swaps, across
// 8 and 4 lanes by two bank-masked DPP row shifts, the last two levels (a plain butterfly on
// one value) by quad permutations. Every exchange stays in registers; the pairs and the order
// of every addition are fixed, so the sums are deterministic.
template <int NV>
__device__ __forceinline__ double wave_reduce_scatter(const double (&v)[NV], int lane) {
  static_assert(NV >= 1 && NV <= 16, "1 to 16 values");
  constexpr int L1 = NV < 8 ? NV : 8, L2 = L1 < 4 ? L1 : 4, L3 = L2 < 2 ? L2 : 2;
  double a[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) a[q] = q < NV ? v[q] : 0.0;
  // lane ^ 32, pairs (i, 8 + i); a pair of zeros stays zero
#pragma unroll
  for (int i = 0; i < L1; ++i) {
    double x = a[i], y = a[8 + i];
    row_swap<32>(x, y);
    a[i] = x + y;
  }
  // lane ^ 16, pairs (i, 4 + i)
#pragma unroll
  for (int i = 0; i < L2; ++i) {
    double x = a[i], y = a[4 + i];
    row_swap<16>(x, y);
    a[i] = x + y;
  }
  // lane ^ 8, pairs (i, 2 + i): lanes 0-7 of a row read lane + 8's a[i], lanes 8-15 lane - 8's
  // a[2 + i]
  const bool up8 = (lane & 8) != 0;
#pragma unroll
  for (int i = 0; i < L3; ++i) {
    const double t = dpp_f64<0x118, 0xC>(dpp_f64<0x108, 0x3>(0.0, a[i]), a[2 + i]);
    a[i] = (up8 ? a[2 + i] : a[i]) + t;
  }
  // lane ^ 4, the pair (0, 1): banks 0 and 2 read lane + 4, banks 1 and 3 lane - 4
  const bool up4 = (lane & 4) != 0;
  const double t = dpp_f64<0x114, 0xA>(dpp_f64<0x104, 0x5>(0.0, a[0]), a[1]);
  double s = (up4 ? a[1] : a[0]) + t;
  s += dpp_f64<0x4E, 0xF>(s, s);  // lane ^ 2
  s += dpp_f64<0xB1, 0xF>(s, s);  // lane ^ 1
  return s;
}

__device__ __forceinline__ int scatter_index(int lane) {
  return ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
}

// Sum SoA block partials part[q * nb + b] over b in a fixed order; one block of 1024 threads,
// every thread receives the totals.
template <int NV>
__device__ __forceinline__ void reduce_soa(const double* __restrict__ part, int nb,
                                           double (&out)[NV]) {
  __shared__ double red[NV * 16];
  __shared__ double tot[NV];
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.0;
  for (int bb = threadIdx.x; bb < nb; bb += 1024) {
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += part[(int64_t)q * nb + bb];
  }
  wave_reduce<NV>(acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) red[wave * NV + q] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = red[threadIdx.x];
    for (int w = 1; w < 16; ++w) s += red[w * NV + threadIdx.x];
    tot[threadIdx.x] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) out[q] = tot[q];
  __syncthreads();
}

}  // namespace mm
