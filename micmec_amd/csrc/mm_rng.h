// Counter-based Gaussian noise for the stochastic hooks (Langevin thermostat and barostat).
//
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11): a keyed bijection of a 256-bit counter, ten rounds of two 64x64->128-bit products
// with the round keys bumped by Weyl constants. Every draw is a pure function of
// (key, counter), so a node's noise needs no generator state, no ordering between threads and
// no storage: the counter names the node (or system), the integration step, the hook call
// and the block, the key names the seed and the stream (thermostat or barostat).
//
// Each 64-bit output w gives a uniform u = ((w >> 12) + 1/2) 2^-52 in (0, 1) (exact in
// double precision), and two uniforms give two normals by the Box-Muller transform
// r = sqrt(-2 ln u_a), z = (r cos 2 pi u_b, r sin 2 pi u_b).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mm {

constexpr uint64_t kPhiloxM0 = 0xD2E7470EE14C6C93ull;
constexpr uint64_t kPhiloxM1 = 0xCA5A826395121157ull;
constexpr uint64_t kPhiloxW0 = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kPhiloxW1 = 0xBB67AE8584CAA73Bull;

// stream tags in the second key word
constexpr uint64_t kStreamThermostat = 1;
constexpr uint64_t kStreamBarostat = 2;

__device__ __forceinline__ void philox4x64_10(const uint64_t ctr_in[4], uint64_t k0, uint64_t k1,
                                              uint64_t out[4]) {
  uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t hi0 = __umul64hi(kPhiloxM0, c0), lo0 = kPhiloxM0 * c0;
    const uint64_t hi1 = __umul64hi(kPhiloxM1, c2), lo1 = kPhiloxM1 * c2;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

__device__ __forceinline__ double open_uniform(uint64_t w) {
  return ((double)(w >> 12) + 0.5) * 0x1.0p-52;
}

// Four standard normals from one Philox block.
__device__ __forceinline__ void normals4(const uint64_t ctr[4], uint64_t k0, uint64_t k1,
                                         double z[4]) {
#pragma clang fp contract(off)
  uint64_t w[4];
  philox4x64_10(ctr, k0, k1, w);
  const double two_pi = 6.283185307179586;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const double r = sqrt(-2.0 * log(open_uniform(w[2 * p])));
    const double t = two_pi * open_uniform(w[2 * p + 1]);
    z[2 * p] = r * cos(t);
    z[2 * p + 1] = r * sin(t);
  }
}

// 64-bit key word stored as two doubles holding its 32-bit halves.
__device__ __forceinline__ uint64_t seed_from(const double* s) {
  return (uint64_t)s[0] | ((uint64_t)s[1] << 32);
}

}  // namespace mm
