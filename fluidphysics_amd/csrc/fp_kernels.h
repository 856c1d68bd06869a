// fp_kernels.h -- device helpers shared by the solver kernels (gfx950, wave64).
//
// Exactness rules carried over from the reference CPU solver (src/solver.js):
//  * cell coordinates are floor(double(p) * (1/cellSize)) in IEEE double
//    (solver.js:167-171, :254, :300);
//  * the hash is ((x*92837111) ^ (y*689287499) ^ (z*283923481)) >>> 0 & (T-1)
//    with the products taken in double and ToInt32 applied at `^`
//    (solver.js:162-164).  While |c*K| < 2^53 that equals the wrapped 32-bit
//    product; beyond it the double product is reduced modulo 2^32 exactly as
//    ECMAScript ToUint32 does, so the result is the reference's for every
//    input, NaN and infinities included (both hash to 0 on that axis).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FP_K1 92837111u
#define FP_K2 689287499u
#define FP_K3 283923481u

// |c| below this: c*K (K < 2^30) is exact in double, use 32-bit wrap arithmetic.
#define FP_FAST_CELL 8.0e6

// ToUint32(c * K) for an integer-valued double c.
__device__ __forceinline__ uint32_t js_axis_hash(double c, uint32_t K) {
  if (fabs(c) < FP_FAST_CELL) return (uint32_t)(int32_t)c * K;
  double t = c * (double)K;  // rounded like the JS number product
  if (!isfinite(t)) return 0u;
  double m = fmod(trunc(t), 4294967296.0);
  if (m < 0) m += 4294967296.0;
  return (uint32_t)m;
}

// The three axis terms of a cell and its two neighbours on that axis
// (offsets -1, 0, +1), exact for every c.
__device__ __forceinline__ void axis_terms(double c, uint32_t K, uint32_t out[3]) {
  if (fabs(c) < FP_FAST_CELL) {
    uint32_t b = (uint32_t)(int32_t)c * K;
    out[0] = b - K;
    out[1] = b;
    out[2] = b + K;
  } else {
    out[0] = js_axis_hash(c - 1.0, K);
    out[1] = js_axis_hash(c, K);
    out[2] = js_axis_hash(c + 1.0, K);
  }
}

struct Cell {
  double x, y, z;
};

__device__ __forceinline__ Cell cell_of(float px, float py, float pz, double inv) {
  Cell c;
  c.x = floor((double)px * inv);
  c.y = floor((double)py * inv);
  c.z = floor((double)pz * inv);
  return c;
}

__device__ __forceinline__ uint32_t cell_hash(const Cell& c, uint32_t mask) {
  return (js_axis_hash(c.x, FP_K1) ^ js_axis_hash(c.y, FP_K2) ^ js_axis_hash(c.z, FP_K3)) & mask;
}

// The 27 probe buckets of a cell in the reference's loop order
// (dx outer, dy, dz inner; solver.js:257-260).
__device__ __forceinline__ void probe_hashes(const Cell& c, uint32_t mask, uint32_t h[27]) {
  uint32_t X[3], Y[3], Z[3];
  axis_terms(c.x, FP_K1, X);
  axis_terms(c.y, FP_K2, Y);
  axis_terms(c.z, FP_K3, Z);
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int d = 0; d < 3; d++) h[a * 9 + b * 3 + d] = (X[a] ^ Y[b] ^ Z[d]) & mask;
}

// Bit k set when probe k hits the same bucket as an earlier probe.
__device__ __forceinline__ uint32_t probe_dup_mask(const uint32_t h[27]) {
  uint32_t m = 0;
#pragma unroll
  for (int k = 1; k < 27; k++) {
    bool d = false;
#pragma unroll
    for (int l = 0; l < k; l++) d |= (h[k] == h[l]);
    m |= d ? (1u << k) : 0u;
  }
  return m;
}

// Number of the cell's 27 probes that hash to bucket `target`.
__device__ __forceinline__ int probe_multiplicity(const Cell& c, uint32_t mask, uint32_t target) {
  uint32_t X[3], Y[3], Z[3];
  axis_terms(c.x, FP_K1, X);
  axis_terms(c.y, FP_K2, Y);
  axis_terms(c.z, FP_K3, Z);
  int n = 0;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
      for (int d = 0; d < 3; d++) n += (((X[a] ^ Y[b] ^ Z[d]) & mask) == target) ? 1 : 0;
  return n;
}

// ---- wave64 scan primitives ------------------------------------------------

// Inclusive prefix sum over the 64 lanes of a wave: DPP row_shr:1,2,4,8 with
// zero fill inside each 16-lane row, then the row totals carried across rows
// with readlane.
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);  // row_shr:8
  const int r0 = __builtin_amdgcn_readlane(v, 15);
  const int r1 = __builtin_amdgcn_readlane(v, 31);
  const int r2 = __builtin_amdgcn_readlane(v, 47);
  const int row = (int)(threadIdx.x & 63) >> 4;
  v += (row > 0 ? r0 : 0) + (row > 1 ? r1 : 0) + (row > 2 ? r2 : 0);
  return v;
}

__device__ __forceinline__ int wave_sum(int v) { return __builtin_amdgcn_readlane(wave_incl_scan(v), 63); }
