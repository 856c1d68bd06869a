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

// Does any pair of the 27 probes share a bucket?  Probes (a,b,c) and
// (a',b',c') collide iff (X[a]^X[a']) ^ (Y[b]^Y[b']) ^ (Z[c]^Z[c']) has no
// bits under the mask, so 63 XOR tests over the per-axis differences replace
// the 351 pairwise comparisons.
__device__ __forceinline__ bool probe_any_dup(const Cell& c, uint32_t mask) {
  uint32_t X[3], Y[3], Z[3];
  axis_terms(c.x, FP_K1, X);
  axis_terms(c.y, FP_K2, Y);
  axis_terms(c.z, FP_K3, Z);
  const uint32_t dx[4] = {0u, (X[0] ^ X[1]) & mask, (X[0] ^ X[2]) & mask, (X[1] ^ X[2]) & mask};
  const uint32_t dy[4] = {0u, (Y[0] ^ Y[1]) & mask, (Y[0] ^ Y[2]) & mask, (Y[1] ^ Y[2]) & mask};
  const uint32_t dz[4] = {0u, (Z[0] ^ Z[1]) & mask, (Z[0] ^ Z[2]) & mask, (Z[1] ^ Z[2]) & mask};
  bool hit = false;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int d = 0; d < 4; d++)
        if (a + b + d > 0) hit |= ((dx[a] ^ dy[b] ^ dz[d]) == 0u);
  return hit;
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

// ---- the cell grid of the hot path -------------------------------------------

// The hot path sorts particles by cells CELL_SUB times finer than the
// reference's cellSize along each axis (1: the reference's cells).  The
// interaction radius is at most one reference cell, so a particle's
// neighbours lie in the NB_RUN^3 block of sort cells around its own: NB_ROWS
// rows along z, each a run of NB_RUN consecutive keys.  Half cells (2) test
// half the candidates (a sphere fills 40% of a 5^3 block of half cells, 20%
// of a 3^3 block of whole ones) but visit 2.7x the rows and scan 8x the
// keys; on MI355X whole cells measured faster for every workload here.  The
// reference's own cell (ref_inv) decides hashing either way: buckets, probe
// multiplicities, the moved and duplicate-probe flags.
#ifndef CELL_SUB
#define CELL_SUB 1
#endif
#define NB_R CELL_SUB                 // neighbour radius in sort cells
#define NB_RUN (2 * NB_R + 1)         // sort cells of one neighbour row
#define NB_ROWS (NB_RUN * NB_RUN)     // neighbour rows

// Cell keys of the hot path's counting sort.  Coordinates beyond +-2^30 (and
// NaN) share one sentinel cell, so such particles still meet each other.
struct CellGrid {
  double inv_cell;     // 1 / sort-cell size
  double ref_inv;      // 1 / the reference's cellSize
  uint32_t hmask;      // reference hash table: T - 1
  uint32_t M, kmask;   // cell table size (power of 2), M - 1
  uint32_t mx, my;     // key multipliers of the x and y cell coordinates
};

__device__ __forceinline__ uint32_t cell_coord_u32(double c) {
  return fabs(c) < 1073741824.0 ? (uint32_t)(int32_t)c : 0x40000000u;
}

// key = (cx * mx + cy * my + cz) mod M (32-bit wrap): the z-neighbours of a
// cell are consecutive keys.  With mx = NY*NZ, my = NZ and M covering
// the walls every cell of the basin has its own key; with odd multipliers
// and a smaller M (small scenes) columns of cells land at scattered offsets
// and the few that overlap are told apart by the distance tests.
__device__ __forceinline__ uint32_t cell_key(const CellGrid& g, const Cell& c) {
  return (cell_coord_u32(c.x) * g.mx + cell_coord_u32(c.y) * g.my + cell_coord_u32(c.z)) & g.kmask;
}

// First key of the z-run of neighbour row (a, b) in [-NB_R, NB_R]^2 of key k.
__device__ __forceinline__ uint32_t row_key(const CellGrid& g, uint32_t k, int a, int b) {
  return (k + (uint32_t)a * g.mx + (uint32_t)b * g.my - (uint32_t)NB_R) & g.kmask;
}

// ---- the constraint CSR -----------------------------------------------------
//
// An entry of solid a's row packs the other endpoint as an offset from a
// (24-bit two's complement: lattice neighbours of a cube the 32-bit CSR can
// hold lie within n^2 + n + 1 < 2^23) and max(deg_a, deg_b) in the top byte;
// the constraint's scale 1 / max(deg) is read from a table of correctly
// rounded reciprocals (the host's float division gives the same values).
__device__ __forceinline__ int adj_other(int a, uint32_t pk) { return a + ((int)(pk << 8) >> 8); }
__device__ static const float kInvDeg[32] = {
    0.f,        1.f / 1,  1.f / 2,  1.f / 3,  1.f / 4,  1.f / 5,  1.f / 6,  1.f / 7,
    1.f / 8,    1.f / 9,  1.f / 10, 1.f / 11, 1.f / 12, 1.f / 13, 1.f / 14, 1.f / 15,
    1.f / 16,   1.f / 17, 1.f / 18, 1.f / 19, 1.f / 20, 1.f / 21, 1.f / 22, 1.f / 23,
    1.f / 24,   1.f / 25, 1.f / 26, 1.f / 27, 1.f / 28, 1.f / 29, 1.f / 30, 1.f / 31};
__device__ __forceinline__ float adj_scale(uint32_t pk) { return kInvDeg[(pk >> 24) & 31u]; }

// ---- global-memory access -------------------------------------------------------
//
// The kernels read their arrays' addresses from a device table of scenes, so
// the compiler cannot see that they point to global memory and emits flat
// loads and stores.  A flat access also counts against the counter that
// scalar loads wait on: every reload of a scene parameter inside a loop then
// waits for all the vector loads in flight, prefetches included.  gld / gst
// access through the global address space instead.
#define FP_GLOBAL __attribute__((address_space(1)))
typedef float fp_f4v __attribute__((ext_vector_type(4)));
typedef float fp_f2v __attribute__((ext_vector_type(2)));
typedef int fp_i4v __attribute__((ext_vector_type(4)));
typedef int fp_i2v __attribute__((ext_vector_type(2)));
typedef float fp_f3v __attribute__((ext_vector_type(3)));

template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const FP_GLOBAL T*)p;
}
__device__ __forceinline__ float4 gld(const float4* p) {
  const fp_f4v v = *(const FP_GLOBAL fp_f4v*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 gld(const float2* p) {
  const fp_f2v v = *(const FP_GLOBAL fp_f2v*)p;
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ int2 gld(const int2* p) {
  const fp_i2v v = *(const FP_GLOBAL fp_i2v*)p;
  return make_int2(v.x, v.y);
}
__device__ __forceinline__ int4 gld(const int4* p) {
  const fp_i4v v = *(const FP_GLOBAL fp_i4v*)p;
  return make_int4(v.x, v.y, v.z, v.w);
}
// x, y, z of a float4 (w = 0) with a 12-byte load: a 16-byte load whose w is
// dead still writes its register, and the compiler then waits for the load
// before it can reuse that register.
__device__ __forceinline__ float4 gld_xyz(const float4* p) {
  const fp_f3v v = *(const FP_GLOBAL fp_f3v*)p;
  return make_float4(v.x, v.y, v.z, 0.f);
}
// x, y, z of a float4 with a 12-byte store (w stays as it is in memory).
__device__ __forceinline__ void gst_xyz(float4* p, const float4& v) {
  fp_f3v t;
  t.x = v.x;
  t.y = v.y;
  t.z = v.z;
  *(FP_GLOBAL fp_f3v*)p = t;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(FP_GLOBAL T*)p = v;
}
__device__ __forceinline__ void gst(float4* p, const float4& v) {
  fp_f4v w = {v.x, v.y, v.z, v.w};
  *(FP_GLOBAL fp_f4v*)p = w;
}
__device__ __forceinline__ void gst(int4* p, const int4& v) {
  fp_i4v w = {v.x, v.y, v.z, v.w};
  *(FP_GLOBAL fp_i4v*)p = w;
}
__device__ __forceinline__ void gst(float2* p, const float2& v) {
  fp_f2v w = {v.x, v.y};
  *(FP_GLOBAL fp_f2v*)p = w;
}

// base + a 32-bit byte offset.  With the base the same in every lane (an
// array of the scene table, a wave's block of the pair list) the access is one
// instruction with the base in scalar registers and the offset in one vector
// register: no 64-bit address arithmetic per access.
#define FP_AT(T, base, off) ((const FP_GLOBAL T*)((const FP_GLOBAL char*)(base) + (uint32_t)(off)))
__device__ __forceinline__ float4 gld_at(const float4* base, uint32_t off) {
  const fp_f4v v = *FP_AT(fp_f4v, base, off);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 gld_xyz_at(const float4* base, uint32_t off) {
  const fp_f3v v = *FP_AT(fp_f3v, base, off);
  return make_float4(v.x, v.y, v.z, 0.f);
}
__device__ __forceinline__ float2 gld_at(const float2* base, uint32_t off) {
  const fp_f2v v = *FP_AT(fp_f2v, base, off);
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ uint32_t gld_at(const uint32_t* base, uint32_t off) { return *FP_AT(uint32_t, base, off); }
__device__ __forceinline__ void gst_at(uint32_t* base, uint32_t off, uint32_t v) {
  *(FP_GLOBAL uint32_t*)((FP_GLOBAL char*)base + off) = v;
}

// ---- solids' bounding box ------------------------------------------------------

// Floats as ints whose signed order is the floats' order.
__device__ __forceinline__ int order_float(float f) {
  const int b = __float_as_int(f);
  return b >= 0 ? b : b ^ 0x7fffffff;
}

// Fold the solid positions of this wave into box[0..5] (min xyz, max xyz,
// ordered-int bits): a wave64 min/max, then one atomic per wave.  Every lane
// of the wave must call it.
__device__ __forceinline__ void box_update(bool has, float px, float py, float pz, int* __restrict__ box) {
  if (!__any(has)) return;
  float mnx = has ? px : INFINITY, mny = has ? py : INFINITY, mnz = has ? pz : INFINITY;
  float mxx = has ? px : -INFINITY, mxy = has ? py : -INFINITY, mxz = has ? pz : -INFINITY;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mnx = fminf(mnx, __shfl_xor(mnx, off));
    mny = fminf(mny, __shfl_xor(mny, off));
    mnz = fminf(mnz, __shfl_xor(mnz, off));
    mxx = fmaxf(mxx, __shfl_xor(mxx, off));
    mxy = fmaxf(mxy, __shfl_xor(mxy, off));
    mxz = fmaxf(mxz, __shfl_xor(mxz, off));
  }
  if ((threadIdx.x & 63) == 0) {
    FP_GLOBAL int* gb = (FP_GLOBAL int*)box;
    __hip_atomic_fetch_min(gb + 0, order_float(mnx), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_min(gb + 1, order_float(mny), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_min(gb + 2, order_float(mnz), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_max(gb + 3, order_float(mxx), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_max(gb + 4, order_float(mxy), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_max(gb + 5, order_float(mxz), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The same over a block of NT threads: wave min/max, then the waves' results
// meet in LDS and one thread per box word sends the block's atomic.  Every
// thread of the block must call it (it synchronizes the block).
template <int NT>
__device__ __forceinline__ void box_update_block(bool has, float px, float py, float pz, int* __restrict__ box) {
  __shared__ float red[NT / 64][6];
  if (!__syncthreads_or(has ? 1 : 0)) return;
  float v[6] = {has ? px : INFINITY, has ? py : INFINITY, has ? pz : INFINITY,
                has ? px : -INFINITY, has ? py : -INFINITY, has ? pz : -INFINITY};
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int c = 0; c < 6; c++) {
      const float o = __shfl_xor(v[c], off);
      v[c] = c < 3 ? fminf(v[c], o) : fmaxf(v[c], o);
    }
  }
  const int wid = (int)threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int c = 0; c < 6; c++) red[wid][c] = v[c];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int c = (int)threadIdx.x;
    float m = red[0][c];
#pragma unroll
    for (int w = 1; w < NT / 64; w++) m = c < 3 ? fminf(m, red[w][c]) : fmaxf(m, red[w][c]);
    FP_GLOBAL int* gb = (FP_GLOBAL int*)box;
    if (c < 3)
      __hip_atomic_fetch_min(gb + c, order_float(m), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      __hip_atomic_fetch_max(gb + c, order_float(m), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
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

// A slot in bucket `key` of a counting sort: atomicAdd(&count[key], 1), but
// one atomic per run of equal keys in consecutive lanes (lanes with
// has == false take part in the vote and get nothing).  Every lane of the
// wave must call it.  Slots inside a bucket come out in no fixed order.
// count[key] += 1 for every lane with has, one atomic per run of equal keys
// in consecutive lanes (no result).  Every lane of the wave must call it.
__device__ __forceinline__ void wave_bucket_add(int* __restrict__ count, uint32_t key, bool has) {
  const int lane = (int)(threadIdx.x & 63);
  const uint32_t kprev = (uint32_t)__shfl_up((int)key, 1);
  const bool hprev = __shfl_up(has ? 1 : 0, 1) != 0;
  const bool head = has && (lane == 0 || !hprev || kprev != key);
  const uint64_t heads = __ballot(head);
  const uint64_t stops = heads | ~__ballot(has);
  const uint64_t above = lane == 63 ? 0ull : (stops & (~0ull << (lane + 1)));
  const int next = above ? (int)__builtin_ctzll(above) : 64;
  if (head)
    (void)__hip_atomic_fetch_add((FP_GLOBAL int*)&count[key], next - lane, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_bucket_slot(int* __restrict__ count, uint32_t key, bool has) {
  const int lane = (int)(threadIdx.x & 63);
  const uint32_t kprev = (uint32_t)__shfl_up((int)key, 1);
  const bool hprev = __shfl_up(has ? 1 : 0, 1) != 0;
  const bool head = has && (lane == 0 || !hprev || kprev != key);
  // a run ends before the next head or the next lane without a key
  const uint64_t heads = __ballot(head);
  const uint64_t stops = heads | ~__ballot(has);
  const uint64_t above = lane == 63 ? 0ull : (stops & (~0ull << (lane + 1)));
  const int next = above ? (int)__builtin_ctzll(above) : 64;
  int base = 0;
  if (head)
    base = __hip_atomic_fetch_add((FP_GLOBAL int*)&count[key], next - lane, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
  const int hl = upto ? 63 - (int)__builtin_clzll(upto) : 0;
  const int b = __shfl(base, hl);
  return (uint32_t)(b + (lane - hl));
}
