// fp_grid.hip -- the two grids.
//
// 1. The cell grid of the hot path (every substep).  Key of a cell c:
//      L(c) = (cx * mx + cy * my + cz) mod M      (32-bit wrap, M a power of 2)
//    so the three z-neighbours of a cell are consecutive keys: the 27
//    neighbour cells of a particle are 9 runs of 3 keys, i.e. 9 contiguous
//    ranges of the cell-sorted order.  Far cells that alias to the same key
//    are removed by the distance tests of the passes.  The build is a
//    counting sort: count (fused into integrate) -> exclusive scan ->
//    scatter in arrival order -> stable rank (ascending particle index
//    inside a cell), which also writes the sorted copies the contact pass
//    reads contiguously (sid = particle of each sorted slot, spos = position
//    + inverse mass, sphash = the particle's bucket in the reference's hash
//    table).  After contacts the fluid is sorted again by its post-contact
//    cell (fp_group_resort) for the fluid passes.
//
// 2. The reference's hash grid (_buildGrid, src/solver.js:174-184), built on
//    demand for fp_build_grid / fp_read_grid from the per-particle buckets
//    kept at grid time (phash): count -> the same scan -> scatter -> rank to
//    the reference's order (descending particle index inside a bucket, as
//    filling each bucket from its end for ascending i does, :182-183).
//    Bit-exact cellStart / cellEntries.
//
// Scan: per-tile sums (scan_reduce), then each tile scans its counts with a
// wave64 DPP scan plus the carry of the tiles before it (scan_apply), and
// zeroes the counts for the next build.  Bytes: 12 per table entry.
//
// The hot-path kernels (kg_*) are group kernels: they stride over a work list
// of (scene, block of that scene) items and read the scene's arrays from its
// SceneArgs entry, so one launch builds the grids of every scene of a batch.
#include <cmath>
#include <mutex>
#include <string>

#include "fp_internal.h"
#include "fp_kernels.h"
#include "fp_cons.h"

// ---- scan (shared by both grids) ---------------------------------------------

// Sum of scan tile `tile` of count[0, t) into block_sums[tile].
__device__ __forceinline__ void scan_reduce_body(int t, const int* __restrict__ count, int* __restrict__ block_sums,
                                                 int tile) {
  __shared__ int wsum[BLOCK / 64];
  const int base = tile * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int s = 0;
  if (base + SCAN_ITEMS <= t) {
    const int4* c4 = reinterpret_cast<const int4*>(count + base);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; k++) {
      const int4 v = gld(c4 + k);
      s += v.x + v.y + v.z + v.w;
    }
  } else {
    for (int k = 0; k < SCAN_ITEMS; k++)
      if (base + k < t) s += gld(count + base + k);
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; w++) tot += wsum[w];
    gst(block_sums + tile, tot);
  }
  __syncthreads();
}

// Exclusive scan of tile `tile` plus the carry of the tiles before it; writes
// start[k] and zeroes count[k]; the tile holding the last entry writes
// start[t] = total.
__device__ __forceinline__ void scan_apply_body(int t, int* __restrict__ count, const int* __restrict__ block_sums,
                                                int* __restrict__ start, int tile) {
  __shared__ int wsum[BLOCK / 64];
  __shared__ int carry_s;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int c = 0;
  for (int k = threadIdx.x; k < tile; k += BLOCK) c += gld(block_sums + k);
  c = wave_sum(c);
  if (lane == 0) wsum[wid] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; w++) tot += wsum[w];
    carry_s = tot;
  }
  __syncthreads();
  const int carry = carry_s;
  __syncthreads();

  const int base = tile * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  const bool full = base + SCAN_ITEMS <= t;
  if (full) {
    int4* c4 = reinterpret_cast<int4*>(count + base);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; k++) {
      const int4 q = gld(c4 + k);
      v[4 * k] = q.x;
      v[4 * k + 1] = q.y;
      v[4 * k + 2] = q.z;
      v[4 * k + 3] = q.w;
    }
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; k++) gst(c4 + k, make_int4(0, 0, 0, 0));
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
      v[k] = (base + k < t) ? gld(count + base + k) : 0;
      if (base + k < t) gst(count + base + k, 0);
    }
  }
  int local = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) local += v[k];
  const int incl = wave_incl_scan(local);
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  int wprefix = 0;
#pragma unroll
  for (int w = 0; w < BLOCK / 64; w++) wprefix += (w < wid) ? wsum[w] : 0;
  int run = carry + wprefix + incl - local;
  if (full) {
    int4* s4 = reinterpret_cast<int4*>(start + base);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; k++) {
      int4 o;
      o.x = run;
      run += v[4 * k];
      o.y = run;
      run += v[4 * k + 1];
      o.z = run;
      run += v[4 * k + 2];
      o.w = run;
      run += v[4 * k + 3];
      gst(s4 + k, o);
    }
  } else {
    for (int k = 0; k < SCAN_ITEMS; k++) {
      if (base + k < t) gst(start + base + k, run);
      run += v[k];
    }
  }
  if (base <= t - 1 && t - 1 < base + SCAN_ITEMS) gst(start + t, run);
  __syncthreads();
}

__global__ void __launch_bounds__(BLOCK) k_scan_reduce(int t, const int* __restrict__ count,
                                                       int* __restrict__ block_sums) {
  scan_reduce_body(t, count, block_sums, blockIdx.x);
}

__global__ void __launch_bounds__(BLOCK) k_scan_apply(int t, int* __restrict__ count,
                                                      const int* __restrict__ block_sums, int* __restrict__ start) {
  scan_apply_body(t, count, block_sums, start, blockIdx.x);
}

__global__ void __launch_bounds__(BLOCK) kg_scan_reduce(const SceneArgs* __restrict__ S,
                                                        const int2* __restrict__ W) {
  FOR_ITEMS(S, W) scan_reduce_body((int)A.fp.cg.M, A.lcount, A.lblock_sums, lb);
}

// fluid: the fluid order's scan (starts in lstart_f), else the grid's (lstart)
__global__ void __launch_bounds__(BLOCK) kg_scan_apply(const SceneArgs* __restrict__ S, const int2* __restrict__ W, int fluid) {
  FOR_ITEMS(S, W)
  scan_apply_body((int)A.fp.cg.M, A.lcount, fluid ? A.lblock_f : A.lblock_sums, fluid ? A.lstart_f : A.lstart, lb);
}

// ---- cell grid -----------------------------------------------------------------

// Count of the cell grid when integrate could not do it (grab rounds ran in
// between): cell key + reference bucket of the current position.
__device__ __forceinline__ void cell_count_body(const SceneArgs& A, int lb) {
  const CellGrid& cg = A.fp.cg;
  const int i = lb * BLOCK + threadIdx.x;
  const bool in = i < A.n;
  const bool act = in && A.active[i];
  float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t k = 0;
  if (act) {
    p = A.pos[i];
    A.phash[i] = cell_hash(cell_of(p.x, p.y, p.z, cg.ref_inv), cg.hmask);
    k = cell_key(cg, cell_of(p.x, p.y, p.z, cg.inv_cell));
    A.lkey[i] = k;
  } else if (in) {
    A.phash[i] = 0xffffffffu;
  }
  const uint32_t sl = wave_bucket_slot(A.lcount, k, act);
  if (act) A.lslot[i] = sl;
  box_update(act && i < A.ns, p.x, p.y, p.z, A.sbox);
}

__global__ void __launch_bounds__(BLOCK) kg_cell_count(const SceneArgs* __restrict__ S, const int2* __restrict__ W) {
  FOR_ITEMS(S, W) cell_count_body(A, lb);
}

__device__ __forceinline__ void cell_scatter_body(const SceneArgs& A, int lb) {
  const int i = lb * BLOCK + threadIdx.x;
  if (i >= A.n) return;
  const uint32_t h = gld(A.phash + i), k = gld(A.lkey + i), sl = gld(A.lslot + i);
  if (h == 0xffffffffu) return;
  gst(A.ltmp + gld(A.lstart + k) + (int)sl, i);
}

__global__ void __launch_bounds__(BLOCK) kg_cell_scatter(const SceneArgs* __restrict__ S,
                                                         const int2* __restrict__ W) {
  FOR_ITEMS(S, W) cell_scatter_body(A, lb);
}

// Stable order inside a cell (ascending particle index) and the sorted
// copies.  The cell's first particle also records how many of its particles
// are solids (they come first): lnsol is read only for non-empty cells.
__device__ __forceinline__ void cell_rank_body(const SceneArgs& A, int lb) {
  const int i = lb * BLOCK + threadIdx.x;
  if (i >= A.n) return;
  // the bucket, the key and the position load together
  const uint32_t h = gld(A.phash + i), k = gld(A.lkey + i);
  const float4 p = gld(A.pos + i);
  if (h == 0xffffffffu) return;
  const int s = gld(A.lstart + k), e = gld(A.lstart + k + 1);
  // the cell's entries four at a time (their loads in flight together)
  int r = 0, nsol = 0;
  for (int q0 = s; q0 < e; q0 += 4) {
    int v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = gld(A.ltmp + min(q0 + u, e - 1));
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const bool in = q0 + u < e;
      r += (in && v[u] < i) ? 1 : 0;
      nsol += (in && v[u] < A.ns) ? 1 : 0;
    }
  }
  const int d = s + r;
  gst(A.sid + d, i);
  gst(A.spos + d, p);
  gst(A.sphash + d, h);
  if (r == 0) gst(A.lnsol + k, nsol);
}

__global__ void __launch_bounds__(BLOCK) kg_cell_rank(const SceneArgs* __restrict__ S, const int2* __restrict__ W) {
  FOR_ITEMS(S, W) cell_rank_body(A, lb);
}

int fp_group_cell_grid(fp_group* g, bool count_done) {
  if (!count_done) {
    PassScope ps(g->prof, g->stream, P_HASH_COUNT);
    GROUP_LAUNCH(g, kg_cell_count, W_PARTICLE, BLOCK);
  }
  {
    PassScope ps(g->prof, g->stream, P_SCAN_REDUCE);
    GROUP_LAUNCH(g, kg_scan_reduce, W_SCAN, BLOCK);
  }
  {
    PassScope ps(g->prof, g->stream, P_SCAN_APPLY);
    GROUP_LAUNCH(g, kg_scan_apply, W_SCAN, BLOCK, 0);
  }
  {
    PassScope ps(g->prof, g->stream, P_SCATTER);
    GROUP_LAUNCH(g, kg_cell_scatter, W_PARTICLE, BLOCK);
  }
  {
    PassScope ps(g->prof, g->stream, P_RANK);
    GROUP_LAUNCH(g, kg_cell_rank, W_PARTICLE, BLOCK);
  }
  FP_GTRY(g, hipGetLastError());
  for (fp_solver* s : g->scenes) {
    s->ref_grid_valid = false;
    s->cells_built = true;
  }
  return FP_OK;
}

// ---- fluid order (after contacts) --------------------------------------------
//
// The contact pass counted each active fluid particle into lcount under the
// key of its post-contact cell (tkey, tslot per grid slot; the first scan
// left lcount zero).  Scan, scatter, then the same stable rank (ascending
// particle index inside a cell) give the fluid order the density,
// displacement and XSPH passes gather from.

__device__ __forceinline__ void fluid_scatter_body(const SceneArgs& A, int lb) {
  const int e = lb * BLOCK + threadIdx.x;
  {
    // the tile sums were read by the scan: clear them for the next contact
    // pass (the scene's blocks stride over its tiles)
    const int tiles = (int)((A.fp.cg.M + SCAN_TILE - 1) / SCAN_TILE);
    const int stride = (A.n + BLOCK - 1) / BLOCK * BLOCK;
    for (int t = e; t < tiles; t += stride) gst(A.lblock_f + t, 0);
  }
  if (e >= A.n) return;
  const uint32_t k = gld(A.tkey + e), sl = gld(A.tslot + e);
  const int x = gld(A.sid + e);
  if (k == NO_KEY) return;
  gst(A.ltmp + gld(A.lstart_f + k) + (int)sl, x);
}

__global__ void __launch_bounds__(BLOCK) kg_fluid_scatter(const SceneArgs* __restrict__ S,
                                                          const int2* __restrict__ W) {
  FOR_ITEMS(S, W) fluid_scatter_body(A, lb);
}

__device__ __forceinline__ void fluid_rank_body(const SceneArgs& A, int lb) {
  const int e = lb * BLOCK + threadIdx.x;
  if (e >= A.n) return;
  // everything indexed by the grid slot loads together
  const uint32_t k = gld(A.tkey + e), hb = gld(A.sphash + e);
  const int i = gld(A.sid + e);
  const float4 tp = gld(A.tpos + e);
  if (k == NO_KEY) return;
  const int s = gld(A.lstart_f + k), t = gld(A.lstart_f + k + 1);
  // the cell's entries four at a time (their loads in flight together)
  int r = 0;
  for (int q0 = s; q0 < t; q0 += 4) {
    int v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = gld(A.ltmp + min(q0 + u, t - 1));
#pragma unroll
    for (int u = 0; u < 4; u++) r += (q0 + u < t && v[u] < i) ? 1 : 0;
  }
  const int d = s + r;
  gst(A.spos2 + d, tp);
  gst(A.sphash2 + d, hb);
}

__global__ void __launch_bounds__(BLOCK) kg_fluid_rank(const SceneArgs* __restrict__ S, const int2* __restrict__ W) {
  FOR_ITEMS(S, W) fluid_rank_body(A, lb);
}

// The fluid rank with constraint iteration `it` on workgroups of their own
// (the solid chain of a group that does not fork, fp_fluid.hip).
__global__ void __launch_bounds__(BLOCK) kg_fluid_rank_cons(const SceneArgs* __restrict__ S,
                                                            const int2* __restrict__ WC,
                                                            const int2* __restrict__ WP, int g0, int it) {
  if ((int)blockIdx.x < g0) {
    FOR_ITEMS_AT(S, WC, (int)blockIdx.x, g0) constraints_body(A, lb, it);
  } else {
    FOR_ITEMS_AT(S, WP, (int)blockIdx.x - g0, (int)gridDim.x - g0) fluid_rank_body(A, lb);
  }
}

int fp_group_resort_with(fp_group* g, int it) {
  if (it < 0) return fp_group_resort(g);
  PassScope ps(g->prof, g->stream, P_RESORT);
  GROUP_LAUNCH(g, kg_scan_apply, W_SCAN, BLOCK, 1);
  GROUP_LAUNCH(g, kg_fluid_scatter, W_PARTICLE, BLOCK);
  GROUP_LAUNCH2(g, kg_fluid_rank_cons, W_CONS, W_PARTICLE, BLOCK, it);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_group_resort(fp_group* g) {
  PassScope ps(g->prof, g->stream, P_RESORT);
  // the tile sums come from the contact pass (lblock_f): no reduce pass
  GROUP_LAUNCH(g, kg_scan_apply, W_SCAN, BLOCK, 1);
  GROUP_LAUNCH(g, kg_fluid_scatter, W_PARTICLE, BLOCK);
  GROUP_LAUNCH(g, kg_fluid_rank, W_PARTICLE, BLOCK);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

// ---- the reference's hash grid (on demand) --------------------------------------

__global__ void __launch_bounds__(BLOCK) k_ref_hash(int n, double inv_cell, uint32_t mask,
                                                    const float4* __restrict__ pos,
                                                    const uint8_t* __restrict__ active,
                                                    uint32_t* __restrict__ phash) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  if (!active[i]) {
    phash[i] = 0xffffffffu;
    return;
  }
  const float4 p = pos[i];
  phash[i] = cell_hash(cell_of(p.x, p.y, p.z, inv_cell), mask);
}

__global__ void __launch_bounds__(BLOCK) k_ref_count(int n, const uint32_t* __restrict__ phash,
                                                     int* __restrict__ count, uint32_t* __restrict__ pslot) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t h = phash[i];
  if (h == 0xffffffffu) return;
  pslot[i] = (uint32_t)atomicAdd(&count[h], 1);
}

__global__ void __launch_bounds__(BLOCK) k_ref_scatter(int n, const uint32_t* __restrict__ phash,
                                                       const uint32_t* __restrict__ pslot,
                                                       const int* __restrict__ cell_start, int* __restrict__ tmp) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t h = phash[i];
  if (h == 0xffffffffu) return;
  tmp[cell_start[h] + (int)pslot[i]] = i;
}

__global__ void __launch_bounds__(BLOCK) k_ref_rank(int n, const uint32_t* __restrict__ phash,
                                                    const int* __restrict__ cell_start, const int* __restrict__ tmp,
                                                    int* __restrict__ entries) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t h = phash[i];
  if (h == 0xffffffffu) return;
  const int s = cell_start[h], e = cell_start[h + 1];
  int r = 0;
  for (int k = s; k < e; k++) r += tmp[k] > i ? 1 : 0;
  entries[s + r] = i;
}

// The reference grid from phash; rehash = recompute phash from pos first.
int fp_launch_ref_grid(fp_solver* s, bool rehash) {
  const int n = s->n;
  if (rehash) {
    hipLaunchKernelGGL(k_ref_hash, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, n, 1.0 / s->P.cell_size,
                       (uint32_t)(s->T - 1), s->pos, s->active, s->phash);
  }
  hipLaunchKernelGGL(k_ref_count, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, n, s->phash, s->count, s->pslot);
  hipLaunchKernelGGL(k_scan_reduce, dim3(s->scan_blocks), dim3(BLOCK), 0, s->stream, s->T, s->count, s->block_sums);
  hipLaunchKernelGGL(k_scan_apply, dim3(s->scan_blocks), dim3(BLOCK), 0, s->stream, s->T, s->count, s->block_sums,
                     s->cell_start);
  hipLaunchKernelGGL(k_ref_scatter, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, n, s->phash, s->pslot,
                     s->cell_start, s->entries_tmp);
  hipLaunchKernelGGL(k_ref_rank, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, n, s->phash, s->cell_start,
                     s->entries_tmp, s->cell_entries);
  FP_TRY(s, hipGetLastError());
  s->ref_grid_valid = true;
  return FP_OK;
}

// ---- the reference grid for callers that cannot allocate --------------------------
//
// fp_build_grid_noalloc: the same count -> scan -> scatter -> rank as
// fp_launch_ref_grid, over positions and flags the caller holds in device
// memory, with every intermediate (bucket and arrival slot per particle, the
// arrival-order entries, the bucket counts and the scan's tile sums) in one
// file-scope device array.  Each GPU holds its own copy of the array for the
// life of the process; its device address is looked up once per device (the
// array's name in host code is only a stand-in for it).  Calls that overlap
// on one GPU share the array and race.

namespace {
constexpr size_t kNaMaxN = FP_GRID_NOALLOC_MAX_N;
constexpr size_t kNaMaxT = FP_GRID_NOALLOC_MAX_T;
constexpr size_t kNaTiles = (kNaMaxT + SCAN_TILE - 1) / SCAN_TILE;
// words: phash[N], pslot[N], tmp[N], count[T], tile sums
constexpr size_t kNaWords = 3 * kNaMaxN + kNaMaxT + kNaTiles;
constexpr int kNaDevices = 64;
std::mutex g_na_mutex;
uint32_t* g_na_ptr[kNaDevices] = {nullptr};
}  // namespace

static __device__ __align__(256) uint32_t scratch[kNaWords];

// The device address of this device's scratch array (looked up on the first
// call per device, with that device current).
static int noalloc_scratch(int device, uint32_t** out) {
  std::lock_guard<std::mutex> lk(g_na_mutex);
  if (!g_na_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fp_fail(nullptr, FP_EDEVICE,
                     std::string("fp_build_grid_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_na_ptr[device] = (uint32_t*)p;
  }
  *out = g_na_ptr[device];
  return FP_OK;
}

// Bucket of each particle from reference-layout positions (xyz triples).
__global__ void __launch_bounds__(BLOCK) k_ref_hash_xyz(int n, double inv_cell, uint32_t mask,
                                                        const float* __restrict__ pos,
                                                        const uint8_t* __restrict__ active,
                                                        uint32_t* __restrict__ phash) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  if (!active[i]) {
    phash[i] = 0xffffffffu;
    return;
  }
  const float* p = pos + 3 * (size_t)i;
  phash[i] = cell_hash(cell_of(p[0], p[1], p[2], inv_cell), mask);
}

__global__ void __launch_bounds__(BLOCK) k_zero_words(int n, int* __restrict__ p) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) p[i] = 0;
}

extern "C" int fp_build_grid_noalloc(const float* pos, const uint8_t* active, int32_t n, int32_t table_size,
                                     double cell_size, int32_t* cell_start, int32_t* cell_entries, void* stream) {
  if (!pos || !active || !cell_start || !cell_entries)
    return fp_fail(nullptr, FP_EINVAL, "fp_build_grid_noalloc: NULL argument");
  if (n < 0 || (size_t)n > kNaMaxN)
    return fp_fail(nullptr, FP_EINVAL, "fp_build_grid_noalloc: n out of range [0, FP_GRID_NOALLOC_MAX_N = 1048576]");
  if (table_size < 1 || (size_t)table_size > kNaMaxT || (table_size & (table_size - 1)) != 0)
    return fp_fail(nullptr, FP_EINVAL,
                   "fp_build_grid_noalloc: table_size must be a power of two in [1, FP_GRID_NOALLOC_MAX_T = 2097152]");
  if (!(cell_size > 0.0) || !std::isfinite(cell_size))
    return fp_fail(nullptr, FP_EINVAL, "fp_build_grid_noalloc: cell_size must be positive and finite");
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return fp_fail(nullptr, FP_EDEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (device < 0 || device >= kNaDevices) return fp_fail(nullptr, FP_EDEVICE, "fp_build_grid_noalloc: device ordinal");
  uint32_t* base = nullptr;
  int rc = noalloc_scratch(device, &base);
  if (rc) return rc;
  uint32_t* phash = base;
  uint32_t* pslot = base + kNaMaxN;
  int* tmp = (int*)(base + 2 * kNaMaxN);
  int* count = (int*)(base + 3 * kNaMaxN);
  int* tiles = (int*)(base + 3 * kNaMaxN + kNaMaxT);
  const hipStream_t st = (hipStream_t)stream;
  const int T = table_size, tb = (T + SCAN_TILE - 1) / SCAN_TILE;
  // the counts start from zero whatever the last call left (a fill kernel,
  // not a memset: the call may be captured into a graph)
  hipLaunchKernelGGL(k_zero_words, dim3(fp_blocks(T)), dim3(BLOCK), 0, st, T, count);
  if (n > 0) {
    hipLaunchKernelGGL(k_ref_hash_xyz, dim3(fp_blocks(n)), dim3(BLOCK), 0, st, n, 1.0 / cell_size,
                       (uint32_t)(T - 1), pos, active, phash);
    hipLaunchKernelGGL(k_ref_count, dim3(fp_blocks(n)), dim3(BLOCK), 0, st, n, phash, count, pslot);
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3(tb), dim3(BLOCK), 0, st, T, count, tiles);
  hipLaunchKernelGGL(k_scan_apply, dim3(tb), dim3(BLOCK), 0, st, T, count, tiles, (int*)cell_start);
  if (n > 0) {
    hipLaunchKernelGGL(k_ref_scatter, dim3(fp_blocks(n)), dim3(BLOCK), 0, st, n, phash, pslot, (int*)cell_start,
                       tmp);
    hipLaunchKernelGGL(k_ref_rank, dim3(fp_blocks(n)), dim3(BLOCK), 0, st, n, phash, (int*)cell_start, tmp,
                       (int*)cell_entries);
  }
  e = hipGetLastError();
  if (e != hipSuccess)
    return fp_fail(nullptr, FP_EDEVICE, std::string("fp_build_grid_noalloc: launch: ") + hipGetErrorString(e));
  return FP_OK;
}
