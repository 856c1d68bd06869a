// fp_grid.hip -- the two grids.
//
// 1. The cell grid of the hot path (every substep).  Key of a cell c:
//      L(c) = (cx * NY*NZ + cy * NZ + cz) mod M      (32-bit wrap, M, NY, NZ powers of 2)
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
//    cell (fp_launch_resort) for the fluid passes.
//
// 2. The reference's hash grid (_buildGrid, src/solver.js:174-184), built on
//    demand for fp_build_grid / fp_read_grid from the per-particle buckets
//    kept at grid time (phash): count -> the same scan -> scatter -> rank to
//    the reference's order (descending particle index inside a bucket, as
//    filling each bucket from its end for ascending i does, :182-183).
//    Bit-exact cellStart / cellEntries.
//
// Scan: per-tile sums (k_scan_reduce), then each tile scans its counts with a
// wave64 DPP scan plus the carry of the tiles before it (k_scan_apply), and
// zeroes the counts for the next build.  Bytes: 12 per table entry.
#include "fp_internal.h"
#include "fp_kernels.h"

// ---- scan (shared by both grids) ---------------------------------------------

__global__ void __launch_bounds__(BLOCK) k_scan_reduce(int t, const int* __restrict__ count,
                                                       int* __restrict__ block_sums) {
  __shared__ int wsum[BLOCK / 64];
  const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int s = 0;
  if (base + SCAN_ITEMS <= t) {
    const int4* c4 = reinterpret_cast<const int4*>(count + base);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; k++) {
      const int4 v = c4[k];
      s += v.x + v.y + v.z + v.w;
    }
  } else {
    for (int k = 0; k < SCAN_ITEMS; k++)
      if (base + k < t) s += count[base + k];
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; w++) tot += wsum[w];
    block_sums[blockIdx.x] = tot;
  }
}

// Exclusive scan of one tile plus the carry of the tiles before it; writes
// start[k] and zeroes count[k]; the tile holding the last entry writes
// start[t] = total.  zero_word (may be null) is cleared for the passes that
// follow (a counter, cleared by a kernel so the frame can be graph-captured).
__global__ void __launch_bounds__(BLOCK) k_scan_apply(int t, int* __restrict__ count,
                                                      const int* __restrict__ block_sums, int* __restrict__ start,
                                                      int* __restrict__ zero_word) {
  if (zero_word && blockIdx.x == 0 && threadIdx.x == 0) *zero_word = 0;
  __shared__ int wsum[BLOCK / 64];
  __shared__ int carry_s;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int c = 0;
  for (int k = threadIdx.x; k < (int)blockIdx.x; k += BLOCK) c += block_sums[k];
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

  const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  const bool full = base + SCAN_ITEMS <= t;
  if (full) {
    int4* c4 = reinterpret_cast<int4*>(count + base);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS / 4; k++) {
      const int4 q = c4[k];
      v[4 * k] = q.x;
      v[4 * k + 1] = q.y;
      v[4 * k + 2] = q.z;
      v[4 * k + 3] = q.w;
      c4[k] = make_int4(0, 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
      v[k] = (base + k < t) ? count[base + k] : 0;
      if (base + k < t) count[base + k] = 0;
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
      s4[k] = o;
    }
  } else {
    for (int k = 0; k < SCAN_ITEMS; k++) {
      if (base + k < t) start[base + k] = run;
      run += v[k];
    }
  }
  if (base <= t - 1 && t - 1 < base + SCAN_ITEMS) start[t] = run;
}

// ---- cell grid -----------------------------------------------------------------

// Count of the cell grid when integrate could not do it (grab rounds ran in
// between): cell key + reference bucket of the current position.
__global__ void __launch_bounds__(BLOCK) k_cell_count(CellGrid cg, int n, int ns, const float4* __restrict__ pos,
                                                      const uint8_t* __restrict__ active, int* __restrict__ lcount,
                                                      uint32_t* __restrict__ lkey, uint32_t* __restrict__ lslot,
                                                      uint32_t* __restrict__ phash, int* __restrict__ sbox) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const bool in = i < n;
  const bool act = in && active[i];
  float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t k = 0;
  if (act) {
    p = pos[i];
    const Cell c = cell_of(p.x, p.y, p.z, cg.inv_cell);
    phash[i] = cell_hash(c, cg.hmask);
    k = cell_key(cg, c);
    lkey[i] = k;
  } else if (in) {
    phash[i] = 0xffffffffu;
  }
  const uint32_t sl = wave_bucket_slot(lcount, k, act);
  if (act) lslot[i] = sl;
  box_update(act && i < ns, p.x, p.y, p.z, sbox);
}

__global__ void __launch_bounds__(BLOCK) k_cell_scatter(int n, const uint32_t* __restrict__ phash,
                                                        const uint32_t* __restrict__ lkey,
                                                        const uint32_t* __restrict__ lslot,
                                                        const int* __restrict__ lstart, int* __restrict__ tmp) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n || phash[i] == 0xffffffffu) return;
  tmp[lstart[lkey[i]] + (int)lslot[i]] = i;
}

// Stable order inside a cell (ascending particle index) and the sorted
// copies.  The cell's first particle also records how many of its particles
// are solids (they come first): lnsol is read only for non-empty cells.
__global__ void __launch_bounds__(BLOCK) k_cell_rank(int n, int ns, const uint32_t* __restrict__ phash,
                                                     const uint32_t* __restrict__ lkey,
                                                     const int* __restrict__ lstart, const int* __restrict__ tmp,
                                                     const float4* __restrict__ pos, int* __restrict__ sid,
                                                     float4* __restrict__ spos, uint32_t* __restrict__ sphash,
                                                     int* __restrict__ lnsol) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t h = phash[i];
  if (h == 0xffffffffu) return;
  const uint32_t k = lkey[i];
  const int s = lstart[k], e = lstart[k + 1];
  int r = 0, nsol = 0;
  for (int q = s; q < e; q++) {
    const int j = tmp[q];
    r += j < i ? 1 : 0;
    nsol += j < ns ? 1 : 0;
  }
  const int d = s + r;
  sid[d] = i;
  spos[d] = pos[i];
  sphash[d] = h;
  if (r == 0) lnsol[k] = nsol;
}

int fp_launch_cell_grid(fp_solver* s, bool count_done) {
  const int n = s->n;
  if (!count_done) {
    PassScope ps(s, P_HASH_COUNT);
    hipLaunchKernelGGL(k_cell_count, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, s->cg, n, s->ns, s->pos,
                       s->active, s->lcount, s->lkey, s->lslot, s->phash, s->sbox);
  }
  const int M = (int)s->cg.M;
  {
    PassScope ps(s, P_SCAN_REDUCE);
    hipLaunchKernelGGL(k_scan_reduce, dim3(s->lscan_blocks), dim3(BLOCK), 0, s->stream, M, s->lcount,
                       s->lblock_sums);
  }
  {
    PassScope ps(s, P_SCAN_APPLY);
    hipLaunchKernelGGL(k_scan_apply, dim3(s->lscan_blocks), dim3(BLOCK), 0, s->stream, M, s->lcount,
                       s->lblock_sums, s->lstart, nullptr);
  }
  {
    PassScope ps(s, P_SCATTER);
    hipLaunchKernelGGL(k_cell_scatter, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, n, s->phash, s->lkey,
                       s->lslot, s->lstart, s->ltmp);
  }
  {
    PassScope ps(s, P_RANK);
    hipLaunchKernelGGL(k_cell_rank, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, n, s->ns, s->phash, s->lkey,
                       s->lstart, s->ltmp, s->pos, s->sid, s->spos, s->sphash, s->lnsol);
  }
  FP_TRY(s, hipGetLastError());
  s->ref_grid_valid = false;
  return FP_OK;
}

// ---- fluid order (after contacts) --------------------------------------------
//
// The contact pass counted each active fluid particle into lcount under the
// key of its post-contact cell (tkey, tslot per grid slot; the first scan
// left lcount zero).  Scan, scatter, then the same stable rank (ascending
// particle index inside a cell) give the fluid order the density,
// displacement and XSPH passes gather from.

__global__ void __launch_bounds__(BLOCK) k_fluid_scatter(int n, const int* __restrict__ sid,
                                                         const uint32_t* __restrict__ tkey,
                                                         const uint32_t* __restrict__ tslot,
                                                         const int* __restrict__ lstart, int* __restrict__ tmp) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= n) return;
  const uint32_t k = tkey[e];
  if (k == NO_KEY) return;
  tmp[lstart[k] + (int)tslot[e]] = sid[e];
}

__global__ void __launch_bounds__(BLOCK) k_fluid_rank(int n, const int* __restrict__ sid,
                                                      const uint32_t* __restrict__ tkey,
                                                      const int* __restrict__ lstart, const int* __restrict__ tmp,
                                                      const float4* __restrict__ tpos,
                                                      const uint32_t* __restrict__ sphash,
                                                      int* __restrict__ fsid, int* __restrict__ fslot,
                                                      float4* __restrict__ spos2, uint32_t* __restrict__ sphash2) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= n) return;
  const uint32_t k = tkey[e];
  if (k == NO_KEY) return;
  const int i = sid[e];
  const int s = lstart[k], t = lstart[k + 1];
  int r = 0;
  for (int q = s; q < t; q++) r += tmp[q] < i ? 1 : 0;
  const int d = s + r;
  fsid[d] = i;
  fslot[i] = d;
  spos2[d] = tpos[e];
  sphash2[d] = sphash[e];
}

int fp_launch_resort(fp_solver* s) {
  const int n = s->n;
  const int M = (int)s->cg.M;
  PassScope ps(s, P_RESORT);
  hipLaunchKernelGGL(k_scan_reduce, dim3(s->lscan_blocks), dim3(BLOCK), 0, s->stream, M, s->lcount, s->lblock_sums);
  hipLaunchKernelGGL(k_scan_apply, dim3(s->lscan_blocks), dim3(BLOCK), 0, s->stream, M, s->lcount, s->lblock_sums,
                     s->lstart, nullptr);
  hipLaunchKernelGGL(k_fluid_scatter, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, n, s->sid, s->tkey, s->tslot,
                     s->lstart, s->ltmp);
  hipLaunchKernelGGL(k_fluid_rank, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, n, s->sid, s->tkey, s->lstart,
                     s->ltmp, s->tpos, s->sphash, s->fsid, s->fslot, s->spos2, s->sphash2);
  FP_TRY(s, hipGetLastError());
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
                     s->cell_start, nullptr);
  hipLaunchKernelGGL(k_ref_scatter, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, n, s->phash, s->pslot,
                     s->cell_start, s->entries_tmp);
  hipLaunchKernelGGL(k_ref_rank, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, n, s->phash, s->cell_start,
                     s->entries_tmp, s->cell_entries);
  FP_TRY(s, hipGetLastError());
  s->ref_grid_valid = true;
  return FP_OK;
}
