// fp_fluid.hip -- contacts and the double-density-relaxation fluid passes,
// run over the cell-sorted particles.
//
// Thread e handles the particle in sorted slot e.  The 27 neighbour cells of
// a particle are 9 z-runs of 3 consecutive cell keys, i.e. 9 contiguous slot
// ranges ("rows").  A block of 256 consecutive slots covers a compact piece
// of space, so the union of its threads' rows is 9 short slot ranges: the
// block stages them into LDS with coalesced 16-byte loads, then every thread
// tests its ~190 candidates from LDS.  Blocks whose rows do not fit (sparse
// spray, the table's wrap point) read the candidates from global memory.
//
// contacts  (solver.js:246-279 per-pair formula, Jacobi): every active
//           particle gathers from the grid-time snapshot spos: solids from
//           non-bonded solids and all fluid, fluid from solids (only fluid
//           inside the solids' bounding box, grown by the contact range,
//           can touch one).  Solids write their corrected position to pos;
//           fluid stages it (tpos) and counts itself under the key of its
//           post-contact cell for the fluid order (fp_grid.hip).
// density   (solver.js:284-327) rho = sum w q^2, rhoNear = sum w q^3 over the
//           reference's pair list; P = k(rho - rho0), Pn = kNear rhoNear.
// displace  (solver.js:329-353) sum of w (p_x - p_y) D / d with the per-pair
//           and total clamps -> pos.
// xsph      (solver.js:431-454) velocity blend toward the neighbour mean over
//           the same pairs (frozen at density time: spos2 is kept).
//
// The pair list of the reference.  For active fluid i < j the reference
// enters (i, j) once for each of the 27 probes of i's CURRENT cell (after
// contacts) that hashes to j's GRID bucket (before contacts), if
// |p_i - p_j|^2 < H^2 in double at the current positions.  x therefore pairs with
//   y > x  with weight #{probes of cell(x) hashing to bucket(y)},
//   y < x  with weight #{probes of cell(y) hashing to bucket(x)},
// both 1 unless a probe set repeats a bucket (PF_DUP) or the particle whose
// bucket is probed changed cell in contacts (PF_MOVED): then the 27
// reference hashes are counted.  The passes gather from the fluid order
// (sorted by post-contact cell), so every neighbour within H sits in one of
// the 27 cells around x.  The inclusion test d^2 < H^2 is decided in
// double, from the float32 positions, whenever the float32 d^2 lies within
// 1e-5 of H^2 -- the reference's own test (XSPH counts every pair, so the
// test's outcome matters even where q ~ 0).
#include <climits>

#include "fp_internal.h"
#include "fp_kernels.h"

#define STAGE_CAP 2560            // float4 slots of LDS per block (40 KB)

__device__ __forceinline__ uint32_t fbits(float w) { return __float_as_uint(w); }

// The slot range(s) of the z-run of 3 cells starting at key k0: [a0, a1)
// and, past the table's end, [0, b1).
__device__ __forceinline__ void run_ranges(const CellGrid& g, const int* __restrict__ lstart, uint32_t k0, int& a0,
                                           int& a1, int& b1) {
  if (k0 + 2u <= g.kmask) {
    a0 = lstart[k0];
    a1 = lstart[k0 + 3u];
    b1 = 0;
  } else {
    a0 = lstart[k0];
    a1 = lstart[g.M];
    b1 = lstart[k0 + 3u - g.M];
  }
}

// d^2 < H^2 exactly as the reference decides it (double from float32 inputs).
__device__ __forceinline__ bool within_h(const FluidParams& fp, const float4& a, const float4& b, float d2) {
  if (d2 < fp.H2_lo) return true;
  if (d2 > fp.H2_hi) return false;
  const double dx = (double)b.x - (double)a.x, dy = (double)b.y - (double)a.y, dz = (double)b.z - (double)a.z;
  return dx * dx + dy * dy + dz * dz < fp.H2d;
}

// ---- block staging of the 9 neighbour rows ------------------------------------

struct RowStage {
  int lo[9], hi[9], off[10];
  int ok;
  uint32_t kmin, kmax;
};

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off));
  return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
  return v;
}

// Every thread of the block calls this (it synchronises).  want: the thread
// needs candidates, around cell key k.  On return st.ok says whether buf
// holds the rows: for each of the 9 rows, the slots from the row of the
// smallest key through the row of the largest key among the wanting threads
// (a superset of every wanting thread's row, unless a row wraps the table:
// then st.ok = 0).  The key span is reduced per wave with shuffles, so only
// one lane per wave touches the shared bounds.
template <class T>
__device__ __forceinline__ void stage_rows(RowStage& st, T* buf, const T* __restrict__ src, const CellGrid& g,
                                           const int* __restrict__ lstart, bool want, uint32_t k) {
  if (threadIdx.x == 0) {
    st.kmin = UINT_MAX;
    st.kmax = 0u;
    st.ok = 1;
  }
  __syncthreads();
  const uint32_t wmin = wave_min_u32(want ? k : UINT_MAX);
  const uint32_t wmax = wave_max_u32(want ? k : 0u);
  if ((threadIdx.x & 63u) == 0u && wmin <= wmax) {
    atomicMin(&st.kmin, wmin);
    atomicMax(&st.kmax, wmax);
  }
  __syncthreads();
  if (threadIdx.x < 9) {
    const int r = threadIdx.x;
    if (st.kmin > st.kmax) {
      st.ok = 0;  // no wanting thread
    } else {
      const uint32_t a = row_key(g, st.kmin, r / 3 - 1, r % 3 - 1);
      const uint32_t b = row_key(g, st.kmax, r / 3 - 1, r % 3 - 1);
      if (b < a || b - a != st.kmax - st.kmin || b + 2u > g.kmask) {
        st.ok = 0;
      } else {
        st.lo[r] = lstart[a];
        st.hi[r] = lstart[b + 3u];
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && st.ok) {
    int tot = 0;
    for (int r = 0; r < 9; r++) {
      st.off[r] = tot;
      tot += st.hi[r] - st.lo[r];
    }
    st.off[9] = tot;
    if (tot == 0 || tot > STAGE_CAP) st.ok = 0;
  }
  __syncthreads();
  if (st.ok) {
    const int tot = st.off[9];
    for (int i = threadIdx.x; i < tot; i += BLOCK) {
      int r = 0;
#pragma unroll
      for (int q = 1; q < 9; q++) r += (i >= st.off[q]) ? 1 : 0;
      buf[i] = src[st.lo[r] + (i - st.off[r])];
    }
  }
  __syncthreads();
}

// ---- contacts -----------------------------------------------------------------

// Is the solid pair (x, y) joined by an alive constraint?  Scans x's CSR
// row without an early exit, so the row's loads are independent and issue
// back to back; one dependent load of the alive flag follows.
__device__ __forceinline__ bool is_bonded(int x, int y, const int* __restrict__ adj_off, const int4* __restrict__ adj,
                                          const uint8_t* __restrict__ alive) {
  const int e = adj_off[x + 1];
  int c = -1;
  for (int k = adj_off[x]; k < e; k++) {
    const int4 a = adj[k];
    c = a.x == y ? a.y : c;
  }
  return c >= 0 && alive[c] != 0;
}

struct ContactCtx {
  const float4* __restrict__ spos;  // xyz + inverse mass
  const int* __restrict__ sid;
  const int* __restrict__ adj_off;
  const int4* __restrict__ adj;
  const uint8_t* __restrict__ alive;
  const int* __restrict__ sbox;  // solids' bounding box at grid time (ordered-int bits)
};

__device__ __forceinline__ float unorder(int v) { return __int_as_float(v >= 0 ? v : v ^ 0x7fffffff); }

__device__ __forceinline__ void contact_pair(const FluidParams& fp, const ContactCtx& c, int x, bool solid_x,
                                             const float4& px, float rx, int y, const float4& py, float& ax,
                                             float& ay, float& az) {
  const bool solid_y = y < fp.ns;
  if (!solid_x && !solid_y) return;
  const float dx = px.x - py.x, dy = px.y - py.y, dz = px.z - py.z;
  const float d2 = dx * dx + dy * dy + dz * dz;
  const float rsum = rx + (solid_y ? fp.rs : fp.rf);
  if (d2 >= rsum * rsum || d2 < 1e-12f) return;
  if (solid_x && solid_y && is_bonded(x, y, c.adj_off, c.adj, c.alive)) return;
  const float d = sqrtf(d2);
  const float f = (rsum - d) / (d * (px.w + py.w)) * px.w * fp.contact_scale;
  ax += dx * f;
  ay += dy * f;
  az += dz * f;
}

template <bool LDS>
__device__ __forceinline__ void contact_rows(const FluidParams& fp, const ContactCtx& cc, const int* __restrict__ lstart,
                                             const RowStage& st, const float4* buf, const int* bid, int e, int x,
                                             bool solid_x, const float4& px, float rx, uint32_t k0, float& ax,
                                             float& ay, float& az) {
#pragma unroll 1
  for (int r = 0; r < 9; r++) {
    int r0, r1, w1;
    run_ranges(fp.cg, lstart, row_key(fp.cg, k0, r / 3 - 1, r % 3 - 1), r0, r1, w1);
    const int base = LDS ? st.off[r] - st.lo[r] : 0;
    for (int q = r0; q < r1; q++) {
      const int y = LDS ? bid[base + q] : cc.sid[q];
      if (q == e || (!solid_x && y >= fp.ns)) continue;
      contact_pair(fp, cc, x, solid_x, px, rx, y, LDS ? buf[base + q] : cc.spos[q], ax, ay, az);
    }
    for (int q = 0; q < w1; q++) {
      if (q == e) continue;
      contact_pair(fp, cc, x, solid_x, px, rx, cc.sid[q], cc.spos[q], ax, ay, az);
    }
  }
}

__global__ void __launch_bounds__(BLOCK) k_contacts(FluidParams fp, ContactCtx cc, const int* __restrict__ lstart,
                                                    float4* __restrict__ pos, float4* __restrict__ tpos,
                                                    uint32_t* __restrict__ tkey, uint32_t* __restrict__ tslot,
                                                    int* __restrict__ lcount) {
  __shared__ RowStage st;
  __shared__ float4 buf[STAGE_CAP];
  __shared__ int bid[STAGE_CAP];
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  const bool valid = e < lstart[fp.cg.M];
  int x = 0;
  float4 px = make_float4(0.f, 0.f, 0.f, 0.f);
  bool solid_x = false, want = false;
  Cell c0;
  uint32_t k0 = 0;
  if (valid) {
    x = cc.sid[e];
    px = cc.spos[e];
    solid_x = x < fp.ns;
    want = solid_x;
    if (!solid_x && fp.ns > 0) {
      const float g = fp.rs + fp.rf;
      want = px.x > unorder(cc.sbox[0]) - g && px.x < unorder(cc.sbox[3]) + g && px.y > unorder(cc.sbox[1]) - g &&
             px.y < unorder(cc.sbox[4]) + g && px.z > unorder(cc.sbox[2]) - g && px.z < unorder(cc.sbox[5]) + g;
    }
    c0 = cell_of(px.x, px.y, px.z, fp.cg.inv_cell);
    k0 = cell_key(fp.cg, c0);
  }
  // stage positions and particle ids of the rows
  stage_rows(st, buf, cc.spos, fp.cg, lstart, want, k0);
  if (st.ok) {
    for (int i = threadIdx.x; i < st.off[9]; i += BLOCK) {
      int r = 0;
#pragma unroll
      for (int k = 1; k < 9; k++) r += (i >= st.off[k]) ? 1 : 0;
      bid[i] = cc.sid[st.lo[r] + (i - st.off[r])];
    }
  }
  __syncthreads();
  if (!valid) {
    if (e < fp.n) tkey[e] = NO_KEY;
    return;
  }
  const float rx = solid_x ? fp.rs : fp.rf;
  float ax = 0.f, ay = 0.f, az = 0.f;
  if (want) {
    if (st.ok)
      contact_rows<true>(fp, cc, lstart, st, buf, bid, e, x, solid_x, px, rx, k0, ax, ay, az);
    else
      contact_rows<false>(fp, cc, lstart, st, buf, bid, e, x, solid_x, px, rx, k0, ax, ay, az);
  }
  const float ox = px.x + ax, oy = px.y + ay, oz = px.z + az;
  if (solid_x) {
    pos[x] = make_float4(ox, oy, oz, px.w);
    tkey[e] = NO_KEY;
    return;
  }
  const Cell c1 = cell_of(ox, oy, oz, fp.cg.inv_cell);
  uint32_t fl = 0;
  if (c1.x != c0.x || c1.y != c0.y || c1.z != c0.z) fl |= PF_MOVED;
  if (probe_any_dup(c1, fp.cg.hmask)) fl |= PF_DUP;
  const uint32_t k1 = cell_key(fp.cg, c1);
  tkey[e] = k1;
  tslot[e] = (uint32_t)atomicAdd(&lcount[k1], 1);
  tpos[e] = make_float4(ox, oy, oz, __uint_as_float((uint32_t)x | (fl << PF_SHIFT)));
}

// ---- the reference's fluid pairs of the particle in slot e ------------------

struct PairSource {
  const float4* __restrict__ spos2;     // fluid order: post-contact xyz, w = index | flags
  const uint32_t* __restrict__ sphash;  // fluid order: bucket at grid time
  const int* __restrict__ lstart;       // fluid-order cell starts
};

struct PairSelf {
  int e, x;
  bool moved, dup;
  Cell c;
  uint32_t gh;
  float4 p;
};

// The reference-hash multiplicity (rare path: moved particles, repeated
// buckets), kept out of line so the candidate loop stays small.
__device__ __attribute__((noinline)) int multiplicity(const Cell& c, uint32_t mask, uint32_t target) {
  return probe_multiplicity(c, mask, target);
}

__device__ __attribute__((noinline)) int multiplicity_of(float x, float y, float z, double inv, uint32_t mask,
                                                         uint32_t target) {
  return probe_multiplicity(cell_of(x, y, z, inv), mask, target);
}

// Weight of the candidate (0: not a pair of x), written to keep the common
// case branch-free: one combined predicate, the exact double test and the
// multiplicity count only where they can matter.
__device__ __forceinline__ int pair_weight(const FluidParams& fp, const PairSource& src, const PairSelf& me, int q,
                                           const float4& py, float& dx, float& dy, float& dz, float& d2) {
  const uint32_t by = fbits(py.w);
  const int y = (int)(by & SID_MASK);
  const uint32_t fy = by >> PF_SHIFT;
  dx = me.p.x - py.x;
  dy = me.p.y - py.y;
  dz = me.p.z - py.z;
  d2 = dx * dx + dy * dy + dz * dz;
  bool ok = (q != me.e) & (d2 < fp.H2_hi);
  if (__builtin_expect(ok && d2 >= fp.H2_lo, 0)) ok = within_h(fp, me.p, py, d2);
  if (!ok) return 0;
  // y > x: x's probes against y's grid bucket; y < x: y's probes against x's
  const bool special = (y > me.x) ? (me.dup | ((fy & PF_MOVED) != 0u)) : (me.moved | ((fy & PF_DUP) != 0u));
  if (__builtin_expect(special, 0)) {
    return y > me.x ? multiplicity(me.c, fp.cg.hmask, src.sphash[q])
                    : multiplicity_of(py.x, py.y, py.z, fp.cg.inv_cell, fp.cg.hmask, me.gh);
  }
  return 1;
}

// The pairs of x, in the deterministic order of the fluid order's rows:
// f.pair(q, w, dx, dy, dz, d2).  LDS: the block's rows are staged in buf (a
// separate instantiation, so each path keeps its own address space:
// ds_read for LDS, global loads otherwise).
template <bool LDS, class F>
__device__ __forceinline__ void visit_rows(const FluidParams& fp, const PairSource& src, const PairSelf& me,
                                           const RowStage& st, const float4* buf, F& f) {
  const uint32_t kx = cell_key(fp.cg, me.c);
#pragma unroll 1
  for (int r = 0; r < 9; r++) {
    int r0, r1, w1;
    run_ranges(fp.cg, src.lstart, row_key(fp.cg, kx, r / 3 - 1, r % 3 - 1), r0, r1, w1);
    const int base = LDS ? st.off[r] - st.lo[r] : 0;
    for (int q = r0; q < r1; q++) {
      float dx, dy, dz, d2;
      const int w = pair_weight(fp, src, me, q, LDS ? buf[base + q] : src.spos2[q], dx, dy, dz, d2);
      if (w) f.pair(q, w, dx, dy, dz, d2);
    }
    for (int q = 0; q < w1; q++) {
      float dx, dy, dz, d2;
      const int w = pair_weight(fp, src, me, q, src.spos2[q], dx, dy, dz, d2);
      if (w) f.pair(q, w, dx, dy, dz, d2);
    }
  }
}

// ---- the pair list ----------------------------------------------------------------
//
// The density pass records the pairs it finds for the displacement and XSPH
// passes (the reference freezes its pair list at density time too,
// solver.js:311, :431-445): entry k of fluid slot e lives at
// nbr[k * stride + e], so a wave reading its lanes' k-th pairs reads one
// contiguous run.  An entry packs the neighbour's fluid slot (27 bits) and
// the weight - 1 (5 bits).  A particle with more than pair_slots pairs gets
// the count pair_slots + 1 and is gathered again from the grid by the later
// passes (same pairs, same order).

#define NB_QBITS 27
#define NB_QMASK ((1u << NB_QBITS) - 1u)

struct PairList {
  uint32_t* __restrict__ nbr;
  int* __restrict__ ncount;
  int stride, cap;
};

// Replays the recorded pairs of me into f, in the order they were found.
template <class F>
__device__ __forceinline__ void replay_pairs(const PairSource& src, const PairList& pl, const PairSelf& me, int cnt,
                                             F& f) {
  const uint32_t* __restrict__ row = pl.nbr + me.e;
  for (int k = 0; k < cnt; k++) {
    const uint32_t v = row[(size_t)k * pl.stride];
    const int q = (int)(v & NB_QMASK);
    const int w = (int)(v >> NB_QBITS) + 1;
    const float4 py = src.spos2[q];
    const float dx = me.p.x - py.x, dy = me.p.y - py.y, dz = me.p.z - py.z;
    f.pair(q, w, dx, dy, dz, dx * dx + dy * dy + dz * dz);
  }
}

// Records every pair it is handed, then forwards it.
template <class F>
struct Recorder {
  F& f;
  const PairList& pl;
  int e;
  int cnt = 0;
  __device__ Recorder(F& f_, const PairList& pl_, int e_) : f(f_), pl(pl_), e(e_) {}
  __device__ __forceinline__ void pair(int q, int w, float dx, float dy, float dz, float d2) {
    if (cnt < pl.cap) pl.nbr[(size_t)cnt * pl.stride + e] = (uint32_t)q | ((uint32_t)(w - 1) << NB_QBITS);
    cnt++;
    f.pair(q, w, dx, dy, dz, d2);
  }
};

// The pairs of me for the passes after density: the recorded list, or the
// grid again (global reads) for a particle whose list overflowed.
template <class F>
__device__ __forceinline__ void recorded_pairs(const FluidParams& fp, const PairSource& src, const PairList& pl,
                                               const PairSelf& me, F& f) {
  const int cnt = pl.ncount[me.e];
  if (__builtin_expect(cnt <= pl.cap, 1)) {
    replay_pairs(src, pl, me, cnt, f);
  } else {
    RowStage none;
    visit_rows<false>(fp, src, me, none, nullptr, f);
  }
}

// Who is in fluid slot e (false past the last one).
__device__ __forceinline__ bool pair_self(const FluidParams& fp, const PairSource& src, int e, PairSelf& me) {
  me.e = e;
  if (e >= src.lstart[fp.cg.M]) return false;
  me.p = src.spos2[e];
  const uint32_t b = fbits(me.p.w);
  me.x = (int)(b & SID_MASK);
  me.moved = (b >> PF_SHIFT) & PF_MOVED;
  me.dup = (b >> PF_SHIFT) & PF_DUP;
  me.c = cell_of(me.p.x, me.p.y, me.p.z, fp.cg.inv_cell);
  me.gh = src.sphash[e];
  return true;
}

struct DensityAcc {
  float inv_H;
  float rho = 0.f, rhon = 0.f;
  __device__ __forceinline__ void pair(int, int w, float, float, float, float d2) {
    const float q = 1.f - sqrtf(d2) * inv_H;
    const float q2 = q * q;
    rho += (float)w * q2;
    rhon += (float)w * q2 * q;
  }
};

__global__ void __launch_bounds__(BLOCK) k_density(FluidParams fp, PairSource src, PairList pl,
                                                   float2* __restrict__ sP, float2* __restrict__ rho_out,
                                                   int* __restrict__ sbox, int* __restrict__ dflags) {
  __shared__ RowStage st;
  __shared__ float4 buf[STAGE_CAP];
  // the solids' bounding box is rebuilt by the next integrate
  if (blockIdx.x == 0 && threadIdx.x < 6) sbox[threadIdx.x] = threadIdx.x < 3 ? INT_MAX : INT_MIN;
  PairSelf me;
  const bool valid = pair_self(fp, src, blockIdx.x * BLOCK + threadIdx.x, me);
  stage_rows(st, buf, src.spos2, fp.cg, src.lstart, valid, valid ? cell_key(fp.cg, me.c) : 0u);
  if (!valid) return;
  DensityAcc acc;
  acc.inv_H = fp.inv_H;
  Recorder<DensityAcc> rec(acc, pl, me.e);
  if (st.ok)
    visit_rows<true>(fp, src, me, st, buf, rec);
  else
    visit_rows<false>(fp, src, me, st, buf, rec);
  if (rec.cnt > pl.cap) {
    pl.ncount[me.e] = pl.cap + 1;
    atomicAdd(&dflags[DF_PAIR_OVF], 1);
  } else {
    pl.ncount[me.e] = rec.cnt;
  }
  rho_out[me.x] = make_float2(acc.rho, acc.rhon);
  sP[me.e] = make_float2(fp.stiff * (acc.rho - fp.rest_density), fp.near_stiff * acc.rhon);
}

struct DisplaceAcc {
  const float2* __restrict__ sP;
  float2 Px;
  float inv_H, max_pair;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  __device__ __forceinline__ void pair(int q, int w, float dx, float dy, float dz, float d2) {
    const float d = sqrtf(d2);
    if (d < 1e-9f) return;
    const float u = 1.f - d * inv_H;
    const float2 Py = sP[q];
    float D = 0.25f * ((Px.x + Py.x) * u + (Px.y + Py.y) * u * u);
    D = fminf(fmaxf(D, -max_pair), max_pair);
    const float s = (float)w * D / d;
    sx += dx * s;
    sy += dy * s;
    sz += dz * s;
  }
};

// spos2 + the clamped displacement -> pos for the fluid; the inverse mass
// is kept from pos.w.
__global__ void __launch_bounds__(BLOCK) k_displace(FluidParams fp, PairSource src, PairList pl,
                                                    const float2* __restrict__ sP, float4* __restrict__ pos) {
  PairSelf me;
  if (!pair_self(fp, src, blockIdx.x * BLOCK + threadIdx.x, me)) return;
  DisplaceAcc acc;
  acc.sP = sP;
  acc.Px = sP[me.e];
  acc.inv_H = fp.inv_H;
  acc.max_pair = fp.max_pair;
  recorded_pairs(fp, src, pl, me, acc);
  float sx = acc.sx, sy = acc.sy, sz = acc.sz;
  const float s2 = sx * sx + sy * sy + sz * sz;
  if (s2 > fp.max_total * fp.max_total) {
    const float g = fp.max_total / sqrtf(s2);
    sx *= g;
    sy *= g;
    sz *= g;
  }
  const float w = pos[me.x].w;
  pos[me.x] = make_float4(me.p.x + sx, me.p.y + sy, me.p.z + sz, w);
}

struct XsphAcc {
  const float4* __restrict__ svtmp;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  int m = 0;
  __device__ __forceinline__ void pair(int q, int w, float, float, float, float) {
    const float4 v = svtmp[q];
    const float fw = (float)w;
    sx += fw * v.x;
    sy += fw * v.y;
    sz += fw * v.z;
    m += w;
  }
};

// XSPH: svtmp -> vel for the fluid that is still active (the reference skips
// particles culled since the density pass, whose velocity stays as it was).
__global__ void __launch_bounds__(BLOCK) k_xsph(FluidParams fp, PairSource src, PairList pl,
                                                const float4* __restrict__ svtmp, float4* __restrict__ vel,
                                                const uint8_t* __restrict__ active) {
  PairSelf me;
  if (!pair_self(fp, src, blockIdx.x * BLOCK + threadIdx.x, me)) return;
  if (!active[me.x]) return;
  const float4 v = svtmp[me.e];
  XsphAcc acc;
  acc.svtmp = svtmp;
  recorded_pairs(fp, src, pl, me, acc);
  if (acc.m == 0) {
    vel[me.x] = v;
    return;
  }
  const float inv_m = 1.f / (float)acc.m;
  vel[me.x] = make_float4(v.x + fp.visc * (acc.sx * inv_m - v.x), v.y + fp.visc * (acc.sy * inv_m - v.y),
                          v.z + fp.visc * (acc.sz * inv_m - v.z), 0.f);
}

// ---- launchers ------------------------------------------------------------------

FluidParams fp_fluid_params(const fp_solver* s) {
  FluidParams fp;
  const fp_params& P = s->P;
  fp.cg = s->cg;
  fp.ns = s->ns;
  fp.n = s->n;
  fp.nf = s->nf;
  fp.rs = (float)P.solid_radius;
  fp.rf = (float)P.fluid_radius;
  fp.inv_H = (float)(1.0 / P.fluid_h);
  fp.H2d = P.fluid_h * P.fluid_h;
  fp.H2_lo = (float)(fp.H2d * (1.0 - 1e-5));
  fp.H2_hi = (float)(fp.H2d * (1.0 + 1e-5));
  fp.max_pair = (float)(0.05 * P.fluid_h);
  fp.max_total = (float)(0.35 * P.fluid_h);
  fp.stiff = (float)P.stiffness;
  fp.near_stiff = (float)P.near_stiffness;
  fp.rest_density = (float)P.rest_density;
  fp.contact_scale = (float)P.contact_scale;
  fp.visc = (float)P.viscosity;
  return fp;
}

static PairSource pair_source(const fp_solver* s) {
  PairSource src;
  src.spos2 = s->spos2;
  src.sphash = s->sphash2;
  src.lstart = s->lstart;
  return src;
}

static PairList pair_list(const fp_solver* s) {
  PairList pl;
  pl.nbr = s->nbr;
  pl.ncount = s->ncount;
  pl.stride = s->n;
  pl.cap = s->pair_slots;
  return pl;
}

int fp_launch_contacts(fp_solver* s) {
  const FluidParams fp = fp_fluid_params(s);
  ContactCtx cc{s->spos, s->sid, s->adj_off, s->adj, s->alive[s->cur], s->sbox};
  PassScope ps(s, P_CONTACTS);
  hipLaunchKernelGGL(k_contacts, dim3(fp_blocks(s->n)), dim3(BLOCK), 0, s->stream, fp, cc, s->lstart, s->pos,
                     s->tpos, s->tkey, s->tslot, s->lcount);
  FP_TRY(s, hipGetLastError());
  return FP_OK;
}

int fp_launch_density(fp_solver* s) {
  const FluidParams fp = fp_fluid_params(s);
  PassScope ps(s, P_DENSITY);
  hipLaunchKernelGGL(k_density, dim3(fp_blocks(s->n)), dim3(BLOCK), 0, s->stream, fp, pair_source(s), pair_list(s),
                     s->sP, s->rho, s->sbox, s->dflags);
  FP_TRY(s, hipGetLastError());
  return FP_OK;
}

int fp_launch_displace(fp_solver* s) {
  const FluidParams fp = fp_fluid_params(s);
  PassScope ps(s, P_DISPLACE);
  hipLaunchKernelGGL(k_displace, dim3(fp_blocks(s->n)), dim3(BLOCK), 0, s->stream, fp, pair_source(s),
                     pair_list(s), s->sP, s->pos);
  FP_TRY(s, hipGetLastError());
  return FP_OK;
}

int fp_launch_xsph(fp_solver* s) {
  if (s->nf == 0) return FP_OK;
  const FluidParams fp = fp_fluid_params(s);
  PassScope ps(s, P_XSPH);
  hipLaunchKernelGGL(k_xsph, dim3(fp_blocks(s->n)), dim3(BLOCK), 0, s->stream, fp, pair_source(s), pair_list(s),
                     s->svtmp, s->vel, s->active);
  FP_TRY(s, hipGetLastError());
  return FP_OK;
}
