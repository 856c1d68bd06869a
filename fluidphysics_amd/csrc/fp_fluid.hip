// fp_fluid.hip -- contacts and the double-density-relaxation fluid passes,
// run over the cell-sorted particles.
//
// The 27 neighbour cells of a particle are 9 z-runs of 3 consecutive cell
// keys, i.e. 9 contiguous slot ranges ("rows") of a cell-sorted order.
// Candidates are read straight from the sorted arrays, four per step so
// their loads overlap; neighbouring threads read overlapping rows, which the
// caches serve (staging the rows in LDS measured slower: it costs occupancy
// and the rows of a block that spans several cell rows do not fit).
//
// contacts  (solver.js:246-279 per-pair formula, Jacobi) from the grid-time
//           snapshot spos.  Solids (k_contacts_solid, 4 lanes per solid,
//           rows split between the lanes) against non-bonded solids and all
//           fluid; the solid's alive bonds are held in registers.  Fluid
//           (k_contacts_fluid) against solids only: cells list their
//           particles by ascending index, so a cell's solids come first and
//           the scan of a cell stops at its first fluid particle; only fluid
//           inside the solids' bounding box, grown by the contact range, can
//           touch a solid.  Solids write their corrected position to pos;
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

#define CONTACT_LANES 4  // lanes per solid in the contact pass
#define MAXB 32          // bonds per solid held in registers

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

// ---- contacts -----------------------------------------------------------------

// The alive bonds of solid x, in registers: the other endpoints of x's
// alive constraints (-1 past the end).  Rows longer than MAXB are not held;
// is_bonded scans the CSR row for them.
struct Bonds {
  int id[MAXB];
  bool held;
};

__device__ __forceinline__ void load_bonds(int x, const int* __restrict__ adj_off, const int4* __restrict__ adj,
                                           const uint8_t* __restrict__ alive, Bonds& b) {
  const int k0 = adj_off[x], k1 = adj_off[x + 1];
  b.held = k1 - k0 <= MAXB;
#pragma unroll
  for (int j = 0; j < MAXB; j++) {
    int v = -1;
    if (b.held && k0 + j < k1) {
      const int4 a = adj[k0 + j];
      if (alive[a.y]) v = a.x;
    }
    b.id[j] = v;
  }
}

__device__ __forceinline__ bool is_bonded(int x, int y, const Bonds& b, const int* __restrict__ adj_off,
                                          const int4* __restrict__ adj, const uint8_t* __restrict__ alive) {
  if (b.held) {
    bool hit = false;
#pragma unroll
    for (int j = 0; j < MAXB; j++) hit |= b.id[j] == y;
    return hit;
  }
  const int e = adj_off[x + 1];
  int c = -1;
  for (int k = adj_off[x]; k < e; k++) {
    const int4 a = adj[k];
    c = a.x == y ? a.y : c;
  }
  return c >= 0 && alive[c] != 0;
}

struct ContactCtx {
  const float4* __restrict__ spos;  // grid order: xyz + inverse mass
  const int* __restrict__ sid;
  const int* __restrict__ lstart;   // grid-order cell starts
  const int* __restrict__ adj_off;
  const int4* __restrict__ adj;
  const uint8_t* __restrict__ alive;
  const int* __restrict__ sbox;  // solids' bounding box at grid time (ordered-int bits)
};

__device__ __forceinline__ float unorder(int v) { return __int_as_float(v >= 0 ? v : v ^ 0x7fffffff); }

// The Jacobi correction of x from y (solver.js:267-275 with x's share).
__device__ __forceinline__ void contact_add(const FluidParams& fp, const float4& px, float rx, const float4& py,
                                            float ry, float& ax, float& ay, float& az) {
  const float dx = px.x - py.x, dy = px.y - py.y, dz = px.z - py.z;
  const float d2 = dx * dx + dy * dy + dz * dz;
  const float rsum = rx + ry;
  if (d2 >= rsum * rsum || d2 < 1e-12f) return;
  const float d = sqrtf(d2);
  const float f = (rsum - d) / (d * (px.w + py.w)) * px.w * fp.contact_scale;
  ax += dx * f;
  ay += dy * f;
  az += dz * f;
}

// Solid x: CONTACT_LANES lanes split its 9 rows (row r to lane r % lanes),
// each sums its candidates in slot order, then the lanes' sums are added
// pairwise.  Writes pos[x] = grid-time position + correction.
__global__ void __launch_bounds__(BLOCK) k_contacts_solid(FluidParams fp, ContactCtx cc, const uint8_t* __restrict__ active,
                                                          float4* __restrict__ pos) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  const int x = t / CONTACT_LANES, lane = t % CONTACT_LANES;
  const bool valid = x < fp.ns && active[x];
  float ax = 0.f, ay = 0.f, az = 0.f;
  float4 px = make_float4(0.f, 0.f, 0.f, 0.f);
  if (valid) {
    px = pos[x];
    Bonds bonds;
    load_bonds(x, cc.adj_off, cc.adj, cc.alive, bonds);
    const uint32_t k0 = cell_key(fp.cg, cell_of(px.x, px.y, px.z, fp.cg.inv_cell));
    for (int r = lane; r < 9; r += CONTACT_LANES) {
      int r0, r1, w1;
      run_ranges(fp.cg, cc.lstart, row_key(fp.cg, k0, r / 3 - 1, r % 3 - 1), r0, r1, w1);
      for (int pass = 0; pass < 2; pass++) {
        const int qa = pass ? 0 : r0, qb = pass ? w1 : r1;
        for (int q = qa; q < qb; q++) {
          const int y = cc.sid[q];
          if (y == x) continue;
          const float4 py = cc.spos[q];
          if (y < fp.ns) {
            // a bonded pair does not collide; test the bond only when close
            const float dx = px.x - py.x, dy = px.y - py.y, dz = px.z - py.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            const float rsum = fp.rs + fp.rs;
            if (d2 >= rsum * rsum || d2 < 1e-12f) continue;
            if (is_bonded(x, y, bonds, cc.adj_off, cc.adj, cc.alive)) continue;
            contact_add(fp, px, fp.rs, py, fp.rs, ax, ay, az);
          } else {
            contact_add(fp, px, fp.rs, py, fp.rf, ax, ay, az);
          }
        }
      }
    }
  }
#pragma unroll
  for (int off = 1; off < CONTACT_LANES; off <<= 1) {
    ax += __shfl_xor(ax, off, CONTACT_LANES);
    ay += __shfl_xor(ay, off, CONTACT_LANES);
    az += __shfl_xor(az, off, CONTACT_LANES);
  }
  if (valid && lane == 0) pos[x] = make_float4(px.x + ax, px.y + ay, px.z + az, px.w);
}

// Fluid in grid slot e: corrections from the solids of its 27 cells (rows in
// order, cells of a row in order, solids of a cell by ascending index), then
// the post-contact cell: flags, count, staging for the fluid order.
__global__ void __launch_bounds__(BLOCK) k_contacts_fluid(FluidParams fp, ContactCtx cc, float4* __restrict__ tpos,
                                                          uint32_t* __restrict__ tkey, uint32_t* __restrict__ tslot,
                                                          int* __restrict__ lcount) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= fp.n) return;
  if (e >= cc.lstart[fp.cg.M]) {
    tkey[e] = NO_KEY;
    return;
  }
  const int x = cc.sid[e];
  if (x < fp.ns) {
    tkey[e] = NO_KEY;
    return;
  }
  const float4 px = cc.spos[e];
  const Cell c0 = cell_of(px.x, px.y, px.z, fp.cg.inv_cell);
  float ax = 0.f, ay = 0.f, az = 0.f;
  if (fp.ns > 0) {
    const float g = fp.rs + fp.rf;
    const bool near = px.x > unorder(cc.sbox[0]) - g && px.x < unorder(cc.sbox[3]) + g &&
                      px.y > unorder(cc.sbox[1]) - g && px.y < unorder(cc.sbox[4]) + g &&
                      px.z > unorder(cc.sbox[2]) - g && px.z < unorder(cc.sbox[5]) + g;
    if (near) {
      const uint32_t k0 = cell_key(fp.cg, c0);
#pragma unroll 1
      for (int r = 0; r < 9; r++) {
        const uint32_t kr = row_key(fp.cg, k0, r / 3 - 1, r % 3 - 1);
        for (uint32_t j = 0; j < 3; j++) {
          const uint32_t k = (kr + j) & fp.cg.kmask;
          const int qe = cc.lstart[k + 1];
          for (int q = cc.lstart[k]; q < qe; q++) {
            const int y = cc.sid[q];
            if (y >= fp.ns) break;
            contact_add(fp, px, fp.rf, cc.spos[q], fp.rs, ax, ay, az);
          }
        }
      }
    }
  }
  const float ox = px.x + ax, oy = px.y + ay, oz = px.z + az;
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
// f.pair(q, w, dx, dy, dz, d2).
template <class F>
__device__ __forceinline__ void visit_rows(const FluidParams& fp, const PairSource& src, const PairSelf& me, F& f) {
  const uint32_t kx = cell_key(fp.cg, me.c);
  const float4* __restrict__ row = src.spos2;
#pragma unroll 1
  for (int r = 0; r < 9; r++) {
    int r0, r1, w1;
    run_ranges(fp.cg, src.lstart, row_key(fp.cg, kx, r / 3 - 1, r % 3 - 1), r0, r1, w1);
    int q = r0;
    // four candidates per step: their loads are in flight together
    for (; q + 4 <= r1; q += 4) {
      const float4 c0 = row[q], c1 = row[q + 1], c2 = row[q + 2], c3 = row[q + 3];
      float dx, dy, dz, d2;
      int w = pair_weight(fp, src, me, q, c0, dx, dy, dz, d2);
      if (w) f.pair(q, w, dx, dy, dz, d2);
      w = pair_weight(fp, src, me, q + 1, c1, dx, dy, dz, d2);
      if (w) f.pair(q + 1, w, dx, dy, dz, d2);
      w = pair_weight(fp, src, me, q + 2, c2, dx, dy, dz, d2);
      if (w) f.pair(q + 2, w, dx, dy, dz, d2);
      w = pair_weight(fp, src, me, q + 3, c3, dx, dy, dz, d2);
      if (w) f.pair(q + 3, w, dx, dy, dz, d2);
    }
    for (; q < r1; q++) {
      float dx, dy, dz, d2;
      const int w = pair_weight(fp, src, me, q, row[q], dx, dy, dz, d2);
      if (w) f.pair(q, w, dx, dy, dz, d2);
    }
    for (q = 0; q < w1; q++) {
      float dx, dy, dz, d2;
      const int w = pair_weight(fp, src, me, q, row[q], dx, dy, dz, d2);
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
    visit_rows(fp, src, me, f);
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
  // the solids' bounding box is rebuilt by the next integrate
  if (blockIdx.x == 0 && threadIdx.x < 6) sbox[threadIdx.x] = threadIdx.x < 3 ? INT_MAX : INT_MIN;
  PairSelf me;
  if (!pair_self(fp, src, blockIdx.x * BLOCK + threadIdx.x, me)) return;
  DensityAcc acc;
  acc.inv_H = fp.inv_H;
  Recorder<DensityAcc> rec(acc, pl, me.e);
  visit_rows(fp, src, me, rec);
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
  ContactCtx cc{s->spos, s->sid, s->lstart, s->adj_off, s->adj, s->alive[s->cur], s->sbox};
  PassScope ps(s, P_CONTACTS);
  if (s->ns > 0)
    hipLaunchKernelGGL(k_contacts_solid, dim3(fp_blocks(s->ns * CONTACT_LANES)), dim3(BLOCK), 0, s->stream, fp, cc,
                       s->active, s->pos);
  hipLaunchKernelGGL(k_contacts_fluid, dim3(fp_blocks(s->n)), dim3(BLOCK), 0, s->stream, fp, cc, s->tpos, s->tkey,
                     s->tslot, s->lcount);
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
