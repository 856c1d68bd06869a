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
#include <algorithm>
#include <climits>

#include "fp_internal.h"
#include "fp_kernels.h"
#include "fp_bounds.h"
#include "fp_cons.h"

#ifndef FLUID_BLOCK
#define FLUID_BLOCK 256  // threads per block of the density, displacement and XSPH passes
#endif
// amdgpu_waves_per_eu(n) for a kernel (WAVES_ATTR(0): the allocator's choice)
#define WAVES_ATTR_0
#define WAVES_ATTR_45 __attribute__((amdgpu_waves_per_eu(4, 5)))
#define WAVES_ATTR_5 __attribute__((amdgpu_waves_per_eu(5)))
#define WAVES_ATTR_6 __attribute__((amdgpu_waves_per_eu(6)))
#define WAVES_ATTR_7 __attribute__((amdgpu_waves_per_eu(7)))
#define WAVES_ATTR_8 __attribute__((amdgpu_waves_per_eu(8)))
#define WAVES_ATTR_(n) WAVES_ATTR_##n
#define WAVES_ATTR(n) WAVES_ATTR_(n)

#ifndef REPLAY_BATCH
#define REPLAY_BATCH 4  // pair-list entries per step of the displacement and XSPH replay
#endif
#ifndef DENSITY_STAGE
#define DENSITY_STAGE 16  // pair-list entries per lane the staging density kernel holds in LDS (> 0)
#endif
#ifndef DENSITY_BATCH
#define DENSITY_BATCH 1  // candidates per step of the density loop (1: fastest measured, also in batches)
#endif
#define MAXB 32          // bonds per solid staged in LDS for the contact pass

__device__ __forceinline__ uint32_t fbits(float w) { return __float_as_uint(w); }

// The slot range(s) of the z-run of 3 cells starting at key k0: [a0, a1)
// and, past the table's end, [0, b1).
__device__ __forceinline__ void run_ranges(const CellGrid& g, const int* __restrict__ lstart, uint32_t k0, int& a0,
                                           int& a1, int& b1) {
  if (k0 + 2u <= g.kmask) {
    a0 = gld(lstart + k0);
    a1 = gld(lstart + k0 + 3u);
    b1 = 0;
  } else {
    a0 = gld(lstart + k0);
    a1 = gld(lstart + g.M);
    b1 = gld(lstart + (k0 + 3u - g.M));
  }
}

// The same for a run of len (1..3) keys starting at key k (k <= kmask).
__device__ __forceinline__ void run_ranges_n(const CellGrid& g, const int* __restrict__ lstart, uint32_t k,
                                             uint32_t len, int& a0, int& a1, int& b1) {
  a0 = gld(lstart + k);
  if (k + len <= g.M) {
    a1 = gld(lstart + (k + len));
    b1 = 0;
  } else {
    a1 = gld(lstart + g.M);
    b1 = gld(lstart + (k + len - g.M));
  }
}

// d^2 < H^2 exactly as the reference decides it (double from float32 inputs).
__device__ __forceinline__ bool within_h(const FluidParams& fp, const float4& a, const float4& b, float d2) {
  if (d2 < fp.H2_lo) return true;
  if (d2 > fp.H2_hi) return false;
  const double dx = (double)b.x - (double)a.x, dy = (double)b.y - (double)a.y, dz = (double)b.z - (double)a.z;
  return dx * dx + dy * dy + dz * dz < fp.H2d;
}

// Distances from p to the lower and upper faces of its sort cell c (cell
// size cs): with the cell size, they give the floor of the distance to any
// particle of a neighbouring cell.
struct FaceDist {
  float x0, x1, y0, y1, z0, z1, cs;
};

__device__ __forceinline__ FaceDist face_dist(float px, float py, float pz, const Cell& c, double cs) {
  FaceDist f;
  f.x0 = (float)((double)px - c.x * cs);
  f.x1 = (float)((c.x + 1.0) * cs - (double)px);
  f.y0 = (float)((double)py - c.y * cs);
  f.y1 = (float)((c.y + 1.0) * cs - (double)py);
  f.z0 = (float)((double)pz - c.z * cs);
  f.z1 = (float)((c.z + 1.0) * cs - (double)pz);
  f.cs = (float)cs;
  return f;
}

// Gap along one axis from p to the slab of cells at offset o (|o| <= NB_R).
__device__ __forceinline__ float axis_gap(float f0, float f1, float cs, int o) {
  return o < 0 ? f0 + (float)(-o - 1) * cs : (o > 0 ? f1 + (float)(o - 1) * cs : 0.f);
}

// Cells of neighbour row r, (a, b) = (r / NB_RUN - NB_R, r % NB_RUN - NB_R),
// that can hold a particle closer than sqrt(r2): key offsets lo .. hi-1 of
// the row's NB_RUN keys; false when the whole row is farther.  r2 carries a
// margin over the float rounding of the gaps.  NaN positions keep every cell.
__device__ __forceinline__ bool row_cells(const FaceDist& f, int r, float r2, uint32_t& lo, uint32_t& hi) {
  const int a = r / NB_RUN - NB_R, b = r % NB_RUN - NB_R;
  const float ex = axis_gap(f.x0, f.x1, f.cs, a), ey = axis_gap(f.y0, f.y1, f.cs, b);
  const float dxy2 = ex * ex + ey * ey;
  if (!(dxy2 == dxy2)) {
    lo = 0u;
    hi = (uint32_t)NB_RUN;
    return true;
  }
  if (!(dxy2 < r2)) return false;
  lo = (uint32_t)NB_R;
  hi = (uint32_t)NB_R + 1u;
#pragma unroll
  for (int o = 1; o <= NB_R; o++) {
    const float gl = axis_gap(f.z0, f.z1, f.cs, -o), gh = axis_gap(f.z0, f.z1, f.cs, o);
    if (dxy2 + gl * gl < r2) lo = (uint32_t)(NB_R - o);
    if (dxy2 + gh * gh < r2) hi = (uint32_t)(NB_R + o + 1);
  }
  return true;
}

// ---- contacts -----------------------------------------------------------------

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

// Is y a bonded (alive constraint) neighbour of x?  `held` lists the other
// endpoints of x's alive constraints (-1 past the end) when x has at most
// MAXB of them; otherwise x's CSR row is scanned.
__device__ __forceinline__ bool bonded(int x, int y, bool held, const int* held_ids, const ContactCtx& cc) {
  if (held) {
    bool hit = false;
#pragma unroll 8
    for (int j = 0; j < MAXB; j++) hit |= held_ids[j] == y;
    return hit;
  }
  const int e = gld(cc.adj_off + x + 1);
  bool hit = false;
  for (int k = gld(cc.adj_off + x); k < e; k++) hit |= adj_other(x, gld(cc.adj + k)) == y && gld(cc.ealive + k) != 0;
  return hit;
}

// Solid x: L lanes (4..32) share its neighbour block in units -- single
// cells when the block has at most 32 (CELL_SUB 1: 27), else rows -- unit u
// goes to lane u % L; cells beyond the contact range are skipped -- and x's
// bonds are staged once in LDS and consulted only for a solid candidate
// within contact range.  Each unit's candidates are summed in slot order into
// its own partial; the 32 partials (units past the last empty) are then
// added by one fixed binary tree, pairs differing in the highest bit of the
// unit number first: inside a lane for the bits above log2(L), then across lanes.
// Every L builds the same tree from the same partials, so the correction is
// the same bits whatever L the launch uses.  Writes pos[x] = grid-time
// position + correction.
// contact units: one cell each when the neighbour block has at most 32 cells
#define CONTACT_CELL_UNITS (NB_ROWS * NB_RUN <= 32)
#define CONTACT_UNITS (CONTACT_CELL_UNITS ? NB_ROWS * NB_RUN : NB_ROWS)
static_assert(CONTACT_UNITS <= 32, "the contact reduction tree has 32 leaves");

template <int L>
__device__ __forceinline__ void contacts_solid_body(const SceneArgs& A, int lb, int* __restrict__ lds_bonds) {
  constexpr int M = 32 / L;  // partials per lane
  const FluidParams& fp = A.fp;
  const ContactCtx& cc = A.cc;
  float4* __restrict__ pos = A.pos;
  const int t = lb * BLOCK + threadIdx.x;
  const int x = t / L, lane = t % L;
  int* held_ids = lds_bonds + (threadIdx.x / L) * MAXB;
  // the solid's flag, CSR row and position load together (a thread past the
  // solids reads solid 0's and is dropped)
  const int xc = x < fp.ns ? x : 0;
  const bool valid = x < fp.ns && gld(A.active + xc);
  int k0 = gld(cc.adj_off + xc), k1 = gld(cc.adj_off + xc + 1);
  const float4 px0 = gld(pos + xc);
  if (!valid) k0 = k1 = 0;
  const bool held = k1 - k0 <= MAXB;
#pragma unroll
  for (int j = lane; j < MAXB; j += L) {
    int v = -1;
    if (valid && held && k0 + j < k1) {
      if (gld(cc.ealive + k0 + j)) v = adj_other(x, gld(cc.adj + k0 + j));
    }
    held_ids[j] = v;
  }
  __syncthreads();
  float px_[M], py_[M], pz_[M];
#pragma unroll
  for (int j = 0; j < M; j++) px_[j] = py_[j] = pz_[j] = 0.f;
  float4 px = make_float4(0.f, 0.f, 0.f, 0.f);
  if (valid) {
    px = px0;
    const Cell c0 = cell_of(px.x, px.y, px.z, fp.cg.inv_cell);
    const uint32_t kx = cell_key(fp.cg, c0);
    const FaceDist fd = face_dist(px.x, px.y, px.z, c0, fp.cs);
    const float rr2 = 4.f * fp.rs * fp.rs;
#pragma unroll
    for (int j = 0; j < M; j++) {
      const int u = lane + L * j;
      if (u >= CONTACT_UNITS) continue;
      const int r = CONTACT_CELL_UNITS ? u / NB_RUN : u;
      uint32_t lo, hi;
      if (!row_cells(fd, r, fp.contact_r2, lo, hi)) continue;
      if (CONTACT_CELL_UNITS) {
        const uint32_t z = (uint32_t)(u % NB_RUN);
        if (z < lo || z >= hi) continue;
        lo = z;
        hi = z + 1;
      }
      int q0, q1, w1;
      run_ranges_n(fp.cg, cc.lstart, (row_key(fp.cg, kx, r / NB_RUN - NB_R, r % NB_RUN - NB_R) + lo) & fp.cg.kmask,
                   hi - lo, q0, q1, w1);
      float ax = 0.f, ay = 0.f, az = 0.f;
      // the row's slots, then (past the table's end) its wrapped part
      for (int part = 0; part < 2; part++) {
        const int e0 = part ? 0 : q0, e1 = part ? w1 : q1;
        for (int q = e0; q < e1; q += 2) {
          const int qb = min(q + 1, e1 - 1);
          const int y0 = gld(cc.sid + q), y1 = gld(cc.sid + qb);
          const float4 p0 = gld(cc.spos + q), p1 = gld(cc.spos + qb);
#pragma unroll
          for (int i = 0; i < 2; i++) {
            const int y = i ? y1 : y0;
            const float4& py = i ? p1 : p0;
            if ((i && q + 1 >= e1) || y == x) continue;
            if (y < fp.ns) {
              // with the solid contact cache, solid pairs are resolved in the
              // constraint iterations
              if (A.sm.cache) continue;
              // a bonded pair does not collide; test the bond only when close
              const float dx = px.x - py.x, dy = px.y - py.y, dz = px.z - py.z;
              const float d2 = dx * dx + dy * dy + dz * dz;
              if (d2 >= rr2 || d2 < 1e-12f) continue;
              if (bonded(x, y, held, held_ids, cc)) continue;
              contact_add(fp, px, fp.rs, py, fp.rs, ax, ay, az);
            } else {
              contact_add(fp, px, fp.rs, py, fp.rf, ax, ay, az);
            }
          }
        }
      }
      px_[j] = ax;
      py_[j] = ay;
      pz_[j] = az;
    }
  }
#pragma unroll
  for (int half = M / 2; half >= 1; half /= 2)
#pragma unroll
    for (int j = 0; j < half; j++) {
      px_[j] += px_[j + half];
      py_[j] += py_[j + half];
      pz_[j] += pz_[j + half];
    }
  float ax = px_[0], ay = py_[0], az = pz_[0];
#pragma unroll
  for (int off = L / 2; off >= 1; off /= 2) {
    ax += __shfl_xor(ax, off, L);
    ay += __shfl_xor(ay, off, L);
    az += __shfl_xor(az, off, L);
  }
  if (lane == 0 && x < fp.ns) {
    // the result also goes to spos_c[1], the snapshot the first constraint
    // iteration reads (a single iteration finishes its solids into pos in
    // the same launch); an inactive solid's position is copied unchanged
    const float4 q = valid ? make_float4(px.x + ax, px.y + ay, px.z + az, px.w) : gld(pos + x);
    if (valid) gst(pos + x, q);
    gst(A.spos_c[1] + x, q);
  }
  __syncthreads();  // the block's next item reuses lds_bonds
}

template <int L>
__global__ void __launch_bounds__(BLOCK) kg_contacts_solid(const SceneArgs* __restrict__ S,
                                                           const int2* __restrict__ W) {
  __shared__ int lds_bonds[(BLOCK / L) * MAXB];
  FOR_ITEMS(S, W) contacts_solid_body<L>(A, lb, lds_bonds);
}

// Fluid in grid slot e: corrections from the solids of its 27 cells (rows in
// order, cells of a row in order, solids of a cell by ascending index), then
// the post-contact cell: flags, count, staging for the fluid order.
__device__ __forceinline__ void contacts_fluid_body(const SceneArgs& A, int lb) {
  const FluidParams& fp = A.fp;
  const ContactCtx& cc = A.cc;
  float4* __restrict__ tpos = A.tpos;
  uint32_t* __restrict__ tkey = A.tkey;
  uint32_t* __restrict__ tslot = A.tslot;
  int* __restrict__ lcount = A.lcount;
  const int e = lb * BLOCK + threadIdx.x;
  // every lane stays to the bucket vote at the end
  bool has = false;
  uint32_t k1 = 0;
  float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
  if (e < fp.n) {
    // the slot's particle and position load together (the empty asm takes
    // both here, so the position load is not sunk behind the branch)
    const int occupied = gld(cc.lstart + fp.cg.M);
    int xs = gld(cc.sid + e);
    float4 px = gld(cc.spos + e);
    asm volatile("" : "+v"(xs), "+v"(px.x), "+v"(px.y), "+v"(px.z), "+v"(px.w));
    const int x = e < occupied ? xs : -1;
    if (x < fp.ns) {
      gst(tkey + e, (uint32_t)NO_KEY);
    } else {
      const Cell c0 = cell_of(px.x, px.y, px.z, fp.cg.inv_cell);  // sort cell
      float ax = 0.f, ay = 0.f, az = 0.f;
      if (fp.ns > 0) {
        const float g = fp.rs + fp.rf;
        const bool near = px.x > unorder(cc.sbox[0]) - g && px.x < unorder(cc.sbox[3]) + g &&
                          px.y > unorder(cc.sbox[1]) - g && px.y < unorder(cc.sbox[4]) + g &&
                          px.z > unorder(cc.sbox[2]) - g && px.z < unorder(cc.sbox[5]) + g;
        if (near) {
          const uint32_t k0 = cell_key(fp.cg, c0);
          const FaceDist fd = face_dist(px.x, px.y, px.z, c0, fp.cs);
#pragma unroll 1
          for (int r = 0; r < NB_ROWS; r++) {
            uint32_t lo, hi;
            if (!row_cells(fd, r, fp.contact_r2, lo, hi)) continue;
            const uint32_t kr = row_key(fp.cg, k0, r / NB_RUN - NB_R, r % NB_RUN - NB_R);
            // the row's cell bounds and solid counts, all loads in flight
            // together (lnsol holds a value only for a non-empty cell)
            int a[NB_RUN], b[NB_RUN], m[NB_RUN];
#pragma unroll
            for (uint32_t j = 0; j < NB_RUN; j++) {
              const uint32_t k = (kr + j) & fp.cg.kmask;
              a[j] = gld(cc.lstart + k);
              b[j] = gld(cc.lstart + k + 1);
              m[j] = gld(cc.lnsol + k);
            }
            int total = 0;
#pragma unroll
            for (uint32_t j = 0; j < NB_RUN; j++) {
              if (j < lo || j >= hi || a[j] == b[j]) m[j] = 0;
              total += m[j];
            }
            // the row's solids as one list (cell by cell, slots ascending:
            // the order of the sums), loaded four at a time ahead of them
            for (int t0 = 0; t0 < total; t0 += 4) {
              float4 py[4];
#pragma unroll
              for (int i = 0; i < 4; i++) {
                int t = min(t0 + i, total - 1), q = 0;
#pragma unroll
                for (uint32_t j = 0; j < NB_RUN; j++) {
                  if (t >= 0 && t < m[j]) q = a[j] + t;
                  t -= m[j];
                }
                py[i] = gld(cc.spos + q);
              }
#pragma unroll
              for (int i = 0; i < 4; i++)
                if (t0 + i < total) contact_add(fp, px, fp.rf, py[i], fp.rs, ax, ay, az);
            }
          }
        }
      }
      const float ox = px.x + ax, oy = px.y + ay, oz = px.z + az;
      // the reference's cells before and after the contacts decide the flags
      const Cell r0 = cell_of(px.x, px.y, px.z, fp.cg.ref_inv), r1 = cell_of(ox, oy, oz, fp.cg.ref_inv);
      uint32_t fl = 0;
      if (r1.x != r0.x || r1.y != r0.y || r1.z != r0.z) fl |= PF_MOVED;
      if (probe_any_dup(r1, fp.cg.hmask)) fl |= PF_DUP;
      k1 = cell_key(fp.cg, cell_of(ox, oy, oz, fp.cg.inv_cell));
      out = make_float4(ox, oy, oz, __uint_as_float((uint32_t)x | (fl << PF_SHIFT)));
      has = true;
    }
  }
  // consecutive grid slots mostly share a cell: one atomic per run
  const uint32_t slot = wave_bucket_slot(lcount, k1, has);
  // the fluid order's scan tiles take their sums here (the lanes walk the
  // grid order, so a wave's keys mostly share one tile)
  wave_bucket_add(A.lblock_f, has ? k1 / SCAN_TILE : 0u, has);
  if (has) {
    gst(tkey + e, k1);
    gst(tslot + e, slot);
    gst(tpos + e, out);
  }
}

#ifndef CONTACTS_FLUID_WAVES
#define CONTACTS_FLUID_WAVES 0
#endif
__global__ void __launch_bounds__(BLOCK) WAVES_ATTR(CONTACTS_FLUID_WAVES) kg_contacts_fluid(const SceneArgs* __restrict__ S,
                                                           const int2* __restrict__ W) {
  FOR_ITEMS(S, W) contacts_fluid_body(A, lb);
}

// Lane index of this thread over the fluid order: the scene's L fluid lanes
// (consecutive) share one fluid slot (slot = lane / L).
__device__ __forceinline__ int fluid_lane(int lb) { return lb * FLUID_BLOCK + (int)threadIdx.x; }

// Sum over the L lanes of one slot (pairwise, the same in every lane).
__device__ __forceinline__ float lanes_sum(float v, int L) {
  for (int off = 1; off < L; off <<= 1) v += __shfl_xor(v, off, L);
  return v;
}
__device__ __forceinline__ int lanes_sum(int v, int L) {
  for (int off = 1; off < L; off <<= 1) v += __shfl_xor(v, off, L);
  return v;
}

// ---- the reference's fluid pairs of the particle in slot e ------------------


// The particle of a fluid slot as the pair passes see it.  Only the slot,
// the lane and the position are held; the particle index, its flags (both
// in p.w) and its cell are derived where needed, which keeps the gather
// loops' live registers few.
struct PairSelf {
  int e;
  int v, h;  // lane (over the fluid order) and its index among the slot's lanes
  float4 p;  // w = particle index | flags << PF_SHIFT
  __device__ __forceinline__ int x() const { return (int)(fbits(p.w) & SID_MASK); }
  __device__ __forceinline__ bool moved() const { return ((fbits(p.w) >> PF_SHIFT) & PF_MOVED) != 0u; }
  __device__ __forceinline__ bool dup() const { return ((fbits(p.w) >> PF_SHIFT) & PF_DUP) != 0u; }
};

// The reference-hash multiplicity, out of line (scalar arguments only, so
// the call passes registers): probes of the cell of (x, y, z) against
// bucket target.
__device__ __attribute__((noinline)) int mult_pos(float x, float y, float z, double inv, uint32_t mask,
                                                  uint32_t target) {
  return probe_multiplicity(cell_of(x, y, z, inv), mask, target);
}

// Weight of a candidate that needs more than the float test: d^2 within
// 1e-5 of H^2 (decided in double), or a multiplicity that is not 1.
__device__ __forceinline__ int slow_weight(const FluidParams& fp, const PairSource& src, const PairSelf& me, int q,
                                           const float4& py, float d2) {
  if (d2 >= fp.H2_lo && !within_h(fp, me.p, py, d2)) return 0;
  const uint32_t by = fbits(py.w);
  const int y = (int)(by & SID_MASK);
  const uint32_t fy = by >> PF_SHIFT;
  // y > x: x's probes against y's grid bucket; y < x: y's probes against x's
  const bool special = (y > me.x()) ? (me.dup() | ((fy & PF_MOVED) != 0u)) : (me.moved() | ((fy & PF_DUP) != 0u));
  if (!special) return 1;
  return y > me.x() ? mult_pos(me.p.x, me.p.y, me.p.z, fp.cg.ref_inv, fp.cg.hmask, src.sphash[q])
                    : mult_pos(py.x, py.y, py.z, fp.cg.ref_inv, fp.cg.hmask, src.sphash[me.e]);
}

// Weight of the candidate in slot q (0: not a pair of x).  The common case
// is decided by float compares; one rarely taken branch covers the rest.
__device__ __forceinline__ int pair_weight(const FluidParams& fp, const PairSource& src, const PairSelf& me, int q,
                                           const float4& py, float& dx, float& dy, float& dz, float& d2) {
  const uint32_t by = fbits(py.w);
  const int y = (int)(by & SID_MASK);
  const uint32_t fy = by >> PF_SHIFT;
  dx = me.p.x - py.x;
  dy = me.p.y - py.y;
  dz = me.p.z - py.z;
  d2 = dx * dx + dy * dy + dz * dz;
  const bool cand = q != me.e;
  const bool inside = cand & (d2 < fp.H2_lo);
  const bool band = cand & (d2 >= fp.H2_lo) & (d2 < fp.H2_hi);
  const bool special = (y > me.x()) ? (me.dup() | ((fy & PF_MOVED) != 0u)) : (me.moved() | ((fy & PF_DUP) != 0u));
  int w = inside ? 1 : 0;
  if (__builtin_expect(band | (inside & special), 0)) w = slow_weight(fp, src, me, q, py, d2);
  return w;
}

struct RowRange {
  int r0, r1, w1;  // [r0, r1) and, past the table's end, [0, w1)
};

// Slot range(s) of neighbour row r of x (empty when culled).
__device__ __forceinline__ RowRange row_range(const FluidParams& fp, const int* __restrict__ lstart, uint32_t kx, int r,
                                              const FaceDist& f) {
  RowRange rr{0, 0, 0};
  uint32_t lo, hi;
  if (!row_cells(f, r, fp.cull_r2, lo, hi)) return rr;
  run_ranges_n(fp.cg, lstart, (row_key(fp.cg, kx, r / NB_RUN - NB_R, r % NB_RUN - NB_R) + lo) & fp.cg.kmask, hi - lo,
               rr.r0, rr.r1, rr.w1);
  return rr;
}

// ---- the pair list ----------------------------------------------------------------
//
// The density pass records the pairs it finds for the displacement and XSPH
// passes (the reference freezes its pair list at density time too,
// solver.js:311, :431-445).  The lanes of the fluid order come in blocks of
// 64, one wave of the fluid passes; block b holds cap + 1 rows of 64
// entries, entry k of lane v at nbr[(v / 64) * (cap + 1) * 64 + k * 64 +
// v % 64].  A wave reading its lanes' k-th pairs reads one contiguous run,
// and every access is the wave's block (scalar registers) plus a 32-bit
// lane offset.  An entry packs the neighbour's fluid slot (27 bits) and the
// weight - 1 (5 bits).  A lane with more than cap pairs gets the count
// cap + 1 and its slot is gathered again from the grid by the later passes
// (same pairs, same order); its extra entries go to the block's sink row
// (row cap) so the recording loop needs no bounds branch.

#define NB_QBITS 27
#define NB_QMASK ((1u << NB_QBITS) - 1u)
#define NB_ROWB 256u  // bytes per row of a block

// The block of the wave that lane v belongs to (the same in every lane of a
// wave: the fluid passes' lanes are v = block * FLUID_BLOCK + threadIdx.x).
__device__ __forceinline__ uint32_t* pair_block(const PairList& pl, int v) {
  const int b = __builtin_amdgcn_readfirstlane(v >> 6);
  return pl.nbr + (size_t)b * ((size_t)(pl.cap + 1) * 64);
}
__device__ __forceinline__ uint32_t pair_lane_off(int v) { return (uint32_t)(v & 63) * 4u; }


// v_sqrt_f32 (within 1 ulp); the reference's own square root is a double
// one, so no float32 rounding of it is more faithful than another.
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Density sums and the pair list of x: every candidate of the culled
// neighbour rows of the fluid order, in slot order.  A z-cell of a row, or a whole row,
// whose box lies farther than H from x is skipped: it cannot hold a pair
// (cull_r2 carries a margin over the float and double tests).
//
// Memory latency and branch overhead bound this loop, so: the next row's
// bounds load while the current row runs; a row is read in batches of
// DENSITY_BATCH candidates per lane (the last batch predicated, no serial
// tail), each batch's loads issued before the previous batch is evaluated;
// the common case is decided with selects and a single, rarely taken branch
// per batch covers the d ~ H band and multiplicities other than 1.
struct DensityGather {
  float rho = 0.f, rhon = 0.f;
  int cnt = 0;
  uint32_t* base;  // the wave's block of the pair list
  uint32_t wofs;   // byte offset of entry cnt (the sink entry once cnt >= cap)
  uint32_t wend;   // byte offset of this lane's sink entry
  uint32_t* lds;   // (staging kernel) this thread's first staged entry (entry k at lds[k * FLUID_BLOCK])
  int pb;          // (staging kernel) entries staged in LDS (the rest go straight to the pair list)
};

template <bool kStage>
__device__ __forceinline__ void density_take(const FluidParams& fp, const PairList& pl, int q, int w, float d2,
                                             DensityGather& g) {
  const float u = fmaf(-fast_sqrt(d2), fp.inv_H, 1.f);
  const float u2 = (float)w * (u * u);
  g.rho += w ? u2 : 0.f;
  g.rhon += w ? u2 * u : 0.f;
  if (w) {
    const uint32_t e = (uint32_t)q | ((uint32_t)(w - 1) << NB_QBITS);
    if (kStage && g.cnt < g.pb) {
      // the first pb entries wait in LDS and are written out row by row at
      // the end (whole rows of the wave's block per store)
      g.lds[g.cnt * FLUID_BLOCK] = e;
    } else {
      // entries past pl.cap all land in the sink row pl.cap, which nothing reads
      gst_at(g.base, g.wofs, e);
      g.wofs = min(g.wofs + NB_ROWB, g.wend);
    }
    g.cnt++;
  }
}

// One batch of DENSITY_BATCH candidates in slots q, q+1, ... (those past
// last are not candidates).  Weight 1 for d^2 < H^2 (1 - 1e-5) unless a flag is set on
// either particle; the band and flagged cases go to the one rare branch.
// flag_min: the least w bits of a candidate that make the pair flagged (0
// when me itself is flagged, so every candidate is).
template <bool kStage>
__device__ __forceinline__ void density_batch(const FluidParams& fp, const PairSource& src, const PairList& pl,
                                              const PairSelf& me, uint32_t flag_min, int q, int last, const float4* c,
                                              DensityGather& g) {
  float d2[DENSITY_BATCH];
  int w[DENSITY_BATCH];
  bool rare = false;
#pragma unroll
  for (int i = 0; i < DENSITY_BATCH; i++) {
    const float dx = me.p.x - c[i].x, dy = me.p.y - c[i].y, dz = me.p.z - c[i].z;
    d2[i] = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const bool cand = (q + i <= last) & (q + i != me.e);
    const bool near = d2[i] < fp.H2_lo;
    const bool inside = cand & near;
    const bool band = cand & !near & (d2[i] < fp.H2_hi);  // NaN: neither
    const bool flagged = fbits(c[i].w) >= flag_min;
    w[i] = inside ? 1 : 0;
    rare |= band | (inside & flagged);
  }
  if (__builtin_expect(rare, 0)) {
#pragma unroll
    for (int i = 0; i < DENSITY_BATCH; i++)
      if ((q + i <= last) & (q + i != me.e) & (d2[i] < fp.H2_hi)) w[i] = slow_weight(fp, src, me, q + i, c[i], d2[i]);
  }
#pragma unroll
  for (int i = 0; i < DENSITY_BATCH; i++) density_take<kStage>(fp, pl, q + i, w[i], d2[i], g);
}

__device__ __forceinline__ void load_batch(const float4* __restrict__ row, int q, int last, float4* b) {
#pragma unroll
  for (int i = 0; i < DENSITY_BATCH; i++) b[i] = gld_at(row, (uint32_t)min(q + i, last) << 4);
}

template <bool kStage>
__device__ __forceinline__ void density_gather(const FluidParams& fp, const PairSource& src, const PairList& pl,
                                               const PairSelf& me, DensityGather& g) {
  const Cell c = cell_of(me.p.x, me.p.y, me.p.z, fp.cg.inv_cell);
  const uint32_t kx = cell_key(fp.cg, c);
  const float4* __restrict__ row = src.spos2;
  const uint32_t flag_min = (me.moved() | me.dup()) ? 0u : (1u << PF_SHIFT);
  const FaceDist f = face_dist(me.p.x, me.p.y, me.p.z, c, fp.cs);
  RowRange nx = row_range(fp, src.lstart, kx, 0, f);
#pragma unroll 1
  for (int r = 0; r < NB_ROWS; r++) {
    const RowRange rr = nx;
    if (r < NB_ROWS - 1) nx = row_range(fp, src.lstart, kx, r + 1, f);
    if (rr.r1 > rr.r0) {
      const int last = rr.r1 - 1;
      // this lane's batches of the row: h, h + LANES, ...; the next batch's
      // loads are in flight while the current one is evaluated.  Two
      // buffers alternate (a copy from one to the other would wait for the
      // loads in flight), and the next batch is loaded unconditionally
      // (indices clamp to the row's last slot): loads under a branch make
      // the compiler wait for them as well as for the batch in use.
      const int step = DENSITY_BATCH << fp.fl_log2;
      int q = rr.r0 + DENSITY_BATCH * me.h;
      float4 b0[DENSITY_BATCH], b1[DENSITY_BATCH];
      if (q <= last) load_batch(row, q, last, b0);
      while (q <= last) {
        const int q1 = q + step;
        load_batch(row, q1, last, b1);
        density_batch<kStage>(fp, src, pl, me, flag_min, q, last, b0, g);
        if (q1 > last) break;
        const int q2 = q1 + step;
        load_batch(row, q2, last, b0);
        density_batch<kStage>(fp, src, pl, me, flag_min, q1, last, b1, g);
        q = q2;
      }
    }
    for (int q = 0; q < (me.h == 0 ? rr.w1 : 0); q++) {
      float dx, dy, dz, d2;
      const int w = pair_weight(fp, src, me, q, gld(row + q), dx, dy, dz, d2);
      density_take<kStage>(fp, pl, q, w, d2, g);
    }
  }
}

// The pairs of x from the grid, one candidate at a time (the rare regather
// of a particle whose recorded list overflowed; same order as density_gather).
template <class F>
__device__ __forceinline__ void visit_rows_simple(const FluidParams& fp, const PairSource& src, const PairSelf& me,
                                                  F& f) {
  const Cell c = cell_of(me.p.x, me.p.y, me.p.z, fp.cg.inv_cell);
  const uint32_t kx = cell_key(fp.cg, c);
  const FaceDist fd = face_dist(me.p.x, me.p.y, me.p.z, c, fp.cs);
#pragma unroll 1
  for (int r = 0; r < NB_ROWS; r++) {
    const RowRange rr = row_range(fp, src.lstart, kx, r, fd);
    // the same batches as density_gather gives this lane
    for (int q0 = rr.r0 + DENSITY_BATCH * me.h; q0 < rr.r1; q0 += DENSITY_BATCH << fp.fl_log2) {
      for (int q = q0; q < min(q0 + DENSITY_BATCH, rr.r1); q++) {
        float dx, dy, dz, d2;
        const int w = pair_weight(fp, src, me, q, src.spos2[q], dx, dy, dz, d2);
        if (w) f.pair(q, w, dx, dy, dz, d2);
      }
    }
    for (int q = 0; q < (me.h == 0 ? rr.w1 : 0); q++) {
      float dx, dy, dz, d2;
      const int w = pair_weight(fp, src, me, q, src.spos2[q], dx, dy, dz, d2);
      if (w) f.pair(q, w, dx, dy, dz, d2);
    }
  }
}

// Replays the recorded pairs of me into f, in the order they were found.
// Entries are read one batch of REPLAY_BATCH ahead; a batch's gathers (the
// neighbour's position when F::kNeedPos, and f.fetch(q)) are issued
// together, before the next batch's entries, and before any of them is used.
// The loads a replay pass starts with, indexed by the lane alone and issued
// together with the slot's own loads: the lane's pair count and its first
// REPLAY_BATCH entries (rows past the block's sink are not read; a lane past
// the lane arrays reads their last lane and is dropped by its caller).
struct ReplayHead {
  int cnt;
  uint32_t v[REPLAY_BATCH];
};

__device__ __forceinline__ ReplayHead replay_head(const FluidParams& fp, const PairList& pl, int v) {
  ReplayHead r;
  const int vc = min(v, fp.n * (1 << fp.fl_log2) - 1);
  const uint32_t* base = pair_block(pl, vc);
  const uint32_t lo = pair_lane_off(vc);
  r.cnt = gld(pl.ncount + vc);
#pragma unroll
  for (int i = 0; i < REPLAY_BATCH; i++) r.v[i] = gld_at(base, lo + (uint32_t)min(i, pl.cap) * NB_ROWB);
  return r;
}

template <class F>
__device__ __forceinline__ void replay_pairs(const PairSource& src, const PairList& pl, const PairSelf& me,
                                             const ReplayHead& rh, F& f) {
  const int cnt = rh.cnt;
  const uint32_t* base = pair_block(pl, me.v);
  const uint32_t lo = pair_lane_off(me.v);
  const int last = cnt - 1;
  // entry k of this lane at byte lo + k * NB_ROWB of the wave's block
  const uint32_t olast = lo + (uint32_t)max(last, 0) * NB_ROWB;
  // entries past the count are not this substep's: entry 0 stands in (a
  // valid slot to gather from; its result is not used)
  uint32_t v[REPLAY_BATCH];
#pragma unroll
  for (int i = 0; i < REPLAY_BATCH; i++) v[i] = i < cnt ? rh.v[i] : 0u;
  for (int k = 0; k < cnt; k += REPLAY_BATCH) {
    // this batch's gathers first, then the next batch's entries: the loads
    // complete in issue order, so waiting for the gathers leaves the
    // entries in flight
    float4 py[REPLAY_BATCH];
    typename F::Extra ex[REPLAY_BATCH];
#pragma unroll
    for (int i = 0; i < REPLAY_BATCH; i++) {
      const int q = (int)(v[i] & NB_QMASK);
      if (F::kNeedPos) py[i] = gld_xyz_at(src.spos2, (uint32_t)q << 4);
      ex[i] = f.fetch(q);
    }
    uint32_t u[REPLAY_BATCH];
#pragma unroll
    for (int i = 0; i < REPLAY_BATCH; i++) u[i] = v[i];
    if (k + REPLAY_BATCH < cnt) {
#pragma unroll
      for (int i = 0; i < REPLAY_BATCH; i++) v[i] = gld_at(base, min(lo + (uint32_t)(k + REPLAY_BATCH + i) * NB_ROWB, olast));
    }
#pragma unroll
    for (int i = 0; i < REPLAY_BATCH; i++) {
      if (k + i <= last) {
        const int q = (int)(u[i] & NB_QMASK);
        const int w = (int)(u[i] >> NB_QBITS) + 1;
        float dx = 0.f, dy = 0.f, dz = 0.f, d2 = 0.f;
        if (F::kNeedPos) {
          dx = me.p.x - py[i].x;
          dy = me.p.y - py[i].y;
          dz = me.p.z - py[i].z;
          d2 = dx * dx + dy * dy + dz * dz;
        }
        f.use(q, w, dx, dy, dz, d2, ex[i]);
      }
    }
  }
}

// The pairs of me for the passes after density: the recorded list, or the
// grid again (global reads) for a particle whose list overflowed.
template <class F>
__device__ __forceinline__ void recorded_pairs(const FluidParams& fp, const PairSource& src, const PairList& pl,
                                               const PairSelf& me, const ReplayHead& rh, F& f) {
  // a slot whose lanes did not all fit is gathered again by all its lanes
  if (__builtin_expect(lanes_sum(rh.cnt > pl.cap ? 1 : 0, 1 << fp.fl_log2) == 0, 1)) {
    replay_pairs(src, pl, me, rh, f);
  } else {
    visit_rows_simple(fp, src, me, f);
  }
}

// Who is in the fluid slot of lane v (false past the last slot).
__device__ __forceinline__ bool pair_self(const FluidParams& fp, const PairSource& src, int v, PairSelf& me) {
  me.v = v;
  me.h = v & ((1 << fp.fl_log2) - 1);
  me.e = v >> fp.fl_log2;
  // the position loads with the fluid order's size (a lane past the last
  // slot reads an in-bounds one and is dropped)
  const int occupied = gld(src.lstart + fp.cg.M);
  me.p = gld(src.spos2 + min(me.e, fp.n - 1));
  return me.e < occupied;
}

// The outputs of a gathered slot (every lane of the slot calls this).
__device__ __forceinline__ void density_store(const FluidParams& fp, const PairList& pl, const PairSelf& me,
                                              const DensityGather& g, float2* __restrict__ sP,
                                              float2* __restrict__ rho_out, int* __restrict__ dflags) {
  const int L = 1 << fp.fl_log2;
  const bool over = lanes_sum(g.cnt > pl.cap ? 1 : 0, L) != 0;
  gst(pl.ncount + me.v, g.cnt > pl.cap ? pl.cap + 1 : g.cnt);
  const float rho = lanes_sum(g.rho, L), rhon = lanes_sum(g.rhon, L);
  if (me.h != 0) return;
  if (over) atomicAdd(&dflags[DF_PAIR_OVF], 1);
  gst(rho_out + me.x(), make_float2(rho, rhon));
  gst(sP + me.e, make_float2(fp.stiff * (rho - fp.rest_density), fp.near_stiff * rhon));
}

// kStage: the first fp.stage entries of each lane wait in LDS (lds,
// DENSITY_STAGE per thread) and are written out row by row at the end, so
// a store covers whole rows of the wave's block instead of one entry per
// lane scattered over several rows.
template <bool kStage>
__device__ __forceinline__ void density_body(const SceneArgs& A, int lb, uint32_t* __restrict__ lds) {
  const FluidParams& fp = A.fp;
  const PairSource& src = A.src;
  const PairList& pl = A.pl;
  PairSelf me;
  if (!pair_self(fp, src, fluid_lane(lb), me)) return;
  DensityGather g;
  g.base = pair_block(pl, me.v);
  g.wofs = pair_lane_off(me.v);
  g.wend = g.wofs + (uint32_t)pl.cap * NB_ROWB;
  g.lds = lds + threadIdx.x;
  g.pb = kStage ? min(fp.stage, pl.cap) : 0;
  const uint32_t lane_off = g.wofs;
  g.wofs += (uint32_t)g.pb * NB_ROWB;
  density_gather<kStage>(fp, src, pl, me, g);
  if (kStage) {
    // the staged entries, row by row: the lanes of a wave that have entry
    // k write their part of row k together
    const int n = min(g.cnt, g.pb);
    for (int k = 0; k < n; k++) gst_at(g.base, lane_off + (uint32_t)k * NB_ROWB, g.lds[k * FLUID_BLOCK]);
  }
  density_store(fp, pl, me, g, A.sP, A.rho, A.dflags);
}

// Work placement of the fluid passes.  A scene's W_FLUID items are one run
// of the work list, item j at list position o + j, and list position b runs
// on XCD b % 8 (one workgroup per item).  The scene's first `a` items hold
// its fluid (the fluid order's occupied slots; the rest is idle capacity).
// Each XCD gets one contiguous run of those a items: the neighbour rows a
// particle gathers (the next z and y rows are one row and one layer of the
// fluid order away) are then mostly in the L2 of the XCD that reads them,
// and every XCD still holds an eighth of each scene's fluid.  A bijection on
// [0, a) for every o; items past a keep their place.
__device__ __forceinline__ int fluid_block(int j, int o, int a) {
  if (j >= a) return j;
  const int c = o & 7, q = a >> 3, r = a & 7;
  const int x = (j + c) & 7;  // the XCD of item j's workgroup
  int first = x * q;          // items of the XCDs before x
  for (int y = 0; y < x; y++) first += ((y - c) & 7) < r ? 1 : 0;
  return first + ((j - ((x - c) & 7)) >> 3);
}

#define FLUID_ITEMS(S, W) FLUID_ITEMS_AT(S, W, (int)blockIdx.x, (int)gridDim.x)
#define FLUID_ITEMS_AT(S, W, b0, G)                                                              \
  for (int w_ = (b0), nw_ = items_count(W); w_ < nw_; w_ += (G))                                  \
    if (const int2 it_ = (W)[1 + w_]; true)                                                          \
      if (const SceneArgs& A = (S)[it_.x]; true)                                                     \
        if (const int lb = fluid_block(it_.y, w_ - it_.y,                                            \
                                       (int)((((int64_t)gld(A.src.lstart + A.fp.cg.M)) << A.fp.fl_log2) + \
                                             FLUID_BLOCK - 1) / FLUID_BLOCK);                        \
            true)

// Minimum waves per SIMD asked of the register allocator for the three
// gather passes (0: the allocator's choice).  These loops are bound by
// memory latency and issue, so resident waves matter more than registers:
// 8 (64 VGPRs, no vector spills) measured 1% faster than the allocator's 7.
#ifndef DENSITY_WAVES
#define DENSITY_WAVES 8
#endif
#ifndef DISPLACE_WAVES
#define DISPLACE_WAVES 8
#endif
#ifndef XSPH_WAVES
#define XSPH_WAVES 8
#endif
#define DENSITY_ATTR WAVES_ATTR(DENSITY_WAVES)

template <bool kStage>
__global__ void __launch_bounds__(FLUID_BLOCK) DENSITY_ATTR kg_density(const SceneArgs* __restrict__ S,
                                                          const int2* __restrict__ W) {
  __shared__ uint32_t lds[kStage ? DENSITY_STAGE * FLUID_BLOCK : 1];
  FLUID_ITEMS(S, W) density_body<kStage>(A, lb, lds);
}

struct DisplaceAcc {
  static constexpr bool kNeedPos = true;
  typedef float2 Extra;
  const float2* __restrict__ sP;
  float2 Px;
  float inv_H, max_pair;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  __device__ __forceinline__ float2 fetch(int q) const { return gld_at(sP, (uint32_t)q << 3); }
  __device__ __forceinline__ void use(int, int w, float dx, float dy, float dz, float d2, float2 Py) {
    const float d = fast_sqrt(d2);
    if (d < 1e-9f) return;
    const float u = 1.f - d * inv_H;
    float D = 0.25f * ((Px.x + Py.x) * u + (Px.y + Py.y) * u * u);
    D = fminf(fmaxf(D, -max_pair), max_pair);
    const float s = (float)w * D / d;
    sx += dx * s;
    sy += dy * s;
    sz += dz * s;
  }
  __device__ __forceinline__ void pair(int q, int w, float dx, float dy, float dz, float d2) {
    use(q, w, dx, dy, dz, d2, fetch(q));
  }
};

// spos2 + the clamped displacement -> pos for the fluid (x, y, z only: the
// inverse mass in pos.w is not read or written).
__device__ __forceinline__ void displace_body(const SceneArgs& A, int lb) {
  const FluidParams& fp = A.fp;
  const PairSource& src = A.src;
  const PairList& pl = A.pl;
  const float2* __restrict__ sP = A.sP;
  PairSelf me;
  const ReplayHead rh = replay_head(fp, pl, fluid_lane(lb));
  const float2 Px = gld(sP + min(fluid_lane(lb) >> fp.fl_log2, fp.n - 1));
  if (!pair_self(fp, src, fluid_lane(lb), me)) return;
  DisplaceAcc acc;
  acc.sP = sP;
  acc.Px = Px;
  acc.inv_H = fp.inv_H;
  acc.max_pair = fp.max_pair;
  recorded_pairs(fp, src, pl, me, rh, acc);
  const int L = 1 << fp.fl_log2;
  float sx = lanes_sum(acc.sx, L), sy = lanes_sum(acc.sy, L), sz = lanes_sum(acc.sz, L);
  if (me.h != 0) return;
  const float s2 = sx * sx + sy * sy + sz * sz;
  if (s2 > fp.max_total * fp.max_total) {
    const float g = fp.max_total / sqrtf(s2);
    sx *= g;
    sy *= g;
    sz *= g;
  }
  // the particle's substep ends here: bounds, then its velocity for XSPH
  fluid_finish(A, me.x(), me.e, make_float4(me.p.x + sx, me.p.y + sy, me.p.z + sz, 0.f));
}

__global__ void __launch_bounds__(FLUID_BLOCK) WAVES_ATTR(DISPLACE_WAVES) kg_displace(const SceneArgs* __restrict__ S,
                                                           const int2* __restrict__ W) {
  FLUID_ITEMS(S, W) displace_body(A, lb);
}

struct XsphAcc {
  static constexpr bool kNeedPos = false;
  typedef float4 Extra;
  const float4* __restrict__ svtmp;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  int m = 0;
  __device__ __forceinline__ float4 fetch(int q) const { return gld_xyz_at(svtmp, (uint32_t)q << 4); }
  __device__ __forceinline__ void use(int, int w, float, float, float, float, float4 v) {
    const float fw = (float)w;
    sx += fw * v.x;
    sy += fw * v.y;
    sz += fw * v.z;
    m += w;
  }
  __device__ __forceinline__ void pair(int q, int w, float dx, float dy, float dz, float d2) {
    use(q, w, dx, dy, dz, d2, fetch(q));
  }
};

// XSPH: svtmp -> vel for the fluid that is still active (the reference skips
// particles culled since the density pass, whose velocity stays as it was).
__device__ __forceinline__ void xsph_body(const SceneArgs& A, int lb) {
  const FluidParams& fp = A.fp;
  const PairSource& src = A.src;
  const PairList& pl = A.pl;
  const float4* __restrict__ svtmp = A.svtmp;
  float4* __restrict__ vel = A.vel;
  const uint8_t* __restrict__ active = A.active;
  // the solids' bounding box is rebuilt by the next integrate; both contact
  // passes, its readers, are done
  if (lb == 0 && threadIdx.x < 6) gst(A.sbox + threadIdx.x, threadIdx.x < 3 ? INT_MAX : INT_MIN);
  // every particle has ended the substep (the displacement and the last
  // constraint iteration ran before this pass): a raised non-finite flag
  // goes to the scene's host-mapped word, which the host reads after a sync
  if (lb == 0 && threadIdx.x == 0 && gld(A.dflags + DF_NONFINITE)) A.tear[1] = 1;
  PairSelf me;
  const ReplayHead rh = replay_head(fp, pl, fluid_lane(lb));
  const float4 v = gld(svtmp + min(fluid_lane(lb) >> fp.fl_log2, fp.n - 1));
  if (!pair_self(fp, src, fluid_lane(lb), me)) return;
  // a particle culled since the density pass keeps its velocity: its flag is
  // read beside the replay and decides only the store
  const bool act = gld(active + me.x()) != 0;
  XsphAcc acc;
  acc.svtmp = svtmp;
  recorded_pairs(fp, src, pl, me, rh, acc);
  const int L = 1 << fp.fl_log2;
  const int m = lanes_sum(acc.m, L);
  const float sx = lanes_sum(acc.sx, L), sy = lanes_sum(acc.sy, L), sz = lanes_sum(acc.sz, L);
  if (me.h != 0 || !act) return;
  if (m == 0) {
    gst(vel + me.x(), v);
    return;
  }
  const float inv_m = 1.f / (float)m;
  gst(vel + me.x(), make_float4(v.x + fp.visc * (sx * inv_m - v.x), v.y + fp.visc * (sy * inv_m - v.y),
                                v.z + fp.visc * (sz * inv_m - v.z), 0.f));
}

__global__ void __launch_bounds__(FLUID_BLOCK) WAVES_ATTR(XSPH_WAVES) kg_xsph(const SceneArgs* __restrict__ S, const int2* __restrict__ W) {
  FLUID_ITEMS(S, W) xsph_body(A, lb);
}

// ---- the solid chain in the fluid chain's launches ----------------------------
//
// After the cell grid, a substep's solid chain (solid contacts, then the
// constraint iterations) and its fluid chain write disjoint data: each pass
// writes its own particles' positions (and velocities, active flags), the
// constraints their own CSR entries and ping-pong buffers; the only shared
// words (the non-finite flag, a tear flag) take the same store from every
// writer.  A group that does not fork runs the solid passes on workgroups of
// their own inside the fluid passes' launches: the solid contacts beside the
// fluid contacts, the constraint iterations beside the fluid rank, the density
// and the displacement.  Each particle's arithmetic is unchanged, so the
// results are the same bits as on separate launches; a small scene's
// substep takes 12 launches instead of 16.  Solid workgroups come first.

template <int L>
__global__ void __launch_bounds__(BLOCK) kg_contacts_both(const SceneArgs* __restrict__ S,
                                                          const int2* __restrict__ WS,
                                                          const int2* __restrict__ WP, int g0) {
  __shared__ int lds_bonds[(BLOCK / L) * MAXB];
  if ((int)blockIdx.x < g0) {
    FOR_ITEMS_AT(S, WS, (int)blockIdx.x, g0) contacts_solid_body<L>(A, lb, lds_bonds);
  } else {
    FOR_ITEMS_AT(S, WP, (int)blockIdx.x - g0, (int)gridDim.x - g0) contacts_fluid_body(A, lb);
  }
}

template <bool kStage>
__global__ void __launch_bounds__(FLUID_BLOCK) DENSITY_ATTR kg_density_cons(const SceneArgs* __restrict__ S,
                                                                           const int2* __restrict__ WC,
                                                                           const int2* __restrict__ WF, int g0,
                                                                           int it) {
  __shared__ uint32_t lds[kStage ? DENSITY_STAGE * FLUID_BLOCK : 1];
  if ((int)blockIdx.x < g0) {
    FOR_ITEMS_AT(S, WC, (int)blockIdx.x, g0) constraints_body(A, lb, it);
  } else {
    FLUID_ITEMS_AT(S, WF, (int)blockIdx.x - g0, (int)gridDim.x - g0) density_body<kStage>(A, lb, lds);
  }
}

__global__ void __launch_bounds__(FLUID_BLOCK) WAVES_ATTR(DISPLACE_WAVES)
    kg_displace_cons(const SceneArgs* __restrict__ S, const int2* __restrict__ WC, const int2* __restrict__ WF,
                     int g0, int it) {
  if ((int)blockIdx.x < g0) {
    FOR_ITEMS_AT(S, WC, (int)blockIdx.x, g0) constraints_body(A, lb, it);
  } else {
    FLUID_ITEMS_AT(S, WF, (int)blockIdx.x - g0, (int)gridDim.x - g0) displace_body(A, lb);
  }
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
  fp.cs = P.cell_size / CELL_SUB;
  fp.cull_r2 = (float)(fp.H2d * (1.0 + 1e-4));
  {
    const double rc = P.solid_radius + std::max(P.solid_radius, P.fluid_radius);
    fp.contact_r2 = (float)(rc * rc * (1.0 + 1e-4));
  }
  fp.max_pair = (float)(0.05 * P.fluid_h);
  fp.max_total = (float)(0.35 * P.fluid_h);
  fp.stiff = (float)P.stiffness;
  fp.near_stiff = (float)P.near_stiffness;
  fp.rest_density = (float)P.rest_density;
  fp.contact_scale = (float)P.contact_scale;
  fp.fl_log2 = s->fluid_lanes == 8 ? 3 : s->fluid_lanes == 4 ? 2 : s->fluid_lanes == 2 ? 1 : 0;
  // entries the staging density kernel holds in LDS per lane (whether the
  // group runs that kernel is the group's choice, fp_group::stage_on)
  fp.stage = DENSITY_STAGE;

  fp.visc = (float)P.viscosity;
  return fp;
}

// Blocks of one scene in a pass of the given kind (fp_internal.h WorkKind).
int fp_scene_blocks(const fp_solver* s, int kind, int contact_lanes) {
  switch (kind) {
    case W_PARTICLE:
      return fp_blocks(s->n);
    case W_SCAN:
      return s->lscan_blocks;
    case W_SOLID:
      return fp_blocks(s->ns * contact_lanes);
    case W_FLUID:
      return (int)(((int64_t)s->n * s->fluid_lanes + FLUID_BLOCK - 1) / FLUID_BLOCK);
    case W_CONS:
      return (int)(((int64_t)s->ns * FP_CLANES + BLOCK - 1) / BLOCK);
    case W_TET:
      return s->tets_on ? fp_blocks(s->NT) : 0;
  }
  return 0;
}

// grid: the workgroups (a multiple of 8), items: a work list with its count.
static void launch_contacts_solid(fp_group* g, hipStream_t st, int lanes, const int2* items, int grid) {
  if (grid <= 0) return;
  switch (lanes) {
    case 32:
      hipLaunchKernelGGL(kg_contacts_solid<32>, dim3(grid), dim3(BLOCK), 0, st, g->d_args, items);
      break;
    case 16:
      hipLaunchKernelGGL(kg_contacts_solid<16>, dim3(grid), dim3(BLOCK), 0, st, g->d_args, items);
      break;
    case 8:
      hipLaunchKernelGGL(kg_contacts_solid<8>, dim3(grid), dim3(BLOCK), 0, st, g->d_args, items);
      break;
    default:
      hipLaunchKernelGGL(kg_contacts_solid<4>, dim3(grid), dim3(BLOCK), 0, st, g->d_args, items);
  }
}

// Solid contacts on stream st (the group's stream, or the side stream of a
// captured frame).  A tuning frame launches this substep's candidate from its
// own work list, between events (tuning frames are never captured).
int fp_group_contacts_solid(fp_group* g, hipStream_t st) {
  if (g->tune_record) {
    const int c = g->tune_cand;
    const size_t need = 2 * (size_t)(g->tune_used + 1);
    while (g->tune_ev.size() < need) {
      hipEvent_t e;
      FP_GTRY(g, hipEventCreate(&e));
      g->tune_ev.push_back(e);
    }
    hipEvent_t* ev = &g->tune_ev[2 * (size_t)g->tune_used++];
    g->tune_of.push_back(c);
    FP_GTRY(g, hipEventRecord(ev[0], st));
    launch_contacts_solid(g, st, kTuneLanes[c], g->d_tune_items + g->tune_off[c], fp_grid8(g->tune_n[c]));
    FP_GTRY(g, hipEventRecord(ev[1], st));
  } else {
    launch_contacts_solid(g, st, g->clanes, g->d_items[W_SOLID],
                          g->capturing ? g->grid_cap[W_SOLID] : fp_grid8(g->n_items[W_SOLID]));
  }
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_group_contacts_fluid(fp_group* g) {
  GROUP_LAUNCH(g, kg_contacts_fluid, W_PARTICLE, BLOCK);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

// Both contact passes, one after the other (timed as one pass).
int fp_group_contacts(fp_group* g) {
  PassScope ps(g->prof, g->stream, P_CONTACTS);
  int rc = fp_group_contacts_solid(g, g->stream);
  if (rc) return rc;
  return fp_group_contacts_fluid(g);
}

// The solid and the fluid contacts in one launch (the group's lane count).
int fp_group_contacts_both(fp_group* g) {
  PassScope ps(g->prof, g->stream, P_CONTACTS);
  switch (g->clanes) {
    case 32:
      GROUP_LAUNCH2(g, kg_contacts_both<32>, W_SOLID, W_PARTICLE, BLOCK);
      break;
    case 16:
      GROUP_LAUNCH2(g, kg_contacts_both<16>, W_SOLID, W_PARTICLE, BLOCK);
      break;
    case 8:
      GROUP_LAUNCH2(g, kg_contacts_both<8>, W_SOLID, W_PARTICLE, BLOCK);
      break;
    default:
      GROUP_LAUNCH2(g, kg_contacts_both<4>, W_SOLID, W_PARTICLE, BLOCK);
  }
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_group_density_with(fp_group* g, int it) {
  if (it < 0) return fp_group_density(g);
  PassScope ps(g->prof, g->stream, P_DENSITY);
  if (g->stage_on)
    GROUP_LAUNCH2(g, kg_density_cons<true>, W_CONS, W_FLUID, FLUID_BLOCK, it);
  else
    GROUP_LAUNCH2(g, kg_density_cons<false>, W_CONS, W_FLUID, FLUID_BLOCK, it);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_group_displace_with(fp_group* g, int it) {
  if (it < 0) return fp_group_displace(g);
  PassScope ps(g->prof, g->stream, P_DISPLACE);
  GROUP_LAUNCH2(g, kg_displace_cons, W_CONS, W_FLUID, FLUID_BLOCK, it);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_group_density(fp_group* g) {
  PassScope ps(g->prof, g->stream, P_DENSITY);
  if (g->stage_on)
    GROUP_LAUNCH(g, kg_density<true>, W_FLUID, FLUID_BLOCK);
  else
    GROUP_LAUNCH(g, kg_density<false>, W_FLUID, FLUID_BLOCK);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_group_displace(fp_group* g) {
  PassScope ps(g->prof, g->stream, P_DISPLACE);
  GROUP_LAUNCH(g, kg_displace, W_FLUID, FLUID_BLOCK);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_group_xsph(fp_group* g) {
  PassScope ps(g->prof, g->stream, P_XSPH);
  GROUP_LAUNCH(g, kg_xsph, W_FLUID, FLUID_BLOCK);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}
