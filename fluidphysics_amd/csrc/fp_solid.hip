// fp_solid.hip -- distance constraints, the knife, pick.
//
// constraints  one Jacobi XPBD iteration over the static CSR adjacency
//              (solver.js:356-381 per-constraint formula).  FP_CLANES lanes
//              work on one solid, each over every FP_CLANES-th incident
//              constraint; the corrections are summed across the lanes with
//              xor shuffles.  Each constraint's correction is scaled by
//              omega / max(deg_a, deg_b) -- the same factor at both
//              endpoints, so the pair's corrections stay equal and opposite.
//              Tearing (d > rest * breakStrain) and plasticity (|C/rest| >
//              threshold -> rest += C * rate) are decided from the iteration's
//              snapshot, so both endpoints agree; each endpoint updates its
//              own CSR entry's copy of the rest length and alive flag.
// cut          the knife wedge test of solver.js:458-486 in IEEE double with
//              the reference's operation order (the library is built with
//              -ffp-contract=off), one thread per constraint.
// pick         pickSolid (solver.js:502-515): one workgroup, lexicographic
//              minimum of (t, index) over the solids that pass both tests.
#include "fp_internal.h"
#include "fp_kernels.h"
#include "fp_bounds.h"
#include "fp_cons.h"

__global__ void __launch_bounds__(BLOCK) kg_constraints(const SceneArgs* __restrict__ S, const int2* __restrict__ W, int it) {
  FOR_ITEMS(S, W) constraints_body(A, lb, it);
}

// ---- the GPU twin's optional solid modes --------------------------------------
//
// Volume tets (wgsl.js:423-519) and the solid contact cache (wgsl.js:532-606)
// as this library formulates them: the cache is collected once per substep
// from the grid (the solids within contact distance + margin of each active
// solid, not bonded to it: ascending index, FP_CACHE_K kept); in every
// constraint iteration, after the distance constraints, each active solid
// gathers from that iteration's positions its share of every alive incident
// tet's zero-compliance volume projection (scaled by 1 / max(incident tets
// of the tet's vertices), each tet's share and the sum capped at half a
// solid radius) and of every cached contact closer than the contact distance
// (half the overlap, split by inverse mass, the sum capped the same way);
// after the iterations, tets with an edge longer than its lattice rest + the
// tear gap die.

// Lattice neighbours (a bond between them could exist): the contact distance
// of a cut face is just under the spacing.
__device__ __forceinline__ bool lattice_adjacent(int i, int j, int n) {
  const int dx = i / (n * n) - j / (n * n), dy = (i / n) % n - (j / n) % n, dz = i % n - j % n;
  return abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1;
}

__device__ __forceinline__ bool bonded_csr(const SceneArgs& A, int x, int y) {
  const int e = gld(A.adj_off + x + 1);
  bool hit = false;
  for (int k = gld(A.adj_off + x); k < e; k++) hit |= adj_other(x, gld(A.adj + k)) == y && gld(A.ealive + k) != 0;
  return hit;
}

// The contact cache of solid i from the grid (positions after integrate and
// grab; run before the contact pass moves the solids).
__device__ __forceinline__ void solid_collect_body(const SceneArgs& A, int lb) {
  const SolidModes& sm = A.sm;
  const int i = lb * BLOCK + (int)threadIdx.x;
  if (!sm.cache || i >= A.ns) return;
  int* ids = sm.cache_ids + (size_t)i * FP_CACHE_K;
  int cnt = 0, over = 0;
  if (gld(A.active + i)) {
    const FluidParams& fp = A.fp;
    const float4 p = gld(A.pos + i);
    const uint32_t kx = cell_key(fp.cg, cell_of(p.x, p.y, p.z, fp.cg.inv_cell));
    for (int r = 0; r < NB_ROWS; r++) {
      const uint32_t k0 = row_key(fp.cg, kx, r / NB_RUN - NB_R, r % NB_RUN - NB_R);
      for (uint32_t j3 = 0; j3 < NB_RUN; j3++) {
        const uint32_t k = (k0 + j3) & fp.cg.kmask;
        const int e1 = gld(A.lstart + k + 1);
        for (int q = gld(A.lstart + k); q < e1; q++) {
          const int y = gld(A.sid + q);
          if (y >= A.ns || y == i) continue;
          const float4 py = gld(A.spos + q);
          const float dx = p.x - py.x, dy = p.y - py.y, dz = p.z - py.z;
          const float lim = lattice_adjacent(i, y, sm.n) ? sm.lim_adj : sm.lim;
          if (dx * dx + dy * dy + dz * dz >= lim * lim) continue;
          if (bonded_csr(A, i, y)) continue;
          // sorted insert, keeping the FP_CACHE_K smallest indices (a cell
          // two neighbour keys alias to is seen twice: no duplicates)
          int pos = cnt;
          while (pos > 0 && ids[pos - 1] > y) pos--;
          if (pos > 0 && ids[pos - 1] == y) continue;
          if (cnt == FP_CACHE_K) {
            over++;
            if (pos == FP_CACHE_K) continue;
          } else {
            cnt++;
          }
          for (int m = cnt - 1; m > pos; m--) ids[m] = ids[m - 1];
          ids[pos] = y;
        }
      }
    }
  }
  gst(sm.cache_n + i, cnt);
  if (over) atomicAdd(&A.dflags[DF_CACHE_OVF], over);
}

__global__ void __launch_bounds__(BLOCK) kg_solid_collect(const SceneArgs* __restrict__ S, const int2* __restrict__ W) {
  FOR_ITEMS(S, W) solid_collect_body(A, lb);
}

__device__ __forceinline__ void cap3(float& x, float& y, float& z, float lim) {
  const float m2 = x * x + y * y + z * z;
  if (m2 > lim * lim) {
    const float f = lim / sqrtf(m2);
    x *= f;
    y *= f;
    z *= f;
  }
}

// The extra step of iteration it for solid i: from spos_c[it & 1] (after the
// distance constraints) into spos_x; the last iteration also finishes the
// solid's substep (bounds, velocity).
__device__ __forceinline__ void solid_extra_body(const SceneArgs& A, int lb, int it) {
  const SolidModes& sm = A.sm;
  const int i = lb * BLOCK + (int)threadIdx.x;
  if (!(sm.tets | sm.cache) || i >= A.ns) return;
  const float4* __restrict__ x = A.spos_c[it & 1];
  float4 p = gld(x + i);
  if (gld(A.active + i)) {
    float tx = 0.f, ty = 0.f, tz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    if (sm.tets) {
      const int k1 = gld(sm.tet_adj_off + i + 1);
      for (int k = gld(sm.tet_adj_off + i); k < k1; k++) {
        const int t = gld(sm.tet_adj + k);
        if (!gld(sm.tet_alive + t)) continue;
        const int4 v = gld(sm.tet_v + t);
        const float4 q0 = gld(x + v.x), q1 = gld(x + v.y), q2 = gld(x + v.z), q3 = gld(x + v.w);
        const float bx = q1.x - q0.x, by = q1.y - q0.y, bz = q1.z - q0.z;
        const float ccx = q2.x - q0.x, ccy = q2.y - q0.y, ccz = q2.z - q0.z;
        const float dx = q3.x - q0.x, dy = q3.y - q0.y, dz = q3.z - q0.z;
        // volume gradients: g1 = (c x d)/6, g2 = (d x b)/6, g3 = (b x c)/6, g0 = -(g1 + g2 + g3)
        const float g1x = (ccy * dz - ccz * dy) / 6.f, g1y = (ccz * dx - ccx * dz) / 6.f, g1z = (ccx * dy - ccy * dx) / 6.f;
        const float g2x = (dy * bz - dz * by) / 6.f, g2y = (dz * bx - dx * bz) / 6.f, g2z = (dx * by - dy * bx) / 6.f;
        const float g3x = (by * ccz - bz * ccy) / 6.f, g3y = (bz * ccx - bx * ccz) / 6.f, g3z = (bx * ccy - by * ccx) / 6.f;
        const float g0x = -(g1x + g2x + g3x), g0y = -(g1y + g2y + g3y), g0z = -(g1z + g2z + g3z);
        const float V = g3x * dx + g3y * dy + g3z * dz;
        const float C = V - gld(sm.tet_v0 + t);
        const float denom = q0.w * (g0x * g0x + g0y * g0y + g0z * g0z) + q1.w * (g1x * g1x + g1y * g1y + g1z * g1z) +
                            q2.w * (g2x * g2x + g2y * g2y + g2z * g2z) + q3.w * (g3x * g3x + g3y * g3y + g3z * g3z);
        if (!(denom >= 1e-30f)) continue;
        float gx = g0x, gy = g0y, gz = g0z;
        if (v.y == i) {
          gx = g1x, gy = g1y, gz = g1z;
        } else if (v.z == i) {
          gx = g2x, gy = g2y, gz = g2z;
        } else if (v.w == i) {
          gx = g3x, gy = g3y, gz = g3z;
        }
        const float f = -C / denom * p.w * gld(sm.tet_scale + t);
        float ux = gx * f, uy = gy * f, uz = gz * f;
        cap3(ux, uy, uz, sm.cap);
        tx += ux;
        ty += uy;
        tz += uz;
      }
      cap3(tx, ty, tz, sm.cap);
    }
    if (sm.cache) {
      const int n = gld(sm.cache_n + i);
      const int* ids = sm.cache_ids + (size_t)i * FP_CACHE_K;
      for (int k = 0; k < n; k++) {
        const int j = gld(ids + k);
        const float4 q = gld(x + j);
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float rsum = lattice_adjacent(i, j, sm.n) ? sm.rsum_adj : sm.rsum;
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 >= rsum * rsum || d2 < 1e-12f) continue;
        const float wsum = p.w + q.w;
        if (wsum == 0.f) continue;
        const float d = sqrtf(d2);
        const float f = (rsum - d) / d * p.w / wsum * 0.5f;
        cx += dx * f;
        cy += dy * f;
        cz += dz * f;
      }
      cap3(cx, cy, cz, sm.cap);
    }
    p.x += tx + cx;
    p.y += ty + cy;
    p.z += tz + cz;
  }
  gst(sm.spos_x + i, p);  // the last iteration's: the positions the tet tear check reads
  if (it == A.cp.iters - 1) solid_finish(A, i, p);
}

__global__ void __launch_bounds__(BLOCK)
    kg_solid_extra(const SceneArgs* __restrict__ S, const int2* __restrict__ W, int it) {
  FOR_ITEMS(S, W) solid_extra_body(A, lb, it);
}

// Tets with an edge longer than its lattice rest + the tear gap die
// (tetUpdate, wgsl.js:471-489); lengths in double from the positions after
// the last iteration (before bounds).
__device__ __forceinline__ void tet_update_body(const SceneArgs& A, int lb) {
  const SolidModes& sm = A.sm;
  const int t = lb * BLOCK + (int)threadIdx.x;
  if (!sm.tets || t >= sm.NT || !gld(sm.tet_alive + t)) return;
  const int4 v4 = gld(sm.tet_v + t);
  const int v[4] = {v4.x, v4.y, v4.z, v4.w};
  const int n = sm.n;
  for (int q = 0; q < 3; q++)
    for (int r = q + 1; r < 4; r++) {
      const float4 a = gld(sm.spos_x + v[q]), b = gld(sm.spos_x + v[r]);
      const double dx = (double)a.x - b.x, dy = (double)a.y - b.y, dz = (double)a.z - b.z;
      const double ix = v[q] / (n * n) - v[r] / (n * n), iy = (v[q] / n) % n - (v[r] / n) % n,
                   iz = v[q] % n - v[r] % n;
      const double rest = sm.spacing * sqrt(ix * ix + iy * iy + iz * iz);
      if (sqrt(dx * dx + dy * dy + dz * dz) > rest + sm.gap) {
        gst(sm.tet_alive + t, (uint8_t)0);
        return;
      }
    }
}

__global__ void __launch_bounds__(BLOCK) kg_tet_update(const SceneArgs* __restrict__ S, const int2* __restrict__ W) {
  FOR_ITEMS(S, W) tet_update_body(A, lb);
}

// Does segment a-b cross the knife wedge (solver.js:466-483 in IEEE double,
// the reference's operation order)?
struct Wedge {
  double ox, oy, oz, d1x, d1y, d1z, d2x, d2y, d2z, nx, ny, nz;
};

__device__ __forceinline__ bool seg_in_wedge(const Wedge& w, const float4& fa, const float4& fb) {
  const double ax = fa.x, ay = fa.y, az = fa.z, bx = fb.x, by = fb.y, bz = fb.z;
  const double sa = (ax - w.ox) * w.nx + (ay - w.oy) * w.ny + (az - w.oz) * w.nz;
  const double sb = (bx - w.ox) * w.nx + (by - w.oy) * w.ny + (bz - w.oz) * w.nz;
  if (sa * sb >= 0) return false;
  const double t = sa / (sa - sb);
  const double px = ax + (bx - ax) * t - w.ox;
  const double py = ay + (by - ay) * t - w.oy;
  const double pz = az + (bz - az) * t - w.oz;
  const double c1 = (w.d1y * pz - w.d1z * py) * w.nx + (w.d1z * px - w.d1x * pz) * w.ny + (w.d1x * py - w.d1y * px) * w.nz;
  const double c2 = (py * w.d2z - pz * w.d2y) * w.nx + (pz * w.d2x - px * w.d2z) * w.ny + (px * w.d2y - py * w.d2x) * w.nz;
  return !(c1 < 0 || c2 < 0);
}

__global__ void __launch_bounds__(BLOCK) k_cut(int nc, const float4* __restrict__ pos, const int2* __restrict__ ab,
                                               const int2* __restrict__ cent, uint8_t* __restrict__ ealive, Wedge w,
                                               int* __restrict__ ncut) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  if (c >= nc) return;
  const int2 ke = cent[c];
  if (!ealive[ke.x]) return;
  const int2 e = ab[c];
  if (!seg_in_wedge(w, pos[e.x], pos[e.y])) return;
  ealive[ke.x] = 0;
  ealive[ke.y] = 0;
  atomicAdd(ncut, 1);
}

// The tets with an edge across the wedge stop holding volume (the GPU twin's
// cutTets, wgsl.js:492-519, with the predicate above).
__global__ void __launch_bounds__(BLOCK) k_cut_tets(int nt, const float4* __restrict__ pos,
                                                    const int4* __restrict__ tv, uint8_t* __restrict__ alive, Wedge w) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= nt || !alive[t]) return;
  const int4 v4 = tv[t];
  const int v[4] = {v4.x, v4.y, v4.z, v4.w};
  for (int q = 0; q < 3; q++)
    for (int r = q + 1; r < 4; r++)
      if (seg_in_wedge(w, pos[v[q]], pos[v[r]])) {
        alive[t] = 0;
        return;
      }
}

__global__ void __launch_bounds__(BLOCK) k_pick(int ns, const float4* __restrict__ pos, double ox, double oy,
                                                double oz, double dx, double dy, double dz, double max_perp2,
                                                double* __restrict__ out_t, int* __restrict__ out_i) {
  __shared__ double st[BLOCK];
  __shared__ int si[BLOCK];
  double bt = INFINITY;
  int bi = -1;
  for (int i = threadIdx.x; i < ns; i += BLOCK) {
    const float4 p = pos[i];
    const double px = (double)p.x - ox, py = (double)p.y - oy, pz = (double)p.z - oz;
    const double t = px * dx + py * dy + pz * dz;
    if (t < 0 || t >= bt) continue;
    const double qx = px - dx * t, qy = py - dy * t, qz = pz - dz * t;
    if (qx * qx + qy * qy + qz * qz > max_perp2) continue;
    bt = t;
    bi = i;
  }
  st[threadIdx.x] = bt;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double t2 = st[threadIdx.x + s];
      const int i2 = si[threadIdx.x + s];
      const double t1 = st[threadIdx.x];
      const int i1 = si[threadIdx.x];
      if (i2 >= 0 && (i1 < 0 || t2 < t1 || (t2 == t1 && i2 < i1))) {
        st[threadIdx.x] = t2;
        si[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *out_t = st[0];
    *out_i = si[0];
  }
}

// The grab selection of the reference's caller (src/main.js:293-301): solids
// of component g (any when g < 0) within radius of c, in ascending index
// order; the distance is decided in double from the float32 positions, as
// the reference's JS does.  One workgroup walks the solids in chunks of
// BLOCK: a wave ballot and the waves' counts give each match its place.
// *n receives the total number of matches (ids past max are not written).
// same >= 0: only the solids whose label equals solid same's.
__global__ void __launch_bounds__(BLOCK) k_select(int ns, const float4* __restrict__ pos,
                                                  const int* __restrict__ group, int same, double cx, double cy,
                                                  double cz, double r2, int* __restrict__ ids, int max_ids,
                                                  int* __restrict__ n) {
  __shared__ int wave_count[BLOCK / 64];
  __shared__ int base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const int g = same >= 0 ? group[same] : -1;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  for (int c0 = 0; c0 < ns; c0 += BLOCK) {
    const int i = c0 + (int)threadIdx.x;
    bool hit = false;
    if (i < ns && (g < 0 || group[i] == g)) {
      const float4 p = pos[i];
      const double dx = (double)p.x - cx, dy = (double)p.y - cy, dz = (double)p.z - cz;
      hit = dx * dx + dy * dy + dz * dz < r2;
    }
    const uint64_t m = __ballot(hit);
    if (lane == 0) wave_count[wave] = __popcll(m);
    __syncthreads();
    int before = base;
    for (int w = 0; w < wave; w++) before += wave_count[w];
    before += __popcll(m & ((1ull << lane) - 1ull));
    if (hit && before < max_ids) ids[before] = i;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 0; w < BLOCK / 64; w++) base += wave_count[w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *n = base;
}

int fp_launch_select(fp_solver* s, const double c[3], double r2, int same, int* ids, int max_ids, int* n) {
  hipLaunchKernelGGL(k_select, dim3(1), dim3(BLOCK), 0, s->stream, s->ns, s->pos, s->dgroup, same, c[0], c[1], c[2],
                     r2, ids, max_ids, n);
  FP_TRY(s, hipGetLastError());
  return FP_OK;
}

// ---- connected components of the solids (solver.js:488-499) -------------------
//
// The partition _recomputeGroups computes, on the device: a lock-free
// union-find over the alive constraints.  parent[x] <= x always; a hook
// links the larger of two roots under the smaller with a compare-and-swap,
// and a failed swap means that root gained a parent meanwhile, so the loop
// continues from a strictly smaller root and ends.  Finding halves the path
// (each write stores an ancestor, so the forest stays valid whatever the
// interleaving).  The root of a component is its smallest solid index, which
// becomes every member's label.  (The reference labels a component by the
// root its sequential union order leaves, which the host relabel keeps for
// the group array; the partition is the same.)

__device__ __forceinline__ int cc_load(const int* p) {
  return __hip_atomic_load((const FP_GLOBAL int*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void cc_store(int* p, int v) {
  __hip_atomic_store((FP_GLOBAL int*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int cc_find(int* parent, int x) {
  int cur = cc_load(parent + x);
  if (cur == x) return x;
  int prev = x, next;
  while (cur > (next = cc_load(parent + cur))) {
    cc_store(parent + prev, next);
    prev = cur;
    cur = next;
  }
  return cur;
}

__global__ void __launch_bounds__(BLOCK) k_cc_init(int ns, int* __restrict__ parent) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < ns) parent[i] = i;
}

__global__ void __launch_bounds__(BLOCK) k_cc_hook(int C, const int2* __restrict__ cons_ab,
                                                   const int2* __restrict__ cent,
                                                   const uint8_t* __restrict__ ealive, int* __restrict__ parent) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  if (c >= C) return;
  if (!ealive[cent[c].x]) return;
  const int2 ab = cons_ab[c];
  int ra = cc_find(parent, ab.x), rb = cc_find(parent, ab.y);
  while (ra != rb) {
    const int hi = max(ra, rb), lo = min(ra, rb);
    int expect = hi;
    if (__hip_atomic_compare_exchange_strong((FP_GLOBAL int*)(parent + hi), &expect, lo, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      break;
    // hi gained the parent `expect` (< hi) meanwhile: continue from the roots
    ra = cc_find(parent, expect);
    rb = cc_find(parent, lo);
  }
}

__global__ void __launch_bounds__(BLOCK) k_cc_final(int ns, int* __restrict__ parent, int* __restrict__ ncomp) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const int r = i < ns ? cc_find(parent, i) : -1;
  if (i < ns) cc_store(parent + i, r);
  const int roots = (int)__popcll(__ballot(i < ns && r == i));
  if ((threadIdx.x & 63) == 0 && roots) atomicAdd(ncomp, roots);
}

// Component labels of the solids into dgroup (and their count into
// dflags[DF_NCOMP]), queued on the solver's stream.
int fp_launch_components(fp_solver* s) {
  if (s->ns <= 0) return FP_OK;
  const int nb = fp_blocks(s->ns);
  hipLaunchKernelGGL(k_cc_init, dim3(nb), dim3(BLOCK), 0, s->stream, s->ns, s->dgroup);
  if (s->C > 0)
    hipLaunchKernelGGL(k_cc_hook, dim3(fp_blocks(s->C)), dim3(BLOCK), 0, s->stream, s->C, s->cons_ab, s->cent,
                       s->ealive, s->dgroup);
  FP_TRY(s, hipMemsetAsync(s->dflags + DF_NCOMP, 0, sizeof(int), s->stream));
  hipLaunchKernelGGL(k_cc_final, dim3(nb), dim3(BLOCK), 0, s->stream, s->ns, s->dgroup, s->dflags + DF_NCOMP);
  FP_TRY(s, hipGetLastError());
  return FP_OK;
}

// iters Jacobi iterations for every scene of the group (the solids' final
// positions end in spos_c[(iters - 1) & 1]; see fp_group_frame).
int fp_group_constraints(fp_group* g, int iters, hipStream_t st) {
  if (iters < 1) return FP_OK;
  if (!g->any_extra) {
    PassScope ps(g->prof, st, P_CONSTRAINTS);
    for (int it = 0; it < iters; it++) GROUP_LAUNCH_ON(g, st, kg_constraints, W_CONS, BLOCK, it);
    FP_GTRY(g, hipGetLastError());
    return FP_OK;
  }
  // with the optional solid modes, each iteration is the distance
  // constraints, then the modes' extra step (the profile counts both per
  // iteration)
  for (int it = 0; it < iters; it++) {
    {
      PassScope ps(g->prof, st, P_CONSTRAINTS);
      GROUP_LAUNCH_ON(g, st, kg_constraints, W_CONS, BLOCK, it);
    }
    int rc = fp_group_solid_extra(g, it, st);
    if (rc) return rc;
  }
  FP_GTRY(g, hipGetLastError());
  return fp_group_tet_update(g, st);
}

int fp_group_solid_collect(fp_group* g, hipStream_t st) {
  if (!g->any_cache) return FP_OK;
  PassScope ps(g->prof, st, P_SOLID_EXTRA);
  GROUP_LAUNCH_ON(g, st, kg_solid_collect, W_PARTICLE, BLOCK);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_group_solid_extra(fp_group* g, int it, hipStream_t st) {
  if (!g->any_extra) return FP_OK;
  PassScope ps(g->prof, st, P_SOLID_EXTRA);
  GROUP_LAUNCH_ON(g, st, kg_solid_extra, W_PARTICLE, BLOCK, it);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_group_tet_update(fp_group* g, hipStream_t st) {
  if (!g->any_tets) return FP_OK;
  PassScope ps(g->prof, st, P_SOLID_EXTRA);
  GROUP_LAUNCH_ON(g, st, kg_tet_update, W_TET, BLOCK);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_launch_cut(fp_solver* s, const double o[3], const double d1[3], const double d2[3], double nx, double ny,
                  double nz) {
  PassScope ps(s, P_CUT);
  const Wedge w{o[0], o[1], o[2], d1[0], d1[1], d1[2], d2[0], d2[1], d2[2], nx, ny, nz};
  if (s->C > 0)
    hipLaunchKernelGGL(k_cut, dim3(fp_blocks(s->C)), dim3(BLOCK), 0, s->stream, s->C, s->pos, s->cons_ab, s->cent,
                       s->ealive, w, s->dflags + DF_CUT);
  if (s->tets_on && s->NT > 0)
    hipLaunchKernelGGL(k_cut_tets, dim3(fp_blocks(s->NT)), dim3(BLOCK), 0, s->stream, s->NT, s->pos, s->tet_v,
                       s->tet_alive, w);
  FP_TRY(s, hipGetLastError());
  return FP_OK;
}

int fp_launch_pick(fp_solver* s, const double o[3], const double d[3], double max_perp2) {
  PassScope ps(s, P_PICK);
  hipLaunchKernelGGL(k_pick, dim3(1), dim3(BLOCK), 0, s->stream, s->ns, s->pos, o[0], o[1], o[2], d[0], d[1], d[2],
                     max_perp2, s->dpick_t, s->dpick_i);
  FP_TRY(s, hipGetLastError());
  return FP_OK;
}
