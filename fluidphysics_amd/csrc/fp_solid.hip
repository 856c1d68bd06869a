// fp_solid.hip -- distance constraints, the knife, pick.
//
// constraints  one Jacobi XPBD iteration over the static CSR adjacency
//              (solver.js:356-381 per-constraint formula).  32 lanes work on
//              one solid: lane k takes the solid's k-th incident constraint,
//              the corrections are summed across the 32 lanes with xor
//              shuffles.  Each constraint's correction is scaled by
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

#define CLANES 32

// One Jacobi iteration `it` of one scene: positions from pos (it = 0) or
// spos_c[(it - 1) & 1] into spos_c[it & 1].  Each CSR entry holds its own
// copy of the constraint's rest length and alive flag; both endpoints
// compute the same update from the same snapshot (d is the same bits either
// way round), so each lane updates its own entry in place, and only when
// the value changes (a tear, a plastic step).
__device__ __forceinline__ void constraints_body(const SceneArgs& A, int lb, int it) {
  const ConsParams& cp = A.cp;
  const float4* __restrict__ sin = it == 0 ? A.pos : A.spos_c[(it - 1) & 1];
  float4* __restrict__ sout = A.spos_c[it & 1];
  float* __restrict__ erest = A.erest;
  uint8_t* __restrict__ ealive = A.ealive;
  const int lane = threadIdx.x & (CLANES - 1);
  const int a = (lb * BLOCK + (int)threadIdx.x) / CLANES;
  const bool valid = a < cp.ns;
  float4 pa = make_float4(0.f, 0.f, 0.f, 0.f);
  int k0 = 0, k1 = 0;
  if (valid) {
    pa = gld(sin + a);
    k0 = gld(A.adj_off + a);
    k1 = gld(A.adj_off + a + 1);
  }
  float cx = 0.f, cy = 0.f, cz = 0.f;
  bool tore = false;
  for (int k = k0 + lane; k < k1; k += CLANES) {
    // the entry's flag, neighbour and rest length load together, then the
    // neighbour's position (a dead entry's neighbour index is still valid):
    // two dependent loads instead of three
    const bool alive = gld(ealive + k) != 0;
    const int2 e = gld(A.adj + k);
    const float rest = gld(erest + k);
    const float4 pb = gld(sin + e.x);
    const float w = pa.w + pb.w;
    const float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
    const float d = sqrtf(dx * dx + dy * dy + dz * dz);
    if (alive && w != 0.f && d >= 1e-9f) {
      if (d > rest * cp.break_strain) {
        gst(ealive + k, (uint8_t)0);
        tore = true;
      } else {
        const float C = d - rest;
        const float dl = -C / (w + cp.alpha);
        const float s = dl / d * pa.w * (cp.omega * __int_as_float(e.y));
        cx += dx * s;
        cy += dy * s;
        cz += dz * s;
        if (fabsf(C / rest) > cp.plastic_thr) gst(erest + k, rest + C * cp.plastic_rate);
      }
    }
  }
#pragma unroll
  for (int off = CLANES / 2; off > 0; off >>= 1) {
    cx += __shfl_xor(cx, off, CLANES);
    cy += __shfl_xor(cy, off, CLANES);
    cz += __shfl_xor(cz, off, CLANES);
  }
  if (valid && lane == 0) {
    const float4 p = make_float4(pa.x + cx, pa.y + cy, pa.z + cz, pa.w);
    // the last iteration finishes the solid's substep (bounds, velocity)
    if (it == A.cp.iters - 1)
      solid_finish(A, a, p);
    else
      gst(sout + a, p);
  }
  if (tore) *A.tear = 1;
}

__global__ void __launch_bounds__(BLOCK) kg_constraints(const SceneArgs* __restrict__ S, const int2* __restrict__ W, int it) {
  FOR_ITEMS(S, W) constraints_body(A, lb, it);
}

__global__ void __launch_bounds__(BLOCK) k_cut(int nc, const float4* __restrict__ pos, const int2* __restrict__ ab,
                                               const int2* __restrict__ cent, uint8_t* __restrict__ ealive,
                                               double ox, double oy, double oz,
                                               double d1x, double d1y, double d1z, double d2x, double d2y,
                                               double d2z, double nx, double ny, double nz, int* __restrict__ ncut) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  if (c >= nc) return;
  const int2 ke = cent[c];
  if (!ealive[ke.x]) return;
  const int2 e = ab[c];
  const float4 fa = pos[e.x], fb = pos[e.y];
  const double ax = fa.x, ay = fa.y, az = fa.z, bx = fb.x, by = fb.y, bz = fb.z;
  const double sa = (ax - ox) * nx + (ay - oy) * ny + (az - oz) * nz;
  const double sb = (bx - ox) * nx + (by - oy) * ny + (bz - oz) * nz;
  if (sa * sb >= 0) return;
  const double t = sa / (sa - sb);
  const double px = ax + (bx - ax) * t - ox;
  const double py = ay + (by - ay) * t - oy;
  const double pz = az + (bz - az) * t - oz;
  const double c1 = (d1y * pz - d1z * py) * nx + (d1z * px - d1x * pz) * ny + (d1x * py - d1y * px) * nz;
  const double c2 = (py * d2z - pz * d2y) * nx + (pz * d2x - px * d2z) * ny + (px * d2y - py * d2x) * nz;
  if (c1 < 0 || c2 < 0) return;
  ealive[ke.x] = 0;
  ealive[ke.y] = 0;
  atomicAdd(ncut, 1);
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
  PassScope ps(g->prof, st, P_CONSTRAINTS);
  for (int it = 0; it < iters; it++) GROUP_LAUNCH_ON(g, st, kg_constraints, W_CONS, BLOCK, it);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_launch_cut(fp_solver* s, const double o[3], const double d1[3], const double d2[3], double nx, double ny,
                  double nz) {
  PassScope ps(s, P_CUT);
  hipLaunchKernelGGL(k_cut, dim3(fp_blocks(s->C)), dim3(BLOCK), 0, s->stream, s->C, s->pos, s->cons_ab, s->cent,
                     s->ealive, o[0], o[1], o[2], d1[0], d1[1], d1[2], d2[0], d2[1], d2[2], nx, ny, nz,
                     s->dflags + DF_CUT);
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
