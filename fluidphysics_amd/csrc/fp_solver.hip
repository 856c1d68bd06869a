// fp_solver.hip -- MI355X (gfx950) kernels of the unified-particle solver and
// the C ABI declared in include/fluidphysics.h.
//
// One substep (the reference's step body, src/solver.js:192-237) is this
// launch sequence on the solver's stream:
//
//   integrate            v.y += g h; prev = p; p += v h            (:194-202)
//   grab                 p += (target - p) k, per unique-index round (:205-210)
//   hash_count           cell (FP64) -> bucket, atomic count        (:177-178)
//   scan_reduce/apply    exclusive scan of the T counts, wave64 DPP
//                        scan + block carry; zeroes the counts       (:179-181)
//   scatter + rank       stable scatter: descending particle index
//                        inside each bucket                          (:182-183)
//   contacts             Jacobi gather of the per-pair contact
//                        correction (:246-279), pos -> pos2
//   density              DDR density gather + neighbour list with the
//                        reference's bucket-duplicate multiplicity   (:284-327)
//   displace             DDR displacement gather, clamps, pos2 -> pos (:329-353)
//   constraints x iters  Jacobi XPBD over the static CSR adjacency with
//                        compliance, tearing and plasticity           (:356-381)
//   bounds_velocity      floor/walls/cull, then v = (p - prev)/h,
//                        damping, speed clamp                         (:389-427, :219-234)
//   xsph                 XSPH over the frozen neighbour list          (:431-454)
//
// Layout in HBM (structure of arrays, 16-byte vectors):
//   pos, pos2: float4[N] (xyz, w = inverse mass)    prev, vel, vtmp: float4[N]
//   active: u8[N]; phash: u32[N] grid bucket of the last build; pflags: u8[N]
//   count: i32[T] (zero between builds); cell_start: i32[T+1];
//   entries_tmp, cell_entries: i32[N]
//   press: float2[N] (P, Pnear); rho: float2[N] (rho, rhoNear)
//   nbr: u32[K * Nf] neighbour-major (entry n of fluid f at n*Nf + f), each
//        entry = particle index | multiplicity << 27; nbr_count: i32[Nf]
//   cons_ab: int2[C]; rest[2]: f32[C]; alive[2]: u8[C] (ping-pong per
//        iteration); adj_off: i32[Ns+1]; adj: int2[2C] (other endpoint, id)
//   spos[2]: float4[Ns] solid positions between constraint iterations
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fluidphysics.h"
#include "fp_kernels.h"

#define BLOCK 256
#define SCAN_ITEMS 16
#define SCAN_TILE (BLOCK * SCAN_ITEMS)
#define NBR_W_SHIFT 27
#define NBR_IDX_MASK ((1u << NBR_W_SHIFT) - 1u)
#define PF_MOVED 1
#define PF_DUP 2

// ============================================================================
// kernels
// ============================================================================

__global__ void __launch_bounds__(BLOCK) k_integrate(int n, float4* __restrict__ pos, float4* __restrict__ prev,
                                                     float4* __restrict__ vel,
                                                     const uint8_t* __restrict__ active, float gh, float h) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n || !active[i]) return;
  float4 v = vel[i];
  float4 p = pos[i];
  v.y += gh;
  prev[i] = make_float4(p.x, p.y, p.z, 0.f);
  p.x = fmaf(v.x, h, p.x);
  p.y = fmaf(v.y, h, p.y);
  p.z = fmaf(v.z, h, p.z);
  pos[i] = p;
  vel[i] = v;
}

// One round of grab pulls; indices inside a round are distinct.
__global__ void __launch_bounds__(BLOCK) k_grab(int n, const int* __restrict__ idx, const float4* __restrict__ tgt,
                                                float4* __restrict__ pos, float k) {
  const int g = blockIdx.x * BLOCK + threadIdx.x;
  if (g >= n) return;
  const int i = idx[g];
  float4 p = pos[i];
  const float4 t = tgt[g];
  p.x = fmaf(t.x - p.x, k, p.x);
  p.y = fmaf(t.y - p.y, k, p.y);
  p.z = fmaf(t.z - p.z, k, p.z);
  pos[i] = p;
}

__global__ void __launch_bounds__(BLOCK) k_hash_count(int n, const float4* __restrict__ pos,
                                                      const uint8_t* __restrict__ active, double inv,
                                                      uint32_t mask, int* __restrict__ count,
                                                      uint32_t* __restrict__ phash, uint32_t* __restrict__ pslot) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  if (!active[i]) {
    phash[i] = 0xffffffffu;
    return;
  }
  const float4 p = pos[i];
  const uint32_t h = cell_hash(cell_of(p.x, p.y, p.z, inv), mask);
  phash[i] = h;
  pslot[i] = (uint32_t)atomicAdd(&count[h], 1);
}

// Block b sums counts[b*SCAN_TILE, (b+1)*SCAN_TILE).
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

// Exclusive scan of one tile with the carry of all earlier tiles; writes
// cell_start[k] and zeroes count[k]; the last tile also writes
// cell_start[t] = total.
__global__ void __launch_bounds__(BLOCK) k_scan_apply(int t, int* __restrict__ count,
                                                      const int* __restrict__ block_sums,
                                                      int* __restrict__ cell_start) {
  __shared__ int wsum[BLOCK / 64];
  __shared__ int carry_s;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // carry = sum of block_sums[0 .. blockIdx.x)
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
  int run = carry + wprefix + incl - local;  // exclusive prefix of this thread's first item
  if (full) {
    int4* s4 = reinterpret_cast<int4*>(cell_start + base);
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
      if (base + k < t) cell_start[base + k] = run;
      run += v[k];
    }
  }
  // the thread holding the table's last entry writes the total
  if (base <= t - 1 && t - 1 < base + SCAN_ITEMS) cell_start[t] = run;
}

__global__ void __launch_bounds__(BLOCK) k_scatter(int n, const uint32_t* __restrict__ phash,
                                                   const uint32_t* __restrict__ pslot,
                                                   const int* __restrict__ cell_start, int* __restrict__ tmp) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t h = phash[i];
  if (h == 0xffffffffu) return;
  tmp[cell_start[h] + (int)pslot[i]] = i;
}

// Stable order: inside bucket h the entries run from the largest particle
// index down (solver.js:182-183 fills each bucket from its end, ascending i).
__global__ void __launch_bounds__(BLOCK) k_rank(int n, const uint32_t* __restrict__ phash,
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

struct SolveParams {
  double inv_cell;
  uint32_t mask;
  int ns, n, nf, K;
  float rs, rf;
  float H, inv_H, H2, max_pair, max_total;
  float stiff, near_stiff, rest_density;
  float contact_scale;
};

// Is the solid pair (x, y) joined by an alive constraint?  Scans x's CSR row.
__device__ __forceinline__ bool is_bonded(int x, int y, const int* __restrict__ adj_off, const int2* __restrict__ adj,
                                          const uint8_t* __restrict__ alive) {
  const int e = adj_off[x + 1];
  for (int k = adj_off[x]; k < e; k++) {
    const int2 a = adj[k];
    if (a.x == y && alive[a.y]) return true;
  }
  return false;
}

// Jacobi contacts (solver.js:246-279 per-pair formula): every active particle
// gathers from the snapshot `pos`: solids from non-bonded solids and fluids,
// fluids from solids; each distinct bucket of the 27 probes once.  Writes
// pos2 for every particle, and pflags for the density pass (PF_MOVED: the
// particle changed cell; PF_DUP: its new cell's probes repeat a bucket).
__global__ void __launch_bounds__(BLOCK) k_contacts(SolveParams sp, const float4* __restrict__ pos,
                                                    float4* __restrict__ pos2, const uint8_t* __restrict__ active,
                                                    const int* __restrict__ cell_start,
                                                    const int* __restrict__ entries, const int* __restrict__ adj_off,
                                                    const int2* __restrict__ adj, const uint8_t* __restrict__ alive,
                                                    uint8_t* __restrict__ pflags) {
  const int x = blockIdx.x * BLOCK + threadIdx.x;
  if (x >= sp.n) return;
  const float4 px = pos[x];
  if (!active[x]) {
    pos2[x] = px;
    pflags[x] = 0;
    return;
  }
  const bool solid_x = x < sp.ns;
  const Cell c0 = cell_of(px.x, px.y, px.z, sp.inv_cell);
  uint32_t hs[27];
  probe_hashes(c0, sp.mask, hs);
  const uint32_t dup = probe_dup_mask(hs);
  const float rx = solid_x ? sp.rs : sp.rf;
  const float wx = px.w;
  float ax = 0.f, ay = 0.f, az = 0.f;
  for (int k = 0; k < 27; k++) {
    if (dup & (1u << k)) continue;
    const uint32_t b = hs[k];
    const int e1 = cell_start[b + 1];
    for (int e = cell_start[b]; e < e1; e++) {
      const int y = entries[e];
      const bool solid_y = y < sp.ns;
      if (y == x || (!solid_x && !solid_y)) continue;
      const float4 py = pos[y];
      const float dx = px.x - py.x, dy = px.y - py.y, dz = px.z - py.z;
      const float d2 = dx * dx + dy * dy + dz * dz;
      const float rsum = rx + (solid_y ? sp.rs : sp.rf);
      if (d2 >= rsum * rsum || d2 < 1e-12f) continue;
      if (solid_x && solid_y && is_bonded(x, y, adj_off, adj, alive)) continue;
      const float d = sqrtf(d2);
      const float f = (rsum - d) / (d * (wx + py.w)) * wx * sp.contact_scale;
      ax += dx * f;
      ay += dy * f;
      az += dz * f;
    }
  }
  const float4 out = make_float4(px.x + ax, px.y + ay, px.z + az, px.w);
  pos2[x] = out;
  uint8_t fl = 0;
  const Cell c1 = cell_of(out.x, out.y, out.z, sp.inv_cell);
  if (c1.x != c0.x || c1.y != c0.y || c1.z != c0.z) {
    fl |= PF_MOVED;
    uint32_t h1[27];
    probe_hashes(c1, sp.mask, h1);
    if (probe_dup_mask(h1)) fl |= PF_DUP;
  } else if (dup) {
    fl |= PF_DUP;
  }
  pflags[x] = fl;
}

// DDR density (solver.js:284-327) as a per-particle gather, plus the
// neighbour list reused by displacement and XSPH.  Pair (i, j), i < j,
// appears in the reference's pair list once per probe of i's cell that hashes
// to j's bucket; a particle x therefore counts
//   y > x: once per visit over x's own 27 probes;
//   y < x: the multiplicity of x's bucket among y's probes, on the first visit.
__global__ void __launch_bounds__(BLOCK) k_density(SolveParams sp, const float4* __restrict__ pos2,
                                                   const uint8_t* __restrict__ active,
                                                   const int* __restrict__ cell_start,
                                                   const int* __restrict__ entries, const uint32_t* __restrict__ phash,
                                                   const uint8_t* __restrict__ pflags, uint32_t* __restrict__ nbr,
                                                   int* __restrict__ nbr_count, float2* __restrict__ press,
                                                   float2* __restrict__ rho_out, unsigned long long* __restrict__ overflow) {
  const int f = blockIdx.x * BLOCK + threadIdx.x;
  if (f >= sp.nf) return;
  const int x = sp.ns + f;
  if (!active[x]) {
    nbr_count[f] = 0;
    rho_out[x] = make_float2(0.f, 0.f);
    press[x] = make_float2(sp.stiff * (0.f - sp.rest_density), 0.f);
    return;
  }
  const float4 px = pos2[x];
  const Cell c0 = cell_of(px.x, px.y, px.z, sp.inv_cell);
  uint32_t hs[27];
  probe_hashes(c0, sp.mask, hs);
  const uint32_t dup = probe_dup_mask(hs);
  const bool moved = pflags[x] & PF_MOVED;
  const uint32_t gh = phash[x];
  float rho = 0.f, rhon = 0.f;
  int cnt = 0;
  bool over = false;
  for (int k = 0; k < 27; k++) {
    const uint32_t b = hs[k];
    const bool dupk = dup & (1u << k);
    const int e1 = cell_start[b + 1];
    for (int e = cell_start[b]; e < e1; e++) {
      const int y = entries[e];
      if (y < sp.ns || y == x) continue;  // fluid neighbours only
      if (y < x && dupk) continue;
      const float4 py = pos2[y];
      const float dx = py.x - px.x, dy = py.y - px.y, dz = py.z - px.z;
      const float d2 = dx * dx + dy * dy + dz * dz;
      if (d2 >= sp.H2) continue;
      int w = 1;
      if (y < x && (moved || (pflags[y] & PF_DUP)))
        w = probe_multiplicity(cell_of(py.x, py.y, py.z, sp.inv_cell), sp.mask, gh);
      if (w == 0) continue;
      const float q = 1.f - sqrtf(d2) * sp.inv_H;
      const float q2 = q * q;
      rho += (float)w * q2;
      rhon += (float)w * q2 * q;
      if (cnt < sp.K)
        nbr[(size_t)cnt * sp.nf + f] = (uint32_t)y | ((uint32_t)w << NBR_W_SHIFT);
      else
        over = true;
      cnt++;
    }
  }
  if (over) atomicAdd(overflow, 1ull);
  nbr_count[f] = cnt < sp.K ? cnt : sp.K;
  rho_out[x] = make_float2(rho, rhon);
  press[x] = make_float2(sp.stiff * (rho - sp.rest_density), sp.near_stiff * rhon);
}

// DDR displacement (solver.js:329-353): pos2 -> pos for every particle.
__global__ void __launch_bounds__(BLOCK) k_displace(SolveParams sp, const float4* __restrict__ pos2,
                                                    float4* __restrict__ pos, const uint8_t* __restrict__ active,
                                                    const uint32_t* __restrict__ nbr,
                                                    const int* __restrict__ nbr_count,
                                                    const float2* __restrict__ press) {
  const int x = blockIdx.x * BLOCK + threadIdx.x;
  if (x >= sp.n) return;
  const float4 px = pos2[x];
  if (x < sp.ns || !active[x]) {
    pos[x] = px;
    return;
  }
  const int f = x - sp.ns;
  const float2 Px = press[x];
  const int cnt = nbr_count[f];
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (int n = 0; n < cnt; n++) {
    const uint32_t e = nbr[(size_t)n * sp.nf + f];
    const int y = (int)(e & NBR_IDX_MASK);
    const float w = (float)(e >> NBR_W_SHIFT);
    const float4 py = pos2[y];
    const float dx = px.x - py.x, dy = px.y - py.y, dz = px.z - py.z;
    const float d = sqrtf(dx * dx + dy * dy + dz * dz);
    if (d < 1e-9f) continue;
    const float q = 1.f - d * sp.inv_H;
    const float2 Py = press[y];
    float D = 0.25f * ((Px.x + Py.x) * q + (Px.y + Py.y) * q * q);
    D = fminf(fmaxf(D, -sp.max_pair), sp.max_pair);
    const float s = w * D / d;
    sx += dx * s;
    sy += dy * s;
    sz += dz * s;
  }
  const float s2 = sx * sx + sy * sy + sz * sz;
  if (s2 > sp.max_total * sp.max_total) {
    const float g = sp.max_total / sqrtf(s2);
    sx *= g;
    sy *= g;
    sz *= g;
  }
  pos[x] = make_float4(px.x + sx, px.y + sy, px.z + sz, px.w);
}

// One Jacobi XPBD iteration (solver.js:356-381 per-constraint formula) over
// the CSR adjacency.  Each solid gathers its constraints' corrections from the
// snapshot (sin, rest_in, alive_in), scaled by omega / max(deg_a, deg_b) so
// both endpoints apply the same factor; the constraint's first endpoint
// (consA) writes the next rest / alive (tear: d > rest * breakStrain;
// plasticity: |C/rest| > thr -> rest += C * rate).
__global__ void __launch_bounds__(BLOCK) k_constraints(int ns, const float4* __restrict__ sin,
                                                       float4* __restrict__ sout, const int* __restrict__ adj_off,
                                                       const int2* __restrict__ adj,
                                                       const float* __restrict__ rest_in,
                                                       const uint8_t* __restrict__ alive_in,
                                                       float* __restrict__ rest_out, uint8_t* __restrict__ alive_out,
                                                       float alpha, float omega, float break_strain,
                                                       float plastic_thr, float plastic_rate,
                                                       int* __restrict__ tear_flag) {
  const int a = blockIdx.x * BLOCK + threadIdx.x;
  if (a >= ns) return;
  const float4 pa = sin[a];
  const int k0 = adj_off[a], k1 = adj_off[a + 1];
  const int deg_a = k1 - k0;
  float cx = 0.f, cy = 0.f, cz = 0.f;
  bool tore = false;
  for (int k = k0; k < k1; k++) {
    const int2 e = adj[k];
    const int b = e.x, c = e.y;
    const bool owner = a < b;  // consA < consB for every lattice constraint
    const float rest = rest_in[c];
    const uint8_t al = alive_in[c];
    if (!al) {
      if (owner) {
        rest_out[c] = rest;
        alive_out[c] = 0;
      }
      continue;
    }
    const float4 pb = sin[b];
    const float w = pa.w + pb.w;
    const float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
    const float d = sqrtf(dx * dx + dy * dy + dz * dz);
    float nrest = rest;
    uint8_t nal = 1;
    if (w != 0.f && d >= 1e-9f) {
      if (d > rest * break_strain) {
        nal = 0;
        tore = true;
      } else {
        const float C = d - rest;
        const float dl = -C / (w + alpha);
        const int deg_b = adj_off[b + 1] - adj_off[b];
        const float s = dl / d * pa.w * (omega / (float)max(deg_a, deg_b));
        cx += dx * s;
        cy += dy * s;
        cz += dz * s;
        if (fabsf(C / rest) > plastic_thr) nrest = rest + C * plastic_rate;
      }
    }
    if (owner) {
      rest_out[c] = nrest;
      alive_out[c] = nal;
    }
  }
  sout[a] = make_float4(pa.x + cx, pa.y + cy, pa.z + cz, pa.w);
  if (tore) *tear_flag = 1;
}

__device__ __forceinline__ float js_sign_or_one(float q) { return q < 0.f ? -1.f : 1.f; }

struct BoundsParams {
  int ns, n;
  float rs, rf, mu_s, mu_f;
  float wall_x, wall_z, wall_h, wall_t;
  float cull_r2;
  float inv_h, damp_s, damp_f, vmax;
};

// Bounds (solver.js:389-427) then velocity + damping + clamp (:219-234).
// Solids read their positions from `solid_src` (the last constraint
// iteration).  Fluid velocities go to vtmp (XSPH reads them), solids' to vel.
__global__ void __launch_bounds__(BLOCK) k_bounds_velocity(BoundsParams bp, float4* __restrict__ pos,
                                                           const float4* __restrict__ solid_src,
                                                           const float4* __restrict__ prev, float4* __restrict__ vel,
                                                           float4* __restrict__ vtmp, uint8_t* __restrict__ active) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= bp.n) return;
  const bool solid = i < bp.ns;
  if (!active[i]) {
    if (solid)
      pos[i] = solid_src[i];  // the reference's constraint sweep moves inactive solids too
    else
      vtmp[i] = vel[i];
    return;
  }
  float4 p = solid ? solid_src[i] : pos[i];
  const float4 pv = prev[i];
  const float r = solid ? bp.rs : bp.rf;
  if (p.y < r) {
    p.y = r;
    const float mu = solid ? bp.mu_s : bp.mu_f;
    p.x -= (p.x - pv.x) * mu;
    p.z -= (p.z - pv.z) * mu;
  }
  if (p.y < bp.wall_h + r) {
    const float lim = r + bp.wall_t;
    if (fabsf(fabsf(p.x) - bp.wall_x) < lim) {
      const float side = js_sign_or_one(p.x);
      p.x = fabsf(p.x) < bp.wall_x ? side * (bp.wall_x - lim) : side * (bp.wall_x + lim);
    }
    if (fabsf(fabsf(p.z) - bp.wall_z) < lim) {
      const float side = js_sign_or_one(p.z);
      p.z = fabsf(p.z) < bp.wall_z ? side * (bp.wall_z - lim) : side * (bp.wall_z + lim);
    }
  }
  pos[i] = p;
  if (!solid && (p.y < -3.f || p.x * p.x + p.z * p.z > bp.cull_r2)) {
    active[i] = 0;
    vtmp[i] = vel[i];
    return;
  }
  float vx = (p.x - pv.x) * bp.inv_h;
  float vy = (p.y - pv.y) * bp.inv_h;
  float vz = (p.z - pv.z) * bp.inv_h;
  const float damp = solid ? bp.damp_s : bp.damp_f;
  vx *= damp;
  vy *= damp;
  vz *= damp;
  const float sp2 = vx * vx + vy * vy + vz * vz;
  if (sp2 > bp.vmax * bp.vmax) {
    const float g = bp.vmax / sqrtf(sp2);
    vx *= g;
    vy *= g;
    vz *= g;
  }
  const float4 v = make_float4(vx, vy, vz, 0.f);
  if (solid)
    vel[i] = v;
  else
    vtmp[i] = v;
}

// XSPH (solver.js:431-454) over the frozen neighbour list: vtmp -> vel.
__global__ void __launch_bounds__(BLOCK) k_xsph(int ns, int nf, float visc, const float4* __restrict__ vtmp,
                                                float4* __restrict__ vel, const uint8_t* __restrict__ active,
                                                const uint32_t* __restrict__ nbr,
                                                const int* __restrict__ nbr_count) {
  const int f = blockIdx.x * BLOCK + threadIdx.x;
  if (f >= nf) return;
  const int x = ns + f;
  const float4 v = vtmp[x];
  const int cnt = nbr_count[f];
  if (!active[x] || cnt == 0) {
    vel[x] = v;
    return;
  }
  float sx = 0.f, sy = 0.f, sz = 0.f, m = 0.f;
  for (int n = 0; n < cnt; n++) {
    const uint32_t e = nbr[(size_t)n * nf + f];
    const float w = (float)(e >> NBR_W_SHIFT);
    const float4 vy = vtmp[e & NBR_IDX_MASK];
    sx += w * vy.x;
    sy += w * vy.y;
    sz += w * vy.z;
    m += w;
  }
  const float inv_m = 1.f / m;
  vel[x] = make_float4(v.x + visc * (sx * inv_m - v.x), v.y + visc * (sy * inv_m - v.y),
                       v.z + visc * (sz * inv_m - v.z), 0.f);
}

// Knife (solver.js:458-486) in IEEE double, the reference's operation order
// (the library is compiled with -ffp-contract=off).
__global__ void __launch_bounds__(BLOCK) k_cut(int nc, const float4* __restrict__ pos, const int2* __restrict__ ab,
                                               uint8_t* __restrict__ alive, double ox, double oy, double oz,
                                               double d1x, double d1y, double d1z, double d2x, double d2y,
                                               double d2z, double nx, double ny, double nz, int* __restrict__ ncut) {
  const int c = blockIdx.x * BLOCK + threadIdx.x;
  if (c >= nc || !alive[c]) return;
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
  alive[c] = 0;
  atomicAdd(ncut, 1);
}

// pickSolid (solver.js:502-515): one workgroup, (t, index) lexicographic min
// over the candidates that pass both tests; double arithmetic.
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

__global__ void __launch_bounds__(BLOCK) k_count_active(int begin, int end, const uint8_t* __restrict__ active,
                                                        int* __restrict__ out) {
  const int i = begin + blockIdx.x * BLOCK + threadIdx.x;
  const int v = (i < end) ? (int)active[i] : 0;
  const int s = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

// emitFluid scatter (solver.js:139-149): slot k of the batch -> particle idx[k].
__global__ void __launch_bounds__(BLOCK) k_emit(int n, const int* __restrict__ idx, const float4* __restrict__ p_in,
                                                const float4* __restrict__ v_in, float4* __restrict__ pos,
                                                float4* __restrict__ prev, float4* __restrict__ vel,
                                                uint8_t* __restrict__ active) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= n) return;
  const int i = idx[k];
  const float4 p = p_in[k];
  pos[i] = p;
  prev[i] = make_float4(p.x, p.y, p.z, 0.f);
  vel[i] = v_in[k];
  active[i] = 1;
}

// ============================================================================
// host side
// ============================================================================

namespace {

thread_local std::string g_last_error = "";

enum PassId {
  P_INTEGRATE,
  P_GRAB,
  P_HASH_COUNT,
  P_SCAN_REDUCE,
  P_SCAN_APPLY,
  P_SCATTER,
  P_RANK,
  P_CONTACTS,
  P_DENSITY,
  P_DISPLACE,
  P_CONSTRAINTS,
  P_BOUNDS_VELOCITY,
  P_XSPH,
  P_CUT,
  P_PICK,
  P_COUNT,
  P_EMIT,
  P_NUM
};
const char* kPassNames[P_NUM] = {"integrate", "grab",        "hash_count",      "scan_reduce", "scan_apply",
                                 "scatter",   "rank",        "contacts",        "density",     "displace",
                                 "constraints", "bounds_velocity", "xsph",        "cut",         "pick",
                                 "count_active", "emit"};

}  // namespace

struct fp_solver {
  fp_params P;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int max_fluid = 0, ns = 0, n = 0, nf = 0, T = 0, C = 0, K = 96;
  int fluid_next = 0;
  bool has_cube = false;
  int64_t substeps_done = 0;
  int64_t device_bytes = 0;
  // device arrays
  float4 *pos = nullptr, *pos2 = nullptr, *prev = nullptr, *vel = nullptr, *vtmp = nullptr;
  uint8_t *active = nullptr, *pflags = nullptr;
  uint32_t *phash = nullptr, *pslot = nullptr;
  int *count = nullptr, *cell_start = nullptr, *entries_tmp = nullptr, *cell_entries = nullptr;
  int* block_sums = nullptr;
  int scan_blocks = 0;
  float2 *press = nullptr, *rho = nullptr;
  uint32_t* nbr = nullptr;
  int* nbr_count = nullptr;
  int2* cons_ab = nullptr;
  float* rest[2] = {nullptr, nullptr};
  uint8_t* alive[2] = {nullptr, nullptr};
  int cur = 0;  // which rest/alive buffer is current
  int* adj_off = nullptr;
  int2* adj = nullptr;
  float4* spos[2] = {nullptr, nullptr};
  int* dflags = nullptr;  // [0] tear flag, [1] cut count, [2] scratch count; [4..5] overflow (u64)
  double* dpick_t = nullptr;
  int* dpick_i = nullptr;
  // grabs: rounds of distinct indices
  int* grab_idx = nullptr;
  float4* grab_tgt = nullptr;
  int grab_cap = 0;
  std::vector<int> grab_round_len;
  // emit staging
  int* emit_idx = nullptr;
  float4 *emit_p = nullptr, *emit_v = nullptr;
  int emit_cap = 0;
  // host copies
  std::vector<int> h_cons_a, h_cons_b;
  std::vector<float> h_snap_pos;  // 3*ns lattice positions (resetSolid)
  std::vector<float> h_snap_rest;
  std::vector<int32_t> h_group;
  bool groups_dirty = false;  // recompute at the next sync point
  bool steps_pending = false;
  bool cut_dirty = false;          // a cut killed constraints since the last step
  bool cut_dirty_at_step = false;  // ... and a step has run since: relabel at its end
  // profiling
  bool profile = false;
  struct Ev {
    int pass;
    hipEvent_t a, b;
  };
  std::vector<Ev> ev_live;
  std::vector<hipEvent_t> ev_pool;
  double prof_ms[P_NUM] = {0};
  int64_t prof_n[P_NUM] = {0};
};

#define FP_TRY(s, expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) {                                                              \
      return fail((s), FP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    }                                                                                    \
  } while (0)

static int fail(fp_solver* s, int code, const std::string& msg) {
  if (s)
    s->err = msg;
  else
    g_last_error = msg;
  return code;
}

static inline int nblocks(int n) { return (n + BLOCK - 1) / BLOCK; }

// Host <-> device copies ordered on the solver's (non-blocking) stream.
static hipError_t d2h(fp_solver* s, void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream);
  return e == hipSuccess ? hipStreamSynchronize(s->stream) : e;
}
static hipError_t h2d(fp_solver* s, void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream);
  return e == hipSuccess ? hipStreamSynchronize(s->stream) : e;
}

static hipEvent_t ev_get(fp_solver* s) {
  if (!s->ev_pool.empty()) {
    hipEvent_t e = s->ev_pool.back();
    s->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct PassScope {
  fp_solver* s;
  int pass;
  hipEvent_t a = nullptr, b = nullptr;
  PassScope(fp_solver* s_, int p) : s(s_), pass(p) {
    if (s->profile) {
      a = ev_get(s);
      b = ev_get(s);
      if (a) hipEventRecord(a, s->stream);
    }
  }
  ~PassScope() {
    if (s->profile && a && b) {
      hipEventRecord(b, s->stream);
      s->ev_live.push_back({pass, a, b});
    }
  }
};

static void profile_collect(fp_solver* s) {
  for (auto& e : s->ev_live) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
      s->prof_ms[e.pass] += ms;
      s->prof_n[e.pass] += 1;
    }
    s->ev_pool.push_back(e.a);
    s->ev_pool.push_back(e.b);
  }
  s->ev_live.clear();
}

template <typename T>
static int dalloc(fp_solver* s, T** p, size_t count) {
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) return fail(s, FP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  s->device_bytes += (int64_t)bytes;
  hipMemsetAsync(*p, 0, bytes, s->stream);
  return FP_OK;
}

static void free_all(fp_solver* s) {
  void* ptrs[] = {s->pos,        s->pos2,      s->prev,      s->vel,       s->vtmp,       s->active,
                  s->pflags,     s->phash,     s->pslot,     s->count,     s->cell_start, s->entries_tmp,
                  s->cell_entries, s->block_sums, s->press,  s->rho,       s->nbr,        s->nbr_count,
                  s->cons_ab,    s->rest[0],   s->rest[1],   s->alive[0],  s->alive[1],   s->adj_off,
                  s->adj,        s->spos[0],   s->spos[1],   s->dflags,    s->dpick_t,    s->dpick_i,
                  s->grab_idx,   s->grab_tgt,  s->emit_idx,  s->emit_p,    s->emit_v};
  for (void* p : ptrs)
    if (p) hipFree(p);
}

static int set_device(fp_solver* s) { FP_TRY(s, hipSetDevice(s->device)); return FP_OK; }

// union-find with path halving; parent[ra] = rb (solver.js:488-499), so the
// labels are the reference's roots.
static void recompute_groups_host(fp_solver* s, const std::vector<uint8_t>& alive) {
  const int n = s->ns;
  std::vector<int32_t> parent(n);
  for (int i = 0; i < n; i++) parent[i] = i;
  auto find = [&](int x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (int c = 0; c < s->C; c++) {
    if (!alive[c]) continue;
    const int ra = find(s->h_cons_a[c]), rb = find(s->h_cons_b[c]);
    if (ra != rb) parent[ra] = rb;
  }
  s->h_group.resize(n);
  for (int i = 0; i < n; i++) s->h_group[i] = find(i);
}

// Sync point: wait for the stream, pick up the tear flag of finished steps,
// refresh the component labels when a tear or a cut happened.
static int settle(fp_solver* s) {
  FP_TRY(s, hipStreamSynchronize(s->stream));
  if (s->profile) profile_collect(s);
  if (s->steps_pending && s->C > 0) {
    int tear = 0;
    FP_TRY(s, d2h(s, &tear, s->dflags, sizeof(int)));
    if (tear) {
      s->groups_dirty = true;
      FP_TRY(s, hipMemsetAsync(s->dflags, 0, sizeof(int), s->stream));
    }
    if (s->cut_dirty_at_step) s->groups_dirty = true;
  }
  s->cut_dirty_at_step = false;
  s->steps_pending = false;
  if (s->groups_dirty) {
    std::vector<uint8_t> al(s->C);
    if (s->C) FP_TRY(s, d2h(s, al.data(), s->alive[s->cur], s->C));
    recompute_groups_host(s, al);
    s->groups_dirty = false;
  }
  return FP_OK;
}

static SolveParams solve_params(const fp_solver* s) {
  SolveParams sp;
  const fp_params& P = s->P;
  sp.inv_cell = 1.0 / P.cell_size;
  sp.mask = (uint32_t)(s->T - 1);
  sp.ns = s->ns;
  sp.n = s->n;
  sp.nf = s->nf;
  sp.K = s->K;
  sp.rs = (float)P.solid_radius;
  sp.rf = (float)P.fluid_radius;
  sp.H = (float)P.fluid_h;
  sp.inv_H = (float)(1.0 / P.fluid_h);
  sp.H2 = (float)(P.fluid_h * P.fluid_h);
  sp.max_pair = (float)(0.05 * P.fluid_h);
  sp.max_total = (float)(0.35 * P.fluid_h);
  sp.stiff = (float)P.stiffness;
  sp.near_stiff = (float)P.near_stiffness;
  sp.rest_density = (float)P.rest_density;
  sp.contact_scale = (float)P.contact_scale;
  return sp;
}

static int launch_grid_build(fp_solver* s) {
  const int n = s->n;
  {
    PassScope ps(s, P_HASH_COUNT);
    hipLaunchKernelGGL(k_hash_count, dim3(nblocks(n)), dim3(BLOCK), 0, s->stream, n, s->pos, s->active,
                       1.0 / s->P.cell_size, (uint32_t)(s->T - 1), s->count, s->phash, s->pslot);
  }
  {
    PassScope ps(s, P_SCAN_REDUCE);
    hipLaunchKernelGGL(k_scan_reduce, dim3(s->scan_blocks), dim3(BLOCK), 0, s->stream, s->T, s->count,
                       s->block_sums);
  }
  {
    PassScope ps(s, P_SCAN_APPLY);
    hipLaunchKernelGGL(k_scan_apply, dim3(s->scan_blocks), dim3(BLOCK), 0, s->stream, s->T, s->count,
                       s->block_sums, s->cell_start);
  }
  {
    PassScope ps(s, P_SCATTER);
    hipLaunchKernelGGL(k_scatter, dim3(nblocks(n)), dim3(BLOCK), 0, s->stream, n, s->phash, s->pslot,
                       s->cell_start, s->entries_tmp);
  }
  {
    PassScope ps(s, P_RANK);
    hipLaunchKernelGGL(k_rank, dim3(nblocks(n)), dim3(BLOCK), 0, s->stream, n, s->phash, s->cell_start,
                       s->entries_tmp, s->cell_entries);
  }
  FP_TRY(s, hipGetLastError());
  return FP_OK;
}

static int launch_substep(fp_solver* s, double h) {
  const fp_params& P = s->P;
  const int n = s->n, ns = s->ns, nf = s->nf;
  {
    PassScope ps(s, P_INTEGRATE);
    hipLaunchKernelGGL(k_integrate, dim3(nblocks(n)), dim3(BLOCK), 0, s->stream, n, s->pos, s->prev, s->vel,
                       s->active, (float)(P.gravity * h), (float)h);
  }
  if (!s->grab_round_len.empty()) {
    PassScope ps(s, P_GRAB);
    int off = 0;
    for (int len : s->grab_round_len) {
      hipLaunchKernelGGL(k_grab, dim3(nblocks(len)), dim3(BLOCK), 0, s->stream, len, s->grab_idx + off,
                         s->grab_tgt + off, s->pos, (float)P.grab_stiffness);
      off += len;
    }
  }
  int rc = launch_grid_build(s);
  if (rc) return rc;
  const SolveParams sp = solve_params(s);
  {
    PassScope ps(s, P_CONTACTS);
    hipLaunchKernelGGL(k_contacts, dim3(nblocks(n)), dim3(BLOCK), 0, s->stream, sp, s->pos, s->pos2, s->active,
                       s->cell_start, s->cell_entries, s->adj_off, s->adj, s->alive[s->cur], s->pflags);
  }
  if (nf > 0) {
    PassScope ps(s, P_DENSITY);
    hipLaunchKernelGGL(k_density, dim3(nblocks(nf)), dim3(BLOCK), 0, s->stream, sp, s->pos2, s->active,
                       s->cell_start, s->cell_entries, s->phash, s->pflags, s->nbr, s->nbr_count, s->press, s->rho,
                       (unsigned long long*)(s->dflags + 4));
  }
  {
    PassScope ps(s, P_DISPLACE);
    hipLaunchKernelGGL(k_displace, dim3(nblocks(n)), dim3(BLOCK), 0, s->stream, sp, s->pos2, s->pos, s->active,
                       s->nbr, s->nbr_count, s->press);
  }
  const float4* solid_src = s->pos;
  if (ns > 0 && s->C > 0) {
    const int iters = P.constraint_iters > 0 ? P.constraint_iters : 1;
    const float alpha = (float)(P.compliance / (h * h));
    PassScope ps(s, P_CONSTRAINTS);
    for (int it = 0; it < iters; it++) {
      const float4* in = it == 0 ? s->pos : s->spos[(it - 1) & 1];
      float4* out = s->spos[it & 1];
      hipLaunchKernelGGL(k_constraints, dim3(nblocks(ns)), dim3(BLOCK), 0, s->stream, ns, in, out, s->adj_off,
                         s->adj, s->rest[s->cur], s->alive[s->cur], s->rest[s->cur ^ 1], s->alive[s->cur ^ 1],
                         alpha, (float)P.jacobi_omega, (float)P.break_strain, (float)P.plastic_threshold,
                         (float)P.plastic_rate, s->dflags);
      s->cur ^= 1;
      solid_src = out;
    }
  }
  {
    BoundsParams bp;
    bp.ns = ns;
    bp.n = n;
    bp.rs = (float)P.solid_radius;
    bp.rf = (float)P.fluid_radius;
    bp.mu_s = (float)P.floor_friction_solid;
    bp.mu_f = (float)P.floor_friction_fluid;
    bp.wall_x = (float)P.wall_x;
    bp.wall_z = (float)P.wall_z;
    bp.wall_h = (float)P.wall_height;
    bp.wall_t = (float)P.wall_half_thick;
    bp.cull_r2 = (float)(P.cull_radius * P.cull_radius);
    bp.inv_h = (float)(1.0 / h);
    bp.damp_s = (float)std::max(0.0, 1.0 - P.solid_damping * h);
    bp.damp_f = (float)std::max(0.0, 1.0 - P.fluid_damping * h);
    bp.vmax = (float)P.max_speed;
    PassScope ps(s, P_BOUNDS_VELOCITY);
    hipLaunchKernelGGL(k_bounds_velocity, dim3(nblocks(n)), dim3(BLOCK), 0, s->stream, bp, s->pos, solid_src,
                       s->prev, s->vel, s->vtmp, s->active);
  }
  if (nf > 0) {
    PassScope ps(s, P_XSPH);
    hipLaunchKernelGGL(k_xsph, dim3(nblocks(nf)), dim3(BLOCK), 0, s->stream, ns, nf, (float)P.viscosity, s->vtmp,
                       s->vel, s->active, s->nbr, s->nbr_count);
  }
  FP_TRY(s, hipGetLastError());
  s->substeps_done++;
  return FP_OK;
}

// ============================================================================
// C ABI
// ============================================================================

extern "C" {

void fp_params_default(fp_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->gravity = -10.0;
  p->solid_radius = 0.09;
  p->fluid_radius = 0.075;
  p->compliance = 2.5e-4;
  p->plastic_threshold = 0.32;
  p->plastic_rate = 0.04;
  p->break_strain = 2.6;
  p->fluid_h = 0.19;
  p->rest_density = 6.0;
  p->stiffness = 0.002;
  p->near_stiffness = 0.008;
  p->viscosity = 0.3;
  p->solid_damping = 1.2;
  p->fluid_damping = 0.6;
  p->floor_friction_solid = 0.4;
  p->floor_friction_fluid = 0.04;
  p->wall_x = 3.5;
  p->wall_z = 3.5;
  p->wall_height = 0.92;
  p->wall_half_thick = 0.07;
  p->cull_radius = 7.0;
  p->max_speed = 14.0;
  p->grab_stiffness = 0.3;
  p->cell_size = 0.22;
  p->constraint_iters = 3;
  p->table_size = 0;
  p->max_neighbors = 0;
  p->jacobi_omega = 4.0;
  p->contact_scale = 1.0;
}

static int check_params(fp_solver* s, const fp_params& p) {
  const double rmax = std::max(p.solid_radius, p.fluid_radius);
  if (!(p.cell_size > 0) || !(p.fluid_h > 0) || p.cell_size < p.fluid_h || p.cell_size < 2 * rmax)
    return fail(s, FP_EINVAL, "cell_size must be >= fluid_h and >= 2 * max(solid_radius, fluid_radius)");
  if (p.table_size < 0 || (p.table_size & (p.table_size - 1)))
    return fail(s, FP_EINVAL, "table_size must be 0 or a power of two");
  if (p.max_neighbors < 0 || p.max_neighbors > 4096) return fail(s, FP_EINVAL, "max_neighbors out of range");
  if (p.constraint_iters < 0 || p.constraint_iters > 1000) return fail(s, FP_EINVAL, "constraint_iters out of range");
  return FP_OK;
}

int fp_create(const fp_params* params, int32_t max_fluid, int32_t device, fp_solver** out) {
  if (!out) return fail(nullptr, FP_EINVAL, "fp_create: out is NULL");
  *out = nullptr;
  if (max_fluid < 0) return fail(nullptr, FP_EINVAL, "fp_create: max_fluid < 0");
  fp_solver* s = new (std::nothrow) fp_solver();
  if (!s) return fail(nullptr, FP_ENOMEM, "fp_create: out of host memory");
  if (params)
    s->P = *params;
  else
    fp_params_default(&s->P);
  int rc = check_params(s, s->P);
  if (rc) {
    g_last_error = s->err;
    delete s;
    return rc;
  }
  if (device < 0) {
    if (hipGetDevice(&device) != hipSuccess) device = 0;
  }
  s->device = device;
  s->max_fluid = max_fluid;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_last_error = std::string("fp_create: ") + hipGetErrorString(e);
    delete s;
    return FP_EDEVICE;
  }
  *out = s;
  return FP_OK;
}

int fp_destroy(fp_solver* s) {
  if (!s) return FP_OK;
  hipSetDevice(s->device);
  if (s->stream) hipStreamSynchronize(s->stream);
  free_all(s);
  for (auto& e : s->ev_live) {
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  for (auto e : s->ev_pool) hipEventDestroy(e);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
  return FP_OK;
}

const char* fp_last_error(const fp_solver* s) { return s ? s->err.c_str() : g_last_error.c_str(); }

int fp_set_params(fp_solver* s, const fp_params* p) {
  if (!s || !p) return fail(s, FP_EINVAL, "fp_set_params: NULL argument");
  int rc = check_params(s, *p);
  if (rc) return rc;
  if (s->has_cube) {
    if (p->table_size != s->P.table_size || p->max_neighbors != s->P.max_neighbors)
      return fail(s, FP_ESTATE, "table_size / max_neighbors are fixed once the scene is allocated");
  }
  s->P = *p;
  return FP_OK;
}

int fp_get_params(const fp_solver* s, fp_params* out) {
  if (!s || !out) return FP_EINVAL;
  *out = s->P;
  return FP_OK;
}

int fp_get_info(const fp_solver* s, fp_info* out) {
  if (!s || !out) return FP_EINVAL;
  fp_solver* m = const_cast<fp_solver*>(s);
  out->capacity = s->n;
  out->solid_count = s->ns;
  out->max_fluid = s->max_fluid;
  out->table_size = s->T;
  out->num_constraints = s->C;
  out->fluid_next = s->fluid_next;
  out->max_neighbors = s->K;
  out->device = s->device;
  out->substeps_done = s->substeps_done;
  out->neighbor_overflow = 0;
  out->device_bytes = s->device_bytes;
  if (s->dflags) {
    hipSetDevice(s->device);
    unsigned long long ov = 0;
    if (d2h(m, &ov, s->dflags + 4, sizeof(ov)) == hipSuccess)
      out->neighbor_overflow = (int64_t)ov;
  }
  return FP_OK;
}

int fp_add_solid_cube(fp_solver* s, double cx, double cy, double cz, int32_t n, double spacing,
                      double particle_mass) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_add_solid_cube: NULL solver");
  if (s->has_cube) return fail(s, FP_ESTATE, "addSolidCube may be called once");
  if (n < 1 || n > 1024) return fail(s, FP_EINVAL, "cube dimension out of range");
  const int64_t cap64 = (int64_t)n * n * n + s->max_fluid;
  if (cap64 >= (1ll << 27)) return fail(s, FP_EINVAL, "capacity exceeds 2^27 particles");
  int rc = set_device(s);
  if (rc) return rc;
  const int ns = n * n * n;
  s->ns = ns;
  s->n = (int)cap64;
  s->nf = s->max_fluid;
  if (s->P.table_size > 0) {
    s->T = s->P.table_size;
  } else {
    int t = 1;
    while (t < 2 * s->n) t <<= 1;
    s->T = t;
  }
  s->K = s->P.max_neighbors > 0 ? s->P.max_neighbors : 96;
  s->scan_blocks = (s->T + SCAN_TILE - 1) / SCAN_TILE;

  // lattice (solver.js:74-89), double arithmetic, float32 stores
  const double half = (n - 1) * spacing * 0.5;
  std::vector<float4> hpos(s->n, make_float4(0.f, 0.f, 0.f, 0.f));
  std::vector<uint8_t> hact(s->n, 0);
  s->h_snap_pos.resize(3 * (size_t)ns);
  const float inv_m = (float)(1 / particle_mass);
  int id = 0;
  for (int ix = 0; ix < n; ix++)
    for (int iy = 0; iy < n; iy++)
      for (int iz = 0; iz < n; iz++) {
        const float x = (float)(cx - half + ix * spacing), y = (float)(cy - half + iy * spacing),
                    z = (float)(cz - half + iz * spacing);
        hpos[id] = make_float4(x, y, z, inv_m);
        s->h_snap_pos[3 * id] = x;
        s->h_snap_pos[3 * id + 1] = y;
        s->h_snap_pos[3 * id + 2] = z;
        hact[id] = 1;
        id++;
      }
  // constraints in the reference's loop order (solver.js:97-116)
  const double maxD2 = 3 * spacing * spacing + 1e-6;
  std::vector<float> hrest;
  s->h_cons_a.clear();
  s->h_cons_b.clear();
  auto idx = [n](int ax, int ay, int az) { return (ax * n + ay) * n + az; };
  for (int ix = 0; ix < n; ix++)
    for (int iy = 0; iy < n; iy++)
      for (int iz = 0; iz < n; iz++)
        for (int dx = -1; dx <= 1; dx++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
              const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
              if (jx < 0 || jy < 0 || jz < 0 || jx >= n || jy >= n || jz >= n) continue;
              const int a = idx(ix, iy, iz), b = idx(jx, jy, jz);
              if (b <= a) continue;
              const double d2 = (dx * dx + dy * dy + dz * dz) * spacing * spacing;
              if (d2 > maxD2) continue;
              s->h_cons_a.push_back(a);
              s->h_cons_b.push_back(b);
              hrest.push_back((float)std::sqrt(d2));
            }
  const int C = (int)hrest.size();
  s->C = C;
  s->h_snap_rest = hrest;
  // CSR adjacency over all constraints (static), entries (other, id)
  std::vector<int> hoff(ns + 1, 0);
  for (int c = 0; c < C; c++) {
    hoff[s->h_cons_a[c] + 1]++;
    hoff[s->h_cons_b[c] + 1]++;
  }
  for (int i = 0; i < ns; i++) hoff[i + 1] += hoff[i];
  std::vector<int2> hadj(std::max(1, 2 * C));
  std::vector<int> fill(ns, 0);
  std::vector<int2> hab(std::max(1, C));
  for (int c = 0; c < C; c++) {
    const int a = s->h_cons_a[c], b = s->h_cons_b[c];
    hadj[hoff[a] + fill[a]++] = make_int2(b, c);
    hadj[hoff[b] + fill[b]++] = make_int2(a, c);
    hab[c] = make_int2(a, b);
  }
  s->h_group.assign(ns, 0);

  const int N = s->n, nf = s->nf, T = s->T;
  if ((rc = dalloc(s, &s->pos, N)) || (rc = dalloc(s, &s->pos2, N)) || (rc = dalloc(s, &s->prev, N)) ||
      (rc = dalloc(s, &s->vel, N)) || (rc = dalloc(s, &s->vtmp, N)) || (rc = dalloc(s, &s->active, N)) ||
      (rc = dalloc(s, &s->pflags, N)) || (rc = dalloc(s, &s->phash, N)) || (rc = dalloc(s, &s->pslot, N)) ||
      (rc = dalloc(s, &s->count, T)) || (rc = dalloc(s, &s->cell_start, (size_t)T + 1)) ||
      (rc = dalloc(s, &s->entries_tmp, N)) || (rc = dalloc(s, &s->cell_entries, N)) ||
      (rc = dalloc(s, &s->block_sums, s->scan_blocks)) || (rc = dalloc(s, &s->press, N)) ||
      (rc = dalloc(s, &s->rho, N)) || (rc = dalloc(s, &s->nbr, (size_t)s->K * std::max(nf, 1))) ||
      (rc = dalloc(s, &s->nbr_count, std::max(nf, 1))) || (rc = dalloc(s, &s->cons_ab, std::max(C, 1))) ||
      (rc = dalloc(s, &s->rest[0], std::max(C, 1))) || (rc = dalloc(s, &s->rest[1], std::max(C, 1))) ||
      (rc = dalloc(s, &s->alive[0], std::max(C, 1))) || (rc = dalloc(s, &s->alive[1], std::max(C, 1))) ||
      (rc = dalloc(s, &s->adj_off, ns + 1)) || (rc = dalloc(s, &s->adj, std::max(2 * C, 1))) ||
      (rc = dalloc(s, &s->spos[0], ns)) || (rc = dalloc(s, &s->spos[1], ns)) || (rc = dalloc(s, &s->dflags, 8)) ||
      (rc = dalloc(s, &s->dpick_t, 1)) || (rc = dalloc(s, &s->dpick_i, 1))) {
    return rc;
  }
  std::vector<uint8_t> ones(std::max(C, 1), 1);
  FP_TRY(s, hipMemcpyAsync(s->pos, hpos.data(), N * sizeof(float4), hipMemcpyHostToDevice, s->stream));
  // prev of solids is zero in the reference until the first integrate (solver.js:43)
  std::vector<float4> zero4(N, make_float4(0.f, 0.f, 0.f, 0.f));
  FP_TRY(s, hipMemcpyAsync(s->prev, zero4.data(), N * sizeof(float4), hipMemcpyHostToDevice, s->stream));
  FP_TRY(s, hipMemcpyAsync(s->active, hact.data(), N, hipMemcpyHostToDevice, s->stream));
  if (C) {
    FP_TRY(s, hipMemcpyAsync(s->cons_ab, hab.data(), C * sizeof(int2), hipMemcpyHostToDevice, s->stream));
    FP_TRY(s, hipMemcpyAsync(s->rest[0], hrest.data(), C * sizeof(float), hipMemcpyHostToDevice, s->stream));
    FP_TRY(s, hipMemcpyAsync(s->alive[0], ones.data(), C, hipMemcpyHostToDevice, s->stream));
    FP_TRY(s, hipMemcpyAsync(s->adj, hadj.data(), 2 * C * sizeof(int2), hipMemcpyHostToDevice, s->stream));
  }
  FP_TRY(s, hipMemcpyAsync(s->adj_off, hoff.data(), (ns + 1) * sizeof(int), hipMemcpyHostToDevice, s->stream));
  FP_TRY(s, hipStreamSynchronize(s->stream));
  s->cur = 0;
  s->has_cube = true;
  return FP_OK;
}

int fp_emit(fp_solver* s, const double* pos, const double* vel, const double* mass, int32_t n) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_emit: NULL solver");
  if (!s->has_cube) return fail(s, FP_ESTATE, "fp_emit before fp_add_solid_cube");
  if (n < 0 || (n > 0 && (!pos || !vel))) return fail(s, FP_EINVAL, "fp_emit: bad arguments");
  if (n == 0) return FP_OK;
  if (s->max_fluid == 0) return fail(s, FP_EINVAL, "fp_emit: maxFluid is 0");
  int rc = set_device(s);
  if (rc) return rc;
  // only the last max_fluid emits of a batch survive the ring (solver.js:140-141)
  const int first = n > s->max_fluid ? n - s->max_fluid : 0;
  const int m = n - first;
  if (m > s->emit_cap) {
    FP_TRY(s, hipStreamSynchronize(s->stream));
    if (s->emit_idx) hipFree(s->emit_idx);
    if (s->emit_p) hipFree(s->emit_p);
    if (s->emit_v) hipFree(s->emit_v);
    s->emit_idx = nullptr;
    s->emit_p = s->emit_v = nullptr;
    const int cap = std::max(m, 1024);
    if ((rc = dalloc(s, &s->emit_idx, cap)) || (rc = dalloc(s, &s->emit_p, cap)) || (rc = dalloc(s, &s->emit_v, cap)))
      return rc;
    s->emit_cap = cap;
  }
  std::vector<int> idx(m);
  std::vector<float4> hp(m), hv(m);
  int next = (int)(((int64_t)s->fluid_next + first) % s->max_fluid);
  for (int k = 0; k < m; k++) {
    const int j = first + k;
    const double ms = mass ? mass[j] : 0.03;
    idx[k] = s->ns + next;
    hp[k] = make_float4((float)pos[3 * j], (float)pos[3 * j + 1], (float)pos[3 * j + 2], (float)(1 / ms));
    hv[k] = make_float4((float)vel[3 * j], (float)vel[3 * j + 1], (float)vel[3 * j + 2], 0.f);
    next = (next + 1) % s->max_fluid;
  }
  s->fluid_next = (int)(((int64_t)s->fluid_next + n) % s->max_fluid);
  FP_TRY(s, hipMemcpyAsync(s->emit_idx, idx.data(), m * sizeof(int), hipMemcpyHostToDevice, s->stream));
  FP_TRY(s, hipMemcpyAsync(s->emit_p, hp.data(), m * sizeof(float4), hipMemcpyHostToDevice, s->stream));
  FP_TRY(s, hipMemcpyAsync(s->emit_v, hv.data(), m * sizeof(float4), hipMemcpyHostToDevice, s->stream));
  {
    PassScope ps(s, P_EMIT);
    hipLaunchKernelGGL(k_emit, dim3(nblocks(m)), dim3(BLOCK), 0, s->stream, m, s->emit_idx, s->emit_p, s->emit_v,
                       s->pos, s->prev, s->vel, s->active);
  }
  FP_TRY(s, hipGetLastError());
  // the staging buffers are reused by the next call: wait for the copies
  FP_TRY(s, hipStreamSynchronize(s->stream));
  return FP_OK;
}

int fp_reset_solid(fp_solver* s) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_reset_solid: NULL solver");
  if (!s->has_cube) return fail(s, FP_ESTATE, "fp_reset_solid before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  rc = settle(s);
  if (rc) return rc;
  const int ns = s->ns;
  std::vector<float4> p(ns);
  FP_TRY(s, d2h(s, p.data(), s->pos, ns * sizeof(float4)));
  std::vector<float4> pv(ns), v(ns, make_float4(0.f, 0.f, 0.f, 0.f));
  for (int i = 0; i < ns; i++) {
    p[i] = make_float4(s->h_snap_pos[3 * i], s->h_snap_pos[3 * i + 1], s->h_snap_pos[3 * i + 2], p[i].w);
    pv[i] = make_float4(p[i].x, p[i].y, p[i].z, 0.f);
  }
  FP_TRY(s, h2d(s, s->pos, p.data(), ns * sizeof(float4)));
  FP_TRY(s, h2d(s, s->prev, pv.data(), ns * sizeof(float4)));
  FP_TRY(s, h2d(s, s->vel, v.data(), ns * sizeof(float4)));
  if (s->C) {
    std::vector<uint8_t> ones(s->C, 1);
    FP_TRY(s, h2d(s, s->rest[s->cur], s->h_snap_rest.data(), s->C * sizeof(float)));
    FP_TRY(s, h2d(s, s->alive[s->cur], ones.data(), s->C));
  }
  s->h_group.assign(ns, 0);
  s->groups_dirty = false;
  s->grab_round_len.clear();
  return FP_OK;
}

int fp_clear_fluid(fp_solver* s) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_clear_fluid: NULL solver");
  if (!s->has_cube) return FP_OK;
  int rc = set_device(s);
  if (rc) return rc;
  if (s->nf) FP_TRY(s, hipMemsetAsync(s->active + s->ns, 0, s->nf, s->stream));
  return FP_OK;
}

static int step_impl(fp_solver* s, double dt, int32_t substeps, bool wait) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_step: NULL solver");
  if (!s->has_cube) return fail(s, FP_ESTATE, "fp_step before fp_add_solid_cube");
  if (substeps < 1) return fail(s, FP_EINVAL, "fp_step: substeps < 1");
  int rc = set_device(s);
  if (rc) return rc;
  const double h = dt / substeps;
  if (s->cut_dirty) {
    s->cut_dirty_at_step = true;
    s->cut_dirty = false;
  }
  for (int k = 0; k < substeps; k++) {
    rc = launch_substep(s, h);
    if (rc) return rc;
  }
  s->steps_pending = true;
  return wait ? settle(s) : FP_OK;
}

int fp_step(fp_solver* s, double dt, int32_t substeps) { return step_impl(s, dt, substeps, true); }
int fp_step_async(fp_solver* s, double dt, int32_t substeps) { return step_impl(s, dt, substeps, false); }

int fp_sync(fp_solver* s) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_sync: NULL solver");
  int rc = set_device(s);
  if (rc) return rc;
  return settle(s);
}

int fp_set_grabs(fp_solver* s, const int32_t* idx, const double* xyz, int32_t n) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_set_grabs: NULL solver");
  if (n < 0 || (n > 0 && (!idx || !xyz))) return fail(s, FP_EINVAL, "fp_set_grabs: bad arguments");
  for (int g = 0; g < n; g++)
    if (idx[g] < 0 || idx[g] >= s->n) return fail(s, FP_EINVAL, "fp_set_grabs: particle index out of range");
  int rc = set_device(s);
  if (rc) return rc;
  s->grab_round_len.clear();
  if (n == 0) return FP_OK;
  // split into rounds of distinct indices, keeping each index's order
  std::vector<int> round_of(n);
  std::vector<int> seen_count;
  {
    std::vector<std::pair<int, int>> key(n);
    for (int g = 0; g < n; g++) key[g] = {idx[g], g};
    std::sort(key.begin(), key.end());
    for (int g = 0; g < n; g++) round_of[key[g].second] = (g > 0 && key[g].first == key[g - 1].first) ? round_of[key[g - 1].second] + 1 : 0;
  }
  int rounds = 0;
  for (int g = 0; g < n; g++) rounds = std::max(rounds, round_of[g] + 1);
  std::vector<int> hidx;
  std::vector<float4> htgt;
  for (int r = 0; r < rounds; r++) {
    int len = 0;
    for (int g = 0; g < n; g++)
      if (round_of[g] == r) {
        hidx.push_back(idx[g]);
        htgt.push_back(make_float4((float)xyz[3 * g], (float)xyz[3 * g + 1], (float)xyz[3 * g + 2], 0.f));
        len++;
      }
    s->grab_round_len.push_back(len);
  }
  if (n > s->grab_cap) {
    FP_TRY(s, hipStreamSynchronize(s->stream));
    if (s->grab_idx) hipFree(s->grab_idx);
    if (s->grab_tgt) hipFree(s->grab_tgt);
    s->grab_idx = nullptr;
    s->grab_tgt = nullptr;
    const int cap = std::max(n, 256);
    if ((rc = dalloc(s, &s->grab_idx, cap)) || (rc = dalloc(s, &s->grab_tgt, cap))) return rc;
    s->grab_cap = cap;
  }
  FP_TRY(s, hipMemcpyAsync(s->grab_idx, hidx.data(), n * sizeof(int), hipMemcpyHostToDevice, s->stream));
  FP_TRY(s, hipMemcpyAsync(s->grab_tgt, htgt.data(), n * sizeof(float4), hipMemcpyHostToDevice, s->stream));
  FP_TRY(s, hipStreamSynchronize(s->stream));
  return FP_OK;
}

int fp_cut(fp_solver* s, const double o[3], const double d1[3], const double d2[3], int32_t* ncut) {
  if (!s || !o || !d1 || !d2) return fail(s, FP_EINVAL, "fp_cut: NULL argument");
  if (ncut) *ncut = 0;
  if (!s->has_cube) return fail(s, FP_ESTATE, "fp_cut before fp_add_solid_cube");
  const double nx = d1[1] * d2[2] - d1[2] * d2[1];
  const double ny = d1[2] * d2[0] - d1[0] * d2[2];
  const double nz = d1[0] * d2[1] - d1[1] * d2[0];
  const double nlen = std::sqrt(nx * nx + ny * ny + nz * nz);
  if (nlen < 1e-8 || s->C == 0) return FP_OK;
  int rc = set_device(s);
  if (rc) return rc;
  rc = settle(s);
  if (rc) return rc;
  FP_TRY(s, hipMemsetAsync(s->dflags + 1, 0, sizeof(int), s->stream));
  {
    PassScope ps(s, P_CUT);
    hipLaunchKernelGGL(k_cut, dim3(nblocks(s->C)), dim3(BLOCK), 0, s->stream, s->C, s->pos, s->cons_ab,
                       s->alive[s->cur], o[0], o[1], o[2], d1[0], d1[1], d1[2], d2[0], d2[1], d2[2], nx, ny, nz,
                       s->dflags + 1);
  }
  FP_TRY(s, hipGetLastError());
  int cnt = 0;
  FP_TRY(s, hipMemcpyAsync(&cnt, s->dflags + 1, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  FP_TRY(s, hipStreamSynchronize(s->stream));
  if (s->profile) profile_collect(s);
  if (ncut) *ncut = cnt;
  // the reference marks components dirty and relabels at the end of the
  // next step (solver.js:238-241, :386)
  if (cnt) s->cut_dirty = true;
  return FP_OK;
}

int fp_pick_solid(fp_solver* s, const double o[3], const double d[3], double max_perp, int32_t* id, double* t) {
  if (!s || !o || !d || !id) return fail(s, FP_EINVAL, "fp_pick_solid: NULL argument");
  *id = -1;
  if (t) *t = INFINITY;
  if (!s->has_cube || s->ns == 0) return FP_OK;
  int rc = set_device(s);
  if (rc) return rc;
  {
    PassScope ps(s, P_PICK);
    hipLaunchKernelGGL(k_pick, dim3(1), dim3(BLOCK), 0, s->stream, s->ns, s->pos, o[0], o[1], o[2], d[0], d[1], d[2],
                       max_perp * max_perp, s->dpick_t, s->dpick_i);
  }
  FP_TRY(s, hipGetLastError());
  double ht = INFINITY;
  int hi = -1;
  FP_TRY(s, hipMemcpyAsync(&ht, s->dpick_t, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  FP_TRY(s, hipMemcpyAsync(&hi, s->dpick_i, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  rc = settle(s);
  if (rc) return rc;
  *id = hi;
  if (t) *t = hi >= 0 ? ht : INFINITY;
  return FP_OK;
}

int fp_fluid_count(fp_solver* s, int32_t* count) {
  if (!s || !count) return fail(s, FP_EINVAL, "fp_fluid_count: NULL argument");
  *count = 0;
  if (!s->has_cube || s->nf == 0) return FP_OK;
  int rc = set_device(s);
  if (rc) return rc;
  FP_TRY(s, hipMemsetAsync(s->dflags + 2, 0, sizeof(int), s->stream));
  {
    PassScope ps(s, P_COUNT);
    hipLaunchKernelGGL(k_count_active, dim3(nblocks(s->nf)), dim3(BLOCK), 0, s->stream, s->ns, s->n, s->active,
                       s->dflags + 2);
  }
  FP_TRY(s, hipGetLastError());
  int c = 0;
  FP_TRY(s, hipMemcpyAsync(&c, s->dflags + 2, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  rc = settle(s);
  if (rc) return rc;
  *count = c;
  return FP_OK;
}

int fp_read_state(fp_solver* s, float* pos, float* prev, float* vel, float* inv_mass, uint8_t* active,
                  int32_t* group) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_read_state: NULL solver");
  if (!s->has_cube) return fail(s, FP_ESTATE, "fp_read_state before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  rc = settle(s);
  if (rc) return rc;
  const int N = s->n;
  std::vector<float4> tmp(N);
  if (pos || inv_mass) {
    FP_TRY(s, d2h(s, tmp.data(), s->pos, N * sizeof(float4)));
    for (int i = 0; i < N; i++) {
      if (pos) {
        pos[3 * i] = tmp[i].x;
        pos[3 * i + 1] = tmp[i].y;
        pos[3 * i + 2] = tmp[i].z;
      }
      if (inv_mass) inv_mass[i] = tmp[i].w;
    }
  }
  float4* src[2] = {s->prev, s->vel};
  float* dst[2] = {prev, vel};
  for (int k = 0; k < 2; k++) {
    if (!dst[k]) continue;
    FP_TRY(s, d2h(s, tmp.data(), src[k], N * sizeof(float4)));
    for (int i = 0; i < N; i++) {
      dst[k][3 * i] = tmp[i].x;
      dst[k][3 * i + 1] = tmp[i].y;
      dst[k][3 * i + 2] = tmp[i].z;
    }
  }
  if (active) FP_TRY(s, d2h(s, active, s->active, N));
  if (group) std::memcpy(group, s->h_group.data(), s->ns * sizeof(int32_t));
  return FP_OK;
}

int fp_write_state(fp_solver* s, const float* pos, const float* prev, const float* vel, const float* inv_mass,
                   const uint8_t* active) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_write_state: NULL solver");
  if (!s->has_cube) return fail(s, FP_ESTATE, "fp_write_state before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  rc = settle(s);
  if (rc) return rc;
  const int N = s->n;
  std::vector<float4> tmp(N);
  if (pos || inv_mass) {
    FP_TRY(s, d2h(s, tmp.data(), s->pos, N * sizeof(float4)));
    for (int i = 0; i < N; i++) {
      if (pos) tmp[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], tmp[i].w);
      if (inv_mass) tmp[i].w = inv_mass[i];
    }
    FP_TRY(s, h2d(s, s->pos, tmp.data(), N * sizeof(float4)));
  }
  float4* dstp[2] = {s->prev, s->vel};
  const float* srcp[2] = {prev, vel};
  for (int k = 0; k < 2; k++) {
    if (!srcp[k]) continue;
    for (int i = 0; i < N; i++) tmp[i] = make_float4(srcp[k][3 * i], srcp[k][3 * i + 1], srcp[k][3 * i + 2], 0.f);
    FP_TRY(s, h2d(s, dstp[k], tmp.data(), N * sizeof(float4)));
  }
  if (active) FP_TRY(s, h2d(s, s->active, active, N));
  return FP_OK;
}

int fp_read_constraints(fp_solver* s, int32_t* a, int32_t* b, float* rest, uint8_t* alive) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_read_constraints: NULL solver");
  if (!s->has_cube) return fail(s, FP_ESTATE, "fp_read_constraints before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  rc = settle(s);
  if (rc) return rc;
  if (a) std::memcpy(a, s->h_cons_a.data(), s->C * sizeof(int32_t));
  if (b) std::memcpy(b, s->h_cons_b.data(), s->C * sizeof(int32_t));
  if (rest && s->C) FP_TRY(s, d2h(s, rest, s->rest[s->cur], s->C * sizeof(float)));
  if (alive && s->C) FP_TRY(s, d2h(s, alive, s->alive[s->cur], s->C));
  return FP_OK;
}

int fp_write_constraints(fp_solver* s, const float* rest, const uint8_t* alive) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_write_constraints: NULL solver");
  if (!s->has_cube) return fail(s, FP_ESTATE, "fp_write_constraints before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  rc = settle(s);
  if (rc) return rc;
  if (rest && s->C) FP_TRY(s, h2d(s, s->rest[s->cur], rest, s->C * sizeof(float)));
  if (alive && s->C) FP_TRY(s, h2d(s, s->alive[s->cur], alive, s->C));
  return FP_OK;
}

int fp_build_grid(fp_solver* s) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_build_grid: NULL solver");
  if (!s->has_cube) return fail(s, FP_ESTATE, "fp_build_grid before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  rc = launch_grid_build(s);
  if (rc) return rc;
  return settle(s);
}

int fp_read_grid(fp_solver* s, int32_t* cell_start, int32_t* cell_entries) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_read_grid: NULL solver");
  if (!s->has_cube) return fail(s, FP_ESTATE, "fp_read_grid before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  rc = settle(s);
  if (rc) return rc;
  if (cell_start) FP_TRY(s, d2h(s, cell_start, s->cell_start, ((size_t)s->T + 1) * sizeof(int)));
  if (cell_entries) FP_TRY(s, d2h(s, cell_entries, s->cell_entries, (size_t)s->n * sizeof(int)));
  return FP_OK;
}

int fp_read_density(fp_solver* s, float* rho, float* rho_near) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_read_density: NULL solver");
  if (!s->has_cube) return fail(s, FP_ESTATE, "fp_read_density before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  rc = settle(s);
  if (rc) return rc;
  std::vector<float2> tmp(s->n);
  FP_TRY(s, d2h(s, tmp.data(), s->rho, s->n * sizeof(float2)));
  for (int i = 0; i < s->n; i++) {
    if (rho) rho[i] = tmp[i].x;
    if (rho_near) rho_near[i] = tmp[i].y;
  }
  return FP_OK;
}

int fp_profile_enable(fp_solver* s, int32_t enable) {
  if (!s) return fail(nullptr, FP_EINVAL, "fp_profile_enable: NULL solver");
  int rc = set_device(s);
  if (rc) return rc;
  rc = settle(s);
  if (rc) return rc;
  s->profile = enable != 0;
  for (int k = 0; k < P_NUM; k++) {
    s->prof_ms[k] = 0;
    s->prof_n[k] = 0;
  }
  return FP_OK;
}

int fp_profile_read(fp_solver* s, fp_pass_stat* out, int32_t max, int32_t* n) {
  if (!s || !out || !n) return fail(s, FP_EINVAL, "fp_profile_read: NULL argument");
  int rc = set_device(s);
  if (rc) return rc;
  rc = settle(s);
  if (rc) return rc;
  int m = 0;
  for (int k = 0; k < P_NUM && m < max; k++) {
    if (!s->prof_n[k]) continue;
    std::memset(out[m].name, 0, sizeof(out[m].name));
    std::strncpy(out[m].name, kPassNames[k], sizeof(out[m].name) - 1);
    out[m].launches = s->prof_n[k];
    out[m].total_ms = s->prof_ms[k];
    m++;
  }
  *n = m;
  return FP_OK;
}

}  // extern "C"
