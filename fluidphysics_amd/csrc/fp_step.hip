// fp_step.hip -- the per-particle passes and the order of one substep
// (the body of the reference's step loop, src/solver.js:192-237):
//
//   integrate (+ hash count)   v.y += g h; prev = p; p += v h        (:194-202)
//   grab rounds                p += (target - p) k                    (:205-210)
//   cell grid                  fp_grid.hip (the reference's hash grid of
//                              :212 is rebuilt from the kept buckets only
//                              when read)
//   contacts                   fp_fluid.hip                            (:213)
//   density, displace          fp_fluid.hip                            (:214)
//   constraints x iters        fp_solid.hip                            (:215)
//   bounds + velocity          floor, walls, cull; v = (p - prev)/h,
//                              damping, speed clamp                   (:216-234)
//   xsph                       fp_fluid.hip                            (:236)
//
// When no grab is active the cell-grid count runs inside the integrate
// kernel (the cell of the integrated position is known right there).
#include "fp_internal.h"
#include "fp_kernels.h"

// Integrate (solver.js:194-202); with COUNT the particle is also entered
// into the cell grid at its integrated position (cell key, arrival slot,
// reference bucket).  Inactive particles get rho = 0 (the reference clears
// rho of every fluid slot each substep, :292).
template <bool COUNT>
__global__ void __launch_bounds__(BLOCK) k_integrate(CellGrid cg, int n, int ns, float4* __restrict__ pos,
                                                     float4* __restrict__ prev, float4* __restrict__ vel,
                                                     const uint8_t* __restrict__ active, float gh, float h,
                                                     int* __restrict__ lcount, uint32_t* __restrict__ lkey,
                                                     uint32_t* __restrict__ lslot, uint32_t* __restrict__ phash,
                                                     float2* __restrict__ rho, int* __restrict__ sbox) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const bool in = i < n;
  const bool act = in && active[i];
  float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t k = 0;
  if (act) {
    float4 v = vel[i];
    p = pos[i];
    v.y += gh;
    prev[i] = make_float4(p.x, p.y, p.z, 0.f);
    p.x = fmaf(v.x, h, p.x);
    p.y = fmaf(v.y, h, p.y);
    p.z = fmaf(v.z, h, p.z);
    pos[i] = p;
    vel[i] = v;
    if (COUNT) {
      const Cell c = cell_of(p.x, p.y, p.z, cg.inv_cell);
      phash[i] = cell_hash(c, cg.hmask);
      k = cell_key(cg, c);
      lkey[i] = k;
    }
  } else if (in) {
    if (COUNT) phash[i] = 0xffffffffu;
    rho[i] = make_float2(0.f, 0.f);
  }
  if (COUNT) {
    // neighbouring indices are often in one cell: one atomic per run
    const uint32_t sl = wave_bucket_slot(lcount, k, act);
    if (act) lslot[i] = sl;
    box_update(act && i < ns, p.x, p.y, p.z, sbox);
  }
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

// Math.sign(q) || 1 (solver.js:413): +1 for +0, -0 and positive q.
__device__ __forceinline__ float sign_or_one(float q) { return q < 0.f ? -1.f : 1.f; }

// Bounds (solver.js:389-427) then velocity + damping + clamp (:219-234).
// Solids read their positions from `solid_src` (the last constraint
// iteration) and write vel; fluid velocities go to the fluid-order slot in svtmp,
// where XSPH reads them (a particle culled here leaves its velocity as it
// was, in vel and in svtmp).
__global__ void __launch_bounds__(BLOCK) k_bounds_velocity(BoundsParams bp, float4* __restrict__ pos,
                                                           const float4* __restrict__ solid_src,
                                                           const float4* __restrict__ prev, float4* __restrict__ vel,
                                                           float4* __restrict__ svtmp, const int* __restrict__ slot,
                                                           uint8_t* __restrict__ active) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= bp.n) return;
  const bool solid = i < bp.ns;
  if (!active[i]) {
    if (solid) pos[i] = solid_src[i];  // the reference's constraint sweep moves inactive solids too
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
      const float side = sign_or_one(p.x);
      p.x = fabsf(p.x) < bp.wall_x ? side * (bp.wall_x - lim) : side * (bp.wall_x + lim);
    }
    if (fabsf(fabsf(p.z) - bp.wall_z) < lim) {
      const float side = sign_or_one(p.z);
      p.z = fabsf(p.z) < bp.wall_z ? side * (bp.wall_z - lim) : side * (bp.wall_z + lim);
    }
  }
  pos[i] = p;
  if (!solid && (p.y < -3.f || p.x * p.x + p.z * p.z > bp.cull_r2)) {
    active[i] = 0;
    svtmp[slot[i]] = vel[i];
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
    svtmp[slot[i]] = v;
}

__global__ void __launch_bounds__(BLOCK) k_count_active(int begin, int end, const uint8_t* __restrict__ active,
                                                        int* __restrict__ out) {
  const int i = begin + blockIdx.x * BLOCK + threadIdx.x;
  const int v = (i < end) ? (int)active[i] : 0;
  const int s = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

// emitFluid scatter (solver.js:139-149): batch entry k -> particle idx[k].
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

static BoundsParams bounds_params(const fp_solver* s, double h) {
  const fp_params& P = s->P;
  BoundsParams bp;
  bp.ns = s->ns;
  bp.n = s->n;
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
  return bp;
}

int fp_launch_substep(fp_solver* s, double h) {
  const fp_params& P = s->P;
  const int n = s->n;
  const bool grabs = !s->grab_round_len.empty();
  {
    PassScope ps(s, P_INTEGRATE);
    if (grabs)
      hipLaunchKernelGGL(k_integrate<false>, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, s->cg, n, s->ns, s->pos,
                         s->prev, s->vel, s->active, (float)(P.gravity * h), (float)h, s->lcount, s->lkey, s->lslot,
                         s->phash, s->rho, s->sbox);
    else
      hipLaunchKernelGGL(k_integrate<true>, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, s->cg, n, s->ns, s->pos,
                         s->prev, s->vel, s->active, (float)(P.gravity * h), (float)h, s->lcount, s->lkey, s->lslot,
                         s->phash, s->rho, s->sbox);
  }
  if (grabs) {
    PassScope ps(s, P_GRAB);
    int off = 0;
    for (int len : s->grab_round_len) {
      hipLaunchKernelGGL(k_grab, dim3(fp_blocks(len)), dim3(BLOCK), 0, s->stream, len, s->grab_idx + off,
                         s->grab_tgt + off, s->pos, (float)P.grab_stiffness);
      off += len;
    }
  }
  int rc = fp_launch_cell_grid(s, !grabs);
  if (rc) return rc;
  if ((rc = fp_launch_contacts(s))) return rc;
  if ((rc = fp_launch_resort(s))) return rc;
  if ((rc = fp_launch_density(s))) return rc;
  if ((rc = fp_launch_displace(s))) return rc;
  const float4* solid_src = nullptr;
  if ((rc = fp_launch_constraints(s, h, &solid_src))) return rc;
  {
    const BoundsParams bp = bounds_params(s, h);
    PassScope ps(s, P_BOUNDS_VELOCITY);
    hipLaunchKernelGGL(k_bounds_velocity, dim3(fp_blocks(n)), dim3(BLOCK), 0, s->stream, bp, s->pos, solid_src,
                       s->prev, s->vel, s->svtmp, s->fslot, s->active);
  }
  if ((rc = fp_launch_xsph(s))) return rc;
  FP_TRY(s, hipGetLastError());
  s->substeps_done++;
  return FP_OK;
}

int fp_launch_emit(fp_solver* s, int m) {
  PassScope ps(s, P_EMIT);
  hipLaunchKernelGGL(k_emit, dim3(fp_blocks(m)), dim3(BLOCK), 0, s->stream, m, s->emit_idx, s->emit_p, s->emit_v,
                     s->pos, s->prev, s->vel, s->active);
  FP_TRY(s, hipGetLastError());
  return FP_OK;
}

int fp_launch_count_active(fp_solver* s) {
  PassScope ps(s, P_COUNT);
  hipLaunchKernelGGL(k_count_active, dim3(fp_blocks(s->nf)), dim3(BLOCK), 0, s->stream, s->ns, s->n, s->active,
                     s->dflags + DF_SCRATCH);
  FP_TRY(s, hipGetLastError());
  return FP_OK;
}
