// fp_bounds.h -- the end of a particle's substep: bounds (solver.js:389-427),
// then velocity + damping + clamp (:219-234).  Run where a particle's final
// position is computed: by the last constraint iteration for a solid (its
// position never goes back to memory in between) and by the displacement
// pass for a fluid particle.  A scene without constraint iterations finishes
// its solids in kg_bounds_velocity.
#pragma once

#include "fp_internal.h"
#include "fp_kernels.h"

// Math.sign(q) || 1 (solver.js:413): +1 for +0, -0 and positive q.
__device__ __forceinline__ float sign_or_one(float q) { return q < 0.f ? -1.f : 1.f; }

// Floor (with friction against the previous position pv) and walls.
__device__ __forceinline__ void bounds_apply(const BoundsParams& bp, bool solid, float4& p, const float4& pv) {
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
}

// (p - pv) / h, damped, clamped to vmax.
__device__ __forceinline__ float4 substep_velocity(const BoundsParams& bp, bool solid, const float4& p,
                                                   const float4& pv) {
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
  return make_float4(vx, vy, vz, 0.f);
}

// The scene's non-finite flag (the reference counts NaN positions in its
// metrics, src/gpu/main.js:546-556; NaNs otherwise propagate as there):
// raised, and kept, when a particle ends a substep with a non-finite
// position or velocity.  Every lane that raises it stores the same word.
__device__ __forceinline__ bool finite3(float x, float y, float z) {
  return __builtin_isfinite(x) & __builtin_isfinite(y) & __builtin_isfinite(z);
}
__device__ __forceinline__ void mark_nonfinite(const SceneArgs& A, bool ok) {
  if (!ok) gst(A.dflags + DF_NONFINITE, 1);
}

// Solid i whose position after the constraints is p: pos and vel (an
// inactive solid only moves: the reference's constraint sweep moves it too).
__device__ __forceinline__ void solid_finish(const SceneArgs& A, int i, float4 p) {
  const uint8_t flag = gld(A.active + i);
  const float4 pv = gld(A.prev + i);
  if (!flag) {
    gst(A.pos + i, p);
    return;
  }
  bounds_apply(A.bp, true, p, pv);
  gst(A.pos + i, p);
  const float4 v = substep_velocity(A.bp, true, p, pv);
  gst(A.vel + i, v);
  mark_nonfinite(A, finite3(p.x, p.y, p.z) & finite3(v.x, v.y, v.z));
}

// Active fluid particle i, in fluid slot fs, whose position after the
// displacement is p (x, y, z; its inverse mass stays in pos.w): pos, and its
// velocity into svtmp[fs] where XSPH reads it; a particle culled here is
// deactivated and leaves its velocity as it was, in vel and in svtmp.
__device__ __forceinline__ void fluid_finish(const SceneArgs& A, int i, int fs, float4 p) {
  const BoundsParams& bp = A.bp;
  const float4 pv = gld(A.prev + i);
  bounds_apply(bp, false, p, pv);
  gst_xyz(A.pos + i, p);
  if (p.y < -3.f || p.x * p.x + p.z * p.z > bp.cull_r2) {
    // culled: the velocity stays the substep's integrated one, vel + g h
    // (integrate does not store it; the same sums here)
    gst(A.active + i, (uint8_t)0);
    float4 v = gld(A.vel + i);
    v.y += A.gvy;
    if (A.gvx != 0.f) v.x += A.gvx;
    if (A.gvz != 0.f) v.z += A.gvz;
    gst(A.vel + i, v);
    gst(A.svtmp + fs, v);
    return;
  }
  const float4 v = substep_velocity(bp, false, p, pv);
  gst(A.svtmp + fs, v);
  mark_nonfinite(A, finite3(p.x, p.y, p.z) & finite3(v.x, v.y, v.z));
}
