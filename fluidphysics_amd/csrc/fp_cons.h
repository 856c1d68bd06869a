// fp_cons.h -- one Jacobi iteration of the distance constraints, the body
// of the constraint pass (solver.js:356-381 per-constraint formula): its
// own launch (kg_constraints), or the workgroups of a fluid pass's launch
// that carry the solid chain of a small group (fp_fluid.hip, fp_grid.hip).
#pragma once

#include "fp_internal.h"
#include "fp_kernels.h"
#include "fp_bounds.h"

// One Jacobi iteration `it` of one scene: positions from spos_c[1] (it = 0:
// the solid contact pass's result, kept there as well as in pos) or
// spos_c[(it - 1) & 1] into spos_c[it & 1]; the last iteration finishes the
// solids into pos, which no iteration reads, so a single iteration reads a
// snapshot no workgroup of its launch writes.  Each CSR entry holds its own
// copy of the constraint's rest length and alive flag; both endpoints
// compute the same update from the same snapshot (d is the same bits either
// way round), so each lane updates its own entry in place, and only when
// the value changes (a tear, a plastic step).  FP_CLANES lanes work on one
// solid; lane l takes entries l, l + FP_CLANES, ... of its row, CCHUNK of
// them per step with their loads in flight together, sums them in that
// order, and the lanes' sums meet in a fixed xor-shuffle tree.
#ifndef CCHUNK
#define CCHUNK 4
#endif
__device__ __forceinline__ void constraints_body(const SceneArgs& A, int lb, int it) {
  const ConsParams& cp = A.cp;
  // with the optional solid modes, iteration it - 1 ended in their extra step (spos_x)
  const bool extra = A.sm.tets | A.sm.cache;
  const float4* __restrict__ sin = it == 0 ? A.spos_c[1] : (extra ? A.sm.spos_x : A.spos_c[(it - 1) & 1]);
  float4* __restrict__ sout = A.spos_c[it & 1];
  float* __restrict__ erest = A.erest;
  uint8_t* __restrict__ ealive = A.ealive;
  const int lane = threadIdx.x & (FP_CLANES - 1);
  const int a = (lb * BLOCK + (int)threadIdx.x) / FP_CLANES;
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
  for (int kb = k0 + lane; kb < k1; kb += CCHUNK * FP_CLANES) {
    // the chunk's flags, neighbours and rest lengths load together, then
    // the neighbours' positions (a dead entry's neighbour index is still
    // valid; entries past the row repeat its last one and are not used)
    bool alive[CCHUNK];
    uint32_t e[CCHUNK];
    float rest[CCHUNK];
    float4 pb[CCHUNK];
#pragma unroll
    for (int u = 0; u < CCHUNK; u++) {
      const int k = min(kb + u * FP_CLANES, k1 - 1);
      alive[u] = gld(ealive + k) != 0;
      e[u] = gld(A.adj + k);
      rest[u] = gld(erest + k);
    }
#pragma unroll
    for (int u = 0; u < CCHUNK; u++) pb[u] = gld(sin + adj_other(a, e[u]));
#pragma unroll
    for (int u = 0; u < CCHUNK; u++) {
      const int k = kb + u * FP_CLANES;
      const float w = pa.w + pb[u].w;
      const float dx = pa.x - pb[u].x, dy = pa.y - pb[u].y, dz = pa.z - pb[u].z;
      const float d = sqrtf(dx * dx + dy * dy + dz * dz);
      if (k < k1 && alive[u] && w != 0.f && d >= 1e-9f) {
        if (d > rest[u] * cp.break_strain) {
          gst(ealive + k, (uint8_t)0);
          tore = true;
        } else {
          const float C = d - rest[u];
          const float dl = -C / (w + cp.alpha);
          const float s = dl / d * pa.w * (cp.omega * adj_scale(e[u]));
          cx += dx * s;
          cy += dy * s;
          cz += dz * s;
          if (fabsf(C / rest[u]) > cp.plastic_thr) gst(erest + k, rest[u] + C * cp.plastic_rate);
        }
      }
    }
  }
#pragma unroll
  for (int off = FP_CLANES / 2; off > 0; off >>= 1) {
    cx += __shfl_xor(cx, off, FP_CLANES);
    cy += __shfl_xor(cy, off, FP_CLANES);
    cz += __shfl_xor(cz, off, FP_CLANES);
  }
  if (valid && lane == 0) {
    const float4 p = make_float4(pa.x + cx, pa.y + cy, pa.z + cz, pa.w);
    // the last iteration finishes the solid's substep (bounds, velocity),
    // unless the optional modes' extra step follows it
    if (it == A.cp.iters - 1 && !extra)
      solid_finish(A, a, p);
    else
      gst(sout + a, p);
  }
  if (tore) *A.tear = 1;
}
