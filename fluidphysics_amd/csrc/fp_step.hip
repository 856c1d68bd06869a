// fp_step.hip -- the per-particle passes, the order of one substep (the
// body of the reference's step loop, src/solver.js:192-237) and the launch
// groups that run it for one scene or a batch of scenes:
//
//   integrate (+ hash count)   v += g h; prev = p; p += v h           (:194-202)
//   grab rounds                p += (target - p) k                    (:205-210)
//   cell grid                  fp_grid.hip (the reference's hash grid of
//                              :212 is rebuilt from the kept buckets only
//                              when read)
//   contacts                   fp_fluid.hip                            (:213)
//   fluid order                fp_grid.hip
//   density, displace          fp_fluid.hip                            (:214)
//   constraints x iters        fp_solid.hip                            (:215)
//   bounds + velocity          floor, walls, cull; v = (p - prev)/h,
//                              damping, speed clamp                   (:216-234)
//   xsph                       fp_fluid.hip                            (:236)
//
// Every pass is one launch over all scenes of a group (fp_internal.h): the
// kernels stride over (scene, block) items and read the scene's SceneArgs.
// A frame without grabs is captured once into a HIP graph and replayed.
#include <cstring>

#include <map>
#include <mutex>

#include "fp_internal.h"
#include "fp_kernels.h"
#include "fp_bounds.h"

// Integrate (solver.js:194-202); with COUNT the particle is also entered
// into the cell grid at its integrated position (cell key, arrival slot,
// reference bucket).  Inactive particles get rho = 0 (the reference clears
// rho of every fluid slot each substep, :292).  Gravity acts on y; the x and
// z components (a tilted basin, batched scenes) are added only when set.
template <bool COUNT>
__device__ __forceinline__ void integrate_body(const SceneArgs& A, int lb) {
  const int i = lb * BLOCK + threadIdx.x;
  const bool in = i < A.n;
  // the flag, the velocity and the position load together (lanes past the
  // end read slot 0)
  const int j = in ? i : 0;
  const uint8_t flag = gld(A.active + j);
  float4 v = gld(A.vel + j);
  float4 p = gld(A.pos + j);
  const bool act = in && flag;
  uint32_t k = 0;
  if (act) {
    v.y += A.gvy;
    if (A.gvx != 0.f) v.x += A.gvx;
    if (A.gvz != 0.f) v.z += A.gvz;
    gst(A.prev + i, make_float4(p.x, p.y, p.z, 0.f));
    p.x = fmaf(v.x, A.h, p.x);
    p.y = fmaf(v.y, A.h, p.y);
    p.z = fmaf(v.z, A.h, p.z);
    gst(A.pos + i, p);
    if (COUNT) {
      gst(A.phash + i, cell_hash(cell_of(p.x, p.y, p.z, A.fp.cg.ref_inv), A.fp.cg.hmask));
      k = cell_key(A.fp.cg, cell_of(p.x, p.y, p.z, A.fp.cg.inv_cell));
      gst(A.lkey + i, k);
    }
  } else {
    p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (in) {
      if (COUNT) gst(A.phash + i, 0xffffffffu);
      gst(A.rho + i, make_float2(0.f, 0.f));
    }
  }
  if (COUNT) {
    // neighbouring indices are often in one cell: one atomic per run
    const uint32_t sl = wave_bucket_slot(A.lcount, k, act);
    if (act) gst(A.lslot + i, sl);
    box_update_block<BLOCK>(act && i < A.ns, p.x, p.y, p.z, A.sbox);
  }
}

template <bool COUNT>
__global__ void __launch_bounds__(BLOCK) kg_integrate(const SceneArgs* __restrict__ S, const int2* __restrict__ W) {
  FOR_ITEMS(S, W) integrate_body<COUNT>(A, lb);
}

// One round of grab pulls of one scene; indices inside a round are distinct.
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

// The solids of a scene without constraint iterations (their positions are
// final after the contacts); every other particle is finished by the last
// constraint iteration (solids) or by the displacement pass (fluid).
__device__ __forceinline__ void bounds_velocity_body(const SceneArgs& A, int lb) {
  const int i = lb * BLOCK + threadIdx.x;
  if (i >= A.bp.ns || A.cp.iters > 0) return;
  solid_finish(A, i, gld(A.pos + i));
}

__global__ void __launch_bounds__(BLOCK) kg_bounds_velocity(const SceneArgs* __restrict__ S,
                                                            const int2* __restrict__ W) {
  FOR_ITEMS(S, W) bounds_velocity_body(A, lb);
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

// ---- scene arguments and launch groups ---------------------------------------

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

SceneArgs fp_scene_args(const fp_solver* s, double h) {
  const fp_params& P = s->P;
  SceneArgs a;
  std::memset(&a, 0, sizeof(a));
  a.fp = fp_fluid_params(s);
  a.src.spos2 = s->spos2;
  a.src.sphash = s->sphash2;
  a.src.lstart = s->lstart_f;
  a.pl.nbr = s->nbr;
  a.pl.ncount = s->ncount;
  a.pl.cap = s->pair_slots / s->fluid_lanes;
  a.cc = ContactCtx{s->spos, s->sid, s->lstart, s->lnsol, s->adj_off, s->adj, s->ealive, s->sbox};
  a.bp = bounds_params(s, h);
  a.cp.ns = s->ns;
  a.cp.iters = (s->ns > 0 && s->C > 0) ? std::max(1, (int)P.constraint_iters) : 0;
  a.cp.alpha = (float)(P.compliance / (h * h));
  a.cp.omega = (float)P.jacobi_omega;
  a.cp.break_strain = (float)P.break_strain;
  a.cp.plastic_thr = (float)P.plastic_threshold;
  a.cp.plastic_rate = (float)P.plastic_rate;
  a.n = s->n;
  a.ns = s->ns;
  a.C = s->C;
  a.gvx = (float)(P.gravity_x * h);
  a.gvy = (float)(P.gravity * h);
  a.gvz = (float)(P.gravity_z * h);
  a.h = (float)h;
  a.pos = s->pos;
  a.prev = s->prev;
  a.vel = s->vel;
  a.active = s->active;
  a.rho = s->rho;
  a.phash = s->phash;
  a.lcount = s->lcount;
  a.lstart = s->lstart;
  a.lstart_f = s->lstart_f;
  a.lblock_f = s->lblock_f;
  a.lblock_sums = s->lblock_sums;
  a.ltmp = s->ltmp;
  a.lnsol = s->lnsol;
  a.lkey = s->lkey;
  a.lslot = s->lslot;
  a.sid = s->sid;
  a.spos = s->spos;
  a.sphash = s->sphash;
  a.tpos = s->tpos;
  a.tkey = s->tkey;
  a.tslot = s->tslot;
  a.spos2 = s->spos2;
  a.svtmp = s->svtmp;
  a.sphash2 = s->sphash2;
  a.sP = s->sP;
  a.sbox = s->sbox;
  a.dflags = s->dflags;
  a.tear = s->tear_dev;
  a.erest = s->erest;
  a.ealive = s->ealive;
  a.cent = s->cent;
  a.adj_off = s->adj_off;
  a.adj = s->adj;
  a.spos_c[0] = s->spos_c[0];
  a.spos_c[1] = s->spos_c[1];
  // the optional solid modes run in the constraint iterations
  SolidModes& m = a.sm;
  m.tets = s->tets_on && a.cp.iters > 0 && s->NT > 0;
  m.cache = s->cache_on && a.cp.iters > 0;
  m.n = s->cube_n;
  m.NT = s->NT;
  m.spacing = s->spacing;
  const double rsum = 2.0 * P.solid_radius, rsum_adj = std::min(rsum, 0.97 * s->spacing);
  const double margin = std::max(0.0, std::min(rsum, P.cell_size - rsum));
  m.rsum = (float)rsum;
  m.rsum_adj = (float)rsum_adj;
  m.lim = (float)(rsum + margin);
  m.lim_adj = (float)(rsum_adj + margin);
  m.cap = (float)(0.5 * P.solid_radius);
  m.gap = P.tear_gap * rsum;
  m.tet_v = s->tet_v;
  m.tet_v0 = s->tet_v0;
  m.tet_alive = s->tet_alive;
  m.tet_scale = s->tet_scale;
  m.tet_adj_off = s->tet_adj_off;
  m.tet_adj = s->tet_adj;
  m.cache_ids = s->cache_ids;
  m.cache_n = s->cache_n;
  m.spos_x = s->spos_x;
  return a;
}

void fp_group_free(fp_group* g) {
  for (hipEvent_t& e : g->shape_ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  g->shape_pending = false;
  for (hipEvent_t e : g->tune_ev) (void)hipEventDestroy(e);
  g->tune_ev.clear();
  if (g->side) (void)hipStreamSynchronize(g->side);
  if (g->ev_fork) (void)hipEventDestroy(g->ev_fork);
  if (g->ev_join) (void)hipEventDestroy(g->ev_join);
  if (g->side) (void)hipStreamDestroy(g->side);
  g->side = nullptr;
  g->ev_fork = g->ev_join = nullptr;
  if (g->d_tune_items) fp_pool_put(g->device, g->d_tune_items);
  g->d_tune_items = nullptr;
  g->tune_items_cap = 0;
  g->tune_items_ok = false;
  if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->gexec = nullptr;
  g->graph = nullptr;
  if (g->d_args) fp_pool_put(g->device, g->d_args);
  g->d_args = nullptr;
  g->args_cap = 0;
  for (int k = 0; k < W_NUM; k++) {
    if (g->d_items[k]) fp_pool_put(g->device, g->d_items[k]);
    g->d_items[k] = nullptr;
    g->items_cap[k] = 0;
    g->n_items[k] = 0;
  }
}

static void drop_graph(fp_group* g) {
  if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->gexec = nullptr;
  g->graph = nullptr;
  g->graph_sub = 0;
}

// A device buffer of at least `count` elements of T, grown geometrically;
// true when it moved (a graph that captured the old address is dropped).
template <class T>
static hipError_t reserve(fp_group* g, T*& p, int& cap, int count, bool& moved) {
  moved = false;
  if (count <= cap && p) return hipSuccess;
  hipError_t e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) return e;
  if (p) fp_pool_put(g->device, p);
  p = nullptr;
  // headroom as large as the growth a captured frame absorbs (1.5x its grids)
  const int n = std::max(count + count / 2 + 16, 2 * cap);
  e = fp_pool_get(g->device, sizeof(T) * (size_t)n, (void**)&p, nullptr);
  cap = e == hipSuccess ? n : 0;
  moved = true;
  return e;
}

// 1: captured frames of groups with FP_SIDE_MIN particle slots or more run
// the solid chain on the group's side stream.
#ifndef FP_SIDE_STREAM
#define FP_SIDE_STREAM 1
#endif
#ifndef FP_SIDE_MIN
#define FP_SIDE_MIN 131072
#endif
// Scene table and work lists for substep h: rewritten when the members,
// their parameters or h changed.  A captured frame survives a rewrite (its
// kernels read the table and the lists' counts on the device); it is dropped
// only when a buffer has to move.
int fp_group_prepare(fp_group* g, double h) {
  bool members = g->dirty || g->versions.size() != g->scenes.size();
  if (!members)
    for (size_t k = 0; k < g->scenes.size(); k++)
      if (g->versions[k] != g->scenes[k]->version) members = true;
  g->members_changed = members;
  const bool changed = members || g->built_h != h || g->items_clanes != g->clanes ||
                       (g->tune_record && !g->tune_items_ok);
  g->any_grabs = false;
  for (fp_solver* s : g->scenes) g->any_grabs |= !s->grab_round_len.empty();
  if (!changed) return FP_OK;
  const int ns = (int)g->scenes.size();
  std::vector<SceneArgs> args(ns);
  g->iters = -1;
  for (int k = 0; k < ns; k++) {
    args[k] = fp_scene_args(g->scenes[k], h);
    if (g->iters < 0) g->iters = args[k].cp.iters;
    if (args[k].cp.iters != g->iters && args[k].cp.iters > 0) {
      if (g->iters == 0) {
        g->iters = args[k].cp.iters;
      } else {
        g->err = "the scenes of a batch must use the same constraint_iters";
        return FP_EINVAL;
      }
    }
  }
  if (g->iters < 0) g->iters = 0;
  g->any_unconstrained = false;
  g->any_cache = g->any_tets = g->stage_rule = false;
  for (int k = 0; k < ns; k++) {
    g->any_unconstrained |= args[k].cp.ns > 0 && args[k].cp.iters == 0;
    // staging pays where a lane records many pairs (2 lanes per particle: the
    // 256-scene batch's density 5.65 -> 4.84 ms per frame); at 4 lanes it
    // cost the large scene 2%
    g->stage_rule |= args[k].fp.fl_log2 <= 1;
    g->any_cache |= args[k].sm.cache != 0;
    g->any_tets |= args[k].sm.tets != 0;
  }
  g->any_extra = g->any_cache || g->any_tets;
  {
    int64_t slots = 0;
    for (fp_solver* s : g->scenes) slots += s->n;
    g->ride = !g->any_extra && !(FP_SIDE_STREAM && slots >= FP_SIDE_MIN);
  }
  int64_t snap = 0;
  for (int k = 0; k < ns; k++) {
    args[k].snap_off = snap;
    snap += 3 * (int64_t)g->scenes[k]->n;
  }
  bool moved = false;
  FP_GTRY(g, reserve(g, g->d_args, g->args_cap, std::max(ns, 1), moved));
  if (moved) drop_graph(g);
  if (ns) {
    FP_GTRY(g, hipMemcpyAsync(g->d_args, args.data(), sizeof(SceneArgs) * ns, hipMemcpyHostToDevice, g->stream));
  }
  // work lists (count first): rewritten for a new membership, or when a
  // parameter change (contact_lanes) changed a scene's block count
  std::vector<int2> items[W_NUM];
  bool rebuild = g->dirty || g->versions.size() != g->scenes.size() || g->items_clanes != g->clanes;
  for (int kind = 0; kind < W_NUM; kind++) {
    items[kind].push_back(make_int2(0, g->xcd_run));
    for (int k = 0; k < ns; k++) {
      const int nb = fp_scene_blocks(g->scenes[k], kind, g->clanes);
      for (int b = 0; b < nb; b++) items[kind].push_back(make_int2(k, b));
    }
    items[kind][0].x = (int)items[kind].size() - 1;
    rebuild |= items[kind][0].x != g->n_items[kind];
  }
  if (rebuild) {
    for (int kind = 0; kind < W_NUM; kind++) {
      const int n = (int)items[kind].size() - 1;
      FP_GTRY(g, reserve(g, g->d_items[kind], g->items_cap[kind], n + 1, moved));
      if (moved) drop_graph(g);
      FP_GTRY(g, hipMemcpyAsync(g->d_items[kind], items[kind].data(), sizeof(int2) * (n + 1), hipMemcpyHostToDevice,
                                g->stream));
      g->n_items[kind] = n;
    }
    g->items_clanes = g->clanes;
    g->tune_items_ok = false;
  }
  // a tuning frame launches the solid contacts from each candidate's list
  // (count first, then the items; tune_off[c] is the count's place)
  std::vector<int2> titems;
  if (g->tune_record && !g->tune_items_ok) {
    for (int c = 0; c < 4; c++) {
      g->tune_off[c] = (int)titems.size();
      titems.push_back(make_int2(0, g->xcd_run));
      for (int k = 0; k < ns; k++) {
        const int nb = fp_scene_blocks(g->scenes[k], W_SOLID, kTuneLanes[c]);
        for (int b = 0; b < nb; b++) titems.push_back(make_int2(k, b));
      }
      g->tune_n[c] = (int)titems.size() - g->tune_off[c] - 1;
      titems[g->tune_off[c]].x = g->tune_n[c];
    }
    const int n = (int)titems.size();
    FP_GTRY(g, reserve(g, g->d_tune_items, g->tune_items_cap, n, moved));
    FP_GTRY(g, hipMemcpyAsync(g->d_tune_items, titems.data(), sizeof(int2) * n, hipMemcpyHostToDevice, g->stream));
    g->tune_items_ok = true;
  }
  // the host vectors are read by the copies above before they go out of scope
  FP_GTRY(g, hipStreamSynchronize(g->stream));
  g->versions.resize(ns);
  for (int k = 0; k < ns; k++) g->versions[k] = g->scenes[k]->version;
  g->built_h = h;
  g->dirty = false;
  return FP_OK;
}

// 1: the first frame after the members (or their parameters) changed runs
// eagerly; capture waits for a frame whose members match the last one's.
#ifndef FP_GRAPH_DEFER
#define FP_GRAPH_DEFER 1
#endif

// One substep of every scene of the group.
static int group_substep(fp_group* g) {
  int rc;
  {
    PassScope ps(g->prof, g->stream, P_INTEGRATE);
    if (g->any_grabs)
      GROUP_LAUNCH(g, kg_integrate<false>, W_PARTICLE, BLOCK);
    else
      GROUP_LAUNCH(g, kg_integrate<true>, W_PARTICLE, BLOCK);
  }
  if (g->any_grabs) {
    PassScope ps(g->prof, g->stream, P_GRAB);
    for (fp_solver* s : g->scenes) {
      int off = 0;
      for (int len : s->grab_round_len) {
        hipLaunchKernelGGL(k_grab, dim3(fp_blocks(len)), dim3(BLOCK), 0, g->stream, len, s->grab_idx + off,
                           s->grab_tgt + off, s->pos, (float)s->P.grab_stiffness);
        off += len;
      }
    }
  }
  if ((rc = fp_group_cell_grid(g, !g->any_grabs))) return rc;
  if (g->ride && !g->tune_record && !(g->prof && g->prof->on)) {
    // a group that does not fork: its solid chain rides in the fluid chain's
    // launches (fp_fluid.hip), the last three constraint iterations beside
    // the fluid rank, the density and the displacement, any earlier ones on
    // launches of their own after the contacts (a tuning frame times the
    // solid contacts alone, and per-pass profiling times every pass on its
    // own launch: both take the serial order below)
    const int I = g->iters, own = I > 3 ? I - 3 : 0, k = I - own;
    if ((rc = fp_group_contacts_both(g))) return rc;
    if ((rc = fp_group_constraints(g, own, g->stream))) return rc;
    if ((rc = fp_group_resort_with(g, k >= 3 ? own : -1))) return rc;
    if ((rc = fp_group_density_with(g, k >= 2 ? I - 2 : -1))) return rc;
    if ((rc = fp_group_displace_with(g, k >= 1 ? I - 1 : -1))) return rc;
  } else if (g->capturing && g->fork) {
    // the solid chain on the side stream, the fluid chain here; they join
    // before XSPH (each chain ends its particles' substep: the last
    // constraint iteration its solids', the displacement its fluid's).  The
    // fluid chain's first kernel is captured before the side stream's: the
    // replayed graph then keeps the fluid chain -- the longer one -- on the
    // launch's hardware queue and puts the solid chain on another, so only
    // the solid chain pays the cross-queue start (about 10 us at each end
    // of a branch on another queue; large scene: frame 1.19 -> 1.15 ms)
    FP_GTRY(g, hipEventRecord(g->ev_fork, g->stream));
    if ((rc = fp_group_contacts_fluid(g))) return rc;
    FP_GTRY(g, hipStreamWaitEvent(g->side, g->ev_fork, 0));
    if ((rc = fp_group_solid_collect(g, g->side))) return rc;
    if ((rc = fp_group_contacts_solid(g, g->side))) return rc;
    if ((rc = fp_group_constraints(g, g->iters, g->side))) return rc;
    FP_GTRY(g, hipEventRecord(g->ev_join, g->side));
    if ((rc = fp_group_resort(g))) return rc;
    if ((rc = fp_group_density(g))) return rc;
    if ((rc = fp_group_displace(g))) return rc;
    FP_GTRY(g, hipStreamWaitEvent(g->stream, g->ev_join, 0));
  } else {
    if ((rc = fp_group_solid_collect(g, g->stream))) return rc;
    if ((rc = fp_group_contacts(g))) return rc;
    if ((rc = fp_group_resort(g))) return rc;
    if ((rc = fp_group_density(g))) return rc;
    if ((rc = fp_group_displace(g))) return rc;
    if ((rc = fp_group_constraints(g, g->iters, g->stream))) return rc;
  }
  if (g->any_unconstrained) {
    PassScope ps(g->prof, g->stream, P_BOUNDS_VELOCITY);
    GROUP_LAUNCH(g, kg_bounds_velocity, W_PARTICLE, BLOCK);
  }
  if ((rc = fp_group_xsph(g))) return rc;
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

// One frame (substeps of dt / substeps) of every scene of the group, queued
// on the group's stream.  Without grabs or profiling the frame is captured
// into a HIP graph once and replayed.
// ---- contact-lane autotuning ---------------------------------------------------

// Tuning results of this process, by group signature: the contact lanes
// (consulted only once a cache file has been loaded: a fresh process tunes)
// and the launch shape (always consulted).  fp_tune_cache_save / _load
// (fp_api.hip) keep them in a JSON file between processes.
static std::mutex g_tune_mu;
static std::map<uint64_t, int> g_lanes_cache;
static std::map<uint64_t, int> g_shape_cache;
static bool g_lanes_cache_on = false;

// The shape a contact-lane choice is kept for: device, scene count and
// solid count to the next power of two.
static uint64_t lanes_signature(const fp_group* g, int64_t solids) {
  const uint64_t a = 64 - __builtin_clzll((uint64_t)g->scenes.size() | 1u);
  const uint64_t b = 64 - __builtin_clzll((uint64_t)solids | 1u);
  return a | (b << 8) | ((uint64_t)(g->device & 255) << 24) | (1ull << 40);
}

std::vector<std::pair<uint64_t, int>> fp_tune_cache_entries(bool lanes) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const std::map<uint64_t, int>& m = lanes ? g_lanes_cache : g_shape_cache;
  return std::vector<std::pair<uint64_t, int>>(m.begin(), m.end());
}

void fp_tune_cache_merge(const std::vector<std::pair<uint64_t, int>>& lanes,
                         const std::vector<std::pair<uint64_t, int>>& shape) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (const auto& e : lanes) g_lanes_cache[e.first] = e.second;
  for (const auto& e : shape) g_shape_cache[e.first] = e.second;
  g_lanes_cache_on = true;
}

void fp_tune_caches_clear() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_lanes_cache.clear();
  g_shape_cache.clear();
  g_lanes_cache_on = false;
}

// Before a frame: decide whether it tunes (the candidates' lane counts are
// then picked per substep), else use the tuned lane count.
void fp_group_tune_begin(fp_group* g) {
  int64_t solids = 0;
  for (fp_solver* s : g->scenes) solids += s->ns;
  if (g->clanes_best && !g->clanes_pinned && (solids > 2 * g->tuned_solids || 2 * solids < g->tuned_solids)) {
    g->clanes_best = 0;  // the work changed size: tune again
    g->tune_k = 0;
  }
  if (!g->clanes_best && solids == 0) {
    g->clanes_best = 4;
    g->tuned_solids = 0;
  }
  g->tune_record = false;
  if (!g->clanes_best && !g->clanes_pinned && g->tune_k == 0) {
    // a choice loaded from a tuning cache file for a group of this shape
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_lanes_cache.find(lanes_signature(g, solids));
    if (g_lanes_cache_on && it != g_lanes_cache.end()) {
      g->clanes_best = it->second;
      g->tuned_solids = solids;
    }
  }
  if (g->clanes_best) {
    g->clanes = g->clanes_best;
    return;
  }
  if (g->tune_k == 0)
    for (int c = 0; c < 4; c++) {
      g->tune_sum[c] = 0.0;
      g->tune_cnt[c] = 0;
    }
  g->tune_record = true;
  g->tune_used = 0;
  g->tune_of.clear();
  g->tuned_solids = solids;
}

// After a tuning frame has been queued: its events are read at the next
// frame (or sync), once they have completed.
// A candidate within this factor of the fastest mean counts as tied.
#ifndef TUNE_TIE
#define TUNE_TIE 1.25
#endif

int fp_group_tune_end(fp_group* g) {
  if (!g->tune_pending) return FP_OK;
  g->tune_pending = false;
  for (int i = 0; i < g->tune_used; i++) {
    float ms = 0.f;
    FP_GTRY(g, hipEventSynchronize(g->tune_ev[2 * i + 1]));
    FP_GTRY(g, hipEventElapsedTime(&ms, g->tune_ev[2 * i], g->tune_ev[2 * i + 1]));
    g->tune_sum[g->tune_of[i]] += ms;
    g->tune_cnt[g->tune_of[i]]++;
  }
  if (++g->tune_k == fp_group::TUNE_FRAMES) {
    int best = -1;
    for (int c = 0; c < 4; c++)
      if (g->tune_cnt[c] &&
          (best < 0 || g->tune_sum[c] / g->tune_cnt[c] < g->tune_sum[best] / g->tune_cnt[best]))
        best = c;
    // near-ties go to more lanes: the pass is timed alone, but a replayed
    // frame runs it beside the fluid chain, where a shorter chain per solid
    // ends the side stream sooner (256 batched scenes: graph frame p50 16.3
    // ms at 4 lanes, 15.7 at 8, 15.9 at 16; timed alone, 8 lanes are 6-12%
    // slower than 4, 16 about 30%)
    if (best >= 0) {
      const double lim = TUNE_TIE * g->tune_sum[best] / g->tune_cnt[best];
      for (int c = 0; c < 4; c++)
        if (g->tune_cnt[c] && kTuneLanes[c] > kTuneLanes[best] && g->tune_sum[c] / g->tune_cnt[c] <= lim) best = c;
    }
    g->clanes_best = kTuneLanes[best < 0 ? 0 : best];
    g->tune_k = 0;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_lanes_cache[lanes_signature(g, g->tuned_solids)] = g->clanes_best;
    }
    // the candidates' lists are idle: every launch that read them has been
    // waited for above
    if (g->d_tune_items) fp_pool_put(g->device, g->d_tune_items);
    g->d_tune_items = nullptr;
    g->tune_items_cap = 0;
    g->tune_items_ok = false;
  }
  return FP_OK;
}

// ---- launch-shape tuning (batches) -----------------------------------------------
//
// Two launch shapes leave every result unchanged: whether the density kernel
// stages its pair-list entries in LDS, and the XCD run length of the group
// kernels' work lists (their W[0].y).  A batch times them over whole
// captured frames -- the solid chain beside the fluid chain, as in the frames
// it will run -- once its contact lanes are tuned and while its members stay
// the same: each candidate runs FP_SHAPE_REPS timed frames (the first
// candidate of each staging after one untimed frame, which captures its
// graph).  The fastest mean is kept if it beats the default (staging by the
// 2-lane rule, runs of FP_XCD_CHUNK) by the factor FP_SHAPE_GAIN, else the
// default, and the choice is cached for later groups of the same signature
// (device, scene count and slot count to the next power of two, the 2-lane
// rule, contact lanes), so a serving loop's changing batches do not tune again.
#define FP_SHAPE_CANDS 6
#ifndef FP_SHAPE_REPS
#define FP_SHAPE_REPS 2
#endif
#ifndef FP_SHAPE_GAIN
#define FP_SHAPE_GAIN 0.98
#endif
static const int kShapeRun[3] = {FP_XCD_CHUNK, 4, 1};  // candidate c: run kShapeRun[c % 3], staging flipped for c >= 3

static uint64_t shape_signature(const fp_group* g) {
  uint64_t slots = 0;
  for (const fp_solver* s : g->scenes) slots += (uint64_t)s->n;
  const uint64_t a = 64 - __builtin_clzll((uint64_t)g->scenes.size() | 1u);
  const uint64_t b = 64 - __builtin_clzll(slots | 1u);
  return a | (b << 8) | ((uint64_t)g->stage_rule << 16) | ((uint64_t)(g->clanes & 63) << 17) |
         ((uint64_t)(g->device & 255) << 24);
}

__global__ void k_set_run(int2* __restrict__ w, int run) { w[0].y = run; }

// The shape of candidate c: the group's staging flag and, if it changed, the
// run length in every work list's header (written on the group's stream, in
// order before the next frame's kernels read it).
static int shape_apply(fp_group* g, int c) {
  g->stage_on = (c < 3) ? (g->stage_rule ? 1 : 0) : (g->stage_rule ? 0 : 1);
  const int run = kShapeRun[c % 3];
  if (run == g->xcd_run) return FP_OK;
  g->xcd_run = run;
  for (int k = 0; k < W_NUM; k++)
    if (g->d_items[k]) hipLaunchKernelGGL(k_set_run, dim3(1), dim3(1), 0, g->stream, g->d_items[k], run);
  if (g->d_tune_items && g->tune_items_ok)
    for (int k = 0; k < 4; k++)
      hipLaunchKernelGGL(k_set_run, dim3(1), dim3(1), 0, g->stream, g->d_tune_items + g->tune_off[k], run);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

// Before a frame: the shape it runs with; timed: it is a tuning frame.
static int shape_begin(fp_group* g, bool graphable, bool& timed) {
  timed = false;
  int c = 0;
  if (g->shape_tune) {
    const uint64_t sig = shape_signature(g);
    if (sig != g->shape_sig || (g->members_changed && g->shape_state == 1)) {
      g->shape_sig = sig;
      int cached = -1;
      {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_shape_cache.find(sig);
        if (it != g_shape_cache.end()) cached = it->second;
      }
      g->shape_state = cached >= 0 ? 2 : 0;
      g->shape_cand = cached >= 0 ? cached : 0;
      g->shape_pending = false;  // a frame timed for the old members does not count
    }
    if (g->shape_state == 0 && graphable && g->clanes_best && !g->members_changed) {
      g->shape_state = 1;
      g->shape_cand = 0;
      g->shape_done = 0;
      for (int k = 0; k < FP_SHAPE_CANDS; k++) {
        g->shape_ms[k] = 0.0;
        g->shape_n[k] = 0;
      }
    }
    if (g->shape_state == 1 && graphable) {
      c = g->shape_cand;
      if (g->shape_done == 0 && c % 3 == 0) {
        g->shape_done = 1;  // untimed: captures the candidate's graph
      } else {
        timed = true;
      }
    } else if (g->shape_state == 2) {
      c = g->shape_cand;
    }
  }
  return shape_apply(g, c);
}

// After a tuning frame has completed (read at the next frame): its time.
static int shape_end(fp_group* g) {
  if (!g->shape_pending) return FP_OK;
  g->shape_pending = false;
  float ms = 0.f;
  FP_GTRY(g, hipEventSynchronize(g->shape_ev[1]));
  FP_GTRY(g, hipEventElapsedTime(&ms, g->shape_ev[0], g->shape_ev[1]));
  const int c = g->shape_pending_cand;
  if (g->shape_state != 1 || c != g->shape_cand) return FP_OK;
  g->shape_ms[c] += ms;
  if (++g->shape_n[c] < FP_SHAPE_REPS) return FP_OK;
  g->shape_cand++;
  g->shape_done = 0;
  if (g->shape_cand < FP_SHAPE_CANDS) return FP_OK;
  int best = 0;
  for (int k = 1; k < FP_SHAPE_CANDS; k++)
    if (g->shape_ms[k] / g->shape_n[k] < g->shape_ms[best] / g->shape_n[best]) best = k;
  if (g->shape_ms[best] / g->shape_n[best] > FP_SHAPE_GAIN * g->shape_ms[0] / g->shape_n[0]) best = 0;
  g->shape_state = 2;
  g->shape_cand = best;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_shape_cache[g->shape_sig] = best;
  return FP_OK;
}

int fp_group_frame(fp_group* g, double dt, int substeps) {
  const double h = dt / substeps;
  int rc = fp_group_tune_end(g);
  if (rc) return rc;
  if ((rc = shape_end(g))) return rc;
  fp_group_tune_begin(g);
  rc = fp_group_prepare(g, h);
  if (rc) return rc;
  const bool profiling = g->prof && g->prof->on;
  // a frame whose members just changed and that would have to capture a new
  // graph runs eagerly instead
  const bool graphable = g->use_graph && !g->any_grabs && !profiling && !g->tune_record &&
                         !(FP_GRAPH_DEFER && g->members_changed && !g->gexec);
  bool shape_timed = false;
  if ((rc = shape_begin(g, graphable, shape_timed))) return rc;
  if (graphable) {
    // the captured frame replays whatever the scene table and the work lists
    // hold; it is captured again when its structure no longer fits: the
    // substep or iteration count, the contact pass's lane count (a kernel
    // template), the bounds pass of unconstrained scenes, the optional solid
    // modes' passes, the density kernel (staging or not), or work lists that
    // outgrew (or shrank far below) the grids captured for them
    const int extra = (g->any_cache ? 1 : 0) | (g->any_tets ? 2 : 0) | (g->stage_on ? 4 : 0);
    bool fits = g->gexec && g->graph_sub == substeps && g->graph_iters == g->iters &&
                g->graph_clanes == g->clanes && g->graph_unconstrained == g->any_unconstrained &&
                g->graph_extra == extra;
    for (int k = 0; k < W_NUM && fits; k++)
      fits = 2 * g->n_items[k] <= 3 * g->grid_cap[k] && 4 * fp_grid8(g->n_items[k]) >= g->grid_cap[k];
    if (!fits) {
      drop_graph(g);
      int64_t slots = 0;
      for (fp_solver* s : g->scenes) slots += s->n;
      g->fork = FP_SIDE_STREAM && slots >= FP_SIDE_MIN;
      for (int k = 0; k < W_NUM; k++) g->grid_cap[k] = fp_grid8(g->n_items[k]);
      if (g->fork && !g->side) {
        FP_GTRY(g, hipStreamCreateWithFlags(&g->side, hipStreamNonBlocking));
        FP_GTRY(g, hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming));
        FP_GTRY(g, hipEventCreateWithFlags(&g->ev_join, hipEventDisableTiming));
      }
      FP_GTRY(g, hipStreamBeginCapture(g->stream, hipStreamCaptureModeRelaxed));
      int crc = FP_OK;
      g->capturing = true;
      for (int k = 0; k < substeps && crc == FP_OK; k++) crc = group_substep(g);
      g->capturing = false;
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(g->stream, &graph);
      if (crc) {
        if (graph) (void)hipGraphDestroy(graph);
        return crc;
      }
      FP_GTRY(g, e);
      g->graph = graph;
      FP_GTRY(g, hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0));
      g->graph_captures++;
      g->graph_sub = substeps;
      g->graph_iters = g->iters;
      g->graph_clanes = g->clanes;
      g->graph_unconstrained = g->any_unconstrained;
      g->graph_extra = extra;
    }
    if (shape_timed) {
      if (!g->shape_ev[0]) {
        FP_GTRY(g, hipEventCreate(&g->shape_ev[0]));
        FP_GTRY(g, hipEventCreate(&g->shape_ev[1]));
      }
      FP_GTRY(g, hipEventRecord(g->shape_ev[0], g->stream));
    }
    FP_GTRY(g, hipGraphLaunch(g->gexec, g->stream));
    if (shape_timed) {
      FP_GTRY(g, hipEventRecord(g->shape_ev[1], g->stream));
      g->shape_pending = true;
      g->shape_pending_cand = g->shape_cand;
    }
    g->graph_launches++;
  } else {
    for (int k = 0; k < substeps; k++) {
      g->tune_cand = (k + g->tune_k) & 3;
      if ((rc = group_substep(g))) return rc;
    }
    if (g->tune_record) {
      g->tune_record = false;
      g->tune_pending = true;
    }
  }
  for (fp_solver* s : g->scenes) s->substeps_done += substeps;
  return FP_OK;
}

// Positions of every scene into one packed xyz buffer at each scene's offset.
__global__ void __launch_bounds__(BLOCK)
    kg_snapshot(const SceneArgs* __restrict__ S, const int2* __restrict__ W, float* __restrict__ out) {
  FOR_ITEMS(S, W) {
    const int i = lb * BLOCK + (int)threadIdx.x;
    if (i < A.n) {
      const float4 p = A.pos[i];
      float* o = out + A.snap_off + 3 * (int64_t)i;
      o[0] = p.x;
      o[1] = p.y;
      o[2] = p.z;
    }
  }
}

int fp_group_snapshot(fp_group* g, float* out) {
  GROUP_LAUNCH(g, kg_snapshot, W_PARTICLE, BLOCK, out);
  FP_GTRY(g, hipGetLastError());
  return FP_OK;
}

int fp_launch_emit(fp_solver* s, int m) {
  PassScope ps(s, P_EMIT);
  hipLaunchKernelGGL(k_emit, dim3(fp_blocks(m)), dim3(BLOCK), 0, s->stream, m, (const int*)(s->emit_buf + 2 * m),
                     s->emit_buf, s->emit_buf + m,
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
