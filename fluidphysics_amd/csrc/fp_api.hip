// fp_api.hip -- the C ABI of include/fluidphysics.h: scene set-up, state
// transfers, sync points, component labels, profiling.  Kernels live in
// fp_grid.hip, fp_fluid.hip, fp_solid.hip and fp_step.hip.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "fp_internal.h"

namespace {

thread_local std::string g_last_error = "";

const char* kPassNames[] = {"integrate",   "grab",        "hash_count",      "scan_reduce", "scan_apply",
                            "scatter",     "rank",        "contacts",        "resort",      "density",
                            "displace",    "constraints", "bounds_velocity", "xsph",        "cut",
                            "pick",        "count_active", "emit",           "solid_extra"};
static_assert(sizeof(kPassNames) / sizeof(kPassNames[0]) == P_NUM, "one name per PassId");

// Synchronous host <-> device copies ordered on the solver's (non-blocking)
// stream.  Copies up to kStageBytes go through one pinned buffer per device,
// shared by that device's solvers under a lock: a pageable copy takes the
// runtime's own staging path, 4-15x slower between 16 KiB and 1 MiB (192 KiB
// device to host: ~90 us, ~12 us from pinned memory).  Larger copies stay
// pageable (there the runtime's path is as fast as a staged copy plus a host
// memcpy).
constexpr size_t kStageBytes = 1u << 20;
constexpr size_t kWarmBytes = 256u << 10;
constexpr int kStageDevices = 64;
struct Stage {
  std::mutex m;
  bool done = false;
  void* buf = nullptr;  // kStageBytes of pinned host memory, or null (copies pageable)
};
Stage g_stage[kStageDevices];

// With the device's first solver (the device current): allocates its buffer,
// then copies kWarmBytes each way.  The first device-to-host copy above
// ~32 KiB in a process takes ~6 ms (every later one ~10 us); paid here, it
// does not land in a step loop (a cut's component relabel reads 2 bytes per
// constraint).
void stage_init(int device, hipStream_t stream) {
  if (device < 0 || device >= kStageDevices) return;
  Stage& st = g_stage[device];
  std::lock_guard<std::mutex> lk(st.m);
  if (st.done) return;
  st.done = true;
  if (hipHostMalloc(&st.buf, kStageBytes, 0) != hipSuccess) {
    st.buf = nullptr;
    (void)hipGetLastError();  // the device's solvers copy pageable instead
    return;
  }
  void* tmp = nullptr;
  if (hipMalloc(&tmp, kWarmBytes) == hipSuccess) {
    (void)hipMemcpyAsync(st.buf, tmp, kWarmBytes, hipMemcpyDeviceToHost, stream);
    (void)hipMemcpyAsync(tmp, st.buf, kWarmBytes, hipMemcpyHostToDevice, stream);
    (void)hipStreamSynchronize(stream);
    (void)hipFree(tmp);
  }
  (void)hipGetLastError();  // a failed warm-up only leaves the first large copy slow
}

hipError_t copy_sync(fp_solver* s, void* dst, const void* src, size_t bytes, bool to_host) {
  const hipMemcpyKind kind = to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
  if (bytes <= kStageBytes && s->device >= 0 && s->device < kStageDevices) {
    Stage& st = g_stage[s->device];
    std::lock_guard<std::mutex> lk(st.m);
    if (st.buf) {
      if (!to_host) std::memcpy(st.buf, src, bytes);
      hipError_t e = hipMemcpyAsync(to_host ? st.buf : dst, to_host ? src : st.buf, bytes, kind, s->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
      if (e == hipSuccess && to_host) std::memcpy(dst, st.buf, bytes);
      return e;
    }
  }
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s->stream);
  return e == hipSuccess ? hipStreamSynchronize(s->stream) : e;
}

hipError_t d2h(fp_solver* s, void* dst, const void* src, size_t bytes) { return copy_sync(s, dst, src, bytes, true); }

hipError_t h2d(fp_solver* s, void* dst, const void* src, size_t bytes) { return copy_sync(s, dst, src, bytes, false); }

void profile_collect(Prof* p) {
  for (auto& e : p->live) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
      p->ms[e.pass] += ms;
      p->n[e.pass] += 1;
    }
    p->pool.push_back(e.a);
    p->pool.push_back(e.b);
  }
  p->live.clear();
}

void prof_free(Prof* p) {
  for (auto& e : p->live) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  for (auto e : p->pool) (void)hipEventDestroy(e);
  p->live.clear();
  p->pool.clear();
}

// A growing staging buffer (emit, grab): one pool block of its own.
template <typename T>
int dalloc(fp_solver* s, T** p, size_t count) {
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  size_t cb = 0;
  hipError_t e = fp_pool_get(s->device, bytes, (void**)p, &cb);
  if (e != hipSuccess) return fp_fail(s, FP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  s->device_bytes += (int64_t)cb;
  e = hipMemsetAsync(*p, 0, bytes, s->stream);
  if (e != hipSuccess) return fp_fail(s, FP_EDEVICE, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
  return FP_OK;
}

template <typename T>
void dfree(fp_solver* s, T*& p) {
  if (p) s->device_bytes -= (int64_t)fp_pool_put(s->device, (void*)p);
  p = nullptr;
}

// The fixed buffers of a scene, in arena order.  `take(ptr, count)` is
// called once per buffer: sizing (Carve without a base) and placement
// (Carve with one) walk the same list.
template <class F>
void scene_buffers(fp_solver* s, F& take) {
  const size_t N = (size_t)s->n, T = (size_t)s->T, M = s->cg.M, ns = (size_t)s->ns;
  const size_t C = (size_t)std::max(s->C, 1);
  take(s->pos, N), take(s->prev, N), take(s->vel, N), take(s->active, N), take(s->rho, N), take(s->phash, N);
  take(s->lcount, M), take(s->lstart, M + 1), take(s->lstart_f, M + 1), take(s->lnsol, M);
  take(s->lblock_sums, (size_t)s->lscan_blocks), take(s->lblock_f, (size_t)s->lscan_blocks);
  take(s->ltmp, N), take(s->lkey, N), take(s->lslot, N), take(s->sid, N), take(s->spos, N), take(s->spos2, N);
  take(s->svtmp, N), take(s->sphash, N), take(s->sP, N), take(s->tpos, N), take(s->tkey, N), take(s->tslot, N);
  take(s->sphash2, N), take(s->sbox, 6);
  const size_t lanes = N * (size_t)s->fluid_lanes, lanes64 = (lanes + 63) / 64 * 64;
  take(s->nbr, (size_t)(s->pair_slots / s->fluid_lanes + 1) * lanes64), take(s->ncount, lanes);
  take(s->pslot, N), take(s->count, T), take(s->cell_start, T + 1), take(s->entries_tmp, N);
  take(s->cell_entries, N), take(s->block_sums, (size_t)s->scan_blocks);
  take(s->cons_ab, C), take(s->cent, C), take(s->erest, 2 * C), take(s->ealive, 2 * C), take(s->ealive_snap, 2 * C);
  take(s->dgroup, std::max<size_t>(ns, 1));
  take(s->adj_off, ns + 1), take(s->adj, 2 * C), take(s->spos_c[0], std::max<size_t>(ns, 1));
  take(s->spos_c[1], std::max<size_t>(ns, 1)), take(s->dflags, (size_t)DF_WORDS), take(s->dpick_t, 1);
  // the optional solid modes (empty when off)
  const size_t NT = s->tets_on ? (size_t)s->NT : 0, nsx = (s->tets_on || s->cache_on) ? ns : 0;
  take(s->tet_v, NT), take(s->tet_v0, NT), take(s->tet_alive, NT), take(s->tet_scale, NT);
  take(s->tet_adj_off, s->tets_on ? ns + 1 : 0), take(s->tet_adj, 4 * NT);
  take(s->cache_ids, s->cache_on ? ns * FP_CACHE_K : 0), take(s->cache_n, s->cache_on ? ns : 0);
  take(s->spos_x, nsx);
  take(s->dpick_i, 1);
}

struct Carve {
  char* base = nullptr;
  size_t off = 0;
  template <typename T>
  void operator()(T*& p, size_t count) {
    off = (off + 255) & ~(size_t)255;
    p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += std::max<size_t>(count, 1) * sizeof(T);
  }
};

void free_all(fp_solver* s) {
  Carve none;  // null every arena pointer
  scene_buffers(s, none);
  if (s->arena) s->device_bytes -= (int64_t)fp_pool_put(s->device, s->arena);
  s->arena = nullptr;
  for (auto& cp : s->checkpoints)
    if (cp.block) s->device_bytes -= (int64_t)fp_pool_put(s->device, cp.block);
  s->checkpoints.clear();
  dfree(s, s->grab_idx), dfree(s, s->grab_tgt), dfree(s, s->emit_buf);
}

uint32_t next_pow2(double v, uint32_t lo) {
  uint32_t t = lo;
  while ((double)t < v) t <<= 1;
  return t;
}

// Cell-grid dimensions (sort cells: the reference's cellSize / CELL_SUB).
// Direct layout: NZ and NY cover the basin (walls + 1 m) and 4 m of height,
// the table holds at least 4 z-y planes and the reference table size, at most
// 2^25 keys (far cells beyond that alias and are rejected by the distance
// tests).  A scene whose particles would fill under a sixteenth of that
// table gets a compact one instead, about one key per particle slot
// (nextpow2(slots), at least 4096), with hashed column offsets: scans of the
// counting sort cost 12 B per key, so small scenes (batches of them) keep
// their tables small.  Columns that alias share rows and cost candidates,
// which the distance tests reject; on the 256-scene batch 1 key per slot
// beat 2 (by 0.8%), 4 (by 3.8%) and 8 (by 14%).
void setup_cell_grid(fp_solver* s) {
  const fp_params& P = s->P;
  const double cs = P.cell_size / CELL_SUB;
  const double half = std::max(std::fabs(P.wall_x), std::fabs(P.wall_z)) + 1.0;
  const uint32_t NZ = next_pow2(std::ceil(2.0 * half / cs), 4);
  const uint32_t NY = next_pow2(std::ceil(4.0 / cs), 4);
  const uint64_t plane = (uint64_t)NZ * NY;
  uint64_t M = std::max<uint64_t>(plane * NZ, (uint64_t)s->T);
  M = std::max<uint64_t>(M, 4 * plane);
  M = std::min<uint64_t>(M, 1ull << 25);
  const uint32_t direct = next_pow2((double)M, 4);
  const uint32_t compact = next_pow2((double)s->n, 4096);
  s->cg.inv_cell = 1.0 / cs;
  s->cg.ref_inv = 1.0 / P.cell_size;
  s->cg.hmask = (uint32_t)(s->T - 1);
  if ((uint64_t)next_pow2(4.0 * s->n, 4096) * 4 <= direct) {
    s->cg.M = compact;
    s->cg.mx = 0x9E3779B1u;  // odd: every x step moves the column
    s->cg.my = 0x85EBCA6Bu;
  } else {
    s->cg.M = direct;
    s->cg.mx = (uint32_t)plane;
    s->cg.my = NZ;
  }
  s->cg.kmask = s->cg.M - 1;
  s->lscan_blocks = (int)((s->cg.M + SCAN_TILE - 1) / SCAN_TILE);
}

// Distance constraints of an n^3 lattice (every neighbour within sqrt(3)
// spacing, once): axis, face-diagonal and body-diagonal pairs.
int64_t cube_constraints(int64_t n) {
  return 3 * n * n * (n - 1) + 6 * n * (n - 1) * (n - 1) + 4 * (n - 1) * (n - 1) * (n - 1);
}

// A cube the 32-bit CSR can hold: capacity below 2^27 particles and both
// endpoint entries of every constraint (2C) below INT_MAX (n <= 436).
bool cube_fits(int64_t n, int64_t max_fluid) {
  if (n < 1 || n > 1024 || max_fluid < 0) return false;
  if (n * n * n + max_fluid >= (1ll << 27)) return false;
  return 2 * cube_constraints(n) < (int64_t)INT_MAX;
}

// Particle counts and the reference table of a cube of side n.
void scene_dims(fp_solver* s, int n) {
  s->cube_n = n;
  s->tets_on = s->P.tet_volume != 0;
  s->cache_on = s->P.solid_contact_cache != 0;
  s->NT = n > 1 ? 5 * (n - 1) * (n - 1) * (n - 1) : 0;
  s->ns = n * n * n;
  s->n = s->ns + s->max_fluid;
  s->nf = s->max_fluid;
  if (s->P.table_size > 0) {
    s->T = s->P.table_size;
  } else {
    int t = 1;
    while (t < 2 * s->n) t <<= 1;  // solver.js:51
    s->T = t;
  }
  s->scan_blocks = (s->T + SCAN_TILE - 1) / SCAN_TILE;
}

// Cell grid and pair-list width, once the counts are known.
void scene_tables(fp_solver* s) {
  setup_cell_grid(s);
  s->fluid_lanes = s->P.fluid_lanes > 0 ? s->P.fluid_lanes
                                         : (s->max_fluid >= FLUID_LANES_AUTO ? FLUID_LANES_BIG : FLUID_LANES_SMALL);
  s->pair_slots = s->P.pair_slots > 0 ? s->P.pair_slots : 64 * s->fluid_lanes;
}

// Per-constraint rest / alive <-> the two CSR-entry copies.
int constraints_to_host(fp_solver* s, float* rest, uint8_t* alive) {
  const int C = s->C;
  if (!C) return FP_OK;
  std::vector<float> er;
  std::vector<uint8_t> ea;
  if (rest) {
    er.resize(2 * (size_t)C);
    FP_TRY(s, d2h(s, er.data(), s->erest, 2 * (size_t)C * sizeof(float)));
  }
  if (alive) {
    ea.resize(2 * (size_t)C);
    FP_TRY(s, d2h(s, ea.data(), s->ealive, 2 * (size_t)C));
  }
  for (int c = 0; c < C; c++) {
    if (rest) rest[c] = er[s->h_cent[c].x];
    if (alive) alive[c] = ea[s->h_cent[c].x];
  }
  return FP_OK;
}

int constraints_to_device(fp_solver* s, const float* rest, const uint8_t* alive) {
  const int C = s->C;
  if (!C) return FP_OK;
  if (rest) {
    std::vector<float> er(2 * (size_t)C);
    for (int c = 0; c < C; c++) er[s->h_cent[c].x] = er[s->h_cent[c].y] = rest[c];
    FP_TRY(s, h2d(s, s->erest, er.data(), 2 * (size_t)C * sizeof(float)));
  }
  if (alive) {
    std::vector<uint8_t> ea(2 * (size_t)C);
    for (int c = 0; c < C; c++) ea[s->h_cent[c].x] = ea[s->h_cent[c].y] = alive[c] ? 1 : 0;
    FP_TRY(s, h2d(s, s->ealive, ea.data(), 2 * (size_t)C));
  }
  return FP_OK;
}

int set_device(fp_solver* s) {
  FP_TRY(s, hipSetDevice(s->device));
  return FP_OK;
}

// union-find with path halving, parent[ra] = rb (solver.js:488-499): the
// labels are the reference's roots.
void recompute_groups_host(fp_solver* s, const std::vector<uint8_t>& alive) {
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

// The component relabel the reference runs at the end of a step that tore
// or followed a cut (solver.js:238-241) is due: run it now, from the alive
// flags as the step left them.  It is run lazily -- when group is read, or
// before a cut or a constraint write changes the flags -- which gives the
// same labels as running it at the end of every such step.
int materialize_groups(fp_solver* s) {
  if (!s->groups_pending) return FP_OK;
  std::vector<uint8_t> al(s->C);
  if (s->groups_from_snap) {
    // the flags as they were when the labels fell due (kept on the device
    // before a cut changed them)
    std::vector<uint8_t> ea(2 * (size_t)s->C);
    FP_TRY(s, d2h(s, ea.data(), s->ealive_snap, ea.size()));
    for (int c = 0; c < s->C; c++) al[c] = ea[s->h_cent[c].x];
  } else {
    int rc = constraints_to_host(s, nullptr, al.data());
    if (rc) return rc;
  }
  recompute_groups_host(s, al);
  s->groups_pending = false;
  s->groups_from_snap = false;
  s->alive_changed = false;
  return FP_OK;
}

// Before a cut changes the alive flags while relabelled groups are due: keep
// the flags the labels are due from in a device copy (no host round trip;
// the labels are computed from it if the caller reads them before the next
// step, whose end makes new labels due from the flags then).
int defer_groups(fp_solver* s) {
  if (!s->groups_pending || s->groups_from_snap || s->C == 0) return FP_OK;
  FP_TRY(s, hipMemcpyAsync(s->ealive_snap, s->ealive, 2 * (size_t)s->C, hipMemcpyDeviceToDevice, s->stream));
  s->groups_from_snap = true;
  return FP_OK;
}

// Sync point: wait for the stream, pick up the tear flag of finished steps,
// mark the components for relabelling when a tear or a cut happened.
int settle(fp_solver* s) {
  FP_TRY(s, hipStreamSynchronize(s->stream));
  if (s->prof.on) profile_collect(&s->prof);
  if (s->steps_pending && s->tear_host[1]) s->nonfinite = true;  // sticky: word 1 is never cleared by a step
  if (s->steps_pending && s->C > 0) {
    volatile int* tear = s->tear_host;
    if (*tear) {
      s->groups_dirty = true;
      *tear = 0;  // the stream is idle: no kernel writes it now
    }
    if (s->cut_dirty_at_step) s->groups_dirty = true;
  }
  s->cut_dirty_at_step = false;
  s->steps_pending = false;
  if (s->groups_dirty) {
    s->groups_pending = true;
    s->groups_from_snap = false;
    s->groups_dirty = false;
    s->alive_changed = false;  // the due labels are the current flags' partition
  }
  return FP_OK;
}

int check_params(fp_solver* s, const fp_params& p) {
  const double rmax = std::max(p.solid_radius, p.fluid_radius);
  if (!(p.cell_size > 0) || !(p.fluid_h > 0) || p.cell_size < p.fluid_h || p.cell_size < 2 * rmax)
    return fp_fail(s, FP_EINVAL, "cell_size must be >= fluid_h and >= 2 * max(solid_radius, fluid_radius)");
  if (p.table_size < 0 || (p.table_size & (p.table_size - 1)))
    return fp_fail(s, FP_EINVAL, "table_size must be 0 or a power of two");
  if (p.constraint_iters < 0 || p.constraint_iters > 1000)
    return fp_fail(s, FP_EINVAL, "constraint_iters out of range");
  if (p.pair_slots < 0 || p.pair_slots > 1024) return fp_fail(s, FP_EINVAL, "pair_slots out of range [0, 1024]");
  if (p.fluid_lanes != 0 && p.fluid_lanes != 1 && p.fluid_lanes != 2 && p.fluid_lanes != 4 && p.fluid_lanes != 8)
    return fp_fail(s, FP_EINVAL, "fluid_lanes must be 0, 1, 2, 4 or 8");
  return FP_OK;
}

}  // namespace

int fp_fail(fp_solver* s, int code, const std::string& msg) {
  if (s)
    s->err = msg;
  else
    g_last_error = msg;
  return code;
}

hipEvent_t fp_event_get(Prof* p) {
  if (!p->pool.empty()) {
    hipEvent_t e = p->pool.back();
    p->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

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
  p->jacobi_omega = 8.0;  // with 3 iterations: the reference's 300-frame statistics at its own constants
  p->contact_scale = 1.0;
  p->tet_volume = 0;
  p->solid_contact_cache = 0;
  p->tear_gap = 1.3;
}

int fp_create(const fp_params* params, int32_t max_fluid, int32_t device, fp_solver** out) {
  if (!out) return fp_fail(nullptr, FP_EINVAL, "fp_create: out is NULL");
  *out = nullptr;
  if (max_fluid < 0) return fp_fail(nullptr, FP_EINVAL, "fp_create: max_fluid < 0");
  fp_solver* s = new (std::nothrow) fp_solver();
  if (!s) return fp_fail(nullptr, FP_ENOMEM, "fp_create: out of host memory");
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
  if (device < 0 && hipGetDevice(&device) != hipSuccess) device = 0;
  s->device = device;
  s->max_fluid = max_fluid;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) stage_init(device, s->stream);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&s->tear_host, 64, hipHostMallocMapped | hipHostMallocCoherent);
  if (e == hipSuccess) {
    // every word (pinned blocks are recycled: a freed solver's words would remain)
    std::memset(s->tear_host, 0, 64);
    e = hipHostGetDevicePointer((void**)&s->tear_dev, s->tear_host, 0);
  }
  if (e != hipSuccess) {
    if (s->tear_host) (void)hipHostFree(s->tear_host);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    g_last_error = std::string("fp_create: ") + hipGetErrorString(e);
    delete s;
    return FP_EDEVICE;
  }
  s->own_stream = s->stream;
  s->group = new fp_group();
  s->group->device = device;
  s->group->stream = s->stream;
  s->group->prof = &s->prof;
  s->group->scenes.push_back(s);
  *out = s;
  return FP_OK;
}

static void upload_free(fp_solver* s);

int fp_destroy(fp_solver* s) {
  if (!s) return FP_OK;
  (void)hipSetDevice(s->device);
  if (s->batch) (void)fp_batch_remove(s->batch, s);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->group) {
    fp_group_free(s->group);
    delete s->group;
  }
  free_all(s);
  upload_free(s);
  prof_free(&s->prof);
  if (s->own_stream) (void)hipStreamDestroy(s->own_stream);
  if (s->tear_host) (void)hipHostFree(s->tear_host);
  delete s;
  return FP_OK;
}

const char* fp_last_error(const fp_solver* s) { return s ? s->err.c_str() : g_last_error.c_str(); }

int fp_set_params(fp_solver* s, const fp_params* p) {
  if (!s || !p) return fp_fail(s, FP_EINVAL, "fp_set_params: NULL argument");
  int rc = check_params(s, *p);
  if (rc) return rc;
  if (s->has_cube &&
      (p->table_size != s->P.table_size || p->pair_slots != s->P.pair_slots || p->fluid_lanes != s->P.fluid_lanes ||
       (p->tet_volume != 0) != (s->P.tet_volume != 0) ||
       (p->solid_contact_cache != 0) != (s->P.solid_contact_cache != 0)))
    return fp_fail(s, FP_ESTATE,
                   "table_size / pair_slots / fluid_lanes / tet_volume / solid_contact_cache are fixed once the scene "
                   "is allocated");
  s->P = *p;
  s->version++;
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
  out->device = s->device;
  out->substeps_done = s->substeps_done;
  out->device_bytes = s->device_bytes;
  out->pair_overflow = 0;
  out->pair_slots = s->pair_slots;
  out->nonfinite = s->nonfinite ? 1 : 0;
  out->reserved = 0;
  if (s->dflags) {
    (void)hipSetDevice(s->device);
    int ov = 0;
    if (d2h(m, &ov, s->dflags + DF_PAIR_OVF, sizeof(ov)) == hipSuccess) out->pair_overflow = ov;
  }
  return FP_OK;
}

int fp_add_solid_cube(fp_solver* s, double cx, double cy, double cz, int32_t n, double spacing,
                      double particle_mass) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_add_solid_cube: NULL solver");
  if (s->has_cube) return fp_fail(s, FP_ESTATE, "addSolidCube may be called once");
  if (!cube_fits(n, s->max_fluid))
    return fp_fail(s, FP_EINVAL,
                   "cube dimension out of range (n^3 + maxFluid must stay below 2^27 and 2 x constraints below "
                   "INT_MAX: n <= 436)");
  int rc = set_device(s);
  if (rc) return rc;
  const int ns = n * n * n;
  scene_dims(s, n);

  // lattice (solver.js:74-89): double arithmetic, float32 stores
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
  // static CSR adjacency: entries (other endpoint, bits of 1/max(deg_a,
  // deg_b)), each with its own copy of the constraint's rest and alive flag
  std::vector<int> hoff(ns + 1, 0);
  for (int c = 0; c < C; c++) {
    hoff[s->h_cons_a[c] + 1]++;
    hoff[s->h_cons_b[c] + 1]++;
  }
  for (int i = 0; i < ns; i++) hoff[i + 1] += hoff[i];
  std::vector<uint32_t> hadj(std::max(1, 2 * C));
  std::vector<float> herest(std::max(1, 2 * C));
  std::vector<int> fill(ns, 0);
  std::vector<int2> hab(std::max(1, C));
  s->h_cent.assign(C, make_int2(0, 0));
  auto deg = [&](int i) { return hoff[i + 1] - hoff[i]; };
  for (int c = 0; c < C; c++) {
    const int a = s->h_cons_a[c], b = s->h_cons_b[c];
    // packed entries: a lattice neighbour lies within n^2 + n + 1 of a (n <=
    // 436, so under 2^23) and a solid has at most 26 of them
    if (std::abs(b - a) >= (1 << 23) || std::max(deg(a), deg(b)) > 31)
      return fp_fail(s, FP_EINVAL, "fp_add_solid_cube: constraint outside the packed CSR's range");
    const uint32_t md = (uint32_t)std::max(deg(a), deg(b)) << 24;
    const int ka = hoff[a] + fill[a]++, kb = hoff[b] + fill[b]++;
    hadj[ka] = ((uint32_t)(b - a) & 0xFFFFFFu) | md;
    hadj[kb] = ((uint32_t)(a - b) & 0xFFFFFFu) | md;
    herest[ka] = herest[kb] = hrest[c];
    s->h_cent[c] = make_int2(ka, kb);
    hab[c] = make_int2(a, b);
  }
  s->h_group.assign(ns, 0);
  s->spacing = spacing;

  // volume tets (the GPU twin's split, sim.js:151-166): five per lattice
  // cell, mirrored on checkerboard parity; the last two vertices swapped
  // where that makes the rest volume (double, from the float32 lattice)
  // positive -- the twin keeps its order and so treats half its tets as dead
  std::vector<int4> htv;
  std::vector<float> htv0, hts;
  std::vector<int> htoff, htadj;
  if (s->tets_on) {
    htv.reserve(s->NT);
    auto vol = [&](int a, int b, int c, int d) {
      const double abx = (double)hpos[b].x - hpos[a].x, aby = (double)hpos[b].y - hpos[a].y,
                   abz = (double)hpos[b].z - hpos[a].z;
      const double acx = (double)hpos[c].x - hpos[a].x, acy = (double)hpos[c].y - hpos[a].y,
                   acz = (double)hpos[c].z - hpos[a].z;
      const double adx = (double)hpos[d].x - hpos[a].x, ady = (double)hpos[d].y - hpos[a].y,
                   adz = (double)hpos[d].z - hpos[a].z;
      return ((aby * acz - abz * acy) * adx + (abz * acx - abx * acz) * ady + (abx * acy - aby * acx) * adz) / 6.0;
    };
    for (int ix = 0; ix + 1 < n; ix++)
      for (int iy = 0; iy + 1 < n; iy++)
        for (int iz = 0; iz + 1 < n; iz++) {
          const int A = idx(ix, iy, iz), B = idx(ix + 1, iy, iz), C = idx(ix + 1, iy + 1, iz), D = idx(ix, iy + 1, iz);
          const int E = idx(ix, iy, iz + 1), F = idx(ix + 1, iy, iz + 1), G = idx(ix + 1, iy + 1, iz + 1),
                    H = idx(ix, iy + 1, iz + 1);
          const int even[5][4] = {{A, C, F, H}, {A, B, C, F}, {A, C, D, H}, {A, E, F, H}, {C, F, G, H}};
          const int odd[5][4] = {{B, D, E, G}, {A, B, D, E}, {B, C, D, G}, {D, E, H, G}, {B, E, F, G}};
          const int(*five)[4] = (ix + iy + iz) % 2 == 0 ? even : odd;
          for (int k = 0; k < 5; k++) {
            int4 v = make_int4(five[k][0], five[k][1], five[k][2], five[k][3]);
            double V = vol(v.x, v.y, v.z, v.w);
            if (V < 0) {
              std::swap(v.z, v.w);
              V = -V;
            }
            htv.push_back(v);
            htv0.push_back((float)V);
          }
        }
    htoff.assign(ns + 1, 0);
    for (const int4& v : htv) {
      htoff[v.x + 1]++, htoff[v.y + 1]++, htoff[v.z + 1]++, htoff[v.w + 1]++;
    }
    for (int i = 0; i < ns; i++) htoff[i + 1] += htoff[i];
    htadj.assign(4 * (size_t)s->NT, 0);
    std::vector<int> tfill(ns, 0);
    for (int t = 0; t < (int)htv.size(); t++) {
      const int q[4] = {htv[t].x, htv[t].y, htv[t].z, htv[t].w};
      int m = 1;
      for (int u = 0; u < 4; u++) {
        htadj[htoff[q[u]] + tfill[q[u]]++] = t;
        m = std::max(m, htoff[q[u] + 1] - htoff[q[u]]);
      }
      hts.push_back(1.f / (float)m);
    }
  }

  scene_tables(s);
  const int N = s->n;
  {
    Carve size;
    scene_buffers(s, size);
    size_t cb = 0;
    const hipError_t e = fp_pool_get(s->device, size.off, &s->arena, &cb);
    if (e != hipSuccess) return fp_fail(s, FP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    s->device_bytes += (int64_t)cb;
    FP_TRY(s, hipMemsetAsync(s->arena, 0, size.off, s->stream));
    Carve place;
    place.base = (char*)s->arena;
    scene_buffers(s, place);
  }
  {
    // phash = "not in the grid" until the first build; an empty solid box
    std::vector<uint32_t> none(N, 0xffffffffu);
    FP_TRY(s, h2d(s, s->phash, none.data(), N * sizeof(uint32_t)));
    const int box[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    FP_TRY(s, h2d(s, s->sbox, box, sizeof(box)));
  }
  std::vector<uint8_t> ones(std::max(2 * C, 1), 1);
  FP_TRY(s, h2d(s, s->pos, hpos.data(), N * sizeof(float4)));
  FP_TRY(s, h2d(s, s->active, hact.data(), N));
  if (C) {
    FP_TRY(s, h2d(s, s->cons_ab, hab.data(), C * sizeof(int2)));
    FP_TRY(s, h2d(s, s->cent, s->h_cent.data(), C * sizeof(int2)));
    FP_TRY(s, h2d(s, s->erest, herest.data(), 2 * C * sizeof(float)));
    FP_TRY(s, h2d(s, s->ealive, ones.data(), 2 * C));
    FP_TRY(s, h2d(s, s->adj, hadj.data(), 2 * C * sizeof(uint32_t)));
  }
  FP_TRY(s, h2d(s, s->adj_off, hoff.data(), (ns + 1) * sizeof(int)));
  if (s->tets_on && s->NT) {
    std::vector<uint8_t> tones(s->NT, 1);
    FP_TRY(s, h2d(s, s->tet_v, htv.data(), htv.size() * sizeof(int4)));
    FP_TRY(s, h2d(s, s->tet_v0, htv0.data(), htv0.size() * sizeof(float)));
    FP_TRY(s, h2d(s, s->tet_scale, hts.data(), hts.size() * sizeof(float)));
    FP_TRY(s, h2d(s, s->tet_alive, tones.data(), tones.size()));
    FP_TRY(s, h2d(s, s->tet_adj_off, htoff.data(), htoff.size() * sizeof(int)));
    FP_TRY(s, h2d(s, s->tet_adj, htadj.data(), htadj.size() * sizeof(int)));
  }
  // prev stays zero until the first integrate (solver.js:43)
  s->has_cube = true;
  return FP_OK;
}

// A pinned host buffer of at least `bytes` for an upload on the solver's
// stream; upload_end() once the copies that read it are queued.
static int upload_begin(fp_solver* s, size_t bytes, char** host) {
  fp_solver::Upload& u = s->upload[s->upload_next];
  if (u.pending) {
    FP_TRY(s, hipEventSynchronize(u.ev));
    u.pending = false;
  }
  if (!u.ev) FP_TRY(s, hipEventCreateWithFlags(&u.ev, hipEventDisableTiming));
  if (bytes > u.cap) {
    if (u.host) (void)hipHostFree(u.host);
    u.host = nullptr;
    u.cap = 0;
    const size_t cap = std::max<size_t>(bytes, (size_t)64 << 10);
    FP_TRY(s, hipHostMalloc(&u.host, cap, 0));
    u.cap = cap;
  }
  *host = (char*)u.host;
  return FP_OK;
}

static int upload_end(fp_solver* s) {
  fp_solver::Upload& u = s->upload[s->upload_next];
  FP_TRY(s, hipEventRecord(u.ev, s->stream));
  u.pending = true;
  s->upload_next ^= 1;
  return FP_OK;
}

static void upload_free(fp_solver* s) {
  for (fp_solver::Upload& u : s->upload) {
    if (u.ev) (void)hipEventSynchronize(u.ev);
    if (u.host) (void)hipHostFree(u.host);
    if (u.ev) (void)hipEventDestroy(u.ev);
    u = fp_solver::Upload();
  }
}

// Device staging for an emit of m particles (or a selection of m - 1 ids),
// grown when needed: m entries of 36 bytes.
static int emit_reserve(fp_solver* s, int m) {
  if (m <= s->emit_cap) return FP_OK;
  FP_TRY(s, hipStreamSynchronize(s->stream));
  dfree(s, s->emit_buf);
  const int cap = std::max(m, 1024);
  const int rc = dalloc(s, &s->emit_buf, ((size_t)cap * 9 + 3) / 4);
  if (rc) return rc;
  s->emit_cap = cap;
  return FP_OK;
}

int fp_emit(fp_solver* s, const double* pos, const double* vel, const double* mass, int32_t n) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_emit: NULL solver");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_emit before fp_add_solid_cube");
  if (n < 0 || (n > 0 && (!pos || !vel))) return fp_fail(s, FP_EINVAL, "fp_emit: bad arguments");
  if (n == 0) return FP_OK;
  if (s->max_fluid == 0) return fp_fail(s, FP_EINVAL, "fp_emit: maxFluid is 0");
  int rc = set_device(s);
  if (rc) return rc;
  // only the last maxFluid emits of a batch survive the ring (solver.js:140-141)
  const int first = n > s->max_fluid ? n - s->max_fluid : 0;
  const int m = n - first;
  if ((rc = emit_reserve(s, m))) return rc;
  // staged in pinned memory: the copies run behind the host, which goes on
  // to queue the next step
  char* h = nullptr;
  if ((rc = upload_begin(s, (size_t)m * (2 * sizeof(float4) + sizeof(int)), &h))) return rc;
  float4* hp = (float4*)h;
  float4* hv = hp + m;
  int* idx = (int*)(hv + m);
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
  // hp, hv and idx are back to back, as the device staging expects
  FP_TRY(s, hipMemcpyAsync(s->emit_buf, hp, (size_t)m * (2 * sizeof(float4) + sizeof(int)), hipMemcpyHostToDevice,
                           s->stream));
  if ((rc = upload_end(s))) return rc;
  // the device staging is reused in stream order: the next emit's copies
  // queue behind this launch
  return fp_launch_emit(s, m);
}

int fp_reset_solid(fp_solver* s) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_reset_solid: NULL solver");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_reset_solid before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
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
    if ((rc = constraints_to_device(s, s->h_snap_rest.data(), ones.data()))) return rc;
  }
  if (s->tets_on && s->NT) {
    std::vector<uint8_t> tones(s->NT, 1);
    FP_TRY(s, h2d(s, s->tet_alive, tones.data(), tones.size()));
  }
  s->h_group.assign(ns, 0);
  s->groups_dirty = false;
  s->groups_pending = false;
  s->groups_from_snap = false;
  s->cut_dirty = false;
  s->grab_round_len.clear();
  return FP_OK;
}

int fp_clear_fluid(fp_solver* s) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_clear_fluid: NULL solver");
  if (!s->has_cube) return FP_OK;
  int rc = set_device(s);
  if (rc) return rc;
  if (s->nf) FP_TRY(s, hipMemsetAsync(s->active + s->ns, 0, s->nf, s->stream));
  return FP_OK;
}

static int step_impl(fp_solver* s, double dt, int32_t substeps, bool wait) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_step: NULL solver");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_step before fp_add_solid_cube");
  if (substeps < 1) return fp_fail(s, FP_EINVAL, "fp_step: substeps < 1");
  int rc = set_device(s);
  if (rc) return rc;
  const double h = dt / substeps;
  if (s->cut_dirty) {
    s->cut_dirty_at_step = true;
    s->cut_dirty = false;
  }
  (void)h;
  if ((rc = fp_group_frame(s->group, dt, substeps))) return fp_fail(s, rc, s->group->err);
  s->steps_pending = true;
  return wait ? settle(s) : FP_OK;
}

int fp_step(fp_solver* s, double dt, int32_t substeps) { return step_impl(s, dt, substeps, true); }
int fp_step_async(fp_solver* s, double dt, int32_t substeps) { return step_impl(s, dt, substeps, false); }

int fp_sync(fp_solver* s) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_sync: NULL solver");
  int rc = set_device(s);
  if (rc) return rc;
  return settle(s);
}

int fp_set_grabs(fp_solver* s, const int32_t* idx, const double* xyz, int32_t n) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_set_grabs: NULL solver");
  if (n < 0 || (n > 0 && (!idx || !xyz))) return fp_fail(s, FP_EINVAL, "fp_set_grabs: bad arguments");
  for (int g = 0; g < n; g++)
    if (idx[g] < 0 || idx[g] >= s->n) return fp_fail(s, FP_EINVAL, "fp_set_grabs: particle index out of range");
  int rc = set_device(s);
  if (rc) return rc;
  s->grab_round_len.clear();
  if (n == 0) return FP_OK;
  // rounds of distinct indices; an index listed k times is pulled in rounds 0..k-1 in list order
  std::vector<int> round_of(n);
  {
    std::vector<std::pair<int, int>> key(n);
    for (int g = 0; g < n; g++) key[g] = {idx[g], g};
    std::sort(key.begin(), key.end());
    for (int g = 0; g < n; g++)
      round_of[key[g].second] = (g > 0 && key[g].first == key[g - 1].first) ? round_of[key[g - 1].second] + 1 : 0;
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
    dfree(s, s->grab_idx);
    dfree(s, s->grab_tgt);
    const int cap = std::max(n, 256);
    if ((rc = dalloc(s, &s->grab_idx, cap)) || (rc = dalloc(s, &s->grab_tgt, cap))) return rc;
    s->grab_cap = cap;
  }
  char* h = nullptr;
  if ((rc = upload_begin(s, (size_t)n * (sizeof(float4) + sizeof(int)), &h))) return rc;
  std::memcpy(h, htgt.data(), n * sizeof(float4));
  std::memcpy(h + n * sizeof(float4), hidx.data(), n * sizeof(int));
  FP_TRY(s, hipMemcpyAsync(s->grab_tgt, h, n * sizeof(float4), hipMemcpyHostToDevice, s->stream));
  FP_TRY(s, hipMemcpyAsync(s->grab_idx, h + n * sizeof(float4), n * sizeof(int), hipMemcpyHostToDevice, s->stream));
  return upload_end(s);
}

int fp_cut(fp_solver* s, const double o[3], const double d1[3], const double d2[3], int32_t* ncut) {
  if (!s || !o || !d1 || !d2) return fp_fail(s, FP_EINVAL, "fp_cut: NULL argument");
  if (ncut) *ncut = 0;
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_cut before fp_add_solid_cube");
  const double nx = d1[1] * d2[2] - d1[2] * d2[1];
  const double ny = d1[2] * d2[0] - d1[0] * d2[2];
  const double nz = d1[0] * d2[1] - d1[1] * d2[0];
  const double nlen = std::sqrt(nx * nx + ny * ny + nz * nz);
  if (nlen < 1e-8 || s->C == 0) return FP_OK;
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  if ((rc = defer_groups(s))) return rc;  // labels due from the flags before the cut
  FP_TRY(s, hipMemsetAsync(s->dflags + DF_CUT, 0, sizeof(int), s->stream));
  if ((rc = fp_launch_cut(s, o, d1, d2, nx, ny, nz))) return rc;
  int cnt = 0;
  FP_TRY(s, d2h(s, &cnt, s->dflags + DF_CUT, sizeof(int)));
  if (s->prof.on) profile_collect(&s->prof);
  if (ncut) *ncut = cnt;
  // the reference marks components dirty; they are relabelled at the end
  // of the next step (solver.js:238-241, :386)
  if (cnt) {
    s->cut_dirty = true;
    s->alive_changed = true;
  }
  return FP_OK;
}

int fp_pick_solid(fp_solver* s, const double o[3], const double d[3], double max_perp, int32_t* id, double* t) {
  if (!s || !o || !d || !id) return fp_fail(s, FP_EINVAL, "fp_pick_solid: NULL argument");
  *id = -1;
  if (t) *t = INFINITY;
  if (!s->has_cube || s->ns == 0) return FP_OK;
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = fp_launch_pick(s, o, d, max_perp * max_perp))) return rc;
  double ht = INFINITY;
  int hi = -1;
  FP_TRY(s, hipMemcpyAsync(&ht, s->dpick_t, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  FP_TRY(s, hipMemcpyAsync(&hi, s->dpick_i, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  if ((rc = settle(s))) return rc;
  *id = hi;
  if (t) *t = hi >= 0 ? ht : INFINITY;
  return FP_OK;
}

int fp_components(fp_solver* s, int32_t* labels, int32_t* count) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_components: NULL solver");
  if (count) *count = 0;
  if (!s->has_cube || s->ns == 0) return FP_OK;
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  if ((rc = fp_launch_components(s))) return rc;
  int nc = 0;
  FP_TRY(s, d2h(s, &nc, s->dflags + DF_NCOMP, sizeof(int)));
  if (count) *count = nc;
  if (labels) FP_TRY(s, d2h(s, labels, s->dgroup, (size_t)s->ns * sizeof(int32_t)));
  return FP_OK;
}

int fp_select_solids(fp_solver* s, const double center[3], double radius, int32_t same_group_as, int32_t* ids,
                     int32_t max_ids, int32_t* n) {
  if (!s || !center || !n || max_ids < 0 || (max_ids > 0 && !ids))
    return fp_fail(s, FP_EINVAL, "fp_select_solids: bad arguments");
  *n = 0;
  if (!s->has_cube || s->ns == 0) return FP_OK;
  if (same_group_as >= s->ns) return fp_fail(s, FP_EINVAL, "fp_select_solids: same_group_as is not a solid");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  if (same_group_as >= 0) {
    if (!s->alive_changed) {
      // the labels the caller would read are the partition of the current
      // alive flags: the device computes it, nothing crosses to the host
      if ((rc = fp_launch_components(s))) return rc;
    } else {
      // a cut or a write since the labels were computed: the caller still
      // sees the labels of before it (solver.js relabels at a step's end)
      if ((rc = materialize_groups(s))) return rc;
      FP_TRY(s, h2d(s, s->dgroup, s->h_group.data(), s->ns * sizeof(int32_t)));
    }
  }
  // the ids and the count land in the solver's emit staging (grown if needed)
  const int cap = std::max(max_ids, 1) + 1;
  if ((rc = emit_reserve(s, cap))) return rc;
  int* out = (int*)s->emit_buf;
  if ((rc = fp_launch_select(s, center, radius * radius, same_group_as, out + 1, max_ids, out))) return rc;
  int total = 0;
  FP_TRY(s, d2h(s, &total, out, sizeof(int)));
  *n = total;
  const int k = std::min(total, max_ids);
  if (k > 0) FP_TRY(s, d2h(s, ids, out + 1, k * sizeof(int32_t)));
  return FP_OK;
}

int fp_fluid_count(fp_solver* s, int32_t* count) {
  if (!s || !count) return fp_fail(s, FP_EINVAL, "fp_fluid_count: NULL argument");
  *count = 0;
  if (!s->has_cube || s->nf == 0) return FP_OK;
  int rc = set_device(s);
  if (rc) return rc;
  FP_TRY(s, hipMemsetAsync(s->dflags + DF_SCRATCH, 0, sizeof(int), s->stream));
  if ((rc = fp_launch_count_active(s))) return rc;
  int c = 0;
  FP_TRY(s, hipMemcpyAsync(&c, s->dflags + DF_SCRATCH, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  if ((rc = settle(s))) return rc;
  *count = c;
  return FP_OK;
}

int fp_read_state(fp_solver* s, float* pos, float* prev, float* vel, float* inv_mass, uint8_t* active,
                  int32_t* group) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_read_state: NULL solver");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_read_state before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
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
  if (group) {
    if ((rc = materialize_groups(s))) return rc;
    std::memcpy(group, s->h_group.data(), s->ns * sizeof(int32_t));
  }
  return FP_OK;
}

int fp_write_state(fp_solver* s, const float* pos, const float* prev, const float* vel, const float* inv_mass,
                   const uint8_t* active) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_write_state: NULL solver");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_write_state before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
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
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_read_constraints: NULL solver");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_read_constraints before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  if (a) std::memcpy(a, s->h_cons_a.data(), s->C * sizeof(int32_t));
  if (b) std::memcpy(b, s->h_cons_b.data(), s->C * sizeof(int32_t));
  return constraints_to_host(s, rest, alive);
}

int fp_write_constraints(fp_solver* s, const float* rest, const uint8_t* alive) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_write_constraints: NULL solver");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_write_constraints before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  if ((rc = materialize_groups(s))) return rc;
  if (alive) s->alive_changed = true;
  return constraints_to_device(s, rest, alive);
}

int fp_read_tets(fp_solver* s, int32_t* verts, uint8_t* alive, int32_t* num_tets) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_read_tets: NULL solver");
  if (!num_tets) return fp_fail(s, FP_EINVAL, "fp_read_tets: NULL num_tets");
  const int NT = s->tets_on ? s->NT : 0;
  *num_tets = NT;
  if (NT == 0 || (!verts && !alive)) return FP_OK;
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  if (verts) FP_TRY(s, d2h(s, verts, s->tet_v, (size_t)NT * sizeof(int4)));
  if (alive) FP_TRY(s, d2h(s, alive, s->tet_alive, (size_t)NT));
  return FP_OK;
}

int fp_build_grid(fp_solver* s) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_build_grid: NULL solver");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_build_grid before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = fp_launch_ref_grid(s, true))) return rc;
  return settle(s);
}

int fp_read_grid(fp_solver* s, int32_t* cell_start, int32_t* cell_entries) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_read_grid: NULL solver");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_read_grid before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  if (!s->ref_grid_valid && (rc = fp_launch_ref_grid(s, false))) return rc;
  if ((rc = settle(s))) return rc;
  if (cell_start) FP_TRY(s, d2h(s, cell_start, s->cell_start, ((size_t)s->T + 1) * sizeof(int)));
  if (cell_entries) FP_TRY(s, d2h(s, cell_entries, s->cell_entries, (size_t)s->n * sizeof(int)));
  return FP_OK;
}

int fp_read_cells(fp_solver* s, int32_t which, double* desc, int32_t* start, int32_t* ids, float* pos) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_read_cells: NULL solver");
  if (which != 0 && which != 1) return fp_fail(s, FP_EINVAL, "fp_read_cells: which must be 0 or 1");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_read_cells before fp_add_solid_cube");
  if (!s->cells_built) return fp_fail(s, FP_ESTATE, "fp_read_cells before the first substep");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  const size_t M = s->cg.M;
  const int* dstart = which ? s->lstart_f : s->lstart;
  int32_t slots = 0;
  FP_TRY(s, d2h(s, &slots, dstart + M, sizeof(int32_t)));
  if (desc) {
    desc[0] = s->cg.inv_cell;
    desc[1] = (double)s->cg.M;
    desc[2] = (double)s->cg.mx;
    desc[3] = (double)s->cg.my;
    desc[4] = (double)slots;
  }
  if (start) FP_TRY(s, d2h(s, start, dstart, (M + 1) * sizeof(int32_t)));
  if (ids && !which) FP_TRY(s, d2h(s, ids, s->sid, (size_t)s->n * sizeof(int32_t)));
  if (which && (ids || pos)) {
    // the fluid order keeps each slot's particle in its position's w (the
    // index below PF_SHIFT, flags above)
    std::vector<float4> p2(s->n);
    FP_TRY(s, d2h(s, p2.data(), s->spos2, (size_t)s->n * sizeof(float4)));
    if (ids)
      for (int k = 0; k < s->n; k++) {
        uint32_t w;
        std::memcpy(&w, &p2[k].w, 4);
        ids[k] = (int32_t)(w & SID_MASK);
      }
    if (pos) std::memcpy(pos, p2.data(), (size_t)s->n * sizeof(float4));
  } else if (pos) {
    FP_TRY(s, d2h(s, pos, s->spos, (size_t)s->n * sizeof(float4)));
  }
  return FP_OK;
}

int fp_read_density(fp_solver* s, float* rho, float* rho_near) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_read_density: NULL solver");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_read_density before fp_add_solid_cube");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  std::vector<float2> tmp(s->n);
  FP_TRY(s, d2h(s, tmp.data(), s->rho, s->n * sizeof(float2)));
  for (int i = 0; i < s->n; i++) {
    if (rho) rho[i] = tmp[i].x;
    if (rho_near) rho_near[i] = tmp[i].y;
  }
  return FP_OK;
}

int fp_profile_enable(fp_solver* s, int32_t enable) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_profile_enable: NULL solver");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  s->prof.on = enable != 0;
  for (int k = 0; k < P_NUM; k++) {
    s->prof.ms[k] = 0;
    s->prof.n[k] = 0;
  }
  return FP_OK;
}

int fp_profile_read(fp_solver* s, fp_pass_stat* out, int32_t max, int32_t* n) {
  if (!s || !out || !n) return fp_fail(s, FP_EINVAL, "fp_profile_read: NULL argument");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  int m = 0;
  for (int k = 0; k < P_NUM && m < max; k++) {
    if (!s->prof.n[k]) continue;
    std::memset(out[m].name, 0, sizeof(out[m].name));
    std::strncpy(out[m].name, kPassNames[k], sizeof(out[m].name) - 1);
    out[m].launches = s->prof.n[k];
    out[m].total_ms = s->prof.ms[k];
    m++;
  }
  *n = m;
  return FP_OK;
}


// ---- batches: many scenes, one launch per pass -------------------------------

struct fp_batch {
  fp_group g;
  Prof prof;
  std::string err;
  int64_t frames = 0;
  // position snapshots: a device gather buffer and two pinned host slots
  float* d_snap = nullptr;
  size_t d_snap_cap = 0;
  float* h_snap[2] = {nullptr, nullptr};
  size_t h_snap_cap[2] = {0, 0};
  int snap_slot = 0;
  std::vector<int64_t> snap_off, snap_off_out[2];
};

static int batch_fail(fp_batch* b, int code, const std::string& msg) {
  if (b)
    b->err = msg;
  else
    g_last_error = msg;
  return code;
}

int fp_batch_create(int32_t device, fp_batch** out) {
  if (!out) return batch_fail(nullptr, FP_EINVAL, "fp_batch_create: out is NULL");
  *out = nullptr;
  if (device < 0 && hipGetDevice(&device) != hipSuccess) device = 0;
  fp_batch* b = new (std::nothrow) fp_batch();
  if (!b) return batch_fail(nullptr, FP_ENOMEM, "fp_batch_create: out of host memory");
  b->g.device = device;
  b->g.prof = &b->prof;
  b->g.shape_tune = FP_SHAPE_TUNE != 0;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->g.stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_last_error = std::string("fp_batch_create: ") + hipGetErrorString(e);
    delete b;
    return FP_EDEVICE;
  }
  *out = b;
  return FP_OK;
}

const char* fp_batch_last_error(const fp_batch* b) { return b ? b->err.c_str() : g_last_error.c_str(); }

// The scene leaves the batch: back to its own stream.
static void batch_detach(fp_batch* b, fp_solver* s) {
  (void)hipStreamSynchronize(b->g.stream);
  s->stream = s->own_stream;
  s->group->stream = s->own_stream;
  s->batch = nullptr;
}

int fp_batch_destroy(fp_batch* b) {
  if (!b) return FP_OK;
  (void)hipSetDevice(b->g.device);
  for (fp_solver* s : b->g.scenes) batch_detach(b, s);
  b->g.scenes.clear();
  fp_group_free(&b->g);
  prof_free(&b->prof);
  if (b->d_snap) fp_pool_put(b->g.device, b->d_snap);
  for (float* h : b->h_snap)
    if (h) (void)hipHostFree(h);
  if (b->g.stream) (void)hipStreamDestroy(b->g.stream);
  delete b;
  return FP_OK;
}

int fp_batch_add(fp_batch* b, fp_solver* s) {
  if (!b || !s) return batch_fail(b, FP_EINVAL, "fp_batch_add: NULL argument");
  if (!s->has_cube) return batch_fail(b, FP_ESTATE, "fp_batch_add: the scene has no cube yet");
  if (s->batch) return batch_fail(b, FP_ESTATE, "fp_batch_add: the scene is in a batch already");
  if (s->device != b->g.device) return batch_fail(b, FP_EINVAL, "fp_batch_add: the scene is on another device");
  (void)hipSetDevice(b->g.device);
  // work queued on the scene's own stream finishes before the batch's
  if (hipStreamSynchronize(s->stream) != hipSuccess) return batch_fail(b, FP_EDEVICE, "fp_batch_add: stream sync");
  s->stream = b->g.stream;
  s->group->stream = b->g.stream;
  s->batch = b;
  b->g.scenes.push_back(s);
  b->g.dirty = true;
  return FP_OK;
}

int fp_batch_remove(fp_batch* b, fp_solver* s) {
  if (!b || !s) return batch_fail(b, FP_EINVAL, "fp_batch_remove: NULL argument");
  auto it = std::find(b->g.scenes.begin(), b->g.scenes.end(), s);
  if (it == b->g.scenes.end()) return batch_fail(b, FP_EINVAL, "fp_batch_remove: the scene is not in this batch");
  (void)hipSetDevice(b->g.device);
  batch_detach(b, s);
  b->g.scenes.erase(it);
  b->g.dirty = true;
  return FP_OK;
}

static int batch_step_impl(fp_batch* b, double dt, int32_t substeps, bool wait) {
  if (!b) return batch_fail(nullptr, FP_EINVAL, "fp_batch_step: NULL batch");
  if (substeps < 1) return batch_fail(b, FP_EINVAL, "fp_batch_step: substeps < 1");
  FP_GTRY(&b->g, hipSetDevice(b->g.device));
  for (fp_solver* s : b->g.scenes)
    if (s->cut_dirty) {
      s->cut_dirty_at_step = true;
      s->cut_dirty = false;
    }
  int rc = fp_group_frame(&b->g, dt, substeps);
  if (rc) return batch_fail(b, rc, b->g.err);
  b->frames++;
  for (fp_solver* s : b->g.scenes) s->steps_pending = true;
  if (!wait) return FP_OK;
  return fp_batch_sync(b);
}

int fp_batch_step(fp_batch* b, double dt, int32_t substeps) { return batch_step_impl(b, dt, substeps, true); }
int fp_batch_step_async(fp_batch* b, double dt, int32_t substeps) { return batch_step_impl(b, dt, substeps, false); }

int fp_batch_sync(fp_batch* b) {
  if (!b) return batch_fail(nullptr, FP_EINVAL, "fp_batch_sync: NULL batch");
  FP_GTRY(&b->g, hipSetDevice(b->g.device));
  FP_GTRY(&b->g, hipStreamSynchronize(b->g.stream));
  if (b->prof.on) profile_collect(&b->prof);
  for (fp_solver* s : b->g.scenes) {
    const int rc = settle(s);
    if (rc) return batch_fail(b, rc, s->err);
  }
  return FP_OK;
}

int fp_batch_set_graph(fp_batch* b, int32_t enable) {
  if (!b) return batch_fail(nullptr, FP_EINVAL, "fp_batch_set_graph: NULL batch");
  b->g.use_graph = enable != 0;
  return FP_OK;
}

int fp_batch_set_contact_lanes(fp_batch* b, int32_t lanes) {
  if (!b) return batch_fail(nullptr, FP_EINVAL, "fp_batch_set_contact_lanes: NULL batch");
  if (lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32)
    return batch_fail(b, FP_EINVAL, "fp_batch_set_contact_lanes: lanes must be 0, 4, 8, 16 or 32");
  b->g.clanes_pinned = lanes != 0;
  b->g.clanes_best = lanes;
  b->g.tune_k = 0;
  b->g.tune_pending = false;
  return FP_OK;
}

int fp_batch_get_info(const fp_batch* b, fp_batch_info* out) {
  if (!b || !out) return FP_EINVAL;
  std::memset(out, 0, sizeof(*out));
  out->scenes = (int32_t)b->g.scenes.size();
  out->graph = b->g.gexec != nullptr ? 1 : 0;
  for (const fp_solver* s : b->g.scenes) {
    out->capacity += s->n;
    out->device_bytes += s->device_bytes;
  }
  out->frames = b->frames;
  out->graph_launches = b->g.graph_launches;
  out->contact_lanes = b->g.clanes_best;
  out->graph_captures = (int32_t)b->g.graph_captures;
  out->density_stage = b->g.stage_on;
  out->xcd_run = b->g.xcd_run;
  out->shape_tuned = b->g.shape_state == 2 ? 1 : 0;
  return FP_OK;
}

int fp_batch_profile_enable(fp_batch* b, int32_t enable) {
  if (!b) return batch_fail(nullptr, FP_EINVAL, "fp_batch_profile_enable: NULL batch");
  const int rc = fp_batch_sync(b);
  if (rc) return rc;
  b->prof.on = enable != 0;
  for (int k = 0; k < P_NUM; k++) {
    b->prof.ms[k] = 0;
    b->prof.n[k] = 0;
  }
  return FP_OK;
}

int fp_batch_profile_read(fp_batch* b, fp_pass_stat* out, int32_t max, int32_t* n) {
  if (!b || !out || !n) return batch_fail(b, FP_EINVAL, "fp_batch_profile_read: NULL argument");
  const int rc = fp_batch_sync(b);
  if (rc) return rc;
  int m = 0;
  for (int k = 0; k < P_NUM && m < max; k++) {
    if (!b->prof.n[k]) continue;
    std::memset(out[m].name, 0, sizeof(out[m].name));
    std::strncpy(out[m].name, kPassNames[k], sizeof(out[m].name) - 1);
    out[m].launches = b->prof.n[k];
    out[m].total_ms = b->prof.ms[k];
    m++;
  }
  *n = m;
  return FP_OK;
}

}  // extern "C"

int fp_scene_bytes(const fp_params* p, int32_t max_fluid, int32_t cube_n, int64_t* bytes) {
  if (!p || !bytes || !cube_fits(cube_n, max_fluid)) return FP_EINVAL;
  fp_solver tmp;
  tmp.P = *p;
  tmp.max_fluid = max_fluid;
  scene_dims(&tmp, cube_n);
  tmp.C = (int)cube_constraints(cube_n);
  scene_tables(&tmp);
  Carve size;
  scene_buffers(&tmp, size);
  int64_t total = (int64_t)fp_pool_class(size.off);
  if (max_fluid > 0) {
    // the emit staging, at its largest (one emit of max_fluid; emit_reserve)
    const size_t cap = (size_t)std::max(max_fluid, 1024);
    total += (int64_t)fp_pool_class((cap * 9 + 3) / 4 * sizeof(float4));
  }
  *bytes = total;
  return FP_OK;
}

int fp_batch_snapshot(fp_batch* b, const float** pos, const int64_t** offsets, int32_t* scenes) {
  if (!b || !pos || !offsets || !scenes) return batch_fail(b, FP_EINVAL, "fp_batch_snapshot: NULL argument");
  fp_group* g = &b->g;
  FP_GTRY(g, hipSetDevice(g->device));
  const int ns = (int)g->scenes.size();
  b->snap_off.assign(ns + 1, 0);
  for (int k = 0; k < ns; k++) b->snap_off[k + 1] = b->snap_off[k] + 3 * (int64_t)g->scenes[k]->n;
  const int64_t total = b->snap_off[ns];
  // the scene table holds each scene's offset: rebuild it if the members changed
  int rc = fp_group_prepare(g, g->built_h > 0 ? g->built_h : 1.0 / 300.0);
  if (rc) return batch_fail(b, rc, g->err);
  const int slot = b->snap_slot;
  if ((size_t)total > b->d_snap_cap || (size_t)total > b->h_snap_cap[slot]) {
    FP_GTRY(g, hipStreamSynchronize(g->stream));
    if ((size_t)total > b->d_snap_cap) {
      if (b->d_snap) fp_pool_put(g->device, b->d_snap);
      b->d_snap = nullptr;
      b->d_snap_cap = 0;
      FP_GTRY(g, fp_pool_get(g->device, std::max<int64_t>(total, 1) * sizeof(float), (void**)&b->d_snap, nullptr));
      b->d_snap_cap = (size_t)total;
    }
    if ((size_t)total > b->h_snap_cap[slot]) {
      if (b->h_snap[slot]) (void)hipHostFree(b->h_snap[slot]);
      b->h_snap[slot] = nullptr;
      b->h_snap_cap[slot] = 0;
      const size_t cap = (size_t)total + (size_t)total / 4;  // room to grow
      FP_GTRY(g, hipHostMalloc((void**)&b->h_snap[slot], std::max<size_t>(cap, 1) * sizeof(float), 0));
      b->h_snap_cap[slot] = cap;
    }
  }
  if (total > 0) {
    rc = fp_group_snapshot(g, b->d_snap);
    if (rc) return batch_fail(b, rc, g->err);
    FP_GTRY(g, hipMemcpyAsync(b->h_snap[slot], b->d_snap, total * sizeof(float), hipMemcpyDeviceToHost, g->stream));
  }
  FP_GTRY(g, hipStreamSynchronize(g->stream));
  *pos = b->h_snap[slot];
  b->snap_off_out[slot] = b->snap_off;
  *offsets = b->snap_off_out[slot].data();
  *scenes = ns;
  b->snap_slot ^= 1;
  return FP_OK;
}


// ---- tuning cache file ----------------------------------------------------------
//
// {"version": 1, "contact_lanes": {"<signature>": lanes, ...},
//  "launch_shape": {"<signature>": candidate, ...}}: the two tables of
// fp_step.hip's tuning caches, signatures as decimal strings.

namespace {

void put_table(std::string& out, const char* name, const std::vector<std::pair<uint64_t, int>>& t, bool last) {
  out += "  \"";
  out += name;
  out += "\": {";
  for (size_t k = 0; k < t.size(); k++) {
    out += k ? ", \"" : "\"";
    out += std::to_string(t[k].first) + "\": " + std::to_string(t[k].second);
  }
  out += last ? "}\n" : "},\n";
}

// A reader of the format above: objects, decimal strings as keys, integers.
struct JsonIn {
  const char* p;
  const char* end;
  void ws() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
  }
  bool lit(char c) {
    ws();
    if (p < end && *p == c) {
      p++;
      return true;
    }
    return false;
  }
  bool str(std::string& s) {
    ws();
    if (p >= end || *p != '"') return false;
    const char* q = ++p;
    while (p < end && *p != '"') {
      if (*p == '\\') return false;  // no escapes in this format
      p++;
    }
    if (p >= end) return false;
    s.assign(q, p++);
    return true;
  }
  bool integer(int64_t& v) {
    ws();
    const char* q = p;
    if (p < end && *p == '-') p++;
    while (p < end && *p >= '0' && *p <= '9') p++;
    if (p == q || (p == q + 1 && *q == '-') || p - q > 19) return false;
    v = std::strtoll(std::string(q, p).c_str(), nullptr, 10);
    return true;
  }
  // {"<u64>": int, ...}
  bool table(std::vector<std::pair<uint64_t, int>>& t) {
    if (!lit('{')) return false;
    if (lit('}')) return true;
    do {
      std::string k;
      int64_t v;
      if (!str(k) || k.empty() || k.size() > 20 || !lit(':') || !integer(v)) return false;
      for (char c : k)
        if (c < '0' || c > '9') return false;
      t.emplace_back(std::strtoull(k.c_str(), nullptr, 10), (int)v);
      if (v < INT_MIN || v > INT_MAX) return false;
    } while (lit(','));
    return lit('}');
  }
};

}  // namespace

int fp_tune_cache_save(const char* path) {
  if (!path) return fp_fail(nullptr, FP_EINVAL, "fp_tune_cache_save: NULL path");
  std::string out = "{\n  \"version\": 1,\n";
  put_table(out, "contact_lanes", fp_tune_cache_entries(true), false);
  put_table(out, "launch_shape", fp_tune_cache_entries(false), true);
  out += "}\n";
  FILE* f = std::fopen(path, "wb");
  if (!f) return fp_fail(nullptr, FP_EIO, std::string("fp_tune_cache_save: cannot open ") + path);
  const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  if (std::fclose(f) != 0 || !ok) return fp_fail(nullptr, FP_EIO, std::string("fp_tune_cache_save: cannot write ") + path);
  return FP_OK;
}

int fp_tune_cache_load(const char* path) {
  if (!path) return fp_fail(nullptr, FP_EINVAL, "fp_tune_cache_load: NULL path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fp_fail(nullptr, FP_EIO, std::string("fp_tune_cache_load: cannot open ") + path);
  std::string text;
  char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0 && text.size() < (1u << 24)) text.append(buf, n);
  const bool rerr = std::ferror(f) != 0;
  std::fclose(f);
  if (rerr) return fp_fail(nullptr, FP_EIO, std::string("fp_tune_cache_load: cannot read ") + path);
  JsonIn in{text.data(), text.data() + text.size()};
  std::vector<std::pair<uint64_t, int>> lanes, shape;
  bool ok = in.lit('{'), seen_v = false;
  while (ok) {
    std::string key;
    ok = in.str(key) && in.lit(':');
    if (!ok) break;
    if (key == "version") {
      int64_t v = 0;
      ok = in.integer(v) && v == 1;
      seen_v = true;
    } else if (key == "contact_lanes") {
      ok = in.table(lanes);
    } else if (key == "launch_shape") {
      ok = in.table(shape);
    } else {
      ok = false;
    }
    if (!ok || !in.lit(',')) break;
  }
  ok = ok && seen_v && in.lit('}');
  in.ws();
  ok = ok && in.p == in.end;
  for (const auto& e : lanes) ok = ok && (e.second == 4 || e.second == 8 || e.second == 16 || e.second == 32);
  for (const auto& e : shape) ok = ok && e.second >= 0 && e.second < 6;
  if (!ok) return fp_fail(nullptr, FP_EINVAL, std::string("fp_tune_cache_load: not a tuning cache of version 1: ") + path);
  fp_tune_cache_merge(lanes, shape);
  return FP_OK;
}

int fp_tune_cache_reset(void) {
  fp_tune_caches_clear();
  return FP_OK;
}

// ---- checkpoints and state blobs -------------------------------------------------
//
// The state that decides a scene's future: positions, previous positions,
// velocities (w: inverse mass), active flags, the last densities, each CSR
// entry's rest length and alive flag, the hose ring position, the substep
// count and the component labels with their pending relabel.  The grid, the
// sorted copies and the pair list are rebuilt every substep.  (The reference's
// GPU twin snapshots its buffers the same way, src/gpu/sim.js:494-547.)

namespace {

template <class F>
void state_arrays(fp_solver* s, F&& f) {
  const size_t N = (size_t)s->n, E = 2 * (size_t)s->C;
  f((void*)s->pos, N * sizeof(float4));
  f((void*)s->prev, N * sizeof(float4));
  f((void*)s->vel, N * sizeof(float4));
  f((void*)s->rho, N * sizeof(float2));
  f((void*)s->active, N);
  f((void*)s->erest, E * sizeof(float));
  f((void*)s->ealive, E);
  f((void*)s->tet_alive, s->tets_on ? (size_t)s->NT : 0);
}

size_t state_array_bytes(fp_solver* s) {
  size_t total = 0;
  state_arrays(s, [&](void*, size_t b) { total += (b + 255) & ~(size_t)255; });
  return total;
}

const int kMaxCheckpoints = 8;

struct StateHeader {
  char magic[8];  // "FPSTATE"
  uint32_t version;
  int32_t capacity, solid_count, max_fluid, num_constraints, fluid_next;
  int64_t substeps_done;
  uint32_t flags;  // 1: relabel pending, 2: a cut awaits the next step
  uint32_t reserved;
};

}  // namespace

int fp_checkpoint_save(fp_solver* s, int32_t slot) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_checkpoint_save: NULL solver");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_checkpoint_save before fp_add_solid_cube");
  if (slot < 0 || slot >= kMaxCheckpoints) return fp_fail(s, FP_EINVAL, "fp_checkpoint_save: slot out of range [0, 8)");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  if (s->groups_from_snap && (rc = materialize_groups(s))) return rc;  // saved state keeps no deferred copy
  if ((int)s->checkpoints.size() <= slot) s->checkpoints.resize(slot + 1);
  fp_solver::Checkpoint& cp = s->checkpoints[slot];
  const size_t need = state_array_bytes(s);
  if (!cp.block) {
    size_t cb = 0;
    const hipError_t e = fp_pool_get(s->device, need, &cp.block, &cb);
    if (e != hipSuccess) return fp_fail(s, FP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    cp.bytes = need;
    s->device_bytes += (int64_t)cb;
  }
  size_t off = 0;
  hipError_t e = hipSuccess;
  state_arrays(s, [&](void* p, size_t b) {
    if (e == hipSuccess && b) e = hipMemcpyAsync((char*)cp.block + off, p, b, hipMemcpyDeviceToDevice, s->stream);
    off += (b + 255) & ~(size_t)255;
  });
  FP_TRY(s, e);
  cp.fluid_next = s->fluid_next;
  cp.substeps_done = s->substeps_done;
  cp.group = s->h_group;
  cp.groups_pending = s->groups_pending;
  cp.cut_dirty = s->cut_dirty;
  FP_TRY(s, hipStreamSynchronize(s->stream));
  return FP_OK;
}

// A restored or imported state starts with the non-finite flag down (the
// stream is idle after settle: no kernel writes either word now).
static int clear_nonfinite(fp_solver* s) {
  FP_TRY(s, hipMemsetAsync(s->dflags + DF_NONFINITE, 0, sizeof(int), s->stream));
  FP_TRY(s, hipStreamSynchronize(s->stream));
  s->tear_host[1] = 0;
  s->nonfinite = false;
  return FP_OK;
}

int fp_checkpoint_restore(fp_solver* s, int32_t slot) {
  if (!s) return fp_fail(nullptr, FP_EINVAL, "fp_checkpoint_restore: NULL solver");
  if (slot < 0 || slot >= (int)s->checkpoints.size() || !s->checkpoints[slot].block)
    return fp_fail(s, FP_EINVAL, "fp_checkpoint_restore: no checkpoint in this slot");
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = settle(s))) return rc;
  const fp_solver::Checkpoint& cp = s->checkpoints[slot];
  size_t off = 0;
  hipError_t e = hipSuccess;
  state_arrays(s, [&](void* p, size_t b) {
    if (e == hipSuccess && b) e = hipMemcpyAsync(p, (char*)cp.block + off, b, hipMemcpyDeviceToDevice, s->stream);
    off += (b + 255) & ~(size_t)255;
  });
  FP_TRY(s, e);
  FP_TRY(s, hipStreamSynchronize(s->stream));
  s->fluid_next = cp.fluid_next;
  s->substeps_done = cp.substeps_done;
  s->h_group = cp.group;
  s->groups_pending = cp.groups_pending;
  s->groups_from_snap = false;
  s->groups_dirty = false;
  s->cut_dirty = cp.cut_dirty;
  s->alive_changed = true;  // conservative: the host labels decide selection
  return clear_nonfinite(s);
}

int fp_state_size(fp_solver* s, int64_t* bytes) {
  if (!s || !bytes) return fp_fail(s, FP_EINVAL, "fp_state_size: NULL argument");
  if (!s->has_cube) return fp_fail(s, FP_ESTATE, "fp_state_size before fp_add_solid_cube");
  int64_t total = (int64_t)sizeof(StateHeader);
  state_arrays(s, [&](void*, size_t b) { total += (int64_t)b; });
  *bytes = total + 4 * (int64_t)s->ns;
  return FP_OK;
}

int fp_state_export(fp_solver* s, void* buf, int64_t bytes) {
  int64_t need = 0;
  int rc = fp_state_size(s, &need);
  if (rc) return rc;
  if (!buf || bytes < need) return fp_fail(s, FP_EINVAL, "fp_state_export: buffer smaller than fp_state_size");
  if ((rc = set_device(s))) return rc;
  if ((rc = settle(s))) return rc;
  if (s->groups_from_snap && (rc = materialize_groups(s))) return rc;  // the blob keeps no deferred copy
  StateHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, "FPSTATE", 8);
  h.version = 1;
  h.capacity = s->n;
  h.solid_count = s->ns;
  h.max_fluid = s->max_fluid;
  h.num_constraints = s->C;
  h.fluid_next = s->fluid_next;
  h.substeps_done = s->substeps_done;
  h.flags = (s->groups_pending ? 1u : 0u) | (s->cut_dirty ? 2u : 0u);
  char* out = (char*)buf;
  std::memcpy(out, &h, sizeof(h));
  size_t off = sizeof(h);
  hipError_t e = hipSuccess;
  state_arrays(s, [&](void* p, size_t b) {
    if (e == hipSuccess && b) e = d2h(s, out + off, p, b);
    off += b;
  });
  FP_TRY(s, e);
  if (s->ns) std::memcpy(out + off, s->h_group.data(), 4 * (size_t)s->ns);
  return FP_OK;
}

int fp_state_import(fp_solver* s, const void* buf, int64_t bytes) {
  int64_t need = 0;
  int rc = fp_state_size(s, &need);
  if (rc) return rc;
  if (!buf || bytes < (int64_t)sizeof(StateHeader)) return fp_fail(s, FP_EINVAL, "fp_state_import: no state blob");
  StateHeader h;
  std::memcpy(&h, buf, sizeof(h));
  if (std::memcmp(h.magic, "FPSTATE", 8) != 0 || h.version != 1)
    return fp_fail(s, FP_EINVAL, "fp_state_import: not a state blob of version 1");
  if (h.capacity != s->n || h.solid_count != s->ns || h.max_fluid != s->max_fluid || h.num_constraints != s->C)
    return fp_fail(s, FP_EINVAL, "fp_state_import: the blob is of a scene of another shape");
  if (bytes < need) return fp_fail(s, FP_EINVAL, "fp_state_import: blob shorter than its header says");
  // the ring cursor addresses a fluid slot (fp_emit writes at solids + cursor)
  if (h.fluid_next < 0 || h.fluid_next >= std::max(s->max_fluid, 1) || h.substeps_done < 0)
    return fp_fail(s, FP_EINVAL, "fp_state_import: ring cursor or substep count out of range");
  if ((rc = set_device(s))) return rc;
  if ((rc = settle(s))) return rc;
  const char* in = (const char*)buf;
  size_t off = sizeof(h);
  hipError_t e = hipSuccess;
  std::vector<uint8_t> flags;
  state_arrays(s, [&](void* p, size_t b) {
    if (e == hipSuccess && b) {
      if (p == (void*)s->active || p == (void*)s->ealive || p == (void*)s->tet_alive) {
        // flags enter the device as 0 / 1 only
        flags.assign(in + off, in + off + b);
        for (uint8_t& f : flags) f = f ? 1 : 0;
        e = h2d(s, p, flags.data(), b);
      } else {
        e = h2d(s, p, in + off, b);
      }
    }
    off += b;
  });
  FP_TRY(s, e);
  s->h_group.assign(s->ns, 0);
  if (s->ns) std::memcpy(s->h_group.data(), in + off, 4 * (size_t)s->ns);
  s->fluid_next = h.fluid_next;
  s->substeps_done = h.substeps_done;
  s->groups_pending = (h.flags & 1u) != 0;
  s->groups_from_snap = false;
  s->groups_dirty = false;
  s->cut_dirty = (h.flags & 2u) != 0;
  s->alive_changed = true;  // conservative: the host labels decide selection
  return clear_nonfinite(s);
}

int fp_nonfinite(const fp_solver* s, int32_t* out) {
  if (!s || !out) return fp_fail(const_cast<fp_solver*>(s), FP_EINVAL, "fp_nonfinite: NULL argument");
  *out = s->nonfinite ? 1 : 0;
  return FP_OK;
}
