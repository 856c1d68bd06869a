// fp_internal.h -- solver state shared by the translation units of
// libfluidphysics.so (host side), the per-pass parameter blocks passed to
// the kernels by value, and the launchers each unit exports.
//
// Layout in HBM (structure of arrays, 16-byte vectors; N = capacity,
// Ns = solids, C = constraints, T = reference hash table size, M = cell
// table size):
//   pos, prev, vel  float4[N] particle-index order (pos.w = inverse mass)
//   active u8[N]; rho float2[N]; phash u32[N] reference bucket at grid time
//   lcount i32[M] (zero between builds), lstart i32[M+1], ltmp/lkey/lslot [N],
//   lnsol i32[M] (solids of each non-empty cell of the grid order)
//   grid order (every active particle, sorted by its cell after integrate;
//   read by contacts): sid i32 (particle of the slot), spos float4 (xyz +
//     inverse mass), sphash u32 (bucket)
//   contacts -> fluid order staging, per grid slot: tpos float4 (post-contact
//     xyz, w = bits of particle index | PF_* flags << 30), tkey u32 (cell key
//     of the post-contact position, ~0 for slots without fluid), tslot u32
//   fluid order (active fluid sorted by its post-contact cell; read by
//   density, displacement, XSPH): fsid, fslot i32[N] (inverse), spos2
//     float4 (= tpos), sphash2 u32, sP float2 (P, Pnear), svtmp float4
//     (velocity before XSPH)
//   nbr u32[(K+L)*N] (L = the scene's fluid lanes per slot, each with K/L
//     entries: entry k of lane v at k*N*L + v; row K/L is a sink), ncount i32[N*L]: the
//     pairs the density pass found, replayed by displacement and XSPH
//   count/cell_start/entries_tmp/cell_entries: the reference's hash grid,
//     rebuilt from phash only when read
//   cons_ab int2[C] endpoints; cent int2[C] the constraint's two CSR entries;
//   adj_off i32[Ns+1]; adj int2[2C] = (other endpoint, bits of
//     1/max(deg_a, deg_b)); erest f32[2C], ealive u8[2C]: each CSR entry's
//     copy of its constraint's rest length and alive flag (the two copies are
//     updated identically, in place, each by its own endpoint's lane);
//     spos_c[2] float4[Ns] between iterations (spos_c[1] first holds the
//     solid contacts' result, the first iteration's snapshot)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/fluidphysics.h"
#include "fp_kernels.h"

#ifndef BLOCK
#define BLOCK 256
#endif
#ifndef SCAN_ITEMS
#define SCAN_ITEMS 4  // cells per thread of a scan tile
#endif
#define SCAN_TILE (BLOCK * SCAN_ITEMS)
#define PF_MOVED 1u  // fluid particle changed cell in the contact pass
#define PF_DUP 2u    // its current cell's 27 reference probes repeat a bucket
#define SID_MASK 0x3fffffffu
#define PF_SHIFT 30
#define NO_KEY 0xffffffffu
// Lanes per fluid particle in density, displacement and XSPH: a property of
// the scene (fp_params.fluid_lanes; 0 = FLUID_LANES_BIG for scenes of at
// least FLUID_LANES_AUTO fluid slots, else FLUID_LANES_SMALL).  It orders the
// pair list and the lane sums, so it is part of the scene's arithmetic.
#ifndef FLUID_LANES_SMALL
#define FLUID_LANES_SMALL 2
#endif
#ifndef FLUID_LANES_BIG
#define FLUID_LANES_BIG 4
#endif
#define FLUID_LANES_AUTO 131072

// XCD run length of the group kernels' work lists (the default; a batch may
// tune it) and whether batches tune their launch shape
#ifndef FP_XCD_CHUNK
#define FP_XCD_CHUNK 16
#endif
#ifndef FP_SHAPE_TUNE
#define FP_SHAPE_TUNE 1  // batches tune their launch shape over captured frames
#endif

// device flag words (dflags)
#define DF_UNUSED0 0     // (the tear flag lives in host-mapped memory: fp_solver::tear_host)
#define DF_CUT 1         // knife count
#define DF_SCRATCH 2     // fluid count
#define DF_PAIR_OVF 3    // particle-substeps whose pair list overflowed
#define DF_NCOMP 4       // components counted by fp_launch_components
#define DF_CACHE_OVF 5   // solid contact partners beyond FP_CACHE_K (dropped, cumulative)
#define DF_NONFINITE 6   // a substep ended a particle with a non-finite position or velocity (sticky)
#define DF_WORDS 8

enum PassId {
  P_INTEGRATE,
  P_GRAB,
  P_HASH_COUNT,
  P_SCAN_REDUCE,
  P_SCAN_APPLY,
  P_SCATTER,
  P_RANK,
  P_CONTACTS,
  P_RESORT,
  P_DENSITY,
  P_DISPLACE,
  P_CONSTRAINTS,
  P_BOUNDS_VELOCITY,
  P_XSPH,
  P_CUT,
  P_PICK,
  P_COUNT,
  P_EMIT,
  P_SOLID_EXTRA,  // the optional solid modes: cache collection, tet / cached-contact steps, tet tearing
  P_NUM
};

// Per-pass timing with HIP events (fp_profile_enable / fp_profile_read).
struct Prof {
  bool on = false;
  struct Ev {
    int pass;
    hipEvent_t a, b;
  };
  std::vector<Ev> live;
  std::vector<hipEvent_t> pool;
  double ms[P_NUM] = {0};
  int64_t n[P_NUM] = {0};
};

struct fp_group;

struct fp_solver {
  fp_params P;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int max_fluid = 0, ns = 0, n = 0, nf = 0, T = 0, C = 0;
  int fluid_next = 0;
  // device-resident checkpoints (fp_checkpoint_save / _restore)
  struct Checkpoint {
    void* block = nullptr;
    size_t bytes = 0;
    int fluid_next = 0;
    int64_t substeps_done = 0;
    std::vector<int32_t> group;
    bool groups_pending = false, cut_dirty = false;
  };
  std::vector<Checkpoint> checkpoints;
  bool has_cube = false;
  int64_t substeps_done = 0;
  int64_t device_bytes = 0;
  // canonical per-particle state (particle index order)
  float4 *pos = nullptr, *prev = nullptr, *vel = nullptr;
  uint8_t* active = nullptr;
  float2* rho = nullptr;
  uint32_t* phash = nullptr;  // reference bucket at grid time
  // cell grid + sorted copies (sorted-slot order)
  CellGrid cg;
  int lscan_blocks = 0;
  int *lcount = nullptr, *lstart = nullptr, *lblock_sums = nullptr, *ltmp = nullptr;
  int* lstart_f = nullptr;  // cell starts of the fluid order (lstart: of the grid order)
  int* lblock_f = nullptr;  // per scan tile: fluid counted into it by the contact pass
  int* lnsol = nullptr;  // solids per non-empty cell of the grid order
  uint32_t *lkey = nullptr, *lslot = nullptr;
  int* sid = nullptr;
  float4* spos = nullptr;
  uint32_t* sphash = nullptr;
  float4* tpos = nullptr;
  uint32_t *tkey = nullptr, *tslot = nullptr;
  float4 *spos2 = nullptr, *svtmp = nullptr;
  uint32_t* sphash2 = nullptr;
  float2* sP = nullptr;
  uint32_t* nbr = nullptr;  // pair list: per 64 lanes of the fluid order, cap + 1 rows of 64 (the last a sink)
  int* ncount = nullptr;
  int pair_slots = 0;
  int* sbox = nullptr;  // solids' bounding box at grid time (6 ordered ints)
  // the reference's hash grid, built on demand
  bool ref_grid_valid = false;
  bool cells_built = false;  // a substep has built the two cell orders (fp_read_cells)
  int scan_blocks = 0;
  uint32_t* pslot = nullptr;
  int *count = nullptr, *cell_start = nullptr, *entries_tmp = nullptr, *cell_entries = nullptr;
  int* block_sums = nullptr;
  // constraints
  int2* cons_ab = nullptr;
  int2* cent = nullptr;       // CSR entries (of a, of b) of each constraint
  int* dgroup = nullptr;      // component labels, uploaded for a selection
  float* erest = nullptr;     // per CSR entry
  uint8_t* ealive = nullptr;  // per CSR entry
  uint8_t* ealive_snap = nullptr;  // ealive when due labels were deferred past a cut (groups_from_snap)
  int* adj_off = nullptr;
  uint32_t* adj = nullptr;  // per CSR entry: the other endpoint - this solid (24-bit) | max(deg) << 24
  float4* spos_c[2] = {nullptr, nullptr};
  // the GPU twin's optional solid modes (fp_params.tet_volume / solid_contact_cache)
  bool tets_on = false, cache_on = false;
  int cube_n = 0, NT = 0;
  double spacing = 0.0;
  int4* tet_v = nullptr;        // vertices, ordered so the rest volume is positive
  float* tet_v0 = nullptr;      // rest volume
  uint8_t* tet_alive = nullptr;
  float* tet_scale = nullptr;   // 1 / max(incident tets of its vertices)
  int *tet_adj_off = nullptr, *tet_adj = nullptr;  // tets incident to each solid (CSR)
  int *cache_ids = nullptr, *cache_n = nullptr;    // solid contact partners (FP_CACHE_K per solid)
  float4* spos_x = nullptr;     // solids after an iteration's extra step
  int* dflags = nullptr;
  double* dpick_t = nullptr;
  int* dpick_i = nullptr;
  int* grab_idx = nullptr;
  float4* grab_tgt = nullptr;
  int grab_cap = 0;
  std::vector<int> grab_round_len;
  // device staging of an emit of m particles: m positions, m velocities, m
  // slot indices, back to back (the host staging's layout, one copy); also
  // the output of fp_select_solids (count, then ids, as ints)
  float4* emit_buf = nullptr;
  int emit_cap = 0;
  // pinned host staging of the small uploads (emits, grab lists): two
  // buffers used in turn, each with an event recorded after the copies that
  // read it, so an upload waits only for the copies of the one before last
  struct Upload {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
  };
  Upload upload[2];
  int upload_next = 0;
  std::vector<int> h_cons_a, h_cons_b;
  std::vector<int2> h_cent;  // CSR entries of each constraint (host copy of cent)
  std::vector<float> h_snap_pos, h_snap_rest;
  std::vector<int32_t> h_group;
  bool groups_dirty = false;
  bool groups_pending = false;  // a relabel is due; run when group is read or alive is about to change
  bool groups_from_snap = false;  // ... from the flags in ealive_snap (a cut changed ealive since)
  bool steps_pending = false;
  bool cut_dirty = false;
  bool cut_dirty_at_step = false;
  // alive flags changed (a cut, a constraint write, a restore) since the
  // group labels were last computed or marked due: until then the labels are
  // not the partition of the current flags
  bool alive_changed = false;
  Prof prof;
  hipStream_t own_stream = nullptr;  // the stream fp_create made (stream may be a batch's)
  // "a constraint tore" flag (word 0) and the forwarded non-finite flag
  // (word 1) in pinned host memory, set by kernels through tear_dev and read
  // by the host after a stream sync
  int* tear_host = nullptr;
  int* tear_dev = nullptr;
  void* arena = nullptr;  // one pool block holding every fixed buffer of the scene
  int fluid_lanes = FLUID_LANES_SMALL;  // lanes per fluid particle (resolved from fp_params)
  bool nonfinite = false;            // a finished substep ended a particle non-finite (sticky)
  fp_group* group = nullptr;         // this scene alone, for fp_step
  fp_batch* batch = nullptr;         // the batch it belongs to, if any
  uint64_t version = 1;              // bumped when its parameters change
};

// ---- per-pass parameter blocks ---------------------------------------------

struct FluidParams {
  CellGrid cg;
  int ns, n, nf;
  int fl_log2;  // log2 of the lanes per fluid particle
  int stage;    // pair-list entries per lane the density pass stages in LDS
  float rs, rf;
  float inv_H, H2_lo, H2_hi, max_pair, max_total;
  double H2d;
  double cs;       // sort-cell size (the reference's cellSize / CELL_SUB)
  float cull_r2;   // a cell whose box lies at least this far (squared) holds no pair
  float contact_r2;  // the same for contacts (rs + max(rs, rf), squared, with margin)
  float stiff, near_stiff, rest_density;
  float contact_scale, visc;
};

struct BoundsParams {
  int ns, n;
  float rs, rf, mu_s, mu_f;
  float wall_x, wall_z, wall_h, wall_t;
  float cull_r2;
  float inv_h, damp_s, damp_f, vmax;
};

struct ConsParams {
  int ns, iters;
  float alpha, omega, break_strain, plastic_thr, plastic_rate;
};

// The fluid order as the fluid passes read it.
struct PairSource {
  const float4* __restrict__ spos2;     // fluid order: post-contact xyz, w = index | flags
  const uint32_t* __restrict__ sphash;  // fluid order: bucket at grid time
  const int* __restrict__ lstart;       // fluid-order cell starts
};

// The pair list the density pass records (layout in fp_fluid.hip): blocks
// of cap + 1 rows of 64 entries, one block per 64 lanes of the fluid order.
struct PairList {
  uint32_t* __restrict__ nbr;
  int* __restrict__ ncount;
  int cap;  // pairs per lane; row cap of a block is its sink
};

// The grid order as the contact pass reads it.
struct ContactCtx {
  const float4* __restrict__ spos;  // grid order: xyz + inverse mass
  const int* __restrict__ sid;
  const int* __restrict__ lstart;   // grid-order cell starts
  const int* __restrict__ lnsol;    // solids of each non-empty cell (they come first)
  const int* __restrict__ adj_off;
  const uint32_t* __restrict__ adj;  // packed entries (adj_other, adj_scale)
  const uint8_t* __restrict__ ealive;  // per CSR entry
  const int* __restrict__ sbox;  // solids' bounding box at grid time (ordered-int bits)
};

// The GPU twin's optional solid modes of one scene (none when both flags are 0).
#define FP_CACHE_K 16  // solid contact partners kept per solid
struct SolidModes {
  int tets, cache;      // the modes in use (only with constraint iterations)
  int n, NT;            // lattice side, tets
  double spacing;       // lattice spacing
  double gap;           // tet edge tear margin (tear_gap solid diameters)
  float rsum, rsum_adj; // solid contact distance, and for lattice neighbours (a cut face)
  float lim, lim_adj;   // the same + the collection margin
  float cap;            // per-tet, per-sum correction cap (half a solid radius)
  const int4* __restrict__ tet_v;
  const float* __restrict__ tet_v0;
  uint8_t* tet_alive;
  const float* __restrict__ tet_scale;
  const int* __restrict__ tet_adj_off;
  const int* __restrict__ tet_adj;
  int* cache_ids;
  int* cache_n;
  float4* spos_x;
};

// Everything a substep's kernels need about one scene, in device memory: the
// kernels of a launch group read their scene's entry through a work list of
// (scene, block of that scene) items, so one launch serves every scene.
struct SceneArgs {
  FluidParams fp;
  PairSource src;
  PairList pl;
  ContactCtx cc;
  BoundsParams bp;
  ConsParams cp;
  int n, ns, C;
  float gvx, gvy, gvz, h;  // gravity * h, substep
  int64_t snap_off;        // first float of the scene in a batch position snapshot
  float4 *pos, *prev, *vel;
  uint8_t* active;
  float2* rho;
  uint32_t* phash;
  int *lcount, *lstart, *lblock_sums, *ltmp, *lnsol;
  int* lstart_f;  // the fluid order's cell starts
  int* lblock_f;  // the fluid order's tile sums (added by the contact pass)
  uint32_t *lkey, *lslot;
  int* sid;
  float4* spos;
  uint32_t* sphash;
  float4* tpos;
  uint32_t *tkey, *tslot;
  float4 *spos2, *svtmp;
  uint32_t* sphash2;
  float2* sP;
  int* sbox;
  int* dflags;
  int* tear;  // host-mapped
  float* erest;
  uint8_t* ealive;
  int2* cent;
  int* adj_off;
  uint32_t* adj;
  float4* spos_c[2];
  SolidModes sm;
};

// Lanes per solid of the distance constraint pass (a power of 2, <= 64).
// Fewer lanes, each over several entries, make fewer and fuller waves: the
// 256 batched scenes' pass took 2.76 ms per frame at 32 lanes, 1.21 at 8
// and 1.04 at 4; the large scene's (its solid chain runs beside the fluid
// chain) 17.9, 18.7 and 22.1 us per substep, its frame 1.174, 1.165 and
// 1.160 ms.
#ifndef FP_CLANES
#define FP_CLANES 4
#endif

// Work lists: the blocks one scene needs in a pass of each kind.
enum WorkKind {
  W_PARTICLE,  // one thread per particle slot (BLOCK per block)
  W_SCAN,      // one scan tile of the cell table
  W_SOLID,     // contact lanes (4..32, chosen per group) per solid
  W_FLUID,     // fluid lanes (the scene's) per particle slot (FLUID_BLOCK per block)
  W_CONS,      // FP_CLANES threads per solid
  W_TET,       // one thread per tet (scenes with volume tets)
  W_NUM
};

// lanes per solid of the solid contact pass that a group's tuner tries
constexpr int kTuneLanes[4] = {4, 8, 16, 32};

// A set of scenes stepped together: one launch per pass over all of them
// (a solver alone is a group of one).  Owns the device scene table, the work
// lists and the captured frame graph.
struct fp_group {
  int device = 0;
  hipStream_t stream = nullptr;
  Prof* prof = nullptr;
  std::vector<fp_solver*> scenes;
  std::vector<uint64_t> versions;  // scene versions the table was built from
  SceneArgs* d_args = nullptr;
  int args_cap = 0;
  int2* d_items[W_NUM] = {nullptr};
  int n_items[W_NUM] = {0};
  int items_cap[W_NUM] = {0};
  int grid_cap[W_NUM] = {0};  // grid of each kind in the captured frame (its items then, rounded up to 8)
  double built_h = -1.0;
  bool dirty = true;
  int iters = 0;
  bool any_grabs = false;
  bool any_unconstrained = false;  // a member has solids but no constraint iterations
  bool any_extra = false;          // a member uses the volume tets or the solid contact cache
  bool any_cache = false;          // ... the solid contact cache
  bool any_tets = false;           // ... the volume tets
  bool stage_rule = false;         // a member runs 2 lanes per fluid particle (LDS staging pays there)
  int stage_on = 0;                // the density kernel stages its pair-list entries in LDS
  int xcd_run = FP_XCD_CHUNK;      // items per XCD run of the group kernels' lists (their W[0].y)
  // launch-shape tuner (batches): the density kernel's staging and the XCD
  // run length, timed over whole captured frames of a stable membership and
  // cached per group signature (see fp_step.hip)
  bool shape_tune = false;         // set for batches (FP_SHAPE_TUNE)
  int shape_state = 0;             // 0 untuned, 1 tuning, 2 settled
  int shape_cand = 0, shape_done = 0;
  uint64_t shape_sig = 0;
  double shape_ms[8] = {0};
  int shape_n[8] = {0};
  bool shape_pending = false;
  int shape_pending_cand = -1;
  hipEvent_t shape_ev[2] = {nullptr, nullptr};
  // captured frame
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  double graph_dt = -1.0;
  int graph_sub = 0;
  int graph_iters = -1;
  int graph_clanes = 0;             // contact lanes the captured frame launches
  bool graph_unconstrained = false;  // the captured frame has the bounds pass of unconstrained scenes
  int graph_extra = -1;              // the captured frame's optional passes (bit 0 cache, 1 tets) and staging density kernel (bit 2)
  bool use_graph = true;
  int64_t graph_launches = 0;
  int64_t graph_captures = 0;
  // A captured frame of a group with at least FP_SIDE_MIN particle slots runs
  // each substep's solid chain (solid contacts, the constraint iterations) on
  // a second stream, beside the fluid chain (fluid contacts, the fluid order,
  // density, displacement); they share no buffer they write, and join before
  // the bounds pass.  Smaller groups (whose solid chain is too short to hide
  // anything), and eager frames (tuning, grabs, per-pass timing), run the two
  // chains one after the other.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool capturing = false;
  bool fork = false;  // the frame being captured forks
  bool ride = false;  // the solid chain rides in the fluid chain's launches (a group that never forks)
  // the members or their parameters changed for this frame: it runs eagerly
  // (a graph captured now might be replayed once; continuous batching
  // changes members every few frames)
  bool members_changed = true;
  // autotuned launch shape: lanes per solid of the solid contact pass (the
  // result does not depend on it).  The first TUNE_FRAMES frames time 4, 8,
  // 16 and 32 lanes with events, rotating the candidate from substep to
  // substep (and the starting candidate from frame to frame) so that every
  // candidate is timed on the same states; the fastest mean is kept, or the
  // most lanes among the candidates within TUNE_TIE of it.  A group whose
  // solid count changes by more than 2x tunes again.
  static constexpr int TUNE_FRAMES = 4;
  int clanes = 4;                 // lanes of the work lists' W_SOLID entries
  int items_clanes = 0;           // lanes the work lists were built for
  int clanes_best = 0;            // 0 while tuning
  bool clanes_pinned = false;     // set by the caller: no tuning
  int tune_k = 0;                 // tuning frames done
  bool tune_pending = false;      // a tuning frame's events await reading
  bool tune_record = false;       // this frame tunes
  int tune_cand = 0;              // candidate of the current substep
  int tune_used = 0;              // event pairs recorded in the pending frame
  std::vector<int> tune_of;       // candidate of each event pair
  double tune_sum[4] = {0.0, 0.0, 0.0, 0.0};
  int tune_cnt[4] = {0, 0, 0, 0};
  int64_t tuned_solids = 0;
  std::vector<hipEvent_t> tune_ev;
  // W_SOLID work lists of the four candidates, back to back, while tuning
  int2* d_tune_items = nullptr;
  int tune_items_cap = 0;
  int tune_off[4] = {0, 0, 0, 0};
  int tune_n[4] = {0, 0, 0, 0};
  bool tune_items_ok = false;
  std::string err;
};

// ---- host helpers (fp_api.hip) ---------------------------------------------

int fp_fail(fp_solver* s, int code, const std::string& msg);
// the tuning caches of fp_step.hip (contact lanes, launch shape), by signature
std::vector<std::pair<uint64_t, int>> fp_tune_cache_entries(bool lanes);
void fp_tune_cache_merge(const std::vector<std::pair<uint64_t, int>>& lanes,
                         const std::vector<std::pair<uint64_t, int>>& shape);
void fp_tune_caches_clear();

// ---- per-device memory pool (fp_pool.hip) -----------------------------------
size_t fp_pool_class(size_t bytes);
hipError_t fp_pool_get(int device, size_t bytes, void** out, size_t* class_bytes);
size_t fp_pool_put(int device, void* ptr);  // returns the block's class bytes (0: not a pool block)
int64_t fp_pool_trim_device(int device);
hipEvent_t fp_event_get(Prof* p);

#define FP_TRY(s, expr)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return fp_fail((s), FP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define FP_GTRY(g, expr)                                                                \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      (g)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                     \
      return FP_EDEVICE;                                                                \
    }                                                                                   \
  } while (0)

// Brackets the launches of one pass with events when profiling is on.
struct PassScope {
  Prof* p;
  hipStream_t stream;
  int pass;
  hipEvent_t a = nullptr, b = nullptr;
  PassScope(Prof* p_, hipStream_t st, int ps) : p(p_), stream(st), pass(ps) {
    if (p && p->on) {
      a = fp_event_get(p);
      b = fp_event_get(p);
      if (a) (void)hipEventRecord(a, stream);
    }
  }
  PassScope(fp_solver* s, int ps) : PassScope(&s->prof, s->stream, ps) {}
  ~PassScope() {
    if (p && p->on && a && b) {
      (void)hipEventRecord(b, stream);
      p->live.push_back({pass, a, b});
    }
  }
};

static inline int fp_blocks(int n) { return (n + BLOCK - 1) / BLOCK; }

// Work lists in device memory: W[0].x = the number of items, the (scene,
// local block) items at W[1 ...].  The kernels read the count there, so a
// captured frame stays valid when the group's members change: the host
// rewrites the scene table and the lists before the next replay, and every
// kernel of the graph strides over whatever the lists hold then.
static inline int fp_grid8(int n) { return (n + 7) & ~7; }

// Launch of a group kernel over work kind k.  A direct launch takes one
// workgroup per item (rounded up to a multiple of 8, the XCD count); a launch
// being captured takes the grid fixed for the graph (grid_cap: the items at
// capture, rounded up), whose workgroups stride over the items of later
// replays -- persistent workgroups over the scene table.
#define GROUP_LAUNCH(g, kernel, kind, block, ...) GROUP_LAUNCH_ON(g, (g)->stream, kernel, kind, block, ##__VA_ARGS__)
#define GROUP_LAUNCH_ON(g, st, kernel, kind, block, ...)                                            \
  do {                                                                                               \
    const int grid_ = (g)->capturing ? (g)->grid_cap[kind] : fp_grid8((g)->n_items[kind]);         \
    if (grid_ > 0)                                                                                   \
      hipLaunchKernelGGL(kernel, dim3(grid_), dim3(block), 0, (st), (g)->d_args, (g)->d_items[kind], \
                         ##__VA_ARGS__);                                                             \
  } while (0)

// One launch over two work lists: kind k0's workgroups first, then kind
// k1's (each count as GROUP_LAUNCH would take it); the kernel gets both lists
// and g0, kind k0's workgroup count.  g0 is a multiple of 8, so a workgroup
// keeps its XCD (b % 8) at its place b - g0 in the second list.
#define GROUP_LAUNCH2(g, kernel, k0, k1, block, ...)                                               \
  do {                                                                                               \
    const int g0_ = (g)->capturing ? (g)->grid_cap[k0] : fp_grid8((g)->n_items[k0]);               \
    const int g1_ = (g)->capturing ? (g)->grid_cap[k1] : fp_grid8((g)->n_items[k1]);               \
    if (g0_ + g1_ > 0)                                                                               \
      hipLaunchKernelGGL(kernel, dim3(g0_ + g1_), dim3(block), 0, (g)->stream, (g)->d_args,          \
                         (g)->d_items[k0], (g)->d_items[k1], g0_, ##__VA_ARGS__);                    \
  } while (0)

// The number of items of work list W (the same in every lane).
__device__ __forceinline__ int items_count(const int2* __restrict__ W) {
  return __builtin_amdgcn_readfirstlane(*(const FP_GLOBAL int*)&W[0].x);
}

// The (scene, local block) items of a group kernel: for (w ...) { SceneArgs& A; int lb }
// Workgroups are dispatched to the 8 XCDs in turn (workgroup b to XCD
// b % 8), each XCD with its own L2; grids are multiples of 8, so item
// position w runs on XCD w % 8 whether a workgroup takes one item or strides
// over several.  The work list is dealt to the XCDs in runs of FP_XCD_CHUNK
// consecutive items (K, the list's W[0].y: FP_XCD_CHUNK unless the group's
// launch-shape tuner chose another), so items next to each other
// (neighbouring particles, whose gathers read the same rows) share an L2
// while every XCD still gets a share of each part of the list (idle
// capacity at a fluid order's end, scenes of unequal size).  The items past
// the last whole round of 8 runs keep their own position.  A bijection on
// the list for every K: results do not depend on it.
#define FP_XCDS 8
__device__ __forceinline__ int xcd_item(int b, int n, int K) {
  const int R = FP_XCDS * K;
  if (K <= 1 || b >= n / R * R) return b;
  const int x = b % FP_XCDS, l = b / FP_XCDS;  // XCD, and the position's place in its turn
  return (l / K * FP_XCDS + x) * K + l % K;
}

// The XCD run length of work list W (the same in every lane).
__device__ __forceinline__ int items_run(const int2* __restrict__ W) {
  return __builtin_amdgcn_readfirstlane(*(const FP_GLOBAL int*)&W[0].y);
}

#define FOR_ITEMS(S, W) FOR_ITEMS_AT(S, W, (int)blockIdx.x, (int)gridDim.x)
// The same over list positions b0, b0 + G, ...: the workgroups of a launch
// that serves two work lists (GROUP_LAUNCH2) start their list at b0 and
// stride by their own count G.
#define FOR_ITEMS_AT(S, W, b0, G)                                                          \
  for (int w_ = (b0), nw_ = items_count(W), k_ = items_run(W); w_ < nw_; w_ += (G))        \
    if (const int2 it_ = (W)[1 + xcd_item(w_, nw_, k_)]; true)                             \
      if (const SceneArgs& A = (S)[it_.x]; true)                                           \
        if (const int lb = it_.y; true)

// ---- per-scene set-up (fp_fluid.hip, fp_step.hip, fp_solid.hip) -------------

FluidParams fp_fluid_params(const fp_solver* s);
SceneArgs fp_scene_args(const fp_solver* s, double h);
int fp_scene_blocks(const fp_solver* s, int kind, int contact_lanes);

// ---- group launches (one per pass, all scenes) -----------------------------

// fp_grid.hip
int fp_group_cell_grid(fp_group* g, bool count_done);
int fp_group_resort(fp_group* g);
// fp_fluid.hip
int fp_group_contacts(fp_group* g);
// The solid chain riding in the fluid chain's launches (groups that do not fork):
int fp_group_contacts_both(fp_group* g);              // solid + fluid contacts, one launch
int fp_group_resort_with(fp_group* g, int it);        // the fluid rank + constraint iteration it
int fp_group_density_with(fp_group* g, int it);       // (it < 0: the pass alone)
int fp_group_displace_with(fp_group* g, int it);
int fp_group_contacts_solid(fp_group* g, hipStream_t st);
int fp_group_contacts_fluid(fp_group* g);
int fp_group_density(fp_group* g);
int fp_group_displace(fp_group* g);
int fp_group_xsph(fp_group* g);
// fp_solid.hip
int fp_group_constraints(fp_group* g, int iters, hipStream_t st);
// fp_step.hip
int fp_group_frame(fp_group* g, double dt, int substeps);
int fp_group_prepare(fp_group* g, double h);
void fp_group_tune_begin(fp_group* g);
int fp_group_tune_end(fp_group* g);
int fp_group_snapshot(fp_group* g, float* out);
void fp_group_free(fp_group* g);

// ---- single-scene launches (events, queries) -------------------------------

int fp_launch_ref_grid(fp_solver* s, bool rehash);
int fp_launch_cut(fp_solver* s, const double o[3], const double d1[3], const double d2[3], double nx, double ny,
                  double nz);
int fp_group_solid_collect(fp_group* g, hipStream_t st);
int fp_group_solid_extra(fp_group* g, int it, hipStream_t st);
int fp_group_tet_update(fp_group* g, hipStream_t st);
int fp_launch_pick(fp_solver* s, const double o[3], const double d[3], double max_perp2);
int fp_launch_select(fp_solver* s, const double c[3], double r2, int same, int* ids, int max_ids, int* n);
int fp_launch_components(fp_solver* s);
int fp_launch_emit(fp_solver* s, int m);
int fp_launch_count_active(fp_solver* s);
