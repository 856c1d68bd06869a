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
//   nbr u32[(K+L)*N] (L = FLUID_LANES lanes per fluid slot, each with K/L
//     entries: entry k of lane v at k*N*L + v; row K/L is a sink), ncount i32[N*L]: the
//     pairs the density pass found, replayed by displacement and XSPH
//   count/cell_start/entries_tmp/cell_entries: the reference's hash grid,
//     rebuilt from phash only when read
//   cons_ab int2[C]; rest[2] f32[C]; alive[2] u8[C] (ping-pong per iteration);
//   adj_off i32[Ns+1]; adj int4[2C] = (other endpoint, constraint id, bits
//     of 1/max(deg_a, deg_b), 0); spos_c[2] float4[Ns] between iterations
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/fluidphysics.h"
#include "fp_kernels.h"

#define BLOCK 256
#define SCAN_ITEMS 16
#define SCAN_TILE (BLOCK * SCAN_ITEMS)
#define PF_MOVED 1u  // fluid particle changed cell in the contact pass
#define PF_DUP 2u    // its current cell's 27 reference probes repeat a bucket
#define SID_MASK 0x3fffffffu
#define PF_SHIFT 30
#define NO_KEY 0xffffffffu
#ifndef FLUID_LANES
#define FLUID_LANES 2  // lanes per fluid particle in density, displacement and XSPH
#endif

// device flag words (dflags)
#define DF_TEAR 0        // a constraint tore this step
#define DF_CUT 1         // knife count
#define DF_SCRATCH 2     // fluid count
#define DF_PAIR_OVF 3    // particle-substeps whose pair list overflowed
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
  P_NUM
};

struct fp_solver {
  fp_params P;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int max_fluid = 0, ns = 0, n = 0, nf = 0, T = 0, C = 0;
  int fluid_next = 0;
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
  int* lnsol = nullptr;  // solids per non-empty cell of the grid order
  uint32_t *lkey = nullptr, *lslot = nullptr;
  int* sid = nullptr;
  float4* spos = nullptr;
  uint32_t* sphash = nullptr;
  float4* tpos = nullptr;
  uint32_t *tkey = nullptr, *tslot = nullptr;
  int *fsid = nullptr, *fslot = nullptr;
  float4 *spos2 = nullptr, *svtmp = nullptr;
  uint32_t* sphash2 = nullptr;
  float2* sP = nullptr;
  uint32_t* nbr = nullptr;  // pair list, (pair_slots + 1) x capacity, entry-major; the last row is a sink
  int* ncount = nullptr;
  int pair_slots = 0;
  int* sbox = nullptr;  // solids' bounding box at grid time (6 ordered ints)
  // the reference's hash grid, built on demand
  bool ref_grid_valid = false;
  int scan_blocks = 0;
  uint32_t* pslot = nullptr;
  int *count = nullptr, *cell_start = nullptr, *entries_tmp = nullptr, *cell_entries = nullptr;
  int* block_sums = nullptr;
  // constraints
  int2* cons_ab = nullptr;
  float* rest[2] = {nullptr, nullptr};
  uint8_t* alive[2] = {nullptr, nullptr};
  int cur = 0;
  int* adj_off = nullptr;
  int4* adj = nullptr;
  float4* spos_c[2] = {nullptr, nullptr};
  int* dflags = nullptr;
  double* dpick_t = nullptr;
  int* dpick_i = nullptr;
  int* grab_idx = nullptr;
  float4* grab_tgt = nullptr;
  int grab_cap = 0;
  std::vector<int> grab_round_len;
  int* emit_idx = nullptr;
  float4 *emit_p = nullptr, *emit_v = nullptr;
  int emit_cap = 0;
  std::vector<int> h_cons_a, h_cons_b;
  std::vector<float> h_snap_pos, h_snap_rest;
  std::vector<int32_t> h_group;
  bool groups_dirty = false;
  bool steps_pending = false;
  bool cut_dirty = false;
  bool cut_dirty_at_step = false;
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

// ---- per-pass parameter blocks ---------------------------------------------

struct FluidParams {
  CellGrid cg;
  int ns, n, nf;
  float rs, rf;
  float inv_H, H2_lo, H2_hi, max_pair, max_total;
  double H2d;
  double cs;       // cell size
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
  int ns;
  float alpha, omega, break_strain, plastic_thr, plastic_rate;
};

// ---- host helpers (fp_api.cpp) ---------------------------------------------

int fp_fail(fp_solver* s, int code, const std::string& msg);
hipEvent_t fp_event_get(fp_solver* s);

#define FP_TRY(s, expr)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return fp_fail((s), FP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Brackets the launches of one pass with events when profiling is on.
struct PassScope {
  fp_solver* s;
  int pass;
  hipEvent_t a = nullptr, b = nullptr;
  PassScope(fp_solver* s_, int p) : s(s_), pass(p) {
    if (s->profile) {
      a = fp_event_get(s);
      b = fp_event_get(s);
      if (a) (void)hipEventRecord(a, s->stream);
    }
  }
  ~PassScope() {
    if (s->profile && a && b) {
      (void)hipEventRecord(b, s->stream);
      s->ev_live.push_back({pass, a, b});
    }
  }
};

static inline int fp_blocks(int n) { return (n + BLOCK - 1) / BLOCK; }

// ---- launchers -------------------------------------------------------------

// fp_grid.hip
int fp_launch_cell_grid(fp_solver* s, bool count_done);
int fp_launch_resort(fp_solver* s);
int fp_launch_ref_grid(fp_solver* s, bool rehash);

// fp_fluid.hip
FluidParams fp_fluid_params(const fp_solver* s);
int fp_launch_contacts(fp_solver* s);
int fp_launch_density(fp_solver* s);
int fp_launch_displace(fp_solver* s);
int fp_launch_xsph(fp_solver* s);

// fp_solid.hip
int fp_launch_constraints(fp_solver* s, double h, const float4** solid_src);
int fp_launch_cut(fp_solver* s, const double o[3], const double d1[3], const double d2[3], double nx, double ny,
                  double nz);
int fp_launch_pick(fp_solver* s, const double o[3], const double d[3], double max_perp2);

// fp_step.hip
int fp_launch_substep(fp_solver* s, double h);
int fp_launch_emit(fp_solver* s, int m);
int fp_launch_count_active(fp_solver* s);
