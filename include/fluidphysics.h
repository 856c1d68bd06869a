/*
 * fluidphysics.h -- C ABI of the MI355X unified-particle solver (libfluidphysics.so).
 *
 * This is the drop-in boundary for the per-substep solver of the reference
 * CPU solver `class Solver` (src/solver.js:8-516).  Every entry point below
 * replaces one method or one public field of that class; the citation on
 * each declaration names it.  The N-API addon (fluidphysics_amd/addon) and
 * the ctypes host (fluidphysics_amd/__init__.py) bind exactly these symbols.
 *
 * Conventions
 *   - Every function returns 0 on success, a negative FP_E* code on failure;
 *     fp_last_error() then holds a message.  Cases the reference ignores
 *     silently keep the reference's answer (fp_cut reports 0 cuts for a
 *     degenerate wedge, fp_pick_solid reports id -1 for no hit).
 *   - Plain C types only: pointers + counts.  Positions crossing the ABI are
 *     float32 xyz triples (the reference's Float32Array layout, 3 per
 *     particle); ray/knife/grab inputs are double (JS numbers).
 *   - Device state stays resident in HBM between calls.  fp_step() enqueues
 *     all substeps on the solver's own HIP stream and waits for them, like
 *     the synchronous JS call; fp_step_async() leaves the wait to fp_sync().
 *   - Not thread-safe per solver; different solvers are independent.
 */
#ifndef FLUIDPHYSICS_H
#define FLUIDPHYSICS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FP_OK 0
#define FP_EINVAL (-1)     /* bad argument (null pointer, out-of-range size, wrong call order) */
#define FP_EDEVICE (-2)    /* HIP runtime error */
#define FP_ENOMEM (-3)     /* device or host allocation failed */
#define FP_ESTATE (-4)     /* call not valid in the solver's current state */
#define FP_EIO (-5)        /* a file could not be opened, read or written */

/* Solver constants.  Defaults (fp_params_default) are the reference's
 * tuning fields, src/solver.js:14-37, plus cellSize (:50) and the
 * hard-coded fluid damping (:220).  Mutable between steps like the JS
 * public fields (fp_set_params). */
typedef struct fp_params {
  double gravity;              /* solver.js:15  -10.0 */
  double solid_radius;         /* :16  0.09 */
  double fluid_radius;         /* :17  0.075 */
  double compliance;           /* :18  2.5e-4 */
  double plastic_threshold;    /* :19  0.32 */
  double plastic_rate;         /* :20  0.04 */
  double break_strain;         /* :21  2.6 */
  double fluid_h;              /* :23  0.19 (interaction radius) */
  double rest_density;         /* :24  6.0 */
  double stiffness;            /* :25  0.002 */
  double near_stiffness;       /* :26  0.008 */
  double viscosity;            /* :27  0.3 (XSPH) */
  double solid_damping;        /* :28  1.2 1/s */
  double fluid_damping;        /* :220 0.6 1/s (hard-coded in the reference) */
  double floor_friction_solid; /* :29  0.4 */
  double floor_friction_fluid; /* :30  0.04 */
  double wall_x;               /* :31  3.5 */
  double wall_z;               /* :31  3.5 */
  double wall_height;          /* :31  0.92 */
  double wall_half_thick;      /* :31  0.07 */
  double cull_radius;          /* :32  7.0 */
  double max_speed;            /* :33  14.0 */
  double grab_stiffness;       /* :36  0.3 */
  double cell_size;            /* :50  0.22 (must be >= fluid_h and >= 2*max radius) */
  int32_t constraint_iters;    /* Jacobi XPBD iterations per substep (the reference runs one
                                  Gauss-Seidel sweep, :215); default 3 */
  int32_t table_size;          /* hash table size, power of two; 0 = nextpow2(2*capacity) (:51) */
  int32_t pair_slots;          /* pairs recorded per fluid particle by the density pass for the
                                  displacement and XSPH passes (the reference's pair list,
                                  :284-327); a particle with more is gathered again from the
                                  grid.  0 = 64 per fluid lane */
  int32_t fluid_lanes;         /* lanes per fluid particle in the density, displacement and
                                  XSPH passes: 1, 2, 4 or 8; 0 = 4 for a maxFluid of 131072 or
                                  more, else 2.  It orders the pair list and the lanes' partial
                                  sums: part of the scene's arithmetic (a scene steps
                                  identically alone and in a batch) */
  double jacobi_omega;         /* Jacobi relaxation: each constraint's correction is scaled by
                                  jacobi_omega / max(deg_a, deg_b); default 8.0, which with 3
                                  iterations reproduces the reference's 300-frame statistics at
                                  its own constants (the GPU-scale scenes use 4.0) */
  double contact_scale;        /* scale on each Jacobi contact correction, 1.0 */
  double gravity_x;            /* gravity along x and z (a tilted basin; the reference has y only,
                                  solver.js:15); default 0 */
  double gravity_z;
  /* Optional solid modes of the reference's GPU twin (src/gpu/wgsl.js; src/solver.js has
     neither).  Both are fixed once fp_add_solid_cube has allocated the scene, and both act
     inside the constraint iterations (none without them). */
  int32_t tet_volume;          /* 1: five volume-preserving tetrahedra per lattice cell, projected
                                  Jacobi-style in every constraint iteration after the distance
                                  constraints (wgsl.js:423-468; tets built at sim.js:144-173);
                                  torn when an edge outgrows its lattice rest by tear_gap solid
                                  diameters (wgsl.js:471-489), cut with the constraints
                                  (wgsl.js:492-519).  0 (default): off */
  int32_t solid_contact_cache; /* 1: solid-solid contacts listed once per substep with a margin
                                  of one solid diameter and resolved in every constraint
                                  iteration (wgsl.js:532-596) instead of once in the contact pass.
                                  0 (default): off */
  double tear_gap;             /* tet edge tear margin in solid diameters (sim.js:49: 1.3) */
} fp_params;

/* Opaque solver handle (one scene, one HIP stream). */
typedef struct fp_solver fp_solver;

typedef struct fp_info {
  int32_t capacity;            /* solids + maxFluid slots (solver.js:41) */
  int32_t solid_count;         /* :89 */
  int32_t max_fluid;           /* :10 */
  int32_t table_size;          /* :51 */
  int32_t num_constraints;     /* consA.length */
  int32_t fluid_next;          /* ring cursor of emitFluid (:140-141) */
  int32_t device;              /* HIP device ordinal */
  int32_t pair_slots;          /* pair_slots in use */
  int64_t substeps_done;       /* substeps executed since creation */
  int64_t device_bytes;        /* HBM bytes held by this solver */
  int64_t pair_overflow;       /* particle-substeps whose pairs did not fit pair_slots
                                  (cumulative; those were gathered again, results unchanged) */
  int32_t nonfinite;           /* fp_nonfinite's flag */
  int32_t reserved;
} fp_info;

/* ---- lifetime ------------------------------------------------------- */

/* Fill the reference defaults (solver.js:14-37, :50, :220). */
void fp_params_default(fp_params* out);

/* new Solver({ maxFluid }) (solver.js:9).  params may be NULL (defaults).
 * device < 0 selects the current HIP device. */
int fp_create(const fp_params* params, int32_t max_fluid, int32_t device, fp_solver** out);

/* Frees device memory and the stream.  NULL is accepted. */
int fp_destroy(fp_solver* s);

/* Message of the last failure on s (s == NULL: last failure of fp_create or
 * of a call made with a NULL solver).  Never NULL. */
const char* fp_last_error(const fp_solver* s);

/* Replace / read the tuning fields (the JS public fields, solver.js:14-37). */
int fp_set_params(fp_solver* s, const fp_params* p);
int fp_get_params(const fp_solver* s, fp_params* out);
int fp_get_info(const fp_solver* s, fp_info* out);

/* ---- scene set-up --------------------------------------------------- */

/* addSolidCube(cx, cy, cz, n, spacing, particleMass) (solver.js:72-123):
 * allocates capacity n^3 + maxFluid, the n^3 lattice (id = (ix*n+iy)*n+iz),
 * the 26-neighbour distance constraints (b > a, loop order of :99-116) and
 * their CSR adjacency.  Call once, before stepping. */
int fp_add_solid_cube(fp_solver* s, double cx, double cy, double cz, int32_t n, double spacing,
                      double particle_mass);

/* emitFluid(x, y, z, vx, vy, vz, mass) called n times in order
 * (solver.js:139-149): ring slots solidCount + fluidNext, overwriting live
 * particles when the ring wraps.  pos/vel: 3n doubles; mass: n doubles or
 * NULL for the default 0.03. */
int fp_emit(fp_solver* s, const double* pos, const double* vel, const double* mass, int32_t n);

/* resetSolid() (solver.js:125-137): restore lattice positions and rest
 * lengths, zero solid velocities, revive every constraint, groups = 0,
 * clear the grabs. */
int fp_reset_solid(fp_solver* s);

/* clearFluid() (solver.js:151-153). */
int fp_clear_fluid(fp_solver* s);

/* ---- the hot path --------------------------------------------------- */

/* step(dt, substeps) (solver.js:187-242): substeps x [integrate, grab,
 * hash grid, contacts, double-density relaxation, XPBD constraints, bounds,
 * velocity + damping, XSPH], then the connected-component labels if a
 * constraint tore.  Synchronous. */
int fp_step(fp_solver* s, double dt, int32_t substeps);

/* Same work as fp_step, enqueued without the final wait.  The component
 * labels are refreshed lazily when fp_read_state asks for them. */
int fp_step_async(fp_solver* s, double dt, int32_t substeps);

/* Wait for all work enqueued on s. */
int fp_sync(fp_solver* s);

/* The `grabs` public field (solver.js:35, applied at :205-210): replaces the
 * list of {i, x, y, z} targets pulled toward each substep.  idx: n particle
 * indices, xyz: 3n doubles.  n == 0 clears the list. */
int fp_set_grabs(fp_solver* s, const int32_t* idx, const double* xyz, int32_t n);

/* cut(ox, oy, oz, d1x, d1y, d1z, d2x, d2y, d2z) (solver.js:458-486): kill
 * every alive constraint whose segment crosses the wedge swept between two
 * rays.  The predicate is evaluated in IEEE double from the float32
 * positions, in the reference's operation order.  *ncut = constraints cut
 * (0 for a degenerate wedge, |d1 x d2| < 1e-8). */
int fp_cut(fp_solver* s, const double o[3], const double d1[3], const double d2[3], int32_t* ncut);

/* The grab selection of the reference's caller (src/main.js:293-301, after
 * pickSolid): the solids in the component of solid same_group_as (any
 * component when it is -1) whose position lies within radius of center,
 * tested in double, in ascending index order.  Writes at most max_ids ids;
 * *n receives the number of matches.  The component is the one the caller
 * would read in group (relabelled at a step's end, solver.js:238-241): when
 * no cut or constraint write happened since, it is computed on the device
 * (fp_components), else from the host labels. */
int fp_select_solids(fp_solver* s, const double center[3], double radius, int32_t same_group_as, int32_t* ids,
                     int32_t max_ids, int32_t* n);
/* _recomputeGroups (solver.js:488-499) on the device: the connected
 * components of the solids over the constraints alive now, from a lock-free
 * union-find (the larger of two roots hooked under the smaller).  labels[i]
 * (ns entries, may be NULL) receives the smallest solid index of i's
 * component and *count (may be NULL) the number of components.  The
 * partition is the reference's; its labels are the roots its sequential
 * union order leaves, which fp_read_state's group keeps. */
int fp_components(fp_solver* s, int32_t* labels, int32_t* count);
/* pickSolid(ox, oy, oz, dx, dy, dz, maxPerp) (solver.js:502-515): nearest
 * solid along the ray with perpendicular distance <= max_perp; ties go to the
 * lowest index.  *id = -1 (and *t = +inf) when nothing is hit. */
int fp_pick_solid(fp_solver* s, const double o[3], const double d[3], double max_perp, int32_t* id,
                  double* t);

/* fluidCount() (solver.js:155-159). */
int fp_fluid_count(fp_solver* s, int32_t* count);

/* The scene's non-finite flag: 1 once a finished substep has ended a
 * particle with a non-finite position or velocity (the velocity pass
 * raises it on the device; the host picks it up at the next sync, with no
 * copy).  The reference lets NaNs propagate and counts NaN positions in
 * the GPU page's metrics (src/gpu/main.js:546-556); the solver does the
 * same arithmetic and reports the event, so a serving loop can retire the
 * scene.  Sticky; fp_checkpoint_restore and fp_state_import clear it.
 * Reads host memory only (cheap to call every frame). */
int fp_nonfinite(const fp_solver* s, int32_t* out);

/* ---- state access (the JS public typed arrays, solver.js:42-48) ------ */

/* Copy state to host.  Any pointer may be NULL.  pos/prev/vel: 3*capacity
 * floats; inv_mass: capacity floats; active: capacity bytes; group:
 * solidCount int32 (union-find roots, solver.js:488-499). */
int fp_read_state(fp_solver* s, float* pos, float* prev, float* vel, float* inv_mass,
                  uint8_t* active, int32_t* group);

/* Overwrite state (any pointer may be NULL to keep that array).  Used to
 * start the solver from a given state, as assigning the typed arrays does. */
int fp_write_state(fp_solver* s, const float* pos, const float* prev, const float* vel,
                   const float* inv_mass, const uint8_t* active);

/* Constraint topology and state: a, b (consA/consB), rest (consRest), alive
 * (consAlive), num_constraints entries each; any pointer may be NULL. */
int fp_read_constraints(fp_solver* s, int32_t* a, int32_t* b, float* rest, uint8_t* alive);
int fp_write_constraints(fp_solver* s, const float* rest, const uint8_t* alive);

/* The volume tets of the GPU twin's optional tet mode (the tets of
 * src/gpu/sim.js, five per lattice cell; none without fp_params.tet_volume):
 * *num_tets is set to their count; verts (4 per tet, ordered so the rest
 * volume is positive) and alive (1 per tet) may be NULL. */
int fp_read_tets(fp_solver* s, int32_t* verts, uint8_t* alive, int32_t* num_tets);

/* No counterpart in the reference: read access to the substep's own cell
 * sort, for tests.  The last substep sorted the particles twice by cell
 * (sort cells of cell_size, keyed (cx * mx + cy * my + cz) mod M with the
 * cell coordinates floor(p * inv_cell) as 32-bit integers): which = 0 is the
 * grid order (every active particle by its position after integrate and the
 * grab pull), which = 1 the fluid order (every active fluid particle by its
 * position after the contact pass).  Inside a key, ascending particle index.
 * desc (5 doubles): inv_cell, M, mx, my, filled slots; start (M + 1 int32):
 * first slot of each key; ids (capacity int32): particle of each slot; pos
 * (4 floats per slot, capacity slots): the position the slot was sorted by
 * (w: the inverse mass in the grid order, index and flag bits in the fluid
 * order).  Any output may be NULL; FP_ESTATE before the first substep. */
int fp_read_cells(fp_solver* s, int32_t which, double* desc, int32_t* start, int32_t* ids, float* pos);

/* _buildGrid() (solver.js:174-184) on the current positions, on the device. */
int fp_build_grid(fp_solver* s);

/* cellStart (table_size + 1 int32) and cellEntries (capacity int32; only the
 * first cellStart[table_size] are defined) of the last grid build. */
int fp_read_grid(fp_solver* s, int32_t* cell_start, int32_t* cell_entries);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): _buildGrid (solver.js:174-184) -- the same kernels and
 * bit-exact result as fp_build_grid + fp_read_grid -- over arrays the caller
 * holds in device memory, allocating nothing.  Every intermediate lives in
 * one static device array of fixed size (~21 MB per GPU, held until the
 * process ends), so n <= FP_GRID_NOALLOC_MAX_N and table_size <=
 * FP_GRID_NOALLOC_MAX_T; larger sizes return FP_EINVAL.
 *   pos: 3n floats (x, y, z per particle), active: n bytes (0 = skipped),
 *   cell_start: table_size + 1 int32 (bucket k holds entries
 *   [cell_start[k], cell_start[k+1]), descending particle index),
 *   cell_entries: n int32 (the first cell_start[table_size] defined);
 *   table_size a power of two; cell_size the reference's cellSize.
 * All pointers are device memory of the current device.  The work is
 * enqueued on `stream` (a hipStream_t; NULL = the null stream) and the call
 * returns without waiting.  Calls on one GPU share the static array: calls
 * that overlap in time on that GPU race (serialise them, e.g. on one stream).
 * The first call on each device looks up the array's device address (not a
 * stream operation); nothing else is allocated or freed. */
#define FP_GRID_NOALLOC_MAX_N (1 << 20)
#define FP_GRID_NOALLOC_MAX_T (1 << 21)
int fp_build_grid_noalloc(const float* pos, const uint8_t* active, int32_t n, int32_t table_size, double cell_size,
                          int32_t* cell_start, int32_t* cell_entries, void* stream);

/* rho / rhoNear (capacity floats each; fluid slots) of the last substep's
 * density pass (solver.js:316-317).  Either pointer may be NULL. */
int fp_read_density(fp_solver* s, float* rho, float* rho_near);

/* ---- measurement ---------------------------------------------------- */

/* Per-pass timing with HIP events recorded on the solver's stream. */
#define FP_MAX_PASSES 24
typedef struct fp_pass_stat {
  char name[32];
  int64_t launches;
  double total_ms;
} fp_pass_stat;

/* enable != 0: bracket every kernel launch with events (adds overhead; for
 * profiling only).  Resets the accumulated statistics. */
int fp_profile_enable(fp_solver* s, int32_t enable);
/* Accumulated statistics; *n = number of passes written (<= max). */
int fp_profile_read(fp_solver* s, fp_pass_stat* out, int32_t max, int32_t* n);

/* ---- checkpoints and state blobs ------------------------------------- */

/* The reference's GPU twin snapshots its buffers for exact repro
 * (src/gpu/sim.js:494-547 snapshot / restore; src/gpu/main.js:308-390
 * records the inputs from such a snapshot and replays them).  Here the state
 * that decides a scene's future is positions, previous positions, velocities
 * and inverse masses, active flags, the last densities, each constraint's
 * rest length and alive flag, the hose ring position, the substep count and
 * the component labels.  Stepping on from a restored state gives the same
 * bits as stepping on from the saved one. */

/* Device-resident copies in one of 8 slots the solver owns (no host
 * transfer; a slot keeps its device memory until the solver is destroyed). */
int fp_checkpoint_save(fp_solver* s, int32_t slot);
int fp_checkpoint_restore(fp_solver* s, int32_t slot);
/* A portable blob: header ("FPSTATE", version 1, the scene's shape) + the
 * arrays in device layout.  Import needs a solver of the same shape
 * (capacity, solids, maxFluid, constraints). */
int fp_state_size(fp_solver* s, int64_t* bytes);
int fp_state_export(fp_solver* s, void* buf, int64_t bytes);
int fp_state_import(fp_solver* s, const void* buf, int64_t bytes);

/* ---- batches: many independent scenes, one launch per pass ----------- */

/* No counterpart in the reference (it steps one scene).  A batch steps
 * every scene added to it with one kernel launch per pass: the kernels walk
 * a (scene, block) work list over a device table of the scenes' arrays and
 * parameters, and a frame without grabs is captured once into a HIP graph
 * and replayed.  Scenes keep their own state and API (emit, cut, grabs,
 * read_state ...); while in a batch they run on the batch's stream.  Every
 * scene must use the same constraint_iters. */
typedef struct fp_batch fp_batch;

typedef struct fp_batch_info {
  int32_t scenes;
  int32_t graph;               /* 1 while a captured frame is held */
  int64_t capacity;            /* particle slots over all scenes */
  int64_t device_bytes;        /* HBM held by the scenes */
  int64_t frames;              /* fp_batch_step calls */
  int64_t graph_launches;      /* frames run by replaying the captured graph */
  int32_t contact_lanes;       /* lanes per solid the solid contact pass uses: tuned over the
                                  first frames (4, 8, 16, 32 timed, the fastest kept; 0 while
                                  tuning).  Results do not depend on it. */
  int32_t graph_captures;      /* frames captured into a graph so far: a captured frame keeps
                                  replaying while members join and leave (its kernels read the
                                  scene table and the work lists on the device); it is captured
                                  again only when the substep or iteration count, the contact
                                  lanes, or the work lists' size (by more than 1.5x) change */
  int32_t density_stage;       /* 1: the density pass stages its pair-list entries in LDS */
  int32_t xcd_run;             /* work-list items dealt to each XCD in turn */
  int32_t shape_tuned;         /* 1 once the batch has chosen these two launch shapes: timed
                                  over captured frames after the contact lanes (or taken from
                                  an earlier batch of the same signature).  Results do not
                                  depend on them. */
} fp_batch_info;

int fp_batch_create(int32_t device, fp_batch** out);
int fp_batch_destroy(fp_batch* b);
const char* fp_batch_last_error(const fp_batch* b);
int fp_batch_add(fp_batch* b, fp_solver* s);
int fp_batch_remove(fp_batch* b, fp_solver* s);
/* One frame (substeps of dt/substeps) of every scene; returns after the
 * frame finished and every scene settled (tears, groups), like fp_step. */
int fp_batch_step(fp_batch* b, double dt, int32_t substeps);
int fp_batch_step_async(fp_batch* b, double dt, int32_t substeps);
int fp_batch_sync(fp_batch* b);
/* 0: launch the passes each frame instead of replaying a captured graph. */
int fp_batch_set_graph(fp_batch* b, int32_t enable);
/* Pin the lanes per solid of the solid contact pass (4, 8, 16 or 32), or 0 to
 * tune them over the next frames.  A launch-shape choice only: results are
 * the same for every value. */
int fp_batch_set_contact_lanes(fp_batch* b, int32_t lanes);
int fp_batch_get_info(const fp_batch* b, fp_batch_info* out);
int fp_batch_profile_enable(fp_batch* b, int32_t enable);
int fp_batch_profile_read(fp_batch* b, fp_pass_stat* out, int32_t max, int32_t* n);
/* Position snapshot of every scene of the batch after the frames queued so
 * far (the per-frame output of a serving loop): float32 x, y, z per particle
 * slot -- each scene's capacity, solids first, the layout of fp_read_state's
 * pos -- scenes in the order they were added.  One gather kernel and one
 * device-to-host copy into pinned memory the batch owns; returns when both
 * finished.  *pos stays valid until the second-next fp_batch_snapshot call
 * or fp_batch_destroy; (*offsets)[k] is scene k's first float and
 * (*offsets)[*scenes] the total. */
int fp_batch_snapshot(fp_batch* b, const float** pos, const int64_t** offsets, int32_t* scenes);

/* ---- device memory -------------------------------------------------- */

/* No counterpart in the reference (it allocates typed arrays per solver).
 * Every solver allocates from a per-device pool: its fixed buffers are one
 * block, blocks come in size classes of a quarter octave and return to the
 * pool when the solver is destroyed, so creating and destroying scenes
 * between frames neither allocates from the runtime (once warm) nor waits
 * for the device. */

/* Device bytes a solver with these parameters holds after
 * fp_add_solid_cube(n) and emits of at most max_fluid particles per call
 * (an admission estimate; needs no device). */
int fp_scene_bytes(const fp_params* p, int32_t max_fluid, int32_t cube_n, int64_t* bytes);
/* Bytes the pool of `device` (-1: current) holds from the runtime, and the
 * part of them cached for reuse. */
int fp_pool_info(int32_t device, int64_t* reserved, int64_t* cached);
/* Return the cached blocks to the runtime (waits for the device). */
int fp_pool_trim(int32_t device);

/* ---- tuning cache ---------------------------------------------------- */

/* No counterpart in the reference (its WGSL has compile-time workgroup
 * sizes).  Batches tune launch shapes that leave every result unchanged:
 * the solid contact pass's lanes per solid over their first frames, then
 * the density kernel's LDS staging and the work lists' XCD run length over
 * whole captured frames.  The choices are kept per group signature (device,
 * scene count, solid or slot count to the next power of two) for the
 * process; these calls keep them in a JSON file between processes:
 *   {"version": 1, "contact_lanes": {"<signature>": lanes, ...},
 *    "launch_shape": {"<signature>": candidate, ...}}
 * fp_tune_cache_load merges a file's entries, after which a new group of a
 * cached signature uses the cached lanes without tuning frames (a process
 * that loads no file always tunes its contact lanes).  Errors: FP_EIO for a
 * file that cannot be opened or written, FP_EINVAL for a malformed file or
 * an out-of-range value (nothing is merged then); fp_last_error(NULL) says
 * which.  fp_tune_cache_reset empties both and turns the lane cache off. */
int fp_tune_cache_save(const char* path);
int fp_tune_cache_load(const char* path);
int fp_tune_cache_reset(void);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDPHYSICS_H */
