/*
 * fp_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle / parity checker).
 *
 * A CPU restatement, in plain C, of the reference CPU solver
 * src/solver.js (class Solver).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Arithmetic model.  JS numbers are IEEE doubles and the solver's state
 * lives in Float32Array / Int32Array / Uint8Array; consRest is a plain
 * Array (double).  Every expression here is evaluated in double in the
 * reference's operation order, and every store into a typed array rounds
 * to the storage type ST:
 *   ST = float  -> reproduces node's solver.js bit for bit
 *                  (libfp_oracle32.so, pinned against golden vectors written
 *                  by the real solver.js, tests/golden/);
 *   ST = double -> the "FP64 replay" (libfp_oracle64.so).
 * Build with -ffp-contract=off and no fast-math (oracle/Makefile).
 *
 * Two solve modes:
 *   OR_MODE_GS      the reference's Gauss-Seidel order (solver.js:246-279,
 *                   :356-381); constraint_iters sweeps (1 = solver.js).
 *   OR_MODE_JACOBI  the MI355X formulation: contacts and distance constraints
 *                   gathered Jacobi-style from a snapshot (per-pair formulas of
 *                   solver.js:267-275 and :366-379), constraint_iters
 *                   iterations, symmetric per-pair scale jacobi_omega /
 *                   max(deg_a, deg_b) over the static CSR adjacency.
 *                   Everything else (grid, DDR fluid, bounds, velocity, XSPH)
 *                   is identical to GS mode because the reference already
 *                   evaluates it from a snapshot.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/fluidphysics.h"

#ifndef ST
#define ST float
#endif

#define OR_MODE_GS 0
#define OR_MODE_JACOBI 1

typedef struct Oracle {
  fp_params P;
  int maxFluid, solidCount, capacity, tableSize, fluidNext;
  ST *pos, *prev, *vel, *invMass;
  uint8_t *type, *active;
  int32_t *group;
  int32_t *cellStart, *cellEntries;
  int maxPairs, pairCount;
  int32_t *pairI, *pairJ;
  ST *pairD;
  ST *rho, *rhoNear, *press, *pressNear, *dispX, *dispY, *dispZ;
  int32_t *nbrCnt;
  int numCons;
  int32_t *consA, *consB;
  double *consRest;
  uint8_t *consAlive;
  ST *snapPos;
  double *snapRest;
  int32_t *adjOff, *adjCons; /* CSR: constraints incident to each solid */
  int componentsDirty;
  int nGrabs;
  int32_t *grabI;
  double *grabXYZ;
  ST *jpos;      /* Jacobi snapshot positions */
  double *jrest; /* Jacobi constraint snapshot */
  uint8_t *jalive;
  int mode;
  /* the GPU twin's optional solid modes (Jacobi mode only): volume tets and
     the solid contact cache */
  int cubeN;
  double spacing;
  int numTets;
  int32_t *tetV;      /* 4 vertices per tet, ordered so the rest volume is positive */
  double *tetV0;      /* rest volume */
  uint8_t *tetAlive;
  int32_t *tetAdjOff, *tetAdj; /* CSR: tets incident to each solid */
  double *tetScale;   /* 1 / max(incident tets of its 4 vertices) */
  int32_t *cacheIds, *cacheN;
  int cacheOverflow;
  ST *xpos;           /* snapshot of the solids for the extra step */
} Oracle;

#define OR_CACHE_K 16 /* solid contact partners kept per solid */

/* ---- helpers -------------------------------------------------------- */

/* ECMAScript ToInt32 followed by >>> 0 (ToUint32) of a double. */
static uint32_t js_touint32(double x) {
  /* every double of magnitude >= 2^53 is an integer, so below 2^63 the
   * truncated value fits int64 and wraps to the same residue mod 2^32 */
  if (fabs(x) < 9.2e18) return (uint32_t)(int64_t)x;
  if (!isfinite(x)) return 0u;
  double t = trunc(x);
  double m = fmod(t, 4294967296.0);
  if (m < 0) m += 4294967296.0;
  return (uint32_t)m;
}

/* _hash (solver.js:162-164): products in double, ToInt32 at `^`, >>> 0, & (T-1). */
static uint32_t js_hash(const Oracle* o, double xi, double yi, double zi) {
  uint32_t a = js_touint32(xi * 92837111.0);
  uint32_t b = js_touint32(yi * 689287499.0);
  uint32_t c = js_touint32(zi * 283923481.0);
  return (a ^ b ^ c) & (uint32_t)(o->tableSize - 1);
}

static double inv_cell(const Oracle* o) { return 1.0 / o->P.cell_size; }

/* _hashPos (solver.js:166-172). */
static uint32_t hash_pos(const Oracle* o, const ST* p, int i) {
  double inv = inv_cell(o);
  return js_hash(o, floor((double)p[3 * i] * inv), floor((double)p[3 * i + 1] * inv),
                 floor((double)p[3 * i + 2] * inv));
}

static int nextpow2_int(int v) {
  int t = 1;
  while (t < v) t <<= 1;
  return t;
}

void or_params_default(fp_params* p) {
  memset(p, 0, sizeof(*p));
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
  p->constraint_iters = 1;
  p->table_size = 0;
  p->jacobi_omega = 1.0;
  p->contact_scale = 1.0;
  p->tear_gap = 1.3;
}

/* ---- lifetime ------------------------------------------------------- */

Oracle* or_create(const fp_params* params, int maxFluid) {
  Oracle* o = (Oracle*)calloc(1, sizeof(Oracle));
  if (!o) return NULL;
  if (params)
    o->P = *params;
  else
    or_params_default(&o->P);
  o->maxFluid = maxFluid; /* solver.js:10 */
  return o;
}

static void free_arrays(Oracle* o) {
  void* ptrs[] = {o->pos,       o->prev,     o->vel,       o->invMass,   o->type,    o->active,
                  o->group,     o->cellStart, o->cellEntries, o->pairI,  o->pairJ,   o->pairD,
                  o->rho,       o->rhoNear,  o->press,     o->pressNear, o->dispX,   o->dispY,
                  o->dispZ,     o->nbrCnt,   o->consA,     o->consB,     o->consRest, o->consAlive,
                  o->snapPos,   o->snapRest, o->adjOff,    o->adjCons,   o->grabI,   o->grabXYZ,
                  o->jpos,      o->jrest,    o->jalive,    o->tetV,      o->tetV0,   o->tetAlive,
                  o->tetAdjOff, o->tetAdj,   o->tetScale,  o->cacheIds,  o->cacheN,  o->xpos};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
}

void or_destroy(Oracle* o) {
  if (!o) return;
  free_arrays(o);
  free(o);
}

void or_set_params(Oracle* o, const fp_params* p) { o->P = *p; }
void or_set_mode(Oracle* o, int mode) { o->mode = mode; }

/* _alloc (solver.js:40-70). */
static int alloc_state(Oracle* o, int n) {
  o->capacity = n;
  o->pos = (ST*)calloc(3 * (size_t)n, sizeof(ST));
  o->prev = (ST*)calloc(3 * (size_t)n, sizeof(ST));
  o->vel = (ST*)calloc(3 * (size_t)n, sizeof(ST));
  o->invMass = (ST*)calloc(n, sizeof(ST));
  o->type = (uint8_t*)calloc(n, 1);
  o->active = (uint8_t*)calloc(n, 1);
  o->group = (int32_t*)calloc(n, sizeof(int32_t));
  o->tableSize = o->P.table_size > 0 ? o->P.table_size : nextpow2_int(2 * n);
  o->cellStart = (int32_t*)calloc(o->tableSize + 1, sizeof(int32_t));
  o->cellEntries = (int32_t*)calloc(n, sizeof(int32_t));
  o->fluidNext = 0;
  o->maxPairs = o->maxFluid * 32;
  o->pairI = (int32_t*)calloc(o->maxPairs > 0 ? o->maxPairs : 1, sizeof(int32_t));
  o->pairJ = (int32_t*)calloc(o->maxPairs > 0 ? o->maxPairs : 1, sizeof(int32_t));
  o->pairD = (ST*)calloc(o->maxPairs > 0 ? o->maxPairs : 1, sizeof(ST));
  o->pairCount = 0;
  o->rho = (ST*)calloc(n, sizeof(ST));
  o->rhoNear = (ST*)calloc(n, sizeof(ST));
  o->press = (ST*)calloc(n, sizeof(ST));
  o->pressNear = (ST*)calloc(n, sizeof(ST));
  o->dispX = (ST*)calloc(n, sizeof(ST));
  o->dispY = (ST*)calloc(n, sizeof(ST));
  o->dispZ = (ST*)calloc(n, sizeof(ST));
  o->nbrCnt = (int32_t*)calloc(n, sizeof(int32_t));
  o->jpos = (ST*)calloc(3 * (size_t)n, sizeof(ST));
  return (o->pos && o->prev && o->vel && o->invMass && o->type && o->active && o->group &&
          o->cellStart && o->cellEntries && o->pairI && o->pairJ && o->pairD && o->rho &&
          o->rhoNear && o->press && o->pressNear && o->dispX && o->dispY && o->dispZ && o->nbrCnt &&
          o->jpos)
             ? 0
             : -1;
}

/* ---- the GPU twin's volume tets (src/gpu/sim.js:144-194) ------------- */

static double tet_signed_volume(const ST* p, int a, int b, int c, int d) {
  const double ab[3] = {(double)p[3 * b] - p[3 * a], (double)p[3 * b + 1] - p[3 * a + 1], (double)p[3 * b + 2] - p[3 * a + 2]};
  const double ac[3] = {(double)p[3 * c] - p[3 * a], (double)p[3 * c + 1] - p[3 * a + 1], (double)p[3 * c + 2] - p[3 * a + 2]};
  const double ad[3] = {(double)p[3 * d] - p[3 * a], (double)p[3 * d + 1] - p[3 * a + 1], (double)p[3 * d + 2] - p[3 * a + 2]};
  return ((ab[1] * ac[2] - ab[2] * ac[1]) * ad[0] + (ab[2] * ac[0] - ab[0] * ac[2]) * ad[1] +
          (ab[0] * ac[1] - ab[1] * ac[0]) * ad[2]) / 6.0;
}

/* Five tets per lattice cell, the split mirrored on checkerboard parity so
 * faces match (sim.js:151-166).  The twin keeps its vertex order and so
 * half its rest volumes come out negative, which its solver then treats as
 * dead tets (wgsl.js:432); here the last two vertices are swapped wherever
 * that makes the rest volume positive, so every tet holds its volume. */
static void build_tets(Oracle* o) {
  const int n = o->cubeN, ns = o->solidCount;
  const int cells = n > 1 ? (n - 1) * (n - 1) * (n - 1) : 0;
  o->numTets = 5 * cells;
  o->tetV = (int32_t*)malloc((4 * (size_t)o->numTets + 1) * sizeof(int32_t));
  o->tetV0 = (double*)malloc(((size_t)o->numTets + 1) * sizeof(double));
  o->tetAlive = (uint8_t*)malloc((size_t)o->numTets + 1);
  o->tetScale = (double*)malloc(((size_t)o->numTets + 1) * sizeof(double));
  int t = 0;
#define IDX(ax, ay, az) (((ax) * n + (ay)) * n + (az))
  for (int ix = 0; ix < n - 1; ix++)
    for (int iy = 0; iy < n - 1; iy++)
      for (int iz = 0; iz < n - 1; iz++) {
        const int A = IDX(ix, iy, iz), B = IDX(ix + 1, iy, iz), C = IDX(ix + 1, iy + 1, iz), D = IDX(ix, iy + 1, iz);
        const int E = IDX(ix, iy, iz + 1), F = IDX(ix + 1, iy, iz + 1), G = IDX(ix + 1, iy + 1, iz + 1),
                  H = IDX(ix, iy + 1, iz + 1);
        const int even[5][4] = {{A, C, F, H}, {A, B, C, F}, {A, C, D, H}, {A, E, F, H}, {C, F, G, H}};
        const int odd[5][4] = {{B, D, E, G}, {A, B, D, E}, {B, C, D, G}, {D, E, H, G}, {B, E, F, G}};
        const int(*five)[4] = (ix + iy + iz) % 2 == 0 ? even : odd;
        for (int k = 0; k < 5; k++, t++) {
          int v[4] = {five[k][0], five[k][1], five[k][2], five[k][3]};
          double V = tet_signed_volume(o->pos, v[0], v[1], v[2], v[3]);
          if (V < 0) {
            const int x = v[2];
            v[2] = v[3];
            v[3] = x;
            V = -V;
          }
          for (int q = 0; q < 4; q++) o->tetV[4 * t + q] = v[q];
          o->tetV0[t] = V;
          o->tetAlive[t] = 1;
        }
      }
#undef IDX
  o->tetAdjOff = (int32_t*)calloc(ns + 1, sizeof(int32_t));
  o->tetAdj = (int32_t*)malloc((4 * (size_t)o->numTets + 1) * sizeof(int32_t));
  for (int k = 0; k < 4 * o->numTets; k++) o->tetAdjOff[o->tetV[k] + 1]++;
  for (int i = 0; i < ns; i++) o->tetAdjOff[i + 1] += o->tetAdjOff[i];
  int32_t* fill = (int32_t*)calloc(ns + 1, sizeof(int32_t));
  for (int k = 0; k < o->numTets; k++)
    for (int q = 0; q < 4; q++) {
      const int i = o->tetV[4 * k + q];
      o->tetAdj[o->tetAdjOff[i] + fill[i]++] = k;
    }
  free(fill);
  for (int k = 0; k < o->numTets; k++) {
    int m = 1;
    for (int q = 0; q < 4; q++) {
      const int i = o->tetV[4 * k + q];
      const int deg = o->tetAdjOff[i + 1] - o->tetAdjOff[i];
      if (deg > m) m = deg;
    }
    o->tetScale[k] = 1.0 / m;
  }
}

/* addSolidCube (solver.js:72-123). */
int or_add_solid_cube(Oracle* o, double cx, double cy, double cz, int n, double spacing,
                      double particleMass) {
  if (o->pos) return -1;
  if (alloc_state(o, n * n * n + o->maxFluid)) return -1;
  const double half = (n - 1) * spacing * 0.5;
  int id = 0;
  for (int ix = 0; ix < n; ix++)
    for (int iy = 0; iy < n; iy++)
      for (int iz = 0; iz < n; iz++) {
        o->pos[3 * id] = (ST)(cx - half + ix * spacing);
        o->pos[3 * id + 1] = (ST)(cy - half + iy * spacing);
        o->pos[3 * id + 2] = (ST)(cz - half + iz * spacing);
        o->invMass[id] = (ST)(1 / particleMass);
        o->type[id] = 1;
        o->active[id] = 1;
        o->group[id] = 0;
        id++;
      }
  o->solidCount = id;

  /* count first, then fill in the reference's loop order (:99-116) */
  const double maxD2 = 3 * spacing * spacing + 1e-6;
  int cap = 13 * id + 1;
  o->consA = (int32_t*)malloc(cap * sizeof(int32_t));
  o->consB = (int32_t*)malloc(cap * sizeof(int32_t));
  o->consRest = (double*)malloc(cap * sizeof(double));
  o->consAlive = (uint8_t*)malloc(cap);
  int c = 0;
#define IDX(ax, ay, az) (((ax) * n + (ay)) * n + (az))
  for (int ix = 0; ix < n; ix++)
    for (int iy = 0; iy < n; iy++)
      for (int iz = 0; iz < n; iz++)
        for (int dx = -1; dx <= 1; dx++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
              int jx = ix + dx, jy = iy + dy, jz = iz + dz;
              if (jx < 0 || jy < 0 || jz < 0 || jx >= n || jy >= n || jz >= n) continue;
              int a = IDX(ix, iy, iz), b = IDX(jx, jy, jz);
              if (b <= a) continue;
              double d2 = (dx * dx + dy * dy + dz * dz) * spacing * spacing;
              if (d2 > maxD2) continue;
              o->consA[c] = a;
              o->consB[c] = b;
              o->consRest[c] = sqrt(d2);
              o->consAlive[c] = 1;
              c++;
            }
#undef IDX
  o->numCons = c;

  /* CSR adjacency (the static structure the Jacobi mode gathers over) */
  o->adjOff = (int32_t*)calloc(id + 1, sizeof(int32_t));
  o->adjCons = (int32_t*)malloc((2 * (size_t)c + 1) * sizeof(int32_t));
  for (int k = 0; k < c; k++) {
    o->adjOff[o->consA[k] + 1]++;
    o->adjOff[o->consB[k] + 1]++;
  }
  for (int i = 0; i < id; i++) o->adjOff[i + 1] += o->adjOff[i];
  int32_t* fill = (int32_t*)calloc(id, sizeof(int32_t));
  for (int k = 0; k < c; k++) {
    int a = o->consA[k], b = o->consB[k];
    o->adjCons[o->adjOff[a] + fill[a]++] = k;
    o->adjCons[o->adjOff[b] + fill[b]++] = k;
  }
  free(fill);
  o->jrest = (double*)malloc((c + 1) * sizeof(double));
  o->jalive = (uint8_t*)malloc(c + 1);

  o->snapPos = (ST*)malloc(3 * (size_t)id * sizeof(ST));
  memcpy(o->snapPos, o->pos, 3 * (size_t)id * sizeof(ST));
  o->snapRest = (double*)malloc((c + 1) * sizeof(double));
  memcpy(o->snapRest, o->consRest, c * sizeof(double));
  o->cubeN = n;
  o->spacing = spacing;
  build_tets(o);
  o->cacheIds = (int32_t*)calloc((size_t)id * OR_CACHE_K + 1, sizeof(int32_t));
  o->cacheN = (int32_t*)calloc(id + 1, sizeof(int32_t));
  o->xpos = (ST*)calloc(3 * (size_t)id + 1, sizeof(ST));
  return 0;
}

/* resetSolid (solver.js:125-137). */
void or_reset_solid(Oracle* o) {
  memcpy(o->pos, o->snapPos, 3 * (size_t)o->solidCount * sizeof(ST));
  memcpy(o->prev, o->snapPos, 3 * (size_t)o->solidCount * sizeof(ST));
  memset(o->vel, 0, 3 * (size_t)o->solidCount * sizeof(ST));
  memcpy(o->consRest, o->snapRest, o->numCons * sizeof(double));
  for (int c = 0; c < o->numCons; c++) o->consAlive[c] = 1;
  for (int i = 0; i < o->solidCount; i++) o->group[i] = 0;
  for (int t = 0; t < o->numTets; t++) o->tetAlive[t] = 1;
  o->nGrabs = 0;
}

/* emitFluid (solver.js:139-149). */
void or_emit(Oracle* o, double x, double y, double z, double vx, double vy, double vz, double mass) {
  int i = o->solidCount + o->fluidNext;
  o->fluidNext = (o->fluidNext + 1) % o->maxFluid;
  o->pos[3 * i] = (ST)x;
  o->pos[3 * i + 1] = (ST)y;
  o->pos[3 * i + 2] = (ST)z;
  o->prev[3 * i] = (ST)x;
  o->prev[3 * i + 1] = (ST)y;
  o->prev[3 * i + 2] = (ST)z;
  o->vel[3 * i] = (ST)vx;
  o->vel[3 * i + 1] = (ST)vy;
  o->vel[3 * i + 2] = (ST)vz;
  o->invMass[i] = (ST)(1 / mass);
  o->type[i] = 0;
  o->active[i] = 1;
}

void or_emit_batch(Oracle* o, const double* pos, const double* vel, const double* mass, int n) {
  for (int k = 0; k < n; k++)
    or_emit(o, pos[3 * k], pos[3 * k + 1], pos[3 * k + 2], vel[3 * k], vel[3 * k + 1],
            vel[3 * k + 2], mass ? mass[k] : 0.03);
}

/* clearFluid (solver.js:151-153). */
void or_clear_fluid(Oracle* o) {
  for (int i = o->solidCount; i < o->capacity; i++) o->active[i] = 0;
}

/* fluidCount (solver.js:155-159). */
int or_fluid_count(const Oracle* o) {
  int m = 0;
  for (int i = o->solidCount; i < o->capacity; i++) m += o->active[i];
  return m;
}

void or_set_grabs(Oracle* o, const int32_t* idx, const double* xyz, int n) {
  free(o->grabI);
  free(o->grabXYZ);
  o->grabI = NULL;
  o->grabXYZ = NULL;
  o->nGrabs = n;
  if (n <= 0) {
    o->nGrabs = 0;
    return;
  }
  o->grabI = (int32_t*)malloc(n * sizeof(int32_t));
  o->grabXYZ = (double*)malloc(3 * (size_t)n * sizeof(double));
  memcpy(o->grabI, idx, n * sizeof(int32_t));
  memcpy(o->grabXYZ, xyz, 3 * (size_t)n * sizeof(double));
}

/* ---- spatial hash (solver.js:174-184) -------------------------------- */

void or_build_grid(Oracle* o) {
  int32_t* cs = o->cellStart;
  memset(cs, 0, (o->tableSize + 1) * sizeof(int32_t));
  for (int i = 0; i < o->capacity; i++)
    if (o->active[i]) cs[hash_pos(o, o->pos, i)]++;
  int32_t sum = 0;
  for (int k = 0; k < o->tableSize; k++) {
    sum += cs[k];
    cs[k] = sum;
  }
  cs[o->tableSize] = sum;
  for (int i = 0; i < o->capacity; i++)
    if (o->active[i]) o->cellEntries[--cs[hash_pos(o, o->pos, i)]] = i;
}

/* Is (i, j), i < j, an alive constraint?  (pairSet.has(i*capacity+j), solver.js:265) */
static int bonded(const Oracle* o, int i, int j) {
  if (i >= o->solidCount || j >= o->solidCount) return 0;
  for (int k = o->adjOff[i]; k < o->adjOff[i + 1]; k++) {
    int c = o->adjCons[k];
    if (o->consAlive[c] && ((o->consA[c] == i && o->consB[c] == j) ||
                            (o->consA[c] == j && o->consB[c] == i)))
      return 1;
  }
  return 0;
}

/* ---- contacts -------------------------------------------------------- */

/* _solveContacts (solver.js:246-279): Gauss-Seidel, in place. */
static void solve_contacts_gs(Oracle* o) {
  ST* p = o->pos;
  const int32_t *cs = o->cellStart, *ce = o->cellEntries;
  const double inv = inv_cell(o);
  const double rs = o->P.solid_radius, rf = o->P.fluid_radius;
  for (int i = 0; i < o->solidCount; i++) {
    if (!o->active[i]) continue;
    const int i3 = i * 3;
    const double xi = floor(p[i3] * inv), yi = floor(p[i3 + 1] * inv), zi = floor(p[i3 + 2] * inv);
    const double wi = o->invMass[i];
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) {
          const uint32_t hsh = js_hash(o, xi + dx, yi + dy, zi + dz);
          for (int e = cs[hsh]; e < cs[hsh + 1]; e++) {
            const int j = ce[e];
            if (j <= i || !o->active[j]) continue;
            const int solidJ = o->type[j] == 1;
            if (solidJ && bonded(o, i, j)) continue;
            const int j3 = j * 3;
            const double ddx = (double)p[i3] - p[j3], ddy = (double)p[i3 + 1] - p[j3 + 1],
                         ddz = (double)p[i3 + 2] - p[j3 + 2];
            const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
            const double rsum = rs + (solidJ ? rs : rf);
            if (d2 >= rsum * rsum || d2 < 1e-12) continue;
            const double d = sqrt(d2);
            const double wj = o->invMass[j];
            const double f = (rsum - d) / (d * (wi + wj));
            p[i3] = (ST)(p[i3] + ddx * f * wi);
            p[i3 + 1] = (ST)(p[i3 + 1] + ddy * f * wi);
            p[i3 + 2] = (ST)(p[i3 + 2] + ddz * f * wi);
            p[j3] = (ST)(p[j3] - ddx * f * wj);
            p[j3 + 1] = (ST)(p[j3 + 1] - ddy * f * wj);
            p[j3 + 2] = (ST)(p[j3 + 2] - ddz * f * wj);
          }
        }
  }
}

/* Jacobi contacts: every active particle x gathers the reference's per-pair
 * correction (solver.js:267-275) from its partners in the snapshot: solids
 * from all non-bonded solids and all fluids, fluids from all solids.  Each
 * distinct bucket among the 27 probes is visited once (a second visit of the
 * same bucket finds the pair already resolved in the Gauss-Seidel order). */
/* The optional solid modes run in Jacobi mode, in the constraint iterations
 * of a scene that has distance constraints. */
static int extra_modes(const Oracle* o) {
  return o->mode == OR_MODE_JACOBI && (o->P.tet_volume || o->P.solid_contact_cache) && o->solidCount > 0 &&
         o->numCons > 0;
}

static void solve_contacts_jacobi(Oracle* o) {
  ST* p = o->pos;
  ST* s = o->jpos;
  memcpy(s, p, 3 * (size_t)o->capacity * sizeof(ST));
  const int32_t *cs = o->cellStart, *ce = o->cellEntries;
  const double inv = inv_cell(o);
  const double rs = o->P.solid_radius, rf = o->P.fluid_radius;
  const double scale = o->P.contact_scale;
  for (int x = 0; x < o->capacity; x++) {
    if (!o->active[x]) continue;
    const int solidX = o->type[x] == 1;
    const int x3 = 3 * x;
    const double cxf = floor(s[x3] * inv), cyf = floor(s[x3 + 1] * inv), czf = floor(s[x3 + 2] * inv);
    const double wx = o->invMass[x];
    uint32_t seen[27];
    int nseen = 0;
    double ax = 0, ay = 0, az = 0;
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) {
          const uint32_t hsh = js_hash(o, cxf + dx, cyf + dy, czf + dz);
          int dup = 0;
          for (int q = 0; q < nseen; q++) dup |= seen[q] == hsh;
          if (dup) continue;
          seen[nseen++] = hsh;
          for (int e = cs[hsh]; e < cs[hsh + 1]; e++) {
            const int y = ce[e];
            if (y == x || !o->active[y]) continue;
            const int solidY = o->type[y] == 1;
            if (!solidX && !solidY) continue;
            /* with the contact cache, solid pairs are resolved in the constraint iterations */
            if (solidX && solidY && o->P.solid_contact_cache && extra_modes(o)) continue;
            if (solidX && solidY && bonded(o, x < y ? x : y, x < y ? y : x)) continue;
            const int y3 = 3 * y;
            const double ddx = (double)s[x3] - s[y3], ddy = (double)s[x3 + 1] - s[y3 + 1],
                         ddz = (double)s[x3 + 2] - s[y3 + 2];
            const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
            const double rsum = (solidX ? rs : rf) + (solidY ? rs : rf);
            if (d2 >= rsum * rsum || d2 < 1e-12) continue;
            const double d = sqrt(d2);
            const double wy = o->invMass[y];
            const double f = (rsum - d) / (d * (wx + wy));
            ax += ddx * f * wx * scale;
            ay += ddy * f * wx * scale;
            az += ddz * f * wx * scale;
          }
        }
    p[x3] = (ST)(s[x3] + ax);
    p[x3 + 1] = (ST)(s[x3 + 1] + ay);
    p[x3 + 2] = (ST)(s[x3 + 2] + az);
  }
}

/* ---- double-density relaxation (solver.js:284-354) ------------------- */

static void solve_fluid(Oracle* o) {
  ST* p = o->pos;
  const int32_t *cs = o->cellStart, *ce = o->cellEntries;
  const double inv = inv_cell(o);
  const double H = o->P.fluid_h, H2 = H * H;
  const double maxPair = 0.05 * H;
  const double maxTotal = 0.35 * H;
  ST *rho = o->rho, *rhoNear = o->rhoNear, *press = o->press, *pressNear = o->pressNear;
  ST *dispX = o->dispX, *dispY = o->dispY, *dispZ = o->dispZ;
  for (int i = o->solidCount; i < o->capacity; i++) {
    rho[i] = 0;
    rhoNear[i] = 0;
  }
  int np = 0;
  for (int i = o->solidCount; i < o->capacity; i++) {
    if (!o->active[i]) continue;
    const int i3 = i * 3;
    const double xi = floor(p[i3] * inv), yi = floor(p[i3 + 1] * inv), zi = floor(p[i3 + 2] * inv);
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) {
          const uint32_t hsh = js_hash(o, xi + dx, yi + dy, zi + dz);
          for (int e = cs[hsh]; e < cs[hsh + 1]; e++) {
            const int j = ce[e];
            if (j <= i || !o->active[j] || o->type[j] != 0) continue;
            const int j3 = j * 3;
            const double ddx = (double)p[j3] - p[i3], ddy = (double)p[j3 + 1] - p[i3 + 1],
                         ddz = (double)p[j3 + 2] - p[i3 + 2];
            const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
            if (d2 >= H2 || np >= o->maxPairs) continue;
            const double d = sqrt(d2);
            const double q = 1 - d / H;
            o->pairI[np] = i;
            o->pairJ[np] = j;
            o->pairD[np] = (ST)d;
            np++;
            const double q2 = q * q;
            rho[i] = (ST)(rho[i] + q2);
            rho[j] = (ST)(rho[j] + q2);
            rhoNear[i] = (ST)(rhoNear[i] + q2 * q);
            rhoNear[j] = (ST)(rhoNear[j] + q2 * q);
          }
        }
  }
  o->pairCount = np;

  for (int i = o->solidCount; i < o->capacity; i++) {
    press[i] = (ST)(o->P.stiffness * (rho[i] - o->P.rest_density));
    pressNear[i] = (ST)(o->P.near_stiffness * rhoNear[i]);
    dispX[i] = 0;
    dispY[i] = 0;
    dispZ[i] = 0;
  }

  for (int k = 0; k < np; k++) {
    const int i = o->pairI[k], j = o->pairJ[k];
    const double d = o->pairD[k];
    if (d < 1e-9) continue;
    const double q = 1 - d / H;
    double D = 0.25 * (((double)press[i] + press[j]) * q + ((double)pressNear[i] + pressNear[j]) * q * q);
    if (D > maxPair)
      D = maxPair;
    else if (D < -maxPair)
      D = -maxPair;
    const int i3 = i * 3, j3 = j * 3;
    const double f = D / d;
    const double ddx = ((double)p[j3] - p[i3]) * f, ddy = ((double)p[j3 + 1] - p[i3 + 1]) * f,
                 ddz = ((double)p[j3 + 2] - p[i3 + 2]) * f;
    dispX[i] = (ST)(dispX[i] - ddx);
    dispY[i] = (ST)(dispY[i] - ddy);
    dispZ[i] = (ST)(dispZ[i] - ddz);
    dispX[j] = (ST)(dispX[j] + ddx);
    dispY[j] = (ST)(dispY[j] + ddy);
    dispZ[j] = (ST)(dispZ[j] + ddz);
  }

  for (int i = o->solidCount; i < o->capacity; i++) {
    if (!o->active[i]) continue;
    double sx = dispX[i], sy = dispY[i], sz = dispZ[i];
    const double s2 = sx * sx + sy * sy + sz * sz;
    if (s2 > maxTotal * maxTotal) {
      const double f = maxTotal / sqrt(s2);
      sx *= f;
      sy *= f;
      sz *= f;
    }
    const int i3 = i * 3;
    p[i3] = (ST)(p[i3] + sx);
    p[i3 + 1] = (ST)(p[i3 + 1] + sy);
    p[i3 + 2] = (ST)(p[i3 + 2] + sz);
  }
}

/* ---- distance constraints -------------------------------------------- */

static void kill_constraint(Oracle* o, int c) {
  o->consAlive[c] = 0; /* _killConstraint (solver.js:383-387) */
  o->componentsDirty = 1;
}

/* _solveConstraints (solver.js:356-381): one Gauss-Seidel sweep. */
static void solve_constraints_gs(Oracle* o, double h) {
  ST* p = o->pos;
  const double alpha = o->P.compliance / (h * h);
  for (int c = 0; c < o->numCons; c++) {
    if (!o->consAlive[c]) continue;
    const int a = o->consA[c], b = o->consB[c];
    const int a3 = a * 3, b3 = b * 3;
    const double wa = o->invMass[a], wb = o->invMass[b];
    const double w = wa + wb;
    if (w == 0) continue;
    const double dx = (double)p[a3] - p[b3], dy = (double)p[a3 + 1] - p[b3 + 1],
                 dz = (double)p[a3 + 2] - p[b3 + 2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d < 1e-9) continue;
    const double rest = o->consRest[c];
    if (d > rest * o->P.break_strain) {
      kill_constraint(o, c);
      continue;
    }
    const double C = d - rest;
    const double dl = -C / (w + alpha);
    const double f = dl / d;
    p[a3] = (ST)(p[a3] + dx * f * wa);
    p[a3 + 1] = (ST)(p[a3 + 1] + dy * f * wa);
    p[a3 + 2] = (ST)(p[a3 + 2] + dz * f * wa);
    p[b3] = (ST)(p[b3] - dx * f * wb);
    p[b3 + 1] = (ST)(p[b3 + 1] - dy * f * wb);
    p[b3 + 2] = (ST)(p[b3 + 2] - dz * f * wb);
    const double strain = C / rest;
    if (fabs(strain) > o->P.plastic_threshold) o->consRest[c] += C * o->P.plastic_rate;
  }
}

/* One Jacobi iteration over the CSR adjacency: every solid a gathers
 *   dl = -C/(w + alpha), f = dl/d, delta_a += (p_a - p_b) * f * w_a * s_ab,
 *   s_ab = omega / max(deg_a, deg_b)
 * from the snapshot (positions, rest, alive); tearing (d > rest*breakStrain)
 * and plasticity (|C/rest| > thr -> rest += C*rate) are decided from the same
 * snapshot, so both endpoints agree. */
static void solve_constraints_jacobi_iter(Oracle* o, double h) {
  ST* p = o->pos;
  ST* s = o->jpos;
  const int ns = o->solidCount;
  memcpy(s, p, 3 * (size_t)ns * sizeof(ST));
  memcpy(o->jrest, o->consRest, o->numCons * sizeof(double));
  memcpy(o->jalive, o->consAlive, o->numCons);
  const double alpha = o->P.compliance / (h * h);
  const double omega = o->P.jacobi_omega;
  for (int a = 0; a < ns; a++) {
    const int a3 = 3 * a;
    const double wa = o->invMass[a];
    const int degA = o->adjOff[a + 1] - o->adjOff[a];
    double cx = 0, cy = 0, cz = 0;
    for (int k = o->adjOff[a]; k < o->adjOff[a + 1]; k++) {
      const int c = o->adjCons[k];
      if (!o->jalive[c]) continue;
      const int b = o->consA[c] == a ? o->consB[c] : o->consA[c];
      const int b3 = 3 * b;
      const double wb = o->invMass[b];
      const double w = wa + wb;
      if (w == 0) continue;
      const double dx = (double)s[a3] - s[b3], dy = (double)s[a3 + 1] - s[b3 + 1],
                   dz = (double)s[a3 + 2] - s[b3 + 2];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      if (d < 1e-9) continue;
      const double rest = o->jrest[c];
      if (d > rest * o->P.break_strain) {
        if (o->consA[c] == a) kill_constraint(o, c);
        continue;
      }
      const double C = d - rest;
      const double dl = -C / (w + alpha);
      const double f = dl / d;
      const int degB = o->adjOff[b + 1] - o->adjOff[b];
      const double sc = omega / (double)(degA > degB ? degA : degB);
      cx += dx * f * wa * sc;
      cy += dy * f * wa * sc;
      cz += dz * f * wa * sc;
      const double strain = C / rest;
      if (o->consA[c] == a && fabs(strain) > o->P.plastic_threshold)
        o->consRest[c] = rest + C * o->P.plastic_rate;
    }
    p[a3] = (ST)(s[a3] + cx);
    p[a3 + 1] = (ST)(s[a3 + 1] + cy);
    p[a3 + 2] = (ST)(s[a3 + 2] + cz);
  }
}

/* ---- the GPU twin's optional solid modes (Jacobi mode) --------------- */

static int lattice_adjacent(const Oracle* o, int i, int j) {
  const int n = o->cubeN;
  const int dx = i / (n * n) - j / (n * n), dy = (i / n) % n - (j / n) % n, dz = i % n - j % n;
  return abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1;
}

/* Rest distance of two lattice sites (wgsl.js:374-381). */
static double lattice_rest(const Oracle* o, int a, int b) {
  const int n = o->cubeN;
  const double dx = a / (n * n) - b / (n * n), dy = (a / n) % n - (b / n) % n, dz = a % n - b % n;
  return o->spacing * sqrt(dx * dx + dy * dy + dz * dz);
}

/* Contact distance of two solids: one diameter, or just under the lattice
 * spacing for lattice neighbours whose bond is gone (a cut face, which then
 * rests at the spacing instead of being pushed apart; wgsl.js:537-541). */
static double solid_rsum(const Oracle* o, int i, int j) {
  const double rsum = 2 * o->P.solid_radius;
  return lattice_adjacent(o, i, j) ? fmin(rsum, o->spacing * 0.97) : rsum;
}

/* Collection margin: a solid diameter (wgsl.js:549), but never past the
 * cell size, so the 27 cells around a solid hold every partner in range. */
static double cache_margin(const Oracle* o) {
  return fmax(0.0, fmin(2 * o->P.solid_radius, o->P.cell_size - 2 * o->P.solid_radius));
}

/* collectContacts (wgsl.js:543-570): for every active solid, the solids
 * within contact distance + margin that are not bonded to it, from the grid
 * of the current positions; ascending index, the OR_CACHE_K smallest kept. */
static void collect_contacts(Oracle* o) {
  const ST* p = o->pos;
  const int32_t *cs = o->cellStart, *ce = o->cellEntries;
  const double inv = inv_cell(o), margin = cache_margin(o);
  for (int i = 0; i < o->solidCount; i++) {
    o->cacheN[i] = 0;
    if (!o->active[i]) continue;
    int32_t* ids = o->cacheIds + (size_t)i * OR_CACHE_K;
    int cnt = 0;
    const int i3 = 3 * i;
    const double cxf = floor(p[i3] * inv), cyf = floor(p[i3 + 1] * inv), czf = floor(p[i3 + 2] * inv);
    uint32_t seen[27];
    int nseen = 0;
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) {
          const uint32_t hsh = js_hash(o, cxf + dx, cyf + dy, czf + dz);
          int dup = 0;
          for (int q = 0; q < nseen; q++) dup |= seen[q] == hsh;
          if (dup) continue;
          seen[nseen++] = hsh;
          for (int e = cs[hsh]; e < cs[hsh + 1]; e++) {
            const int j = ce[e];
            if (j == i || j >= o->solidCount) continue;
            const int j3 = 3 * j;
            const double ddx = (double)p[i3] - p[j3], ddy = (double)p[i3 + 1] - p[j3 + 1],
                         ddz = (double)p[i3 + 2] - p[j3 + 2];
            const double lim = solid_rsum(o, i, j) + margin;
            if (ddx * ddx + ddy * ddy + ddz * ddz >= lim * lim) continue;
            if (bonded(o, i < j ? i : j, i < j ? j : i)) continue;
            /* insert j, keeping the smallest OR_CACHE_K indices in order */
            if (cnt == OR_CACHE_K && j > ids[OR_CACHE_K - 1]) {
              o->cacheOverflow++;
              continue;
            }
            if (cnt == OR_CACHE_K) o->cacheOverflow++;
            int k = cnt < OR_CACHE_K ? cnt++ : OR_CACHE_K - 1;
            while (k > 0 && ids[k - 1] > j) {
              ids[k] = ids[k - 1];
              k--;
            }
            ids[k] = j;
          }
        }
    o->cacheN[i] = cnt;
  }
}

static void cap_vec(double* v, double lim) {
  const double m2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  if (m2 > lim * lim) {
    const double f = lim / sqrt(m2);
    v[0] *= f;
    v[1] *= f;
    v[2] *= f;
  }
}

/* The extra step of one constraint iteration, from the snapshot of the
 * solids after the distance constraints, every active solid gathering
 *   tets (solveTets, wgsl.js:423-468): for each alive incident tet, with
 *     V = (pb - pa) x (pc - pa) . (pd - pa) / 6, C = V - V0 and the volume
 *     gradients g (summing to zero), the XPBD zero-compliance projection
 *     dp_i = -C w_i g_i / sum_q(w_q |g_q|^2) scaled by 1 / max(incident tets
 *     of the 4 vertices), each tet's share capped at half a solid radius;
 *   contacts (solidContacts, wgsl.js:573-596): for each cached partner
 *     closer than rsum, dd (rsum - d) / d * w_i / (w_i + w_j) * 0.5;
 * each sum capped at half a solid radius (applySolidContacts,
 * wgsl.js:598-606). */
static void solid_extra_step(Oracle* o) {
  const int ns = o->solidCount;
  ST* p = o->pos;
  ST* x = o->xpos;
  memcpy(x, p, 3 * (size_t)ns * sizeof(ST));
  const double cap = 0.5 * o->P.solid_radius;
  for (int i = 0; i < ns; i++) {
    if (!o->active[i]) continue;
    double dt[3] = {0, 0, 0}, dc[3] = {0, 0, 0};
    if (o->P.tet_volume) {
      for (int k = o->tetAdjOff[i]; k < o->tetAdjOff[i + 1]; k++) {
        const int t = o->tetAdj[k];
        if (!o->tetAlive[t]) continue;
        const int32_t* v = o->tetV + 4 * t;
        double q[4][3];
        for (int m = 0; m < 4; m++)
          for (int c = 0; c < 3; c++) q[m][c] = x[3 * v[m] + c];
        double b[3], c[3], d[3];
        for (int m = 0; m < 3; m++) {
          b[m] = q[1][m] - q[0][m];
          c[m] = q[2][m] - q[0][m];
          d[m] = q[3][m] - q[0][m];
        }
        double g[4][3];
        /* gb = (c x d) / 6, gc = (d x b) / 6, gd = (b x c) / 6, ga = -(gb + gc + gd) */
        g[1][0] = (c[1] * d[2] - c[2] * d[1]) / 6;
        g[1][1] = (c[2] * d[0] - c[0] * d[2]) / 6;
        g[1][2] = (c[0] * d[1] - c[1] * d[0]) / 6;
        g[2][0] = (d[1] * b[2] - d[2] * b[1]) / 6;
        g[2][1] = (d[2] * b[0] - d[0] * b[2]) / 6;
        g[2][2] = (d[0] * b[1] - d[1] * b[0]) / 6;
        g[3][0] = (b[1] * c[2] - b[2] * c[1]) / 6;
        g[3][1] = (b[2] * c[0] - b[0] * c[2]) / 6;
        g[3][2] = (b[0] * c[1] - b[1] * c[0]) / 6;
        for (int m = 0; m < 3; m++) g[0][m] = -(g[1][m] + g[2][m] + g[3][m]);
        const double V = g[3][0] * d[0] + g[3][1] * d[1] + g[3][2] * d[2];
        const double C = V - o->tetV0[t];
        double denom = 0;
        for (int m = 0; m < 4; m++)
          denom += o->invMass[v[m]] * (g[m][0] * g[m][0] + g[m][1] * g[m][1] + g[m][2] * g[m][2]);
        if (denom < 1e-30) continue;
        int me = 0;
        while (v[me] != i) me++;
        const double f = -C / denom * o->invMass[i] * o->tetScale[t];
        double c1[3] = {g[me][0] * f, g[me][1] * f, g[me][2] * f};
        cap_vec(c1, cap);
        dt[0] += c1[0];
        dt[1] += c1[1];
        dt[2] += c1[2];
      }
      cap_vec(dt, cap);
    }
    if (o->P.solid_contact_cache) {
      const double wi = o->invMass[i];
      for (int k = 0; k < o->cacheN[i]; k++) {
        const int j = o->cacheIds[(size_t)i * OR_CACHE_K + k];
        const double ddx = (double)x[3 * i] - x[3 * j], ddy = (double)x[3 * i + 1] - x[3 * j + 1],
                     ddz = (double)x[3 * i + 2] - x[3 * j + 2];
        const double rsum = solid_rsum(o, i, j);
        const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
        if (d2 >= rsum * rsum || d2 < 1e-12) continue;
        const double wsum = wi + o->invMass[j];
        if (wsum == 0) continue;
        const double d = sqrt(d2);
        const double f = (rsum - d) / d * wi / wsum * 0.5;
        dc[0] += ddx * f;
        dc[1] += ddy * f;
        dc[2] += ddz * f;
      }
      cap_vec(dc, cap);
    }
    for (int m = 0; m < 3; m++) p[3 * i + m] = (ST)(x[3 * i + m] + dt[m] + dc[m]);
  }
}

/* tetUpdate (wgsl.js:471-489): a tet dies when an edge outgrows its lattice
 * rest by tear_gap solid diameters. */
static void tet_update(Oracle* o) {
  const ST* p = o->pos;
  const double gap = o->P.tear_gap * 2 * o->P.solid_radius;
  for (int t = 0; t < o->numTets; t++) {
    if (!o->tetAlive[t]) continue;
    const int32_t* v = o->tetV + 4 * t;
    for (int q = 0; q < 3 && o->tetAlive[t]; q++)
      for (int r = q + 1; r < 4; r++) {
        const int a = v[q], b = v[r];
        const double dx = (double)p[3 * a] - p[3 * b], dy = (double)p[3 * a + 1] - p[3 * b + 1],
                     dz = (double)p[3 * a + 2] - p[3 * b + 2];
        if (sqrt(dx * dx + dy * dy + dz * dz) > lattice_rest(o, a, b) + gap) {
          o->tetAlive[t] = 0;
          break;
        }
      }
  }
}


/* ---- bounds (solver.js:389-427) -------------------------------------- */

static double js_sign_or_one(double q) {
  /* Math.sign(q) || 1 : +-1 for nonzero, 1 for +-0 and NaN */
  if (q > 0) return 1;
  if (q < 0) return -1;
  return 1;
}

static void solve_bounds(Oracle* o) {
  ST *p = o->pos, *pv = o->prev;
  const double wx = o->P.wall_x, wz = o->P.wall_z, wh = o->P.wall_height, ht = o->P.wall_half_thick;
  for (int i = 0; i < o->capacity; i++) {
    if (!o->active[i]) continue;
    const int i3 = i * 3;
    const int solid = o->type[i] == 1;
    const double r = solid ? o->P.solid_radius : o->P.fluid_radius;
    if (p[i3 + 1] < r) {
      p[i3 + 1] = (ST)r;
      const double mu = solid ? o->P.floor_friction_solid : o->P.floor_friction_fluid;
      p[i3] = (ST)(p[i3] - ((double)p[i3] - pv[i3]) * mu);
      p[i3 + 2] = (ST)(p[i3 + 2] - ((double)p[i3 + 2] - pv[i3 + 2]) * mu);
    }
    if (p[i3 + 1] < wh + r) {
      const double lim = r + ht;
      for (int ax = 0; ax <= 2; ax += 2) {
        const double wpos = ax == 0 ? wx : wz;
        const double q = p[i3 + ax];
        if (fabs(fabs(q) - wpos) < lim) {
          const double side = js_sign_or_one(q);
          const int inside = fabs(q) < wpos;
          p[i3 + ax] = (ST)(inside ? side * (wpos - lim) : side * (wpos + lim));
        }
      }
    }
    if (!solid) {
      const double dx = p[i3], dz = p[i3 + 2];
      if (p[i3 + 1] < -3 || dx * dx + dz * dz > o->P.cull_radius * o->P.cull_radius) o->active[i] = 0;
    }
  }
}

/* ---- XSPH (solver.js:431-454) ----------------------------------------- */

static void apply_viscosity(Oracle* o) {
  ST* v = o->vel;
  const double k = o->P.viscosity;
  for (int i = o->solidCount; i < o->capacity; i++) {
    o->dispX[i] = 0;
    o->dispY[i] = 0;
    o->dispZ[i] = 0;
    o->nbrCnt[i] = 0;
  }
  for (int q = 0; q < o->pairCount; q++) {
    const int i = o->pairI[q], j = o->pairJ[q];
    const int i3 = i * 3, j3 = j * 3;
    o->dispX[i] = (ST)((double)o->dispX[i] + v[j3]);
    o->dispY[i] = (ST)((double)o->dispY[i] + v[j3 + 1]);
    o->dispZ[i] = (ST)((double)o->dispZ[i] + v[j3 + 2]);
    o->nbrCnt[i]++;
    o->dispX[j] = (ST)((double)o->dispX[j] + v[i3]);
    o->dispY[j] = (ST)((double)o->dispY[j] + v[i3 + 1]);
    o->dispZ[j] = (ST)((double)o->dispZ[j] + v[i3 + 2]);
    o->nbrCnt[j]++;
  }
  for (int i = o->solidCount; i < o->capacity; i++) {
    const int m = o->nbrCnt[i];
    if (!m || !o->active[i]) continue;
    const int i3 = i * 3;
    v[i3] = (ST)(v[i3] + k * ((double)o->dispX[i] / m - v[i3]));
    v[i3 + 1] = (ST)(v[i3 + 1] + k * ((double)o->dispY[i] / m - v[i3 + 1]));
    v[i3 + 2] = (ST)(v[i3 + 2] + k * ((double)o->dispZ[i] / m - v[i3 + 2]));
  }
}

/* ---- connected components (solver.js:488-499) ----------------------- */

void or_recompute_groups(Oracle* o) {
  const int n = o->solidCount;
  int32_t* parent = (int32_t*)malloc((n > 0 ? n : 1) * sizeof(int32_t));
  for (int i = 0; i < n; i++) parent[i] = i;
#define FIND(x0, out)                            \
  do {                                           \
    int x_ = (x0);                               \
    while (parent[x_] != x_) {                   \
      parent[x_] = parent[parent[x_]];           \
      x_ = parent[x_];                           \
    }                                            \
    (out) = x_;                                  \
  } while (0)
  for (int c = 0; c < o->numCons; c++) {
    if (!o->consAlive[c]) continue;
    int ra, rb;
    FIND(o->consA[c], ra);
    FIND(o->consB[c], rb);
    if (ra != rb) parent[ra] = rb;
  }
  for (int i = 0; i < n; i++) {
    int r;
    FIND(i, r);
    o->group[i] = r;
  }
#undef FIND
  free(parent);
}

/* ---- the step (solver.js:187-242) ------------------------------------ */

/* Integrate (solver.js:194-202): v.y += g h; prev = p; p += v h for every
 * active particle.  The x and z components of gravity (a tilted basin; the
 * reference has y only, :15, :197) are added the same way when set -- an
 * unset component adds nothing, so a zero velocity keeps its sign bit and
 * the float32 build stays bit-identical to solver.js. */
static void integrate(Oracle* o, double h) {
  ST *p = o->pos, *pv = o->prev, *v = o->vel;
  const double gx = o->P.gravity_x, gz = o->P.gravity_z;
  for (int i = 0; i < o->capacity; i++) {
    if (!o->active[i]) continue;
    const int i3 = i * 3;
    if (gx != 0) v[i3] = (ST)(v[i3] + gx * h);
    v[i3 + 1] = (ST)(v[i3 + 1] + o->P.gravity * h);
    if (gz != 0) v[i3 + 2] = (ST)(v[i3 + 2] + gz * h);
    pv[i3] = p[i3];
    pv[i3 + 1] = p[i3 + 1];
    pv[i3 + 2] = p[i3 + 2];
    p[i3] = (ST)(p[i3] + (double)v[i3] * h);
    p[i3 + 1] = (ST)(p[i3 + 1] + (double)v[i3 + 1] * h);
    p[i3 + 2] = (ST)(p[i3 + 2] + (double)v[i3 + 2] * h);
  }
}

/* Grab pull (solver.js:205-210): every {i, x, y, z} in list order. */
static void apply_grabs(Oracle* o) {
  ST* p = o->pos;
  for (int g = 0; g < o->nGrabs; g++) {
    const int i3 = o->grabI[g] * 3;
    const double k = o->P.grab_stiffness;
    p[i3] = (ST)(p[i3] + (o->grabXYZ[3 * g] - p[i3]) * k);
    p[i3 + 1] = (ST)(p[i3 + 1] + (o->grabXYZ[3 * g + 1] - p[i3 + 1]) * k);
    p[i3 + 2] = (ST)(p[i3 + 2] + (o->grabXYZ[3 * g + 2] - p[i3 + 2]) * k);
  }
}

void or_substep(Oracle* o, double h) {
  const double invH = 1 / h;
  ST *p = o->pos, *pv = o->prev, *v = o->vel;
  integrate(o, h);
  apply_grabs(o);
  or_build_grid(o);
  if (extra_modes(o) && o->P.solid_contact_cache) collect_contacts(o);
  if (o->mode == OR_MODE_JACOBI)
    solve_contacts_jacobi(o);
  else
    solve_contacts_gs(o);
  solve_fluid(o);
  const int iters = o->P.constraint_iters > 0 ? o->P.constraint_iters : 1;
  for (int it = 0; it < iters; it++) {
    if (o->mode == OR_MODE_JACOBI) {
      solve_constraints_jacobi_iter(o, h);
      if (extra_modes(o)) solid_extra_step(o);
    } else {
      solve_constraints_gs(o, h);
    }
  }
  if (extra_modes(o) && o->P.tet_volume) tet_update(o);
  solve_bounds(o);
  const double dampS = fmax(0, 1 - o->P.solid_damping * h);
  const double dampF = fmax(0, 1 - o->P.fluid_damping * h);
  const double vmax = o->P.max_speed;
  for (int i = 0; i < o->capacity; i++) {
    if (!o->active[i]) continue;
    const int i3 = i * 3;
    v[i3] = (ST)(((double)p[i3] - pv[i3]) * invH);
    v[i3 + 1] = (ST)(((double)p[i3 + 1] - pv[i3 + 1]) * invH);
    v[i3 + 2] = (ST)(((double)p[i3 + 2] - pv[i3 + 2]) * invH);
    const double damp = o->type[i] == 1 ? dampS : dampF;
    v[i3] = (ST)(v[i3] * damp);
    v[i3 + 1] = (ST)(v[i3 + 1] * damp);
    v[i3 + 2] = (ST)(v[i3 + 2] * damp);
    const double sp2 = (double)v[i3] * v[i3] + (double)v[i3 + 1] * v[i3 + 1] + (double)v[i3 + 2] * v[i3 + 2];
    if (sp2 > vmax * vmax) {
      const double f = vmax / sqrt(sp2);
      v[i3] = (ST)(v[i3] * f);
      v[i3 + 1] = (ST)(v[i3 + 1] * f);
      v[i3 + 2] = (ST)(v[i3 + 2] * f);
    }
  }
  apply_viscosity(o);
}

/* Test hook: the first stages of a substep (integrate, grab, grid, contacts),
 * leaving the post-contact positions in pos -- what the fluid passes see. */
void or_substep_until_contacts(Oracle* o, double h) {
  integrate(o, h);
  apply_grabs(o);
  or_build_grid(o);
  if (o->mode == OR_MODE_JACOBI)
    solve_contacts_jacobi(o);
  else
    solve_contacts_gs(o);
}

/* Test hooks: one pass alone on the current state, in the current mode --
 * the contacts over the grid of the current positions, or `iters`
 * constraint sweeps / iterations with compliance for substep h. */
void or_solve_contacts(Oracle* o) {
  or_build_grid(o);
  if (o->mode == OR_MODE_JACOBI)
    solve_contacts_jacobi(o);
  else
    solve_contacts_gs(o);
}

void or_solve_constraints(Oracle* o, double h, int iters) {
  for (int it = 0; it < iters; it++) {
    if (o->mode == OR_MODE_JACOBI) {
      solve_constraints_jacobi_iter(o, h);
      if (extra_modes(o)) solid_extra_step(o);
    } else {
      solve_constraints_gs(o, h);
    }
  }
}

/* Test hooks of the optional modes: the contact cache of the current
 * positions (builds the grid), one extra step alone, the tet tear check, and
 * the tets' state. */
void or_collect_contacts(Oracle* o) {
  or_build_grid(o);
  collect_contacts(o);
}
void or_solid_extra_step(Oracle* o) { solid_extra_step(o); }
void or_tet_update(Oracle* o) { tet_update(o); }
int or_num_tets(const Oracle* o) { return o->numTets; }
void or_read_tets(const Oracle* o, int32_t* v, double* v0, uint8_t* alive) {
  if (v) memcpy(v, o->tetV, 4 * (size_t)o->numTets * sizeof(int32_t));
  if (v0) memcpy(v0, o->tetV0, (size_t)o->numTets * sizeof(double));
  if (alive) memcpy(alive, o->tetAlive, (size_t)o->numTets);
}
void or_write_tet_alive(Oracle* o, const uint8_t* alive) { memcpy(o->tetAlive, alive, (size_t)o->numTets); }
void or_read_cache(const Oracle* o, int32_t* ids, int32_t* n, int32_t* overflow) {
  if (ids) memcpy(ids, o->cacheIds, (size_t)o->solidCount * OR_CACHE_K * sizeof(int32_t));
  if (n) memcpy(n, o->cacheN, (size_t)o->solidCount * sizeof(int32_t));
  if (overflow) *overflow = o->cacheOverflow;
}

void or_step(Oracle* o, double dt, int substeps) {
  const double h = dt / substeps;
  for (int s = 0; s < substeps; s++) or_substep(o, h);
  if (o->componentsDirty) {
    or_recompute_groups(o);
    o->componentsDirty = 0;
  }
}

/* ---- knife (solver.js:458-486) --------------------------------------- */

/* Does segment a-b cross the wedge (solver.js:466-483, double arithmetic in
 * the reference's operation order)? */
static int segment_in_wedge(const ST* p, int a, int b, double ox, double oy, double oz, double d1x, double d1y,
                            double d1z, double d2x, double d2y, double d2z, double nx, double ny, double nz) {
  const int a3 = a * 3, b3 = b * 3;
  const double sa = ((double)p[a3] - ox) * nx + ((double)p[a3 + 1] - oy) * ny + ((double)p[a3 + 2] - oz) * nz;
  const double sb = ((double)p[b3] - ox) * nx + ((double)p[b3 + 1] - oy) * ny + ((double)p[b3 + 2] - oz) * nz;
  if (sa * sb >= 0) return 0;
  const double t = sa / (sa - sb);
  const double px = p[a3] + ((double)p[b3] - p[a3]) * t - ox;
  const double py = p[a3 + 1] + ((double)p[b3 + 1] - p[a3 + 1]) * t - oy;
  const double pz = p[a3 + 2] + ((double)p[b3 + 2] - p[a3 + 2]) * t - oz;
  const double c1 = (d1y * pz - d1z * py) * nx + (d1z * px - d1x * pz) * ny + (d1x * py - d1y * px) * nz;
  const double c2 = (py * d2z - pz * d2y) * nx + (pz * d2x - px * d2z) * ny + (px * d2y - py * d2x) * nz;
  return !(c1 < 0 || c2 < 0);
}

int or_cut(Oracle* o, double ox, double oy, double oz, double d1x, double d1y, double d1z, double d2x,
           double d2y, double d2z) {
  const double nx = d1y * d2z - d1z * d2y;
  const double ny = d1z * d2x - d1x * d2z;
  const double nz = d1x * d2y - d1y * d2x;
  const double nLen = sqrt(nx * nx + ny * ny + nz * nz);
  if (nLen < 1e-8) return 0;
  const ST* p = o->pos;
  int cuts = 0;
  for (int c = 0; c < o->numCons; c++) {
    if (!o->consAlive[c]) continue;
    if (!segment_in_wedge(p, o->consA[c], o->consB[c], ox, oy, oz, d1x, d1y, d1z, d2x, d2y, d2z, nx, ny, nz))
      continue;
    kill_constraint(o, c);
    cuts++;
  }
  /* tets whose edges cross the wedge stop holding volume across the cut
     (the GPU twin's cutTets, wgsl.js:492-519, with this predicate) */
  if (o->P.tet_volume)
    for (int t = 0; t < o->numTets; t++) {
      if (!o->tetAlive[t]) continue;
      const int32_t* v = o->tetV + 4 * t;
      int hit = 0;
      for (int q = 0; q < 3 && !hit; q++)
        for (int r = q + 1; r < 4 && !hit; r++)
          hit = segment_in_wedge(p, v[q], v[r], ox, oy, oz, d1x, d1y, d1z, d2x, d2y, d2z, nx, ny, nz);
      if (hit) o->tetAlive[t] = 0;
    }
  /* the reference recomputes groups at the end of the next step() */
  return cuts;
}

/* ---- pickSolid (solver.js:502-515) ------------------------------------ */

int or_pick_solid(const Oracle* o, double ox, double oy, double oz, double dx, double dy, double dz,
                  double maxPerp, double* tOut) {
  const ST* p = o->pos;
  int best = -1;
  double bestT = INFINITY;
  for (int i = 0; i < o->solidCount; i++) {
    const int i3 = i * 3;
    const double px = p[i3] - ox, py = p[i3 + 1] - oy, pz = p[i3 + 2] - oz;
    const double t = px * dx + py * dy + pz * dz;
    if (t < 0 || t >= bestT) continue;
    const double qx = px - dx * t, qy = py - dy * t, qz = pz - dz * t;
    if (qx * qx + qy * qy + qz * qz > maxPerp * maxPerp) continue;
    best = i;
    bestT = t;
  }
  if (tOut) *tOut = bestT;
  return best;
}

/* ---- state access for the tests --------------------------------------- */

typedef struct or_info {
  int32_t capacity, solid_count, max_fluid, table_size, num_constraints, fluid_next, pair_count,
      components_dirty;
} or_info;

void or_get_info(const Oracle* o, or_info* out) {
  out->capacity = o->capacity;
  out->solid_count = o->solidCount;
  out->max_fluid = o->maxFluid;
  out->table_size = o->tableSize;
  out->num_constraints = o->numCons;
  out->fluid_next = o->fluidNext;
  out->pair_count = o->pairCount;
  out->components_dirty = o->componentsDirty;
}

/* Copy out (float32 views of the storage; pass NULL to skip). */
void or_read_state(const Oracle* o, float* pos, float* prev, float* vel, float* invMass,
                   uint8_t* active, int32_t* group) {
  const size_t n = o->capacity;
  for (size_t k = 0; k < 3 * n; k++) {
    if (pos) pos[k] = (float)o->pos[k];
    if (prev) prev[k] = (float)o->prev[k];
    if (vel) vel[k] = (float)o->vel[k];
  }
  for (size_t i = 0; i < n; i++) {
    if (invMass) invMass[i] = (float)o->invMass[i];
    if (active) active[i] = o->active[i];
  }
  if (group)
    for (int i = 0; i < o->solidCount; i++) group[i] = o->group[i];
}

/* Full-precision copy out (double views). */
void or_read_state64(const Oracle* o, double* pos, double* prev, double* vel) {
  const size_t n = 3 * (size_t)o->capacity;
  for (size_t k = 0; k < n; k++) {
    if (pos) pos[k] = (double)o->pos[k];
    if (prev) prev[k] = (double)o->prev[k];
    if (vel) vel[k] = (double)o->vel[k];
  }
}

void or_write_state(Oracle* o, const float* pos, const float* prev, const float* vel,
                    const float* invMass, const uint8_t* active) {
  const size_t n = o->capacity;
  for (size_t k = 0; k < 3 * n; k++) {
    if (pos) o->pos[k] = (ST)pos[k];
    if (prev) o->prev[k] = (ST)prev[k];
    if (vel) o->vel[k] = (ST)vel[k];
  }
  for (size_t i = 0; i < n; i++) {
    if (invMass) o->invMass[i] = (ST)invMass[i];
    if (active) o->active[i] = active[i];
  }
}

void or_read_grid(const Oracle* o, int32_t* cellStart, int32_t* cellEntries) {
  if (cellStart) memcpy(cellStart, o->cellStart, (o->tableSize + 1) * sizeof(int32_t));
  if (cellEntries) memcpy(cellEntries, o->cellEntries, o->capacity * sizeof(int32_t));
}

void or_read_constraints(const Oracle* o, int32_t* a, int32_t* b, double* rest, uint8_t* alive) {
  for (int c = 0; c < o->numCons; c++) {
    if (a) a[c] = o->consA[c];
    if (b) b[c] = o->consB[c];
    if (rest) rest[c] = o->consRest[c];
    if (alive) alive[c] = o->consAlive[c];
  }
}

void or_write_constraints(Oracle* o, const double* rest, const uint8_t* alive) {
  for (int c = 0; c < o->numCons; c++) {
    if (rest) o->consRest[c] = rest[c];
    if (alive) o->consAlive[c] = alive[c];
  }
}

void or_read_density(const Oracle* o, float* rho, float* rhoNear) {
  for (int i = 0; i < o->capacity; i++) {
    if (rho) rho[i] = (float)o->rho[i];
    if (rhoNear) rhoNear[i] = (float)o->rhoNear[i];
  }
}

/* Pair list of the last substep (pairI, pairJ; pairCount entries). */
void or_read_pairs(const Oracle* o, int32_t* pi, int32_t* pj) {
  for (int k = 0; k < o->pairCount; k++) {
    if (pi) pi[k] = o->pairI[k];
    if (pj) pj[k] = o->pairJ[k];
  }
}

uint32_t or_hash_cell(const Oracle* o, double xi, double yi, double zi) { return js_hash(o, xi, yi, zi); }
