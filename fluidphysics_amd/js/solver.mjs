// solver.mjs -- drop-in for the reference CPU solver's module (src/solver.js):
// `export class Solver`, `FLUID`, `SOLID`, the same method names, defaults,
// return shapes and mutable tuning fields, with every substep solved by the
// HIP kernels of libfluidphysics.so through the N-API addon (fp_napi.node).
//
//   import { Solver } from './fluidphysics_amd/js/solver.mjs';
//   const s = new Solver({ maxFluid: 2600 });
//   s.addSolidCube(0, 2.0, 0, 7, 0.18);
//   s.emitFluid(x, y, z, vx, vy, vz);
//   s.grabs = [{ i, x, y, z }];
//   s.step(1 / 60, 5);
//   const n = s.cut(ox, oy, oz, d1x, d1y, d1z, d2x, d2y, d2z);
//   const hit = s.pickSolid(ox, oy, oz, dx, dy, dz, 0.2);   // { id, t } | null
//   s.pos, s.vel, s.active, s.group                          // typed arrays
//
// Differences from the reference, all at the boundary:
//  * pos / prev / vel / invMass / active / group are refreshed from the
//    device on first access after a step (one device->host copy, then cached
//    until the next mutating call).  Writing into them does not reach the
//    device; use writeState() for that.
//  * emitFluid is buffered and sent in one batch before the next call that
//    needs it.
//  * device errors throw (the reference has none to throw).
import { createRequire } from 'module';
import { fileURLToPath } from 'url';
import path from 'path';

const require = createRequire(import.meta.url);
const HERE = path.dirname(fileURLToPath(import.meta.url));
const addon = require(path.join(HERE, '..', 'addon', 'fp_napi.node'));

export const FLUID = 0;
export const SOLID = 1;

// reference field names (src/solver.js:14-37) forwarded to fp_params
const TUNING = ['gravity', 'solidRadius', 'fluidRadius', 'compliance', 'plasticThreshold',
  'plasticRate', 'breakStrain', 'fluidH', 'restDensity', 'stiffness', 'nearStiffness',
  'viscosity', 'solidDamping', 'fluidDamping', 'floorFrictionSolid', 'floorFrictionFluid',
  'cullRadius', 'maxSpeed', 'grabStiffness', 'cellSize', 'constraintIters', 'jacobiOmega',
  'contactScale', 'gravityX', 'gravityZ', 'tetVolume', 'solidContactCache', 'tearGap'];

export class Solver {
  constructor({ maxFluid = 2600, params = {}, device = -1 } = {}) {
    this.maxFluid = maxFluid;
    this.solidCount = 0;
    this.capacity = 0;
    this._p = Object.assign(addon.defaults(), params);
    this._h = addon.create(maxFluid, this._p, device);
    this._dirtyParams = false;
    this._emits = [];
    this._views = null;
    this._grabsPushed = false;
    this.grabs = [];
    for (const k of TUNING) {
      Object.defineProperty(this, k, {
        get: () => this._p[k],
        set: (v) => { this._p[k] = v; this._dirtyParams = true; },
        enumerable: true,
      });
    }
    const p = this._p;
    const self = this;
    this.wall = {
      get x() { return p.wallX; }, set x(v) { p.wallX = v; self._dirtyParams = true; },
      get z() { return p.wallZ; }, set z(v) { p.wallZ = v; self._dirtyParams = true; },
      get height() { return p.wallHeight; }, set height(v) { p.wallHeight = v; self._dirtyParams = true; },
      get halfThick() { return p.wallHalfThick; }, set halfThick(v) { p.wallHalfThick = v; self._dirtyParams = true; },
    };
  }

  _flush() {
    if (this._dirtyParams) { addon.setParams(this._h, this._p); this._dirtyParams = false; }
    if (this._emits.length) {
      const n = this._emits.length;
      const pos = new Float64Array(3 * n), vel = new Float64Array(3 * n), mass = new Float64Array(n);
      for (let k = 0; k < n; k++) {
        const e = this._emits[k];
        pos[3 * k] = e[0]; pos[3 * k + 1] = e[1]; pos[3 * k + 2] = e[2];
        vel[3 * k] = e[3]; vel[3 * k + 1] = e[4]; vel[3 * k + 2] = e[5];
        mass[k] = e[6];
      }
      this._emits = [];
      addon.emit(this._h, pos, vel, mass, n);
    }
  }

  _stale() { this._views = null; }

  // addSolidCube (src/solver.js:72-123)
  addSolidCube(cx, cy, cz, n = 7, spacing = 0.18, particleMass = 0.08) {
    this._flush();
    addon.addSolidCube(this._h, cx, cy, cz, n, spacing, particleMass);
    const i = addon.info(this._h);
    this.solidCount = i.solidCount;
    this.capacity = i.capacity;
    this.tableSize = i.tableSize;
    this._stale();
  }

  // emitFluid (src/solver.js:139-149)
  emitFluid(x, y, z, vx, vy, vz, mass = 0.03) {
    this._emits.push([x, y, z, vx, vy, vz, mass]);
    this._stale();
  }

  // emitFluid over arrays in one transfer: pos/vel Float64Array(3n), mass number|Float64Array(n)
  emitBatch(pos, vel, mass = 0.03) {
    this._flush();
    const n = pos.length / 3;
    const m = typeof mass === 'number' ? new Float64Array(n).fill(mass) : mass;
    addon.emit(this._h, pos, vel, m, n);
    this._stale();
  }

  resetSolid() { this._flush(); addon.resetSolid(this._h); this.grabs.length = 0; this._stale(); }

  clearFluid() { this._flush(); addon.clearFluid(this._h); this._stale(); }

  fluidCount() { this._flush(); return addon.fluidCount(this._h); }
  // true once a finished step ended a particle with a non-finite position or
  // velocity (fp_nonfinite; sticky); the reference counts NaNs in its GPU
  // page's metrics (src/gpu/main.js:546-556)
  nonFinite() { return addon.nonFinite(this._h); }

  _pushGrabs() {
    const g = this.grabs;
    if (!g.length && !this._grabsPushed) return;  // nothing held now or at the last step
    this._grabsPushed = g.length > 0;
    const idx = new Int32Array(g.length), xyz = new Float64Array(3 * g.length);
    for (let k = 0; k < g.length; k++) {
      idx[k] = g[k].i; xyz[3 * k] = g[k].x; xyz[3 * k + 1] = g[k].y; xyz[3 * k + 2] = g[k].z;
    }
    addon.setGrabs(this._h, idx, xyz, g.length);
  }

  // step (src/solver.js:187-242); wait = false leaves the wait to sync()
  step(dt, substeps = 5, wait = true) {
    this._flush();
    this._pushGrabs();
    addon.step(this._h, dt, substeps, wait ? 1 : 0);
    this._stale();
  }

  sync() { addon.sync(this._h); }

  // cut (src/solver.js:458-486)
  cut(ox, oy, oz, d1x, d1y, d1z, d2x, d2y, d2z) {
    this._flush();
    const n = addon.cut(this._h, ox, oy, oz, d1x, d1y, d1z, d2x, d2y, d2z);
    this._stale();
    return n;
  }

  // pickSolid (src/solver.js:502-515)
  pickSolid(ox, oy, oz, dx, dy, dz, maxPerp = 0.2) {
    this._flush();
    return addon.pickSolid(this._h, ox, oy, oz, dx, dy, dz, maxPerp);
  }

  // the grab selection of src/main.js:293-301 as a device pass: ids of the
  // solids in solid sameGroupAs's component (any when -1) within radius of
  // (cx, cy, cz), ascending
  selectSolids(cx, cy, cz, radius = 0.35, sameGroupAs = -1) {
    this._flush();
    return addon.selectSolids(this._h, cx, cy, cz, radius, sameGroupAs);
  }

  // _recomputeGroups (solver.js:488-499) on the device over the constraints
  // alive now: {labels, count}, labels[i] the smallest solid index of solid
  // i's component (the partition of `group`, whose labels are the
  // reference's union-find roots).
  components() {
    this._flush();
    return addon.components(this._h);
  }

  _state() {
    if (!this._views) {
      this._flush();
      const n = this.capacity;
      const v = {
        pos: new Float32Array(3 * n), prev: new Float32Array(3 * n), vel: new Float32Array(3 * n),
        invMass: new Float32Array(n), active: new Uint8Array(n), group: new Int32Array(this.solidCount),
      };
      addon.readState(this._h, v.pos, v.prev, v.vel, v.invMass, v.active, v.group);
      this._views = v;
    }
    return this._views;
  }

  get pos() { return this._state().pos; }
  get prev() { return this._state().prev; }
  get vel() { return this._state().vel; }
  get invMass() { return this._state().invMass; }
  get active() { return this._state().active; }
  get group() { return this._state().group; }
  get type() {
    const t = new Uint8Array(this.capacity);
    t.fill(SOLID, 0, this.solidCount);
    return t;
  }

  get consAlive() {
    const c = addon.info(this._h).numConstraints;
    const alive = new Uint8Array(c);
    addon.readConstraints(this._h, null, null, null, alive);
    return alive;
  }

  writeState({ pos = null, prev = null, vel = null, invMass = null, active = null } = {}) {
    this._flush();
    addon.writeState(this._h, pos, prev, vel, invMass, active);
    this._stale();
  }

  readGrid() {
    const i = addon.info(this._h);
    const cs = new Int32Array(i.tableSize + 1), ce = new Int32Array(i.capacity);
    addon.readGrid(this._h, cs, ce);
    return { cellStart: cs, cellEntries: ce };
  }

  // the optional tet mode's tets (tetVolume): 4 vertex indices and an alive
  // flag per tet
  readTets() {
    const n = addon.readTets(this._h, null, null);
    const verts = new Int32Array(4 * n), alive = new Uint8Array(n);
    if (n) addon.readTets(this._h, verts, alive);
    return { verts, alive };
  }

  _buildGrid() { this._flush(); addon.buildGrid(this._h); }

  info() { return addon.info(this._h); }

  destroy() { if (this._h) { addon.destroy(this._h); this._h = null; } }
}

// Many independent scenes stepped together (no counterpart in the reference,
// which steps one scene): one kernel launch per pass over every member, the
// frame replayed from a HIP graph.  Members keep their Solver API between
// frames; all use the same constraintIters.
export class Batch {
  constructor({ device = -1 } = {}) {
    this._h = addon.batchCreate(device);
    this.scenes = [];
  }

  add(solver) { solver._flush(); addon.batchAdd(this._h, solver._h); this.scenes.push(solver); }

  remove(solver) {
    addon.batchRemove(this._h, solver._h);
    this.scenes.splice(this.scenes.indexOf(solver), 1);
  }

  // one frame of every member (src/solver.js:187-242 each)
  step(dt, substeps = 5, wait = true) {
    for (const s of this.scenes) { s._flush(); s._pushGrabs(); }
    addon.batchStep(this._h, dt, substeps, wait ? 1 : 0);
    for (const s of this.scenes) s._stale();
  }

  sync() { addon.batchSync(this._h); }

  graph(on = true) { addon.batchSetGraph(this._h, on ? 1 : 0); }

  // pin the solid contact pass's lanes per solid (4..32; 0 = tune); results do not depend on it
  contactLanes(lanes = 0) { addon.batchSetContactLanes(this._h, lanes); }

  info() { return addon.batchInfo(this._h); }

  // positions of every member after the frames queued so far: one device
  // copy; { pos: Float32Array (into `out` if given), offsets } with member k
  // at pos.subarray(offsets[k], offsets[k + 1])
  snapshot(out = undefined) { return out ? addon.batchSnapshot(this._h, out) : addon.batchSnapshot(this._h); }

  destroy() {
    if (this._h) { addon.batchDestroy(this._h); this._h = null; }
    this.scenes = [];
  }
}

// Device bytes a solver with these parameters holds after addSolidCube(n)
// and emits of at most maxFluid particles at once (an admission estimate).
export function sceneBytes(params, maxFluid, n) { return addon.sceneBytes(params, maxFluid, n); }

export function poolInfo(device = -1) { return addon.poolInfo(device); }
