// Writes the golden vectors of tests/golden/ by running the reference CPU
// solver (src/solver.js, class Solver) under node on the scenario inputs that
// gen_inputs.py wrote.  Run in the development container only (the reference
// is not shipped anywhere):  sh tests/golden/make_golden.sh
//
// Outputs, per scenario <name>:
//   <name>_trace.json      per-frame SHA-256 digests of pos / vel / active /
//                          consAlive / consRest / group and the per-frame
//                          statistics (kinetic energy, mean density, alive
//                          constraints, fluid count, pair count), cut counts,
//                          pick results
//   <name>_grid0.bin.gz    cellStart + cellEntries of _buildGrid() on the
//                          initial state
//   <name>_sub1.bin.gz     full state after one substep from the initial state
//   <name>_f<k>.bin.gz     full state after frame k (dumpFrames)
// Binary layout (little endian): pos f32[3N] prev f32[3N] vel f32[3N]
// rho f32[N] rhoNear f32[N] group i32[Ns] consRest f64[C] active u8[N]
// consAlive u8[C].
import fs from 'fs';
import path from 'path';
import crypto from 'crypto';
import zlib from 'zlib';
import { pathToFileURL, fileURLToPath } from 'url';

const REF = process.env.FP_REFERENCE_SOLVER || '/root/reference/src/solver.js';
const HERE = path.dirname(fileURLToPath(import.meta.url));

const sha = (t) => crypto.createHash('sha256')
  .update(Buffer.from(t.buffer, t.byteOffset, t.byteLength)).digest('hex');

function stateBuffers(s) {
  const rest = Float64Array.from(s.consRest);
  const alive = Uint8Array.from(s.consAlive);
  return [s.pos, s.prev, s.vel, s.rho, s.rhoNear,
    s.group.subarray(0, s.solidCount), rest, s.active, alive];
}

function writeState(s, file) {
  const parts = stateBuffers(s).map((t) => Buffer.from(t.buffer, t.byteOffset, t.byteLength));
  fs.writeFileSync(file, zlib.gzipSync(Buffer.concat(parts), { level: 9 }));
}

function frameStats(s) {
  let ke = 0, rhoSum = 0, nf = 0, alive = 0;
  for (let i = 0; i < s.capacity; i++) {
    if (!s.active[i]) continue;
    const i3 = i * 3;
    const v2 = s.vel[i3] * s.vel[i3] + s.vel[i3 + 1] * s.vel[i3 + 1] + s.vel[i3 + 2] * s.vel[i3 + 2];
    ke += 0.5 * (1 / s.invMass[i]) * v2;
    if (i >= s.solidCount) { rhoSum += s.rho[i]; nf++; }
  }
  for (let c = 0; c < s.consAlive.length; c++) alive += s.consAlive[c];
  return { ke, meanRho: nf ? rhoSum / nf : 0, alive, fluid: s.fluidCount(), pairs: s.pairCount };
}

function frameDigests(s) {
  return {
    pos: sha(s.pos), vel: sha(s.vel), active: sha(s.active),
    alive: sha(Uint8Array.from(s.consAlive)), rest: sha(Float64Array.from(s.consRest)),
    group: sha(s.group.subarray(0, s.solidCount)),
  };
}

// d.fields: public tuning fields to set (cellSize only exists once
// addSolidCube has allocated, so it is set after it)
function makeSolver(Solver, d) {
  const s = new Solver({ maxFluid: d.maxFluid });
  s.gravity = d.gravity;
  const f = d.fields || {};
  for (const k of Object.keys(f)) if (k !== 'cellSize') s[k] = f[k];
  s.addSolidCube(...d.cube);
  if (f.cellSize !== undefined) s.cellSize = f.cellSize;
  for (const e of d.emit) s.emitFluid(...e);
  return s;
}

async function run(Solver, d) {
  const out = { name: d.name, reference: 'src/solver.js', substeps: d.substeps, frames: [], cuts: [], picks: [] };
  const base = path.join(HERE, d.name);

  // initial grid
  if (d.dumpInit !== false) {
    const s = makeSolver(Solver, d);
    s._buildGrid();
    const buf = Buffer.concat([Buffer.from(s.cellStart.buffer), Buffer.from(s.cellEntries.buffer)]);
    fs.writeFileSync(base + '_grid0.bin.gz', zlib.gzipSync(buf, { level: 9 }));
  }
  {
    const s = makeSolver(Solver, d);
    out.tableSize = s.tableSize;
    out.capacity = s.capacity;
    out.solidCount = s.solidCount;
    out.numCons = s.consA.length;
    out.init = { digests: frameDigests(s), stats: frameStats(s) };
  }
  // one substep
  if (!d.statsOnly) {
    const s = makeSolver(Solver, d);
    s.step(1 / 60 / d.substeps, 1);
    if (d.dumpInit !== false) writeState(s, base + '_sub1.bin.gz');
    out.sub1 = { digests: frameDigests(s), stats: frameStats(s) };
  }
  // the frame loop
  const s = makeSolver(Solver, d);
  let grabTargets = null;
  for (let f = 0; f < d.frames; f++) {
    for (const c of d.cuts) {
      if (c.frame === f) out.cuts.push({ frame: f, n: s.cut(...c.o, ...c.d1, ...c.d2), alive: sha(Uint8Array.from(s.consAlive)) });
    }
    if (d.grab) {
      const g = d.grab;
      if (f === g.frame) {
        const hit = s.pickSolid(...g.o, ...g.d, g.maxPerp);
        out.picks.push({ frame: f, hit });
        if (hit) {
          const hp = [0, 1, 2].map((k) => g.o[k] + g.d[k] * hit.t);
          const grp = s.group[hit.id];
          grabTargets = [];
          for (let i = 0; i < s.solidCount; i++) {
            if (s.group[i] !== grp) continue;
            const dx = s.pos[i * 3] - hp[0], dy = s.pos[i * 3 + 1] - hp[1], dz = s.pos[i * 3 + 2] - hp[2];
            if (dx * dx + dy * dy + dz * dz < g.radius * g.radius)
              grabTargets.push({ i, x: hp[0] + dx, y: hp[1] + dy, z: hp[2] + dz });
          }
          out.grabCount = grabTargets.length;
        }
      }
      if (grabTargets && f > g.frame && f < g.until) {
        for (const t of grabTargets) { t.x += g.drag[0]; t.y += g.drag[1]; t.z += g.drag[2]; }
      }
      if (grabTargets && f === g.until) grabTargets = [];
      s.grabs = grabTargets ? grabTargets.map((t) => ({ i: t.i, x: t.x, y: t.y, z: t.z })) : [];
    }
    s.step(1 / 60, d.substeps);
    out.frames.push(d.statsOnly ? { f: f + 1, stats: frameStats(s) }
      : { f: f + 1, digests: frameDigests(s), stats: frameStats(s) });
    if (d.statsOnly && (f + 1) % 25 === 0) console.log(d.name, 'frame', f + 1, JSON.stringify(frameStats(s)));
    if (d.dumpFrames.includes(f + 1)) writeState(s, base + '_f' + (f + 1) + '.bin.gz');
  }
  fs.writeFileSync(base + '_trace.json', JSON.stringify(out));
  console.log(d.name, 'frames', out.frames.length, 'final', JSON.stringify(out.frames[out.frames.length - 1].stats));
}

async function main() {
  const { Solver } = await import(pathToFileURL(REF).href);
  const names = process.argv.slice(2);
  const dir = path.join(HERE, 'inputs');
  for (const f of fs.readdirSync(dir).sort()) {
    const d = JSON.parse(fs.readFileSync(path.join(dir, f), 'utf8'));
    if (names.length && !names.includes(d.name)) continue;
    await run(Solver, d);
  }
}

main().catch((e) => { console.error(e); process.exit(1); });
