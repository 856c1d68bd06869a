// Steps several scenes as one Batch through the Node host and writes every
// frame's batch snapshot:
//   node run_batch.mjs <scenes.json> <out.bin> <frames> <substeps> <dt>
// scenes.json = [{params, maxFluid, cube, pos (flat xyz), vel, mass, cuts: [{frame, o, d1, d2}]}]
// out.bin = frames x (float32 snapshot of every scene); stdout = {offsets, info}.
import fs from 'fs';
import { Solver, Batch, sceneBytes } from '../../fluidphysics_amd/js/solver.mjs';

const [inp, out, framesS, subS, dtS] = process.argv.slice(2);
const scenes = JSON.parse(fs.readFileSync(inp, 'utf8'));
const frames = parseInt(framesS, 10), sub = parseInt(subS, 10), dt = parseFloat(dtS);
const b = new Batch();
const solvers = [];
const bytes = [];
for (const d of scenes) {
  const s = new Solver({ maxFluid: d.maxFluid, params: d.params });
  s.addSolidCube(...d.cube);
  s.emitBatch(Float64Array.from(d.pos), Float64Array.from(d.vel), d.mass);
  bytes.push([sceneBytes(d.params, d.maxFluid, d.cube[3]), s.info().deviceBytes]);
  b.add(s);
  solvers.push(s);
}
const chunks = [];
let offsets = null;
for (let f = 0; f < frames; f++) {
  scenes.forEach((d, k) => {
    for (const c of d.cuts) if (c.frame === f) solvers[k].cut(...c.o, ...c.d1, ...c.d2);
  });
  b.step(dt, sub);
  const snap = b.snapshot();
  offsets = snap.offsets;
  chunks.push(Buffer.from(snap.pos.buffer, snap.pos.byteOffset, snap.pos.byteLength));
}
fs.writeFileSync(out, Buffer.concat(chunks));
console.log(JSON.stringify({ offsets, info: b.info(), bytes }));
b.destroy();
for (const s of solvers) s.destroy();
