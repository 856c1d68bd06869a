// Runs a scenario (tests/golden/gen_inputs.py format) through the Node host
// (fluidphysics_amd/js/solver.mjs) and writes the final state:
//   node run_scene.mjs <input.json> <out.bin> <frames> <substeps-per-frame> <dt-per-frame>
// out.bin = pos f32[3N] vel f32[3N] active u8[N] (little endian).
import fs from 'fs';
import { Solver } from '../../fluidphysics_amd/js/solver.mjs';

const [inp, out, framesS, subS, dtS] = process.argv.slice(2);
const d = JSON.parse(fs.readFileSync(inp, 'utf8'));
const s = new Solver({ maxFluid: d.maxFluid, params: d.params || {} });
s.gravity = d.gravity;
s.addSolidCube(...d.cube);
for (const e of d.emit) s.emitFluid(...e);
const frames = parseInt(framesS, 10), sub = parseInt(subS, 10), dt = parseFloat(dtS);
for (let f = 0; f < frames; f++) {
  for (const c of d.cuts || []) if (c.frame === f) s.cut(...c.o, ...c.d1, ...c.d2);
  s.step(dt, sub);
}
const p = s.pos, v = s.vel, a = s.active;
fs.writeFileSync(out, Buffer.concat([Buffer.from(p.buffer), Buffer.from(v.buffer), Buffer.from(a.buffer)]));
if (d.dumpTets) {
  // the optional tet mode's topology and alive flags: verts i32[4T] then alive u8[T]
  const t = s.readTets();
  fs.writeFileSync(out + '.tets', Buffer.concat([Buffer.from(t.verts.buffer), Buffer.from(t.alive.buffer)]));
}
console.log(JSON.stringify({ capacity: s.capacity, fluid: s.fluidCount(), tableSize: s.tableSize }));
s.destroy();
