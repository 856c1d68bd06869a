// Times the reference CPU solver (src/solver.js, class Solver) under node on
// one scene input, single-threaded: node time_reference.mjs <input.json>
// <frames> <repeats>.  Development container only (the reference is read
// from /root/reference, or FP_REFERENCE_SOLVER).  The input is a scenario
// JSON as tests/golden/gen_inputs.py writes it, plus optional per-frame emit
// bursts ("hose": [[x, y, z, vx, vy, vz, mass], ...] per frame).  Prints one
// JSON line: wall ms per frame of each repeat (a fresh solver each time) and
// the active particles of the last frame.
import fs from 'fs';
import { pathToFileURL } from 'url';

const REF = process.env.FP_REFERENCE_SOLVER || '/root/reference/src/solver.js';

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

async function main() {
  const [file, framesArg, repeatsArg] = process.argv.slice(2);
  const d = JSON.parse(fs.readFileSync(file, 'utf8'));
  const frames = parseInt(framesArg || '10', 10), repeats = parseInt(repeatsArg || '3', 10);
  const { Solver } = await import(pathToFileURL(REF).href);
  const ms = [];
  let active = 0;
  for (let r = 0; r < repeats; r++) {
    const s = makeSolver(Solver, d);
    const t0 = process.hrtime.bigint();
    for (let f = 0; f < frames; f++) {
      for (const e of (d.hose && d.hose[f]) || []) s.emitFluid(...e);
      s.step(1 / 60, d.substeps);
    }
    ms.push(Number(process.hrtime.bigint() - t0) / 1e6 / frames);
    active = s.solidCount + s.fluidCount();
  }
  console.log(JSON.stringify({ name: d.name, frames, ms_per_frame: ms, active }));
}

main().catch((e) => { console.error(e); process.exit(1); });
