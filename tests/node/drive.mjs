// Headless version of the reference's browser drive test (test/drive.js),
// driving the MI355X solver through the Node host (fluidphysics_amd/js/solver.mjs)
// with API calls in place of mouse events, on the reference CPU page's scene
// (src/main.js:8-9).  Timeline of test/drive.js:18-48 at 60 frames/s:
//   settle 3.5 s (210 frames) -> two vertical knife strokes -> grab + drag +
//   release -> hose held 6 s (360 frames, src/main.js:224-251 pattern with a
//   seeded jitter).
// Prints one JSON line per phase: {phase, frames, fluid, alive, ke, minY, maxY, nan,
// comps, same}: comps = the device's component count (components()), same =
// its partition equals the one `group` labels.
import { Solver } from '../../fluidphysics_amd/js/solver.mjs';

// the library defaults: 3 Jacobi iterations, relaxation 8
const s = new Solver({ maxFluid: 2600 });
s.addSolidCube(0, 2.0, 0, 7, 0.18);
const cam = [5.2, 3.6, 6.4];
let frame = 0;

function norm(v) { const l = Math.hypot(v[0], v[1], v[2]); return [v[0] / l, v[1] / l, v[2] / l]; }
function toward(p) { return norm([p[0] - cam[0], p[1] - cam[1], p[2] - cam[2]]); }

// splitmix64-style seeded jitter in (-0.2, 0.2), replacing Math.random
let seed = 0x9e3779b97f4a7c15n;
function rnd() {
  seed = (seed + 0x9e3779b97f4a7c15n) & 0xffffffffffffffffn;
  let z = seed;
  z = ((z ^ (z >> 30n)) * 0xbf58476d1ce4e5b9n) & 0xffffffffffffffffn;
  z = ((z ^ (z >> 27n)) * 0x94d049bb133111ebn) & 0xffffffffffffffffn;
  z = z ^ (z >> 31n);
  return Number(z >> 11n) / 9007199254740992;
}

function report(phase, frames) {
  const p = s.pos, v = s.vel, a = s.active, im = s.invMass;
  let ke = 0, nan = 0, minY = Infinity, maxY = -Infinity;
  for (let i = 0; i < s.capacity; i++) {
    if (!a[i]) continue;
    const i3 = 3 * i;
    if (!Number.isFinite(p[i3]) || !Number.isFinite(p[i3 + 1]) || !Number.isFinite(p[i3 + 2])) { nan++; continue; }
    ke += 0.5 / im[i] * (v[i3] * v[i3] + v[i3 + 1] * v[i3 + 1] + v[i3 + 2] * v[i3 + 2]);
    if (i < s.solidCount) { minY = Math.min(minY, p[i3 + 1]); maxY = Math.max(maxY, p[i3 + 1]); }
  }
  let alive = 0;
  for (const x of s.consAlive) alive += x;
  const c = s.components(), g = s.group, byLabel = new Map(), byGroup = new Map();
  let same = true;
  for (let i = 0; i < s.solidCount; i++) {
    const l = c.labels[i], q = g[i];
    if (!byLabel.has(l)) byLabel.set(l, q); else if (byLabel.get(l) !== q) same = false;
    if (!byGroup.has(q)) byGroup.set(q, l); else if (byGroup.get(q) !== l) same = false;
  }
  console.log(JSON.stringify({ phase, frames, fluid: s.fluidCount(), alive, ke, minY, maxY, nan, nonFinite: s.nonFinite(), comps: c.count, same }));
}

function run(n, each) { for (let k = 0; k < n; k++) { if (each) each(k); s.step(1 / 60, 5); frame++; } }

// settle
run(210);
report('settle', frame);

// two vertical strokes across the body (test/drive.js:18-28)
for (const x of [-0.15, 0.2]) {
  let prev = toward([x, 2.0, 0.0]);
  for (let k = 1; k <= 8; k++) {
    const d = toward([x, 2.0 - 0.25 * k, 0.0]);
    s.cut(...cam, ...prev, ...d);
    prev = d;
    run(1);
  }
}
report('slice', frame);

// grab: pick from the camera, take the hit's group within 0.35 (src/main.js:288-305), drag up, release
const hitDir = toward([0, 0.6, 0]);
const hit = s.pickSolid(...cam, ...hitDir, 0.25);
if (hit) {
  const hp = [cam[0] + hitDir[0] * hit.t, cam[1] + hitDir[1] * hit.t, cam[2] + hitDir[2] * hit.t];
  const grp = s.group[hit.id];
  const p = s.pos;
  const offs = [];
  for (let i = 0; i < s.solidCount; i++) {
    if (s.group[i] !== grp) continue;
    const dx = p[3 * i] - hp[0], dy = p[3 * i + 1] - hp[1], dz = p[3 * i + 2] - hp[2];
    if (dx * dx + dy * dy + dz * dz < 0.35 * 0.35) offs.push({ i, dx, dy, dz });
  }
  run(40, (k) => {
    const lift = Math.min(1.2, 0.03 * k);
    s.grabs = offs.map((o) => ({ i: o.i, x: hp[0] + o.dx, y: hp[1] + o.dy + lift, z: hp[2] + o.dz }));
  });
  s.grabs = [];
}
report('grab', frame);

// hose (src/main.js:224-251 pattern: 5 per step on a golden-angle disc, speed 7.5)
const hose = [-3.1, 1.55, -3.1];
let phase = 0;
run(360, () => {
  const dir = norm([1.5 - hose[0], 0 - hose[1], 1.5 - hose[2]]);
  const tip = [hose[0] + dir[0] * 0.6, hose[1] + dir[1] * 0.6, hose[2] + dir[2] * 0.6];
  const side = norm([dir[2], 0, -dir[0]]);
  const up = [side[1] * dir[2] - side[2] * dir[1], side[2] * dir[0] - side[0] * dir[2], side[0] * dir[1] - side[1] * dir[0]];
  const spacing = s.fluidRadius * 2.2;
  for (let k = 0; k < 5; k++) {
    phase += 2.39996;
    const rad = spacing * (0.4 + (k % 2));
    const ox = Math.cos(phase) * rad, oy = Math.sin(phase) * rad;
    s.emitFluid(
      tip[0] + side[0] * ox + up[0] * oy + dir[0] * k * spacing * 0.5,
      tip[1] + side[1] * ox + up[1] * oy + dir[1] * k * spacing * 0.5,
      tip[2] + side[2] * ox + up[2] * oy + dir[2] * k * spacing * 0.5,
      dir[0] * 7.5 + (rnd() - 0.5) * 0.4, dir[1] * 7.5 + (rnd() - 0.5) * 0.4, dir[2] * 7.5 + (rnd() - 0.5) * 0.4);
  }
});
report('hose', frame);
s.destroy();
