'use strict';
// GPU check of the Node-API addon: the reference's published Poiseuille and
// Couette known answers (docs/sim-specs/lattice/lattice-boltzmann-d3q19/equivalence.md:240-243)
// and read-back consistency.  Prints one JSON line; exits non-zero on failure.
const assert = require('assert');
const { Lattice, version } = require('./index.js');

function maxSpeed(u, n) {
  let m = 0;
  for (let c = 0; c < n; ++c) {
    const s = Math.hypot(u[c], u[n + c], u[2 * n + c]);
    if (s > m) m = s;
  }
  return m;
}

const out = { version: version() };

{
  const nx = 64, ny = 32, nz = 3, n = nx * ny * nz;
  const lat = new Lattice({ nx, ny, nz, tau: 0.7, storage: 'f64', boundary: 'ywalls', force: [1e-5, 0, 0] });
  lat.setEquilibrium(null, new Float64Array(3 * n));
  lat.step(1000);
  const { rho, u } = lat.readMoments();
  out.poiseuille_umax = maxSpeed(u, n).toExponential(7);
  const h = lat.health();
  assert.strictEqual(h.checkHealth, 0);
  assert.strictEqual(h.uMaxLat.toExponential(7), out.poiseuille_umax);
  out.poiseuille_rhomax = Math.max(...rho).toExponential(7);
  assert.strictEqual(out.poiseuille_umax, '8.6508320e-3');
  assert.strictEqual(out.poiseuille_rhomax, '1.0000000e+0');
  assert.strictEqual(lat.stepCount(), 1000);
  const f = lat.readPopulations('f64');
  assert.strictEqual(f.length, 19 * n);
  let mass = 0;
  for (let i = 0; i < f.length; ++i) mass += f[i];
  const d = lat.diagnostics();
  assert(Math.abs(mass - d.mass) < 1e-9 * n, 'diagnostics mass vs read-back sum');
  lat.close();
}
{
  const nx = 32, ny = 16, nz = 3, n = nx * ny * nz;
  const lat = new Lattice({ nx, ny, nz, tau: 0.7, storage: 'f64', boundary: 'ywalls', wallVelocityTop: [0.05, 0, 0] });
  lat.setEquilibrium(null, new Float64Array(3 * n));
  lat.step(500);
  out.couette_umax = maxSpeed(lat.readMoments().u, n).toExponential(7);
  assert.strictEqual(out.couette_umax, '5.0000000e-2');
  lat.close();
}
{
  // set/read round trip in f32 storage: a rest state is a fixed point
  const n = 8 * 8 * 8;
  const lat = new Lattice({ nx: 8, ny: 8, nz: 8, tau: 0.6, boundary: 'halfway', wallFaces: ['-x', '+x', '-y', '+y', '-z', '+z'] });
  lat.setEquilibrium(null, new Float32Array(3 * n));
  lat.step(10);
  const f = lat.readPopulations('f32');
  const w = [1 / 3, 1 / 18, 1 / 36];
  let err = 0;
  for (let i = 0; i < 19; ++i) {
    const wi = i === 0 ? w[0] : (i <= 6 ? w[1] : w[2]);
    for (let c = 0; c < n; ++c) err = Math.max(err, Math.abs(f[i * n + c] - wi));
  }
  out.rest_fixed_point_err = err;
  assert(err < 1e-7);
  let threw = false;
  try { lat.step(-1); } catch (e) { threw = /EINVAL|status 1/.test(String(e)); }
  assert(threw, 'negative step count must throw');
  assert.throws(() => lat.step('ten'), TypeError, 'a step count that is not a number must throw');
  lat.close();
}
{
  // moment-encoded storage: ranges from the FP64 trajectory's stored states, then the quantized run
  const nx = 64, ny = 32, nz = 3, n = nx * ny * nz, steps = 1000;
  const cfg = { nx, ny, nz, tau: 0.7, boundary: 'ywalls', force: [1e-5, 0, 0] };
  const ref = new Lattice({ ...cfg, storage: 'f64' });
  ref.setEquilibrium(null, new Float64Array(3 * n));
  let { lo, hi } = ref.momentRange(true);
  assert.throws(() => ref.momentRange('yes'), TypeError, 'momentRange takes a boolean');
  for (let s = 0; s < steps; ++s) {
    ref.step(1);
    const r = ref.momentRange(true);
    for (let k = 0; k < 19; ++k) { lo[k] = Math.min(lo[k], r.lo[k]); hi[k] = Math.max(hi[k], r.hi[k]); }
  }
  const uRef = ref.readMoments().u;
  ref.close();
  for (let k = 0; k < 19; ++k) {
    const span = Math.max(hi[k] - lo[k], 1e-12);
    lo[k] -= 0.25 * span; hi[k] += 0.25 * span;
  }
  const lat = new Lattice({ ...cfg, storage: 'me16' });
  lat.setMomentRanges(lo, hi);
  lat.setEquilibrium(null, new Float64Array(3 * n));
  lat.step(steps);
  const u = lat.readMoments().u;
  let err = 0;
  for (let i = 0; i < u.length; ++i) err = Math.max(err, Math.abs(u[i] - uRef[i]) - 5e-2 * Math.abs(uRef[i]));
  out.me16_poiseuille_umax = maxSpeed(u, n).toExponential(2);
  assert(err <= 1e-5, 'me16 within the lbm-quantized tolerance of the FP64 run');
  assert.strictEqual(out.me16_poiseuille_umax, '8.65e-3');
  lat.close();
}
console.log(JSON.stringify(out));
