/* Node drop-in checks on a GPU: the reference's own usage patterns
 * (new LBM / voxelize / step / force, the worker pool, Voxel.build) run
 * through windtunnel_amd/js against golden vectors from the reference. */
'use strict';
const path = require('path');
const { load, arr, sha, assert, close, applyParams, replay, checkMeshMeta } = require('./fixtures.js');
const JS = path.join(__dirname, '..', '..', 'windtunnel_amd', 'js');
const { LBM, native } = require(path.join(JS, 'lbm.js'));
const { Pool } = require(path.join(JS, 'pool.js'));
const Voxel = require(path.join(JS, 'voxel.js'));

const checks = [];
function check(name, fn) { checks.push([name, fn]); }

check('exact LBM object reproduces lbm_bgk_sphere', () => {
  const fx = load('lbm_bgk_sphere');
  const L = new LBM(...fx.dims, { arith: 'f64' });
  applyParams(L, fx.params);
  L.reset();
  replay(L, fx.voxelize);
  assert(L.frontalArea === fx.voxelize.result.frontalArea, 'frontal area');
  assert(sha(L.solid) === fx.voxelize.after.solid && sha(L.near) === fx.voxelize.after.near, 'solid/near');
  const want = arr(fx.forces);
  for (let s = 0; s < fx.nsteps; s++) {
    L.step();
    close(L.force, want.subarray(3 * s, 3 * s + 3), 1e-11, 1e-14, 'force step ' + s);
  }
  const snap = fx.snaps[String(fx.nsteps)];
  for (const k of ['rho', 'ux', 'uy', 'uz']) assert(sha(L[k]) === snap[k], k);
  assert(sha(L.f) === snap.f && sha(L.g) === snap.g, 'populations');
  assert(L.steps === fx.nsteps && L.parity === snap.parity, 'steps/parity');
  for (const s of fx.samples) {
    const o = [0, 0, 0];
    L.sample(s.p[0], s.p[1], s.p[2], o);
    assert(o[0] === s.u[0] && o[1] === s.u[1] && o[2] === s.u[2], 'sample ' + s.p);
    assert(L.isSolid(s.p[0], s.p[1], s.p[2]) === s.solid, 'isSolid ' + s.p);
  }
});

check('closure voxelize equals the recorded mask (lbm_trt_les)', () => {
  const fx = load('lbm_trt_les');
  const L = new LBM(...fx.dims, { arith: 'f64' });
  applyParams(L, fx.params);
  L.reset();
  const ell = (x, y, z) => {
    const a = (x - 10) / 5, b = (y - 8) / 3, c = (z - 8) / 2.5;
    if (a * a + b * b + c * c <= 1) return true;
    return Math.abs(x - 14) <= 3 && Math.abs(y - 7) <= 2 && Math.abs(z - 8) <= 1.5;
  };
  L.voxelize(ell, { x0: 4, y0: 4, z0: 4, x1: 18, y1: 12, z1: 12 }, false);
  assert(sha(L.solid) === fx.voxelize.after.solid, 'solid');
  const f = L.stepMany(fx.nsteps);
  close(f, arr(fx.forces), 1e-11, 1e-14, 'forces');
  for (const k of ['rho', 'ux', 'uy', 'uz']) assert(sha(L[k]) === fx.snaps[String(fx.nsteps)][k], k);
});

check('worker pool protocol reproduces lbm_pool3', async () => {
  const fx = load('lbm_pool3');
  const L = new LBM(...fx.dims, { arith: 'f64' });
  applyParams(L, fx.main_params);
  L.reset();
  replay(L, fx.voxelize);
  const pool = await new Pool(L, fx.nWorkers).start();
  assert(pool.slabs.map(s => [s.z0, s.z1]).join() === fx.slabs.map(s => [s[0], s[1]]).join(), 'slab split');
  pool.syncParams();                       // tau/les/u0/movingWalls only, as the reference sends
  const want = arr(fx.forces);
  let done = 0;
  while (done < fx.nsteps) {
    const n = Math.min(16, fx.nsteps - done);
    const sf = await pool.step(n);
    for (let s = 0; s < n; s++) close(sf[s], want.subarray(3 * (done + s), 3 * (done + s) + 3), 1e-11, 1e-14, 'pool force');
    done += n;
  }
  pool.dispose();
  for (const k of ['rho', 'ux', 'uy', 'uz']) assert(sha(L[k]) === fx.pooled[k], k);
  assert(L.parity === fx.pooled.parity && L.steps === fx.pooled.steps, 'parity/steps');
});

check('Pool.step leaves the event loop running; its forces equal the synchronous slab step', async () => {
  const mk = () => {
    const L = new LBM(512, 128, 128);
    L.tau = 0.56; L.u0 = 0.05;
    L.reset();
    L.voxelize((x, y, z) => (x - 100) ** 2 + (y - 64) ** 2 + (z - 64) ** 2 < 144,
      { x0: 80, y0: 44, z0: 44, x1: 120, y1: 84, z1: 84 }, false);
    return L;
  };
  const A = mk(), B = mk();
  const pool = await new Pool(A, 3).start();
  pool.syncParams(true);
  let ticks = 0, running = true;
  const tick = () => { ticks++; if (running) setImmediate(tick); };
  setImmediate(tick);
  const got = await pool.step(16);
  running = false;
  pool.dispose();
  assert(ticks > 0, 'the event loop ran while the batch stepped (ticks ' + ticks + ')');
  // the same batch, synchronously on the twin lattice
  B._sync();
  native().setParity(B._h, 0);
  const f = new Float64Array(16 * 3 * 3);
  native().stepSlabs(B._h, 16, 3, f);
  for (let s = 0; s < 16; s++) {
    for (let k = 0; k < 3; k++) {
      const want = f[(s * 3) * 3 + k] + f[(s * 3 + 1) * 3 + k] + f[(s * 3 + 2) * 3 + k];
      assert(got[s][k] === want, 'force step ' + s);
    }
  }
  assert(sha(A.rho) === sha(B.rho) && sha(A.ux) === sha(B.ux), 'fields');
});

check('lbmWorker.js step loop through stepSlab/_boundaries reproduces lbm_pool3', () => {
  const fx = load('lbm_pool3');
  const L = new LBM(...fx.dims, { arith: 'f64' });
  applyParams(L, fx.worker_params);
  L.reset();
  replay(L, fx.voxelize);
  const nz = fx.dims[2], NW = fx.nWorkers, want = arr(fx.forces);
  for (let s = 0; s < fx.nsteps; s++) {
    const tot = [0, 0, 0];
    for (let id = 0; id < NW; id++) {
      const z0 = 1 + Math.floor((nz - 2) * id / NW), z1 = 1 + Math.floor((nz - 2) * (id + 1) / NW);
      const bz0 = id === 0 ? 0 : z0, bz1 = id === NW - 1 ? nz : z1;
      L.stepSlab(z0, z1);
      L._boundaries(L.g, bz0, bz1, id === 0, id === NW - 1);
      for (let k = 0; k < 3; k++) tot[k] += L._sf[k];
    }
    L.parity ^= 1;
    close(tot, want.subarray(3 * s, 3 * s + 3), 1e-11, 1e-14, 'slab force ' + s);
  }
  for (const k of ['rho', 'ux', 'uy', 'uz']) assert(sha(L[k]) === fx.pooled[k], k);
  assert(L.idx(3, 2, 1) === 3 + L.nx * (2 + L.ny), 'idx');
});

check('gray adjoint pair: backward is linear in Fbar and repeatable', () => {
  const L = new LBM(20, 12, 12);
  L.tau = 0.6; L.u0 = 0.05; L.les = false; L.trt = true; L.lambda = 3 / 16;
  L.reset();
  L.voxelize((x, y, z) => (x - 7) ** 2 + (y - 6) ** 2 + (z - 6) ** 2 <= 4, { x0: 4, y0: 3, z0: 3, x1: 10, y1: 9, z1: 9 }, false);
  const w = new Float32Array(L.n);
  for (let z = 0; z < 12; z++) for (let y = 0; y < 12; y++) for (let x = 0; x < 20; x++) {
    const r = Math.sqrt((x + 0.5 - 7) ** 2 + (y + 0.5 - 6) ** 2 + (z + 0.5 - 6) ** 2);
    const i = L.idx(x, y, z);
    w[i] = L.solid[i] ? 0 : Math.min(0.99, Math.max(0, (3.6 - r) / 1.6));
  }
  L.setGray(w);
  const N = 30, f = L.adjointForward(N, 5);
  assert(f.length === 3 * N && L.steps === N, 'forward');
  const Fb = new Float64Array(3 * N);
  for (let k = N - 10; k < N; k++) Fb[3 * k] = 0.1;
  const a = L.adjointBackward(Fb).dJdw, b = L.adjointBackward(Fb.map(v => 2 * v)).dJdw;
  let nz = 0, worst = 0;
  for (let i = 0; i < L.n; i++) {
    if (a[i] !== 0) nz++;
    worst = Math.max(worst, Math.abs(b[i] - 2 * a[i]));
  }
  assert(nz > 0, 'gradient is zero everywhere');
  assert(worst <= 1e-6 * Math.max(...a.map(Math.abs)), 'linearity ' + worst);
  let threw = false;
  try { L.adjointBackward(new Float64Array(3)); } catch (e) { threw = /Fbar/.test(e.message); }
  L.adjointRelease();
  assert(threw, 'short Fbar must be refused');
});

check('voxelizeBody rasterises the catalogue bodies as the reference does (lbm_bodies)', () => {
  const fx = load('lbm_bodies');
  for (const c of fx.cases) {
    const L = new LBM(...fx.dims, { arith: 'f64' });
    L.u0 = 0.06;
    L.reset();
    const bb = c.box_in;
    const body = { parts: c.body.parts, M: c.body.M, pos: c.body.pos, scale: c.body.scale, def: {},
      _updateMat() {}, box() { return { x0: bb[0], y0: bb[1], z0: bb[2], x1: bb[3], y1: bb[4], z1: bb[5] }; } };
    L.voxelizeBody(body, c.ground, c.rolling_u0);
    const w = c.after;
    assert(sha(L.solid) === w.solid && sha(L.near) === w.near, c.key + ' solid/near');
    assert(sha(L.uwx) === w.uwx && sha(L.uwy) === w.uwy && sha(L.uwz) === w.uwz, c.key + ' wall velocities');
    assert(L.frontalArea === c.result.frontalArea && L.rideHeight === c.result.rideHeight, c.key + ' stats');
  }
});

check('slice and sampleMany equal the app\'s own formulas over the downloaded fields', () => {
  // updateSlice (js/main.js:925-965) before its colour maps, and sample() per point
  const fx = load('lbm_bgk_sphere');
  const L = new LBM(...fx.dims);
  applyParams(L, fx.params);
  L.reset();
  replay(L, fx.voxelize);
  L.stepMany(120);
  const { nx, ny, nz, ux, uy, uz, rho, solid } = L;
  const u0 = L.u0;
  for (const axis of ['xy', 'xz']) {
    const xy = axis === 'xy';
    const h = xy ? ny : nz, k = xy ? nz >> 1 : ny >> 1;
    const at = xy ? (x, y) => x + nx * (y + ny * k) : (x, z) => x + nx * (k + ny * z);
    for (const field of ['speed', 'pressure', 'vorticity']) {
      const got = L.slice(axis, k, field);
      const want = new Float32Array(nx * h);
      for (let b = 0; b < h; b++) {
        for (let a = 0; a < nx; a++) {
          const i = at(a, b), o = a + b * nx;
          assert(got.solid[o] === solid[i], 'slice solid map');
          if (solid[i]) continue;
          const ip = a < nx - 1 ? at(a + 1, b) : i, im = a > 0 ? at(a - 1, b) : i;
          const jp = b < h - 1 ? at(a, b + 1) : i, jm = b > 0 ? at(a, b - 1) : i;
          if (field === 'speed') want[o] = Math.hypot(ux[i], uy[i], uz[i]) / u0;
          else if (field === 'pressure') {
            const r = (rho[i] * 2 + rho[ip] + rho[im] + rho[jp] + rho[jm]) / 6;
            want[o] = ((r - 1) / 3) / (0.5 * u0 * u0);
          } else if (xy) want[o] = (uy[ip] - uy[im]) * 0.5 - (ux[jp] - ux[jm]) * 0.5;
          else want[o] = (ux[jp] - ux[jm]) * 0.5 - (uz[ip] - uz[im]) * 0.5;
        }
      }
      // Math.hypot and sqrt of the sum of squares may differ by an ulp of binary64
      const tol = field === 'speed' ? 1.2e-7 : 0;
      for (let o = 0; o < want.length; o++)
        if (Math.abs(got.values[o] - want[o]) > tol * Math.abs(want[o]))
          throw new Error(`${axis} ${field} at ${o}: ${got.values[o]} vs ${want[o]}`);
    }
  }
  const pts = new Float64Array(3 * 500);
  let seed = 7;
  const rnd = () => ((seed = (seed * 1103515245 + 12345) >>> 0) / 4294967296);
  for (let p = 0; p < 500; p++) { pts[3 * p] = rnd() * (nx + 2) - 1; pts[3 * p + 1] = rnd() * ny; pts[3 * p + 2] = rnd() * nz; }
  const many = L.sampleMany(pts), one = [0, 0, 0];
  for (let p = 0; p < 500; p++) {
    L.sample(pts[3 * p], pts[3 * p + 1], pts[3 * p + 2], one);
    for (let c = 0; c < 3; c++) assert(many[3 * p + c] === one[c], 'sampleMany point ' + p);
  }
});

check('boundaryMode selects the channel boundaries', () => {
  const fx = load('lbm_c1');
  const run = mode => {
    const L = new LBM(...fx.dims);
    applyParams(L, fx.params);
    L.boundaryMode = mode;
    L.reset();
    replay(L, fx.voxelize);
    L.stepMany(50);
    return L;
  };
  const A = run('reference'), B = run('channel');
  assert(B.boundaryMode === 'channel' && A.boundaryMode === 'reference', 'mode property');
  assert(sha(A.ux) !== sha(B.ux), 'the channel boundaries change the flow');
  let threw = false;
  try { B.boundaryMode = 'periodic'; } catch (e) { threw = true; }
  assert(threw, 'unknown mode refused');
});

check('Voxel.build is bit-exact (vox_box, vox_hitcap, vox_sphere)', () => {
  for (const name of ['vox_box', 'vox_hitcap', 'vox_sphere']) {
    const fx = load(name);
    const v = Voxel.build(arr(fx.mesh), fx.res);
    assert(sha(v.grid) === fx.grid_sha256, name + ' grid');
    assert(v.stats.solid === fx.stats.solid && v.stats.openness === fx.stats.openness, name + ' stats');
    assert(v.nx === fx.nx && v.cell === fx.cell && v.min.join() === fx.min.join(), name + ' geometry');
  }
  let msg = null;
  try { Voxel.build(new Float32Array(0)); } catch (e) { msg = e.message; }
  assert(msg === load('vox_errors').empty, 'error message');
});

check('MeshIO.meta gives the reference\'s credits for every mesh_io file', () => {
  const n = checkMeshMeta(require(path.join(JS, 'meshio.js')));
  assert(n >= 45, 'mesh_io cases: ' + n);
});

check('ForceStats folds the drop-in\'s step forces into the app\'s run report', () => {
  const { ForceStats } = require(path.join(JS, 'stats.js'));
  const fx = load('lbm_bgk_sphere');
  const L = new LBM(...fx.dims, { arith: 'f64' });
  applyParams(L, fx.params);
  L.reset();
  replay(L, fx.voxelize);
  const st = new ForceStats(5);
  const F = [];
  for (let s = 0; s < 80; s++) {
    if (s === 20) st.setCapture(true, L.steps);
    L.step();
    F.push(Array.from(L.force));
    st.accumulate(L.force, L.u0, L.frontalArea, L.steps);
  }
  st.setCapture(false, L.steps);
  const csv = st.runCSV(L, { name: 'sphere', custom: false, credit: null, ref: 'frontal', scale: 16, ground: false,
    lengthM: 1, windMs: 30, aoa: 0, yaw: 0 }, '2026-01-01T00:00:00.000Z').split('\n');
  const row = (k) => csv.find((l) => l.startsWith(k + ',')).split(',');
  assert(row('capture_from_step')[1] === '20' && row('capture_to_step')[1] === '80', 'capture window');
  assert(row('samples')[1] === '60' && row('timesteps_total')[1] === '80', 'sample count');
  // the captured samples are the pair averages of steps 21..80
  const q = 0.5 * L.u0 * L.u0 * L.frontalArea;
  let m = 0;
  for (let s = 20; s < 80; s++) m += 0.5 * (F[s][0] + F[s - 1][0]) / q;
  m /= 60;
  const cd = Number(row('Cd')[1]);
  assert(Math.abs(cd - m) < 1e-5, 'Cd ' + cd + ' vs ' + m);
});

check('injected device fault surfaces as an Error', () => {
  // WT_FAULT=step:3 is set by the runner for this process
  const L = new LBM(16, 8, 8);
  L.stepMany(3);
  let err = null;
  try { L.step(); } catch (e) { err = e; }
  assert(err && err.code === 'WT-5', 'fault code');
});

(async () => {
  let failed = 0;
  // the injected-fault check needs WT_FAULT in its own process
  const only = process.argv[2] === 'fault';
  for (const [name, fn] of checks) {
    if (only !== name.startsWith('injected')) continue;
    try { await fn(); console.log('ok   ' + name); } catch (e) { failed++; console.log('FAIL ' + name + ': ' + e.message); }
  }
  console.log(failed ? `${failed} failed` : 'all passed');
  process.exit(failed ? 1 : 0);
})();
