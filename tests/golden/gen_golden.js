/* Golden-vector generator: runs the reference solver and voxelizer under Node
 * in THIS container and writes small fixtures (hashes, force series, slices,
 * voxel grids) into tests/golden/fixtures/. Only this script reads the
 * reference; the fixtures it writes are data (inputs and expected outputs).
 *
 *   node tests/golden/gen_golden.js [/root/reference] [case ...]
 *
 * The reference scripts are IIFEs that attach LBM / Voxel / Shapes to
 * `window` or `self`; they are evaluated in a vm context whose global object
 * plays both roles. Nothing from the reference is copied into the outputs.
 */
'use strict';
const fs = require('fs');
const path = require('path');
const vm = require('vm');
const zlib = require('zlib');
const crypto = require('crypto');
const synth = require('./synth.js');

const CASE = a => a.startsWith('lbm_') || a.startsWith('vox_') || a.startsWith('mesh_');
const REF = process.argv[2] && !CASE(process.argv[2]) ? process.argv[2] : '/root/reference';
const ONLY = process.argv.slice(2).filter(CASE);
const OUT = path.join(__dirname, 'fixtures');
fs.mkdirSync(OUT, { recursive: true });

const ctx = { console, TextDecoder };
ctx.window = ctx; ctx.self = ctx;
vm.createContext(ctx);
for (const f of ['js/lbm.js', 'js/voxel.js', 'js/shapes.js', 'js/meshio.js']) {
  vm.runInContext(fs.readFileSync(path.join(REF, f), 'utf8'), ctx, { filename: f });
}
const { LBM, Voxel, Shapes, MeshIO } = ctx;

// ---------------------------------------------------------------- helpers
const bytesOf = a => Buffer.from(a.buffer, a.byteOffset, a.byteLength);
const sha = a => crypto.createHash('sha256').update(bytesOf(a)).digest('hex');
const DT = { Float32Array: 'f4', Float64Array: 'f8', Uint8Array: 'u1', Int32Array: 'i4', Uint32Array: 'u4' };
function pack(a, shape) {
  return { dtype: DT[a.constructor.name], shape: shape || [a.length],
    gz_b64: zlib.gzipSync(bytesOf(a), { level: 9 }).toString('base64') };
}
function bits(u8) {               // LSB-first bit packing of a 0/1 byte grid
  const out = new Uint8Array((u8.length + 7) >> 3);
  for (let i = 0; i < u8.length; i++) if (u8[i]) out[i >> 3] |= 1 << (i & 7);
  return out;
}
function write(name, obj) {
  const file = path.join(OUT, name + '.json');
  fs.writeFileSync(file, JSON.stringify(obj));
  console.log(`${name}: ${(fs.statSync(file).size / 1024).toFixed(1)} KiB`);
}
function want(name) { return !ONLY.length || ONLY.includes(name); }

function stateHashes(L) {
  return { f: sha(L.f), g: sha(L.g), rho: sha(L.rho), ux: sha(L.ux), uy: sha(L.uy), uz: sha(L.uz),
    solid: sha(L.solid), near: sha(L.near), uwx: sha(L.uwx), uwy: sha(L.uwy), uwz: sha(L.uwz),
    parity: L.parity, steps: L.steps };
}

/**
 * Evaluate inside/wallVel over the clamped box in the reference's loop
 * order so the same body can be replayed through another implementation's
 * voxelize: mask 1 = inside, 2 = inside with a wall velocity (wv triple).
 */
function boxMask(L, inside, box, wallVel) {
  const { nx, ny, nz } = L;
  const x0 = Math.max(1, Math.floor(box.x0)), x1 = Math.min(nx - 2, Math.ceil(box.x1));
  const y0 = Math.max(1, Math.floor(box.y0)), y1 = Math.min(ny - 2, Math.ceil(box.y1));
  const z0 = Math.max(1, Math.floor(box.z0)), z1 = Math.min(nz - 2, Math.ceil(box.z1));
  const bx = Math.max(0, x1 - x0 + 1), by = Math.max(0, y1 - y0 + 1), bz = Math.max(0, z1 - z0 + 1);
  const mask = new Uint8Array(bx * by * bz);
  const wv = [];
  const out = [0, 0, 0];
  let k = 0;
  for (let z = z0; z <= z1; z++) for (let y = y0; y <= y1; y++) for (let x = x0; x <= x1; x++, k++) {
    if (!inside(x + 0.5, y + 0.5, z + 0.5)) continue;
    mask[k] = 1;
    if (wallVel && wallVel(x + 0.5, y + 0.5, z + 0.5, out)) { mask[k] = 2; wv.push(out[0], out[1], out[2]); }
  }
  return { box: [x0, y0, z0, x1, y1, z1], mask: pack(mask, [bz, by, bx]), wv: pack(new Float32Array(wv)),
    box_in: [box.x0, box.y0, box.z0, box.x1, box.y1, box.z1] };
}

function voxelizeRecorded(L, inside, box, ground, wallVel, roadVel) {
  const rec = boxMask(L, inside, box, wallVel);
  L.voxelize(inside, box, ground, wallVel, roadVel);
  rec.ground = !!ground; rec.roadVel = roadVel || 0;
  rec.result = { frontalArea: L.frontalArea, planformArea: L.planformArea, solidCells: L.solidCells,
    rideHeight: L.rideHeight, movingWalls: L.movingWalls };
  rec.after = stateHashes(L);
  return rec;
}

const sphere = (cx, cy, cz, r) => (x, y, z) =>
  (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz) <= r * r;
const sphereBox = (cx, cy, cz, r) => ({ x0: cx - r - 1, y0: cy - r - 1, z0: cz - r - 1, x1: cx + r + 1, y1: cy + r + 1, z1: cz + r + 1 });

function params(L) {
  return { tau: L.tau, les: L.les, csm: L.csm, trt: L.trt, lambda: L.lambda, u0: L.u0, movingWalls: L.movingWalls };
}

/** Run steps, recording every force and hashes at the listed step counts. */
function run(L, nsteps, at, rawAt) {
  const forces = new Float64Array(nsteps * 3);
  const snaps = {};
  for (let s = 0; s < nsteps; s++) {
    L.step();
    forces[3 * s] = L.force[0]; forces[3 * s + 1] = L.force[1]; forces[3 * s + 2] = L.force[2];
    if (at.includes(L.steps)) snaps[L.steps] = stateHashes(L);
    if (rawAt && rawAt.steps === L.steps) {
      snaps.raw = {};
      for (const k of rawAt.fields) snaps.raw[k] = pack(L[k].slice(), k === 'f' ? [19, L.nz, L.ny, L.nx] : [L.nz, L.ny, L.nx]);
    }
  }
  return { forces: pack(forces, [nsteps, 3]), snaps };
}

// ------------------------------------------------------------ LBM cases
if (want('lbm_init')) {
  const L = new LBM(12, 10, 8);
  L.u0 = 0.07; L.reset();
  const reset = stateHashes(L);
  const v1 = voxelizeRecorded(L, sphere(5, 5, 4, 2.5), sphereBox(5, 5, 4, 2.5), true,
    (x, y, z, o) => { if (x < 5) { o[0] = 0.01; o[1] = -0.02; o[2] = 0.003; return true; } return false; }, 0.05);
  const v2 = voxelizeRecorded(L, sphere(6, 4, 4, 2), sphereBox(6, 4, 4, 2), false);
  write('lbm_init', { dims: [12, 10, 8], u0: 0.07, reset, voxelize: [v1, v2] });
}

function smallCase(name, dims, setp, body, nsteps, opts) {
  if (!want(name)) return;
  opts = opts || {};
  const t0 = Date.now();
  const L = new LBM(dims[0], dims[1], dims[2]);
  setp(L);
  L.reset();
  const vox = voxelizeRecorded(L, body.inside, body.box, body.ground, body.wallVel, body.roadVel);
  const prm = params(L);
  const res = run(L, nsteps, opts.at || [1, 2, 10, nsteps], opts.raw);
  const out = { dims, params: prm, voxelize: vox, nsteps, forces: res.forces, snaps: res.snaps };
  if (opts.samples) {
    out.samples = opts.samples.map(p => { const o = [0, 0, 0]; L.sample(p[0], p[1], p[2], o); return { p, u: o.slice(), solid: L.isSolid(p[0], p[1], p[2]) }; });
  }
  if (opts.more) opts.more(L, out);
  write(name, out);
  console.log(`  ${name}: ${((Date.now() - t0) / 1000).toFixed(1)} s`);
}

smallCase('lbm_bgk_sphere', [32, 16, 16], L => { L.trt = false; L.les = false; L.tau = 0.56; L.u0 = 0.05; },
  { inside: sphere(11, 8, 8, 4), box: sphereBox(11, 8, 8, 4), ground: false }, 100,
  { raw: { steps: 100, fields: ['rho', 'ux', 'uy', 'uz'] },
    samples: [[10.3, 7.7, 8.1], [20.5, 8, 8], [0.2, 3, 3], [31, 8, 8], [15.25, 12.5, 3.75]] });

const ellBox = (e, b) => (x, y, z) => {
  const a = (x - e.c[0]) / e.s[0], bb = (y - e.c[1]) / e.s[1], c = (z - e.c[2]) / e.s[2];
  if (a * a + bb * bb + c * c <= 1) return true;
  return Math.abs(x - b.c[0]) <= b.s[0] && Math.abs(y - b.c[1]) <= b.s[1] && Math.abs(z - b.c[2]) <= b.s[2];
};
smallCase('lbm_trt_les', [32, 16, 16],
  L => { L.trt = true; L.lambda = 3 / 16; L.les = true; L.csm = 0.1; L.tau = 0.505; L.u0 = 0.08; },
  { inside: ellBox({ c: [10, 8, 8], s: [5, 3, 2.5] }, { c: [14, 7, 8], s: [3, 2, 1.5] }),
    box: { x0: 4, y0: 4, z0: 4, x1: 18, y1: 12, z1: 12 }, ground: false }, 100,
  { raw: { steps: 100, fields: ['rho', 'ux', 'uy', 'uz'] } });

{
  // a wheeled ground vehicle: rolling road + spinning wheels exercise the
  // moving-wall correction and the SOLID_WALL (not force-counted) floor
  const NX = 40, NY = 16, NZ = 18;
  const b = new Shapes.Body('sedan');
  b.fit(NX, NY, NZ);
  b.pos = [NX * 0.36, 1, NZ * 0.5];
  smallCase('lbm_moving', [NX, NY, NZ], L => { },
    { inside: b.makeTest(), box: b.box(), ground: true, wallVel: b.makeWheelVel(0.06), roadVel: 0.06 }, 80,
    { raw: { steps: 80, fields: ['rho', 'ux', 'uy', 'uz'] } });
}

smallCase('lbm_outlet_body', [24, 12, 12], L => { L.trt = false; L.tau = 0.6; L.u0 = 0.05; },
  { inside: (x, y, z) => x > 18 && y > 4 && y < 8 && z > 4 && z < 8, box: { x0: 17, y0: 3, z0: 3, x1: 30, y1: 9, z1: 9 }, ground: false }, 30);

smallCase('lbm_revox', [32, 16, 16], L => { },
  { inside: sphere(10, 8, 8, 3), box: sphereBox(10, 8, 8, 3), ground: false }, 30,
  { at: [1, 30],
    more: (L, out) => {
      out.revox = voxelizeRecorded(L, sphere(12.5, 7.5, 8, 3), sphereBox(12.5, 7.5, 8, 3), true);
      const r = run(L, 30, [31, 60]);
      out.forces2 = r.forces; out.snaps2 = r.snaps;
    } });

if (want('lbm_pool3')) {
  // The worker-pool path (js/pool.js + js/lbmWorker.js) emulated in one
  // thread: with double buffering every worker reads f and writes a disjoint
  // part of g, so running the slabs one after another is equivalent.
  const dims = [32, 16, 16], NW = 3, nsteps = 40;
  const L = new LBM(dims[0], dims[1], dims[2]);
  L.tau = 0.52; L.les = true; L.u0 = 0.07; L.trt = false; L.lambda = 3 / 16; L.csm = 0.1; // not sent to workers
  L.reset();
  const vox = voxelizeRecorded(L, ellBox({ c: [10, 8, 8], s: [5, 3, 2.5] }, { c: [14, 7, 8], s: [3, 2, 1.5] }),
    { x0: 4, y0: 4, z0: 4, x1: 18, y1: 12, z1: 12 }, false);
  const bufs = L.buffers();
  const W = [];
  for (let id = 0; id < NW; id++) {
    const w = new LBM(dims[0], dims[1], dims[2], { buffers: bufs });
    w.tau = L.tau; w.les = L.les; w.u0 = L.u0; w.movingWalls = L.movingWalls;   // the 'params' message
    const lo = 1, hi = dims[2] - 1, span = hi - lo;
    w.z0 = lo + Math.floor(span * id / NW); w.z1 = lo + Math.floor(span * (id + 1) / NW);
    w.bz0 = id === 0 ? 0 : w.z0; w.bz1 = id === NW - 1 ? dims[2] : w.z1;
    W.push(w);
  }
  const forces = new Float64Array(nsteps * 3);
  const slabs = W.map(w => [w.z0, w.z1, w.bz0, w.bz1]);
  const parity0 = L.parity;
  for (let s = 0; s < nsteps; s++) {
    const per = [];
    for (let id = 0; id < NW; id++) {
      const w = W[id];
      w.parity = (parity0 + s) & 1;
      w.stepSlab(w.z0, w.z1);
      w._boundaries(w.g, w.bz0, w.bz1, id === 0, id === NW - 1);
      per.push([w._sf[0], w._sf[1], w._sf[2]]);
    }
    let fx = 0, fy = 0, fz = 0;
    for (let id = 0; id < NW; id++) { fx += per[id][0]; fy += per[id][1]; fz += per[id][2]; }
    forces[3 * s] = fx; forces[3 * s + 1] = fy; forces[3 * s + 2] = fz;
  }
  L.parity ^= (nsteps & 1); L.steps += nsteps;
  const pooled = stateHashes(L);
  // the same run single-threaded with the parameters the workers actually used
  const S = new LBM(dims[0], dims[1], dims[2]);
  S.tau = 0.52; S.les = true; S.u0 = 0.07; S.reset();
  S.voxelize(ellBox({ c: [10, 8, 8], s: [5, 3, 2.5] }, { c: [14, 7, 8], s: [3, 2, 1.5] }),
    { x0: 4, y0: 4, z0: 4, x1: 18, y1: 12, z1: 12 }, false);
  const single = run(S, nsteps, [nsteps]);
  write('lbm_pool3', { dims, nWorkers: NW, slabs, main_params: params(L), worker_params: params(W[0]),
    voxelize: vox, nsteps, forces: pack(forces, [nsteps, 3]), pooled,
    single_forces: single.forces, single: single.snaps[nsteps] });
}

if (want('lbm_c1')) {
  // plain reference method on the CPU-runnable workload: 96x48x48 BGK,
  // sphere r=8 at (32,24,24), u0=0.05, tau=0.56, 2000 steps
  const t0 = Date.now();
  const L = new LBM(96, 48, 48);
  L.trt = false; L.les = false; L.tau = 0.56; L.u0 = 0.05;
  L.reset();
  const vox = voxelizeRecorded(L, sphere(32, 24, 24, 8), sphereBox(32, 24, 24, 8), false);
  const forces = new Float64Array(2000 * 3);
  const snaps = {};
  let tStep = 0;
  for (let s = 0; s < 2000; s++) {
    const ts = Date.now();
    L.step();
    tStep += Date.now() - ts;
    forces.set(L.force, 3 * s);
    if ([1, 100, 500, 2000].includes(L.steps)) snaps[L.steps] = stateHashes(L);
  }
  const z = 24, sl = (a) => a.slice(L.nx * L.ny * z, L.nx * L.ny * (z + 1));
  write('lbm_c1', { dims: [96, 48, 48], params: params(L), voxelize: vox, nsteps: 2000,
    forces: pack(forces, [2000, 3]), snaps,
    slice_z24: { rho: pack(sl(L.rho), [48, 96]), ux: pack(sl(L.ux), [48, 96]), uy: pack(sl(L.uy), [48, 96]) },
    node_ms_per_step: tStep / 2000, node_version: process.version });
  console.log(`  lbm_c1: ${((Date.now() - t0) / 1000).toFixed(1)} s, ${(tStep / 2000).toFixed(2)} ms/step`);
}

if (want('lbm_c2')) {
  // TRT (Lambda = 3/16) + LES (Cs = 0.1) wake of the synthetic mesh body
  // through the app's own mesh pipeline: Shapes.addModel (orient to unit
  // length, Voxel.build at res 200) and the grid-backed Body.makeTest, at
  // scale = 1/k so the body keeps its size in cells. 256x128x128, 5000 steps.
  const t0 = Date.now();
  const NX = 256, NY = 128, NZ = 128, NSTEPS = 5000;
  const P0 = synth.ellipsoidBox();
  const P2 = P0.slice();
  const k = Shapes.orient(P2, 0, 1, true, false);
  const def = Shapes.addModel('c2body', 'c2body', P0.slice(), { forward: 0, up: 1, flip: true, ground: false, res: 200 });
  const L = new LBM(NX, NY, NZ);
  L.trt = true; L.lambda = 3 / 16; L.les = true; L.csm = 0.1; L.tau = 0.505; L.u0 = 0.08;
  L.reset();
  const b = new Shapes.Body('c2body');
  b.scale = 1 / k;
  b.pos = [0.3 * NX, NY / 2, NZ / 2];
  b._updateMat();
  const vox = voxelizeRecorded(L, b.makeTest(), b.box(), false);
  vox.body = { scale: b.scale, pos: b.pos, M: b.M, ext: b.ext };
  const v = def.vox;
  const forces = new Float64Array(NSTEPS * 3);
  const snaps = {};
  let tStep = 0;
  for (let s = 0; s < NSTEPS; s++) {
    const ts = Date.now();
    L.step();
    tStep += Date.now() - ts;
    forces.set(L.force, 3 * s);
    if ([1, 100, 1000, NSTEPS].includes(L.steps)) snaps[L.steps] = stateHashes(L);
    if (L.steps % 500 === 0) console.log(`  lbm_c2: step ${L.steps}, ${(tStep / L.steps).toFixed(1)} ms/step`);
  }
  write('lbm_c2', { dims: [NX, NY, NZ], params: params(L), nsteps: NSTEPS,
    mesh: { sha256: sha(P0), ntri: P0.length / 9, oriented_sha256: sha(def.tris), k,
      addModel: { forward: 0, up: 1, flip: true, ground: false, res: 200 } },
    vox: { nx: v.nx, ny: v.ny, nz: v.nz, cell: v.cell, min: v.min, inv: v.inv, ext: v.ext,
      grid_sha256: sha(v.grid), stats: v.stats },
    voxelize: vox, forces: pack(forces, [NSTEPS, 3]), snaps,
    node_ms_per_step: tStep / NSTEPS, node_version: process.version });
  console.log(`  lbm_c2: ${((Date.now() - t0) / 1000).toFixed(1)} s, ${(tStep / NSTEPS).toFixed(2)} ms/step`);
}

if (want('lbm_bodies')) {
  // the app's analytic bodies (js/shapes.js catalogue) placed as js/main.js:453
  // places them, some pitched and yawed, rasterised through Body.makeTest and
  // (ground vehicles) makeWheelVel with a rolling road (js/main.js:607-608)
  const NX = 96, NY = 44, NZ = 44;
  const cases = [];
  const poses = { sedan: [0, 0.12], sports: [0.03, -0.2], truck: [0, 0], airplane: [0.1, 0.3], wing: [0.14, 0],
    sphere: [0, 0], cube: [0.2, 0.4], cylinder: [0, 0.25], teardrop: [-0.1, 0.15] };
  for (const key of Object.keys(poses)) {
    const b = new Shapes.Body(key);
    b.fit(NX, NY, NZ);
    const ground = !!b.def.ground;
    b.pos = [NX * 0.36, ground ? 1 : NY * 0.5, NZ * 0.5];
    b.pitch = poses[key][0]; b.yaw = poses[key][1];
    b._updateMat();
    const L = new LBM(NX, NY, NZ);
    L.u0 = 0.06;
    L.reset();
    const rec = voxelizeRecorded(L, b.makeTest(), b.box(), ground, ground ? b.makeWheelVel(L.u0) : null,
      ground ? L.u0 : 0);
    rec.key = key;
    rec.body = { M: b.M, pos: b.pos, scale: b.scale, pitch: b.pitch, yaw: b.yaw,
      parts: b.parts.map(p => ({ t: p.t, c: p.c, s: p.s, m: p.m, naca: p.naca || null, wheel: !!p.wheel })) };
    rec.rolling_u0 = ground ? L.u0 : 0;
    cases.push(rec);
  }
  write('lbm_bodies', { dims: [NX, NY, NZ], cases });
}

// test.html procedure on two of its bluff bodies (js body via Shapes, replayed as a mask)
for (const shape of ['cube', 'sphere']) {
  const name = 'lbm_testpage_' + shape;
  if (!want(name)) continue;
  const t0 = Date.now();
  const NX = 96, NY = 44, NZ = 44, WARM = 400, AVG = 300;
  const L = new LBM(NX, NY, NZ);
  L.tau = 0.5 + Math.pow(10, -1.1 - 2.3 * 0.62);
  const b = new Shapes.Body(shape);
  b.fit(NX, NY, NZ);
  b.pos = [NX * 0.36, NY * 0.5, NZ * 0.5];
  const vox = voxelizeRecorded(L, b.makeTest(), b.box(), false);
  const forces = new Float64Array((WARM + AVG) * 3);
  let Fx = 0, Fy = 0, n = 0;
  for (let t = 0; t < WARM + AVG; t++) {
    L.step();
    forces.set(L.force, 3 * t);
    if (t >= WARM) { Fx += L.force[0]; Fy += L.force[1]; n++; }
  }
  const q = 0.5 * L.u0 * L.u0 * L.frontalArea;
  write(name, { dims: [NX, NY, NZ], params: params(L), voxelize: vox, warm: WARM, avg: AVG,
    forces: pack(forces, [WARM + AVG, 3]), cd: Fx / n / q, cl: Fy / n / q, final: stateHashes(L) });
  console.log(`  ${name}: ${((Date.now() - t0) / 1000).toFixed(1)} s, cd=${(Fx / n / q).toFixed(4)}`);
}

// test.html procedure on the rest of its case table (test.html:37-49): the
// shape, whether it stands on the ground and its angle of attack; the body is
// placed, pitched and voxelized as the page does and replayed as a mask
if (want('lbm_testpage_cases')) {
  const t0 = Date.now();
  const NX = 96, NY = 44, NZ = 44, WARM = 400, AVG = 300;
  const table = [['cylinder', false, 0], ['teardrop', false, 0], ['sedan', true, 0], ['truck', true, 0],
    ['wing', false, -8], ['wing', false, 0], ['wing', false, 8], ['wing', false, 16], ['airplane', false, 2]];
  const cases = [];
  for (const [shape, ground, aoa] of table) {
    const L = new LBM(NX, NY, NZ);
    L.tau = 0.5 + Math.pow(10, -1.1 - 2.3 * 0.62);
    const b = new Shapes.Body(shape);
    b.fit(NX, NY, NZ);
    b.pos = [NX * 0.36, ground ? 1 : NY * 0.5, NZ * 0.5];
    b.pitch = -aoa * Math.PI / 180;
    const vox = voxelizeRecorded(L, b.makeTest(), b.box(), ground);
    const planform = b.def.ref === 'planform';
    const A = planform ? L.planformArea : L.frontalArea;
    const forces = new Float64Array((WARM + AVG) * 3);
    let Fx = 0, Fy = 0, n = 0;
    for (let t = 0; t < WARM + AVG; t++) {
      L.step();
      forces.set(L.force, 3 * t);
      if (t >= WARM) { Fx += L.force[0]; Fy += L.force[1]; n++; }
    }
    const q = 0.5 * L.u0 * L.u0 * A;
    cases.push({ shape, ground, aoa, planform, area: A, voxelize: vox, forces: pack(forces, [WARM + AVG, 3]),
      cd: Fx / n / q, cl: Fy / n / q, final: stateHashes(L) });
    console.log(`  ${shape} aoa=${aoa}: cd=${(Fx / n / q).toFixed(4)} cl=${(Fy / n / q).toFixed(4)}`);
  }
  write('lbm_testpage_cases', { dims: [NX, NY, NZ], params: params(new LBM(NX, NY, NZ)), tau: 0.5 + Math.pow(10, -1.1 - 2.3 * 0.62),
    warm: WARM, avg: AVG, cases });
  console.log(`  lbm_testpage_cases: ${((Date.now() - t0) / 1000).toFixed(1)} s`);
}

// ------------------------------------------------------------ voxel cases
function boxMesh(sx, sy, sz) {        // closed box soup, 12 triangles
  const h = [sx / 2, sy / 2, sz / 2], v = [];
  const quads = [
    [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]], [[-1, -1, 1], [-1, 1, 1], [1, 1, 1], [1, -1, 1]],
    [[-1, -1, -1], [-1, 1, -1], [-1, 1, 1], [-1, -1, 1]], [[1, -1, -1], [1, -1, 1], [1, 1, 1], [1, 1, -1]],
    [[-1, -1, -1], [-1, -1, 1], [1, -1, 1], [1, -1, -1]], [[-1, 1, -1], [1, 1, -1], [1, 1, 1], [-1, 1, 1]]];
  for (const f of quads) {
    const p = f.map(c => [c[0] * h[0], c[1] * h[1], c[2] * h[2]]);
    v.push(...p[0], ...p[1], ...p[2], ...p[0], ...p[2], ...p[3]);
  }
  return new Float32Array(v);
}
function sphereMesh(r, seg) {         // UV sphere soup, 2*seg*seg triangles
  const v = [];
  const P = (u, t) => [r * Math.sin(t) * Math.cos(u), r * Math.cos(t), r * Math.sin(t) * Math.sin(u)];
  for (let i = 0; i < seg; i++) for (let j = 0; j < seg; j++) {
    const u0 = i / seg * 2 * Math.PI, u1 = (i + 1) / seg * 2 * Math.PI;
    const t0 = j / seg * Math.PI, t1 = (j + 1) / seg * Math.PI;
    const a = P(u0, t0), b = P(u1, t0), c = P(u1, t1), d = P(u0, t1);
    v.push(...a, ...b, ...c, ...a, ...c, ...d);
  }
  return new Float32Array(v);
}

function voxCase(name, P, res, storeMesh, extra) {
  if (!want(name)) return;
  const t0 = Date.now();
  const inP = new Float32Array(P);                 // build() must not modify its input
  const v = Voxel.build(P, res);
  const out = { res, ntri: P.length / 9, mesh_sha256: sha(inP),
    nx: v.nx, ny: v.ny, nz: v.nz, cell: v.cell, min: v.min, inv: v.inv, ext: v.ext, centre: v.centre,
    bbox: v.bbox, stats: v.stats, grid_sha256: sha(v.grid), grid_bits: pack(bits(v.grid)),
    bounds: Voxel.bounds(P) };
  if (storeMesh) out.mesh = pack(inP, [P.length / 9, 9]);
  if (extra) Object.assign(out, extra(v));
  write(name, out);
  console.log(`  ${name}: ${((Date.now() - t0) / 1000).toFixed(1)} s, solid=${v.stats.solid} openness=${v.stats.openness}`);
}

{
  const P = boxMesh(4, 2, 3);
  Shapes.orient(P, 0, 1, false, false);
  voxCase('vox_box', P, 96, true);
}
{
  const P = sphereMesh(1, 48);
  Shapes.orient(P, 0, 1, false, false);
  voxCase('vox_sphere', P, 96, true);
}
voxCase('vox_c2body', synth.ellipsoidBox(), 96, false);
voxCase('vox_c2body_r200', synth.ellipsoidBox(), 200, false);
voxCase('vox_leaky', synth.noisySphere({ level: 4, R: 20, holes: 0.02, seed: 5 }), 64, false);
voxCase('vox_leaky_l6', synth.noisySphere({ level: 6, R: 56, holes: 0.005, seed: 5 }), 120, false);
// the 640x320x320 scale case: level-7 icosphere (327 680 triangles before the
// 0.5 % deletion), R = 56 cells, voxelized at one lattice cell per voxel
voxCase('vox_leaky_l7', synth.noisySphere({ level: 7, R: 56, holes: 0.005, seed: 5 }), 128, false);
{
  // a small soup dense enough along one ray to hit the 512-crossing cap
  const tris = [];
  for (let k = 0; k < 700; k++) {
    const x = k * 0.01;
    tris.push(x, -1, -1, x, 1, -1, x, 0, 1.5);
  }
  voxCase('vox_hitcap', new Float32Array(tris), 40, true);
}

if (want('vox_errors')) {
  const msg = f => { try { f(); return null; } catch (e) { return e.message; } };
  write('vox_errors', {
    empty: msg(() => Voxel.build(new Float32Array(0), 96)),
    short: msg(() => Voxel.build(new Float32Array(8), 96)),
    degenerate: msg(() => Voxel.build(new Float32Array(9).fill(1.5), 96)),
    too_large: msg(() => Voxel.build(new Float32Array([0, 0, 0, 1, 0, 0, 0, 1, 1]), 400))
  });
}

if (want('vox_meshcube_lookup')) {
  // testmodels.html end-to-end geometry: mesh cube via addModel + grid lookup
  const NX = 80, NY = 36, NZ = 36;
  const P = boxMesh(0.44, 0.44, 0.44);
  Shapes.addModel('meshcube', 'meshcube', P, { forward: 0, up: 1, flip: false, ground: false, res: 96 });
  Shapes.BODIES.meshcube.fill = 1.0;
  const out = {};
  for (const key of ['cube', 'meshcube']) {
    const L = new LBM(NX, NY, NZ);
    const b = new Shapes.Body(key);
    b.fit(NX, NY, NZ);
    b.pos = [NX * 0.36, NY / 2, NZ / 2];
    b._updateMat();
    const rec = voxelizeRecorded(L, b.makeTest(), b.box(), false);
    rec.body = { scale: b.scale, pos: b.pos, M: b.M, ext: b.ext };
    out[key] = rec;
  }
  const v = Shapes.BODIES.meshcube.vox;
  out.vox = { nx: v.nx, ny: v.ny, nz: v.nz, cell: v.cell, min: v.min, inv: v.inv, grid_bits: pack(bits(v.grid)),
    mesh: pack(new Float32Array(Shapes.BODIES.meshcube.tris)), stats: v.stats };
  write('vox_meshcube_lookup', out);
}

if (want('mesh_io')) {
  // MeshIO.parse (js/meshio.js:322-340) and MeshIO.meta on the files of gen_meshio.py, under
  // their own names and, for the ones with a mesh extension, under a name
  // whose extension the dispatch does not know (so the content is sniffed)
  const dir = path.join(OUT, 'meshio');
  const cases = [];
  for (const file of fs.readdirSync(dir).sort()) {
    const names = [file];
    if (/\.(stl|obj|glb)$/.test(file)) names.push(file.replace(/\.[a-z]+$/, '') + '_as.dat');
    for (const name of names) {
      const b = fs.readFileSync(path.join(dir, file));
      const ab = b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength);
      // MeshIO.meta (js/meshio.js:277-314): the credits the file carries
      let meta;
      try { meta = { meta: MeshIO.meta(name, ab) }; } catch (e) { meta = { meta_error: e.message }; }
      try {
        const P = MeshIO.parse(name, ab);
        let nan = 0;
        for (let i = 0; i < P.length; i++) if (P[i] !== P[i]) nan++;
        cases.push(Object.assign({ file, name, ok: true, count: P.length, sha256: sha(P), nan,
          head: Array.from(P.subarray(0, 9)).map(v => (v !== v ? 'NaN' : v)) }, meta));
      } catch (e) {
        cases.push(Object.assign({ file, name, ok: false, error: e.message }, meta));
      }
    }
  }
  write('mesh_io', { cases });
  console.log(`  mesh_io: ${cases.length} cases`);
}

