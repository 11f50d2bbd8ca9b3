/* lbm.js -- drop-in for the reference's solver object, backed by one MI355X.
 *
 * `LBM` keeps the constructor, settable properties, methods and statics of
 * js/lbm.js:51-552 (class LBM): the lattice lives in HBM behind the
 * libwindtunnel C ABI (through the wt_napi N-API addon), and the
 * macroscopic fields, solid map and populations are host typed arrays
 * downloaded on first read after each step (the reference reads them every
 * frame, js/main.js:925-970, so they are cached, never pushed per step).
 *
 * Solver settings are plain properties as in the reference (js/lbm.js:106-111)
 * and are pushed to the device at the start of every step/voxelize/reset.
 * There is no CPU fallback: construction throws without the HIP library or a
 * gfx950 device.
 */
'use strict';
const path = require('path');

let addon = null;
let addonError = null;
try {
  addon = require(path.join(__dirname, '..', 'wt_napi.node'));
} catch (e) {
  addonError = e;
}
function native() {
  if (!addon) throw new Error('windtunnel addon not built: ' + (addonError && addonError.message));
  return addon;
}

const Q = 19;
const EX = new Int32Array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]);
const EY = new Int32Array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1]);
const EZ = new Int32Array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]);
const W = new Float64Array(Q).map((_, q) => (q === 0 ? 1 / 3 : q <= 6 ? 1 / 18 : 1 / 36));
const FLUID = 0, SOLID_OBJ = 1, SOLID_WALL = 2;
const STORE = { f32: 0, f16: 1, f64: 2 }, ARITH = { f32: 0, f64: 1 }, LAYOUT = { ab: 0, esoteric: 1 };

class LBM {
  /**
   * @param opts.storage 'f32' (default) | 'f16' | 'f64'
   * @param opts.arith   'f32' (default) | 'f64' (binary64: the reference's own numbers)
   * @param opts.layout  'ab' (default, two buffers) | 'esoteric' (in place)
   */
  constructor(nx, ny, nz, opts) {
    opts = opts || {};
    this.nx = nx; this.ny = ny; this.nz = nz;
    this.n = nx * ny * nz;
    this.storage = opts.storage || 'f32';
    this.arith = opts.arith || 'f32';
    this.layout = opts.layout || 'ab';
    if (!(this.storage in STORE) || !(this.arith in ARITH) || !(this.layout in LAYOUT)) {
      throw new Error('unknown storage/arith/layout option');
    }
    this._h = native().create(nx, ny, nz, STORE[this.storage], ARITH[this.arith], LAYOUT[this.layout]);
    this.u0 = 0.06;
    this.tau = 0.5035;
    this.les = true;
    this.csm = 0.14;
    this.trt = true;
    this.lambda = 1 / 4;
    this.movingWalls = false;
    this.force = new Float64Array(3);
    this.frontalArea = 1;
    this.planformArea = 1;
    this.solidCells = 0;
    this.ground = false;
    this.steps = 0;
    this.rideHeight = 0;
    this._fields = null;
    this._solid = null;
    this._uw = null;
    this._f3 = new Float64Array(3);
    this._sf = new Float64Array(3);
    this.reset();
  }

  /** Push the solver properties to the device (js/lbm.js:106-111). */
  _sync() {
    native().setParams(this._h, { tau: this.tau, csm: this.csm, lambda: this.lambda, u0: this.u0,
      les: this.les ? 1 : 0, trt: this.trt ? 1 : 0, movingWalls: this.movingWalls ? 1 : 0 });
  }

  _dirty() { this._fields = null; }

  /** Boundaries: 'reference' (js/lbm.js:473-512) or 'channel' (free-slip
   * side walls, equilibrium outflow; wt_set_boundary_mode). */
  get boundaryMode() { return this._bc || 'reference'; }
  set boundaryMode(m) {
    native().setBoundaryMode(this._h, m === 'channel' ? 1 : (m === 'reference' ? 0 : -1));
    this._bc = m;
  }

  get parity() { return native().getState(this._h).parity; }
  set parity(p) { native().setParity(this._h, p & 1); }
  get nu() { return (this.tau - 0.5) / 3; }
  idx(x, y, z) { return x + this.nx * (y + this.ny * z); }

  _macro() {
    if (!this._fields) {
      const T = this.storage === 'f64' ? Float64Array : Float32Array;
      const F = { rho: new T(this.n), ux: new T(this.n), uy: new T(this.n), uz: new T(this.n) };
      native().getFields(this._h, F.rho, F.ux, F.uy, F.uz);
      this._fields = F;
    }
    return this._fields;
  }
  get rho() { return this._macro().rho; }
  get ux() { return this._macro().ux; }
  get uy() { return this._macro().uy; }
  get uz() { return this._macro().uz; }

  _masks() {
    if (!this._solid) {
      this._solid = { solid: new Uint8Array(this.n), near: new Uint8Array(this.n) };
      native().getSolid(this._h, this._solid.solid, this._solid.near);
    }
    return this._solid;
  }
  get solid() { return this._masks().solid; }
  get near() { return this._masks().near; }

  _walls() {
    if (!this._uw) {
      this._uw = { uwx: new Float32Array(this.n), uwy: new Float32Array(this.n), uwz: new Float32Array(this.n) };
      native().getWalls(this._h, this._uw.uwx, this._uw.uwy, this._uw.uwz);
    }
    return this._uw;
  }
  get uwx() { return this._walls().uwx; }
  get uwy() { return this._walls().uwy; }
  get uwz() { return this._walls().uwz; }

  _pops(which) {
    const a = this.storage === 'f64' ? new Float64Array(Q * this.n) : new Float32Array(Q * this.n);
    native().getPops(this._h, which, a);
    return a;
  }
  /** Current-timestep populations (read) and the other buffer (js/lbm.js:131-132). */
  get f() { return this._pops(0); }
  get g() { return this._pops(1); }

  /** The device lattice, for GPU workers attaching to it (js/lbm.js:135-140). */
  buffers() { return { lattice: this._h }; }

  reset() {
    this._sync();
    native().reset(this._h);
    this.steps = 0;
    this.force[0] = this.force[1] = this.force[2] = 0;
    this._dirty();
  }

  _applyStats(s) {
    this.ground = s.ground;
    this.movingWalls = s.movingWalls;
    this.frontalArea = s.frontalArea;
    this.planformArea = s.planformArea;
    this.solidCells = s.solidCells;
    this.rideHeight = s.rideHeight;
    this._solid = null;
    this._uw = null;
    this._dirty();
  }

  _clampBox(box) {
    const { nx, ny, nz } = this;
    return Int32Array.of(
      Math.max(1, Math.floor(box.x0)), Math.max(1, Math.floor(box.y0)), Math.max(1, Math.floor(box.z0)),
      Math.min(nx - 2, Math.ceil(box.x1)), Math.min(ny - 2, Math.ceil(box.y1)), Math.min(nz - 2, Math.ceil(box.z1)));
  }

  /**
   * Rasterise a body (js/lbm.js:191-266). The point tests are host closures,
   * evaluated here over the clamped box in the reference's loop order; the
   * lattice updates run on the device.
   */
  voxelize(inside, box, ground, wallVel, roadVel) {
    const b = this._clampBox(box);
    const bx = Math.max(0, b[3] - b[0] + 1), by = Math.max(0, b[4] - b[1] + 1), bz = Math.max(0, b[5] - b[2] + 1);
    const mask = new Uint8Array(bx * by * bz);
    const wv = [];
    const out = [0, 0, 0];
    let k = 0;
    for (let z = b[2]; z <= b[5]; z++) {
      for (let y = b[1]; y <= b[4]; y++) {
        for (let x = b[0]; x <= b[3]; x++, k++) {
          if (!inside(x + 0.5, y + 0.5, z + 0.5)) continue;
          mask[k] = 1;
          if (wallVel && wallVel(x + 0.5, y + 0.5, z + 0.5, out)) { mask[k] = 2; wv.push(out[0], out[1], out[2]); }
        }
      }
    }
    this._sync();
    this._applyStats(native().voxelize(this._h, b, mask, new Float32Array(wv), !!ground, roadVel || 0));
  }

  /** voxelize with the closures already evaluated: mask over the clamped box. */
  voxelizeMask(b, mask, wv, ground, roadVel) {
    this._sync();
    this._applyStats(native().voxelize(this._h, Int32Array.from(b), mask, wv || null, !!ground, roadVel || 0));
  }

  /**
   * Grid-backed body (a Voxel.build result placed by rotation M, position
   * pos and scale as Body.makeTest does, js/shapes.js:369-388) evaluated on
   * the GPU instead of through a host closure.
   */
  voxelizeGrid(vox, M, pos, scale, box, ground, roadVel) {
    this._sync();
    this._applyStats(native().voxelizeGrid(this._h, vox.grid, vox, Float64Array.from(M), Float64Array.from(pos),
      scale, this._clampBox(box), !!ground, roadVel || 0));
  }

  /**
   * voxelize(body.makeTest(), body.box(), ground, rolling ? body.makeWheelVel(u0) : null,
   * rolling ? u0 : 0) for a Shapes.Body (js/main.js:607-608, js/shapes.js:365-470), with the
   * point tests and wheel velocities evaluated on the GPU. Voxel-backed (uploaded) bodies go
   * through the grid lookup. `rollingU0` > 0 spins the wheels and moves the road.
   */
  voxelizeBody(body, ground, rollingU0) {
    body._updateMat();
    const u0 = rollingU0 > 0 ? rollingU0 : 0;
    if (body.def && body.def.custom) {
      this.voxelizeGrid(body.def.vox, body.M, body.pos, body.scale, body.box(), ground, u0);
      return;
    }
    const parts = body.parts.map(p => ({ t: p.t, c: p.c, s: p.s, m: p.m, naca: p.naca || null, wheel: !!p.wheel }));
    this._sync();
    this._applyStats(native().voxelizeParts(this._h, parts, Float64Array.from(body.M), Float64Array.from(body.pos),
      body.scale, this._clampBox(body.box()), !!ground, u0, u0));
  }

  /** Advance one step (js/lbm.js:285-291). */
  step() {
    this._sync();
    native().step(this._h, 1, this._f3);
    this.force[0] = this._f3[0]; this.force[1] = this._f3[1]; this.force[2] = this._f3[2];
    this.steps++;
    this._dirty();
  }

  /** Advance n steps in one device batch; returns the per-step forces [3n]. */
  stepMany(n) {
    const f = new Float64Array(3 * n);
    this._sync();
    native().step(this._h, n, f);
    this.force.set(f.subarray(3 * n - 3));
    this.steps += n;
    this._dirty();
    return f;
  }

  /** Device milliseconds for n steps (HIP events on the lattice stream). */
  timeSteps(n) {
    this._sync();
    const ms = native().timeSteps(this._h, n);
    this.steps += n;
    this._dirty();
    return ms;
  }

  /**
   * The app's slice view on the device (js/main.js:925-965): axis 'xy'
   * (plane z = k) or 'xz' (plane y = k), field 'speed' (|u|/u0), 'pressure'
   * (C_p of the neighbour-averaged density) or 'vorticity'. Returns
   * { values: Float32Array(nx*h), solid: Uint8Array(nx*h) }, x fastest;
   * colour mapping stays with the caller. u0 is the one of the last step.
   */
  slice(axis, k, field) {
    const ax = axis === 'xz' ? 1 : 0;
    const f = { speed: 0, pressure: 1, vorticity: 2 }[field];
    if (f === undefined) throw new Error('unknown slice field ' + field);
    const cells = this.nx * (ax ? this.nz : this.ny);
    const values = new Float32Array(cells), solid = new Uint8Array(cells);
    native().slice(this._h, ax, k, f, values, solid);
    return { values, solid };
  }

  /** sample() at many points on the device: xyz Float64Array(3m) -> Float64Array(3m). */
  sampleMany(xyz, out) {
    const o = out || new Float64Array(xyz.length);
    native().sampleMany(this._h, xyz, o);
    return o;
  }

  /**
   * sample() at one point (js/lbm.js:515-539): the device's batched sample
   * (wt_sample) with a batch of one, on the fields of the last step.
   */
  sample(x, y, z, out) {
    const p = this._probe || (this._probe = { xyz: new Float64Array(3), uvw: new Float64Array(3) });
    p.xyz[0] = x; p.xyz[1] = y; p.xyz[2] = z;
    this._sync();  // u0 answers for points off the lattice
    native().sampleMany(this._h, p.xyz, p.uvw);
    out[0] = p.uvw[0]; out[1] = p.uvw[1]; out[2] = p.uvw[2];
    return out;
  }

  /** isSolid() (js/lbm.js:541-545): the cell holding the point, by ToInt32 truncation. */
  isSolid(x, y, z) {
    const c = [x | 0, y | 0, z | 0], dims = [this.nx, this.ny, this.nz];
    for (let k = 0; k < 3; k++) if (c[k] < 0 || c[k] >= dims[k]) return false;
    return this.solid[this.idx(c[0], c[1], c[2])] !== 0;
  }

  /**
   * Collide + stream the z rows [z0,z1) (js/lbm.js:293-465) on the device.
   * The x/y boundary planes and the outlet copy of those rows come with the
   * same launch, so _boundaries() only has the z end planes left. Leaves the
   * slab's force in this._sf; no parity flip, as in the reference.
   */
  stepSlab(z0, z1) {
    this._sync();
    native().stepSlab(this._h, z0, z1, this._sf);
    this._dirty();
  }

  /** _boundaries(g, z0, z1, doLo, doHi) (js/lbm.js:467-512); g is the device's next buffer. */
  _boundaries(g, z0, z1, doLo, doHi) {
    const a = z0 === undefined ? 0 : z0, b = z1 === undefined ? this.nz : z1;
    const lo = z0 === undefined ? true : !!doLo, hi = z1 === undefined ? true : !!doHi;
    this._sync();
    native().boundaries(this._h, a, b, lo, hi);
    this._dirty();
  }

  // ---- gray bounce-back and its discrete adjoint (no reference counterpart)

  /** Gray weights, one per cell (0 plain fluid, (0,1) gray); null clears. */
  setGray(w) {
    native().setGray(this._h, w ? Float32Array.from(w) : null);
  }

  /** n steps as stepMany(n) does, recording sqrt(n) checkpoints; returns the forces [3n]. */
  adjointForward(n, interval) {
    const f = new Float64Array(3 * n);
    this._sync();
    native().adjointForward(this._h, n, interval || 0, f);
    this.force.set(f.subarray(3 * n - 3));
    this.steps += n;
    this._adjSteps = n;
    this._dirty();
    return f;
  }

  /** dJ/dw per cell for J = sum_k Fbar[3k..3k+2] . F_k over the recorded forward pass. */
  adjointBackward(Fbar) {
    const n = this._adjSteps || 0;
    const g = new Float64Array(this.n);
    const ms = native().adjointBackward(this._h, n, Float64Array.from(Fbar), g);
    return { dJdw: g, ms };
  }

  /** Id of the recorded forward pass (each adjointForward takes the next one). */
  adjointRecord() {
    return native().adjointRecord(this._h);
  }

  adjointRelease() {
    native().adjointRelease(this._h);
    this._adjSteps = 0;
  }
}

LBM.Q = Q; LBM.EX = EX; LBM.EY = EY; LBM.EZ = EZ; LBM.W = W;
LBM.FLUID = FLUID; LBM.SOLID_OBJ = SOLID_OBJ; LBM.SOLID_WALL = SOLID_WALL;
LBM.available = () => addon !== null;
LBM._native = native;

module.exports = { LBM, native };
if (typeof globalThis !== 'undefined') globalThis.LBM = LBM;
