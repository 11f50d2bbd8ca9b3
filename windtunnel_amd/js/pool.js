/* pool.js -- the reference's worker pool and worker protocol on one GPU.
 *
 * `Pool` keeps the interface of js/pool.js:49-131 (start, syncParams,
 * step(n) -> Promise of per-step [fx,fy,fz], dispose, slabs, stepForces) and
 * `GpuWorker` speaks the message protocol of js/lbmWorker.js:35-80
 * ({cmd:'init'|'params'|'step'} in, {ok:'init'|'step'} out, per-(step,
 * worker) forces at forces[(s*N+id)*3+k]).
 *
 * On one GPU the workers are logical z-slabs of one device lattice, split
 * exactly as the worker 'init' handler splits them (js/lbmWorker.js:43-48).
 * A step of every slab must finish before the next step starts (the
 * reference's barrier), so worker 0 launches the whole batch for all slabs
 * (wt_step_slabs, on a libuv worker thread so the event loop stays free) and
 * answers its 'step' message when the device is done, the other workers at
 * once; the pool totals the per-slab forces in worker order as
 * js/pool.js:108-118 does.
 *
 * The 'params' message carries tau, les, u0 and movingWalls as in the
 * reference; trt, lambda and csm are optional additions, and workers that do
 * not receive them keep the constructor defaults, as reference workers do.
 */
'use strict';
const { LBM, native } = require('./lbm.js');

const WORKER_DEFAULTS = { trt: true, lambda: 1 / 4, csm: 0.14 };

class GpuWorker {
  constructor() {
    this.onmessage = null;
    this.onerror = null;
    this._terminated = false;
  }

  postMessage(m) {
    const fail = (e) => { if (this.onerror) this.onerror({ message: e.message, error: e }); };
    const reply = (r) => { if (r && !this._terminated && this.onmessage) this.onmessage({ data: r }); };
    Promise.resolve().then(() => {
      if (this._terminated) return;
      try {
        const r = this._handle(m);
        // a step batch answers when the device is done with it
        if (r instanceof Promise) r.then(reply, fail);
        else reply(r);
      } catch (e) {
        fail(e);
      }
    });
  }

  terminate() { this._terminated = true; }

  _handle(m) {
    if (m.cmd === 'init') {
      this.id = m.id; this.nWorkers = m.nWorkers;
      this.h = m.buffers.lattice;
      this.forces = new Float64Array(m.forces);
      this.nx = m.nx; this.ny = m.ny; this.nz = m.nz;
      const lo = 1, hi = m.nz - 1, span = hi - lo;
      this.z0 = lo + Math.floor(span * this.id / this.nWorkers);
      this.z1 = lo + Math.floor(span * (this.id + 1) / this.nWorkers);
      this.params = Object.assign({ tau: 0.5035, les: true, u0: 0.06, movingWalls: false }, WORKER_DEFAULTS);
      return { ok: 'init', id: this.id, z0: this.z0, z1: this.z1 };
    }
    if (m.cmd === 'params') {
      for (const k of ['tau', 'les', 'u0', 'movingWalls', 'trt', 'lambda', 'csm']) {
        if (m[k] !== undefined) this.params[k] = m[k];
      }
      return null;
    }
    if (m.cmd === 'step') {
      if (this.id === 0) {
        const p = this.params, N = this.nWorkers;
        native().setParams(this.h, { tau: p.tau, csm: p.csm, lambda: p.lambda, u0: p.u0,
          les: p.les ? 1 : 0, trt: p.trt ? 1 : 0, movingWalls: p.movingWalls ? 1 : 0 });
        native().setParity(this.h, m.parity & 1);
        // the batch runs on a libuv worker thread (stepSlabsAsync), so the
        // event loop stays free, as with the reference's worker threads
        const f = new Float64Array(m.n * N * 3);
        return native().stepSlabsAsync(this.h, m.n, N, f).then(() => {
          this.forces.set(f, 0);
          return { ok: 'step', id: this.id };
        });
      }
      return { ok: 'step', id: this.id };
    }
    return null;
  }
}

/** Why the GPU pool is unavailable, or null when it is available (js/pool.js:20-33). */
function unavailableReason() {
  if (!LBM.available()) return 'windtunnel GPU addon not built';
  return null;
}

/** One GPU: one logical worker is the natural split (js/pool.js:44-47). */
function suggestedWorkers() { return 1; }

class Pool {
  constructor(sim, nWorkers) {
    this.sim = sim;
    this.n = nWorkers;
    this.workers = [];
    this.busy = false;
    this.maxBatch = 16;
    this.forces = new Float64Array(new SharedArrayBuffer(this.maxBatch * nWorkers * 3 * 8));
    this.stepForces = [];
    this.makeWorker = () => new GpuWorker();
  }

  async start() {
    const s = this.sim;
    const buffers = s.buffers();
    const ready = [];
    for (let i = 0; i < this.n; i++) {
      const w = this.makeWorker();
      this.workers.push(w);
      ready.push(new Promise((res, rej) => {
        w.onmessage = (e) => { if (e.data.ok === 'init') res(e.data); };
        w.onerror = (e) => rej(new Error(e.message || 'worker failed'));
      }));
      w.postMessage({ cmd: 'init', id: i, nWorkers: this.n, nx: s.nx, ny: s.ny, nz: s.nz,
        buffers, forces: this.forces.buffer });
    }
    this.slabs = await Promise.all(ready);
    return this;
  }

  /** Push the settings the workers keep their own copy of (js/pool.js:86-94). */
  syncParams(full) {
    const s = this.sim;
    for (const w of this.workers) {
      const m = { cmd: 'params', tau: s.tau, les: s.les, u0: s.u0, movingWalls: s.movingWalls };
      if (full) { m.trt = s.trt; m.lambda = s.lambda; m.csm = s.csm; }
      w.postMessage(m);
    }
  }

  /** Advance n steps across all slabs; resolves with per-step forces (js/pool.js:97-125). */
  step(n) {
    n = Math.max(1, Math.min(this.maxBatch, n | 0));
    this.busy = true;
    return new Promise((resolve, reject) => {
      let left = this.n;
      const parity = this.sim.parity;
      const finish = () => {
        this.stepForces.length = 0;
        for (let s = 0; s < n; s++) {
          let fx = 0, fy = 0, fz = 0;
          for (let w = 0; w < this.n; w++) {
            const o = (s * this.n + w) * 3;
            fx += this.forces[o]; fy += this.forces[o + 1]; fz += this.forces[o + 2];
          }
          this.stepForces.push([fx, fy, fz]);
        }
        // the device lattice flipped its own parity; keep the page's counters in step
        this.sim.steps += n;
        const last = this.stepForces[n - 1];
        this.sim.force[0] = last[0]; this.sim.force[1] = last[1]; this.sim.force[2] = last[2];
        this.sim._dirty();
        this.busy = false;
        resolve(this.stepForces);
      };
      for (const w of this.workers) {
        w.onmessage = () => { if (--left === 0) finish(); };
        w.onerror = (e) => { this.busy = false; reject(e.error || new Error(e.message)); };
        w.postMessage({ cmd: 'step', n, parity });
      }
    });
  }

  dispose() {
    for (const w of this.workers) w.terminate();
    this.workers.length = 0;
  }
}

module.exports = { Pool, GpuWorker, unavailableReason, suggestedWorkers, WORKER_DEFAULTS };
if (typeof globalThis !== 'undefined') globalThis.LBMPool = module.exports;
