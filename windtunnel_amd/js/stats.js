'use strict';
// Force-coefficient statistics and the run report of the wind-tunnel app
// (js/main.js:60-213 and 1228-1322) for a Node host driving the solver:
// settling, the sliding pair average of consecutive step forces, the ring
// of the last HIST samples, the explicit capture window, the blocking-method
// error bar, the convergence state and the CSV report with the app's rows.
// The browser app keeps these in closures over its page state; here they are
// one object, and the report takes the test conditions as an argument.

const HIST = 6000;       // samples kept in the ring (js/main.js:65)
const CAP_MAX = 250000;  // samples a capture keeps (js/main.js:74)
const BLOCKS = 16;       // blocking-method block count (js/main.js:163)
const SETTLE = 150;      // steps skipped after a reset (js/main.js:83)
const RHO_AIR = 1.225;   // kg/m^3 (js/main.js:12)
const NU_AIR = 1.48e-5;  // m^2/s (js/main.js:13)

/** Mean and standard error of x[0..n) from 16 block means; below 4 samples
 * per block, the plain mean and sigma/sqrt(n) (js/main.js:153-179). */
function blockStats(x, n) {
  const per = Math.floor(n / BLOCKS);
  if (per < 4) {
    let m = 0;
    for (let i = 0; i < n; i++) m += x[i];
    m /= n;
    let ss = 0;
    for (let i = 0; i < n; i++) ss += (x[i] - m) * (x[i] - m);
    return { mean: m, sem: Math.sqrt(ss / Math.max(1, n - 1) / n), blocked: false };
  }
  const means = new Float64Array(BLOCKS);
  let mean = 0;
  for (let b = 0; b < BLOCKS; b++) {
    let s = 0;
    for (let i = b * per, e = (b + 1) * per; i < e; i++) s += x[i];
    means[b] = s / per;
    mean += means[b];
  }
  mean /= BLOCKS;
  let ss = 0;
  for (let b = 0; b < BLOCKS; b++) ss += (means[b] - mean) * (means[b] - mean);
  return { mean, sem: Math.sqrt(ss / (BLOCKS - 1) / BLOCKS), blocked: true, used: BLOCKS * per };
}

class ForceStats {
  constructor(settle = SETTLE) {
    this.ring = { cd: new Float32Array(HIST), cl: new Float32Array(HIST), cy: new Float32Array(HIST) };
    this.capture = { on: false, cd: [], cl: [], cy: [], from: 0, to: 0 };
    this.reset(settle);
  }

  /** New test conditions: forget the samples, skip `settle` steps; a running
   * capture ends (js/main.js:76-84). `step` is the solver's step count. */
  reset(settle = SETTLE, step = 0) {
    this.ema = [0, 0, 0];
    this.count = 0;
    this.n = 0;           // samples taken (may exceed HIST)
    this.prev = null;
    if (this.capture.on) { this.capture.on = false; this.capture.to = step; }
    this.settle = settle;
  }

  /** Start (a new) or stop the capture window at solver step `step`
   * (js/main.js:131-135). */
  setCapture(on, step) {
    if (on) this.capture = { on: true, cd: [], cl: [], cy: [], from: step, to: step };
    else { this.capture.on = false; this.capture.to = step; }
  }

  /**
   * Fold in the force of one completed step (js/main.js:105-129): after the
   * settling steps, sample = (F_k + F_{k-1}) / 2 / (u0^2 A / 2), the pair
   * average that cancels the lattice's step-parity oscillation of the
   * momentum-exchange force. `area` is the reference area in cells.
   */
  accumulate(force, u0, area, step) {
    if (this.settle > 0) { this.settle--; return; }
    const a = 1 / Math.min(++this.count, 500);
    for (let k = 0; k < 3; k++) this.ema[k] += (force[k] - this.ema[k]) * a;
    if (this.prev === null) { this.prev = Float64Array.from(force); return; }
    const q = 0.5 * u0 * u0 * area;
    const s = [0, 1, 2].map((k) => 0.5 * (force[k] + this.prev[k]) / q);
    const i = this.n % HIST;
    this.ring.cd[i] = s[0]; this.ring.cl[i] = s[1]; this.ring.cy[i] = s[2];
    this.prev.set(force);
    this.n++;
    const c = this.capture;
    if (c.on && c.cd.length < CAP_MAX) {
      c.cd.push(s[0]); c.cl.push(s[1]); c.cy.push(s[2]);
      c.to = step;
    }
  }

  /** The ring's samples of one coefficient, oldest first (js/main.js:147-153). */
  ordered(which) {
    const src = this.ring[which], n = Math.min(this.n, HIST);
    const start = this.n <= HIST ? 0 : this.n % HIST;
    const out = new Float64Array(n);
    for (let i = 0; i < n; i++) out[i] = src[(start + i) % HIST];
    return out;
  }

  /** Statistics of the ring and the convergence state, or null below 60
   * samples (js/main.js:186-207). */
  runStats() {
    const x = this.ordered('cd'), n = x.length;
    if (n < 60) return null;
    const cd = blockStats(x, n);
    const h = n >> 1;
    let m1 = 0, m2 = 0;
    for (let i = 0; i < h; i++) m1 += x[i];
    for (let i = h; i < n; i++) m2 += x[i];
    const drift = Math.abs(m1 / h - m2 / (n - h));
    const cl = blockStats(this.ordered('cl'), n), cy = blockStats(this.ordered('cy'), n);
    const rel = 1.96 * cd.sem / Math.max(1e-9, Math.abs(cd.mean));
    let state = 'noisy';
    if (n < 400) state = 'averaging';
    else if (drift > 2.5 * cd.sem) state = 'drifting';
    else if (rel < 0.03) state = 'converged';
    return { n, cd, cl, cy, drift, rel, state };
  }

  /**
   * The run report as CSV text (js/main.js:1228-1322): conditions, results
   * with 95 % intervals and the raw series, from the capture when it holds
   * samples, else from the ring. `sim` is the solver object (LBM); `cond`
   * holds what the page knows: { name, custom, credit: {title, author,
   * license, source} | null, ref: 'frontal' | 'planform', scale (cells per
   * body length), ground, lengthM, windMs, aoa, yaw }. `generated` is the
   * ISO time stamp (now when omitted).
   */
  runCSV(sim, cond, generated) {
    const c = this.capture, captured = c.cd.length > 0;
    let series;
    if (captured) series = { cd: c.cd, cl: c.cl, cy: c.cy, n: c.cd.length };
    else {
      const cd = this.ordered('cd');
      series = { cd, cl: this.ordered('cl'), cy: this.ordered('cy'), n: cd.length };
    }
    let st = null;
    if (series.n >= 60) {
      const rs = captured ? null : this.runStats();
      st = {
        n: series.n, cd: blockStats(series.cd, series.n), cl: blockStats(series.cl, series.n),
        cy: blockStats(series.cy, series.n), state: captured ? 'captured window' : (rs || {}).state || 'live window',
      };
    }
    const planform = cond.ref === 'planform';
    const cells = planform ? sim.planformArea : sim.frontalArea;
    const mPerCell = cond.lengthM / cond.scale;
    const area = cells * mPerCell * mPerCell;
    const U = cond.windMs, q = 0.5 * RHO_AIR * U * U;
    const out = ['Wind Tunnel run report,,', 'parameter,value,unit'];
    const row = (k, v, u) => out.push([k, v, u === undefined ? '' : u].join(','));
    row('generated', generated || new Date().toISOString());
    row('body', JSON.stringify(cond.name));
    row('body_source', cond.custom ? 'uploaded' : 'bundled');
    if (cond.credit) {
      for (const k of ['title', 'author', 'license', 'source']) row('model_' + k, JSON.stringify(cond.credit[k] || ''));
    }
    row('grid_nx', sim.nx, 'cells'); row('grid_ny', sim.ny, 'cells'); row('grid_nz', sim.nz, 'cells');
    row('grid_total', sim.n, 'cells');
    row('solid_cells', sim.solidCells, 'cells');
    row('reference_area_type', planform ? 'planform' : 'frontal');
    row('reference_area', cells, 'cells');
    row('reference_area_real', area.toFixed(4), 'm^2');
    row('blockage', (100 * sim.frontalArea / (sim.ny * sim.nz)).toFixed(2), '%');
    row('body_length', cond.lengthM, 'm');
    row('cells_per_body_length', cond.scale.toFixed(2), 'cells');
    row('wind_speed', U, 'm/s');
    row('wind_speed_kmh', (U * 3.6).toFixed(1), 'km/h');
    row('mach', (U / 343).toFixed(4));
    row('dynamic_pressure', q.toFixed(2), 'Pa');
    row('angle_of_attack', cond.aoa, 'deg');
    row('yaw', cond.yaw, 'deg');
    row('ground_plane', cond.ground ? 'yes' : 'no');
    row('lattice_u0', sim.u0);
    row('lattice_tau', sim.tau.toFixed(6));
    row('lattice_nu', sim.nu.toExponential(4));
    row('les_smagorinsky', sim.les ? 'on' : 'off');
    row('smagorinsky_C', sim.csm);
    row('reynolds_simulated', Math.round(sim.u0 * cond.scale / sim.nu));
    row('reynolds_full_scale', Math.round(U * cond.lengthM / NU_AIR));
    row('timesteps_total', sim.steps);
    row('sample_source', captured ? 'explicit capture' : 'rolling live window');
    if (captured) {
      row('capture_from_step', c.from);
      row('capture_to_step', c.to);
      if (c.cd.length >= CAP_MAX) row('capture_truncated_at', CAP_MAX, 'samples');
    }
    row('samples', st ? st.n : 0);
    row('convergence', st ? st.state : 'not started');
    out.push(',,', 'result,mean,ci95,unit');
    if (st) {
      const res = (k, m, sem, u) => out.push([k, m.toFixed(5), (1.96 * sem).toFixed(5), u].join(','));
      res('Cd', st.cd.mean, st.cd.sem, '-');
      res('Cl', st.cl.mean, st.cl.sem, '-');
      res('Cy', st.cy.mean, st.cy.sem, '-');
      // dimensional values: coefficient * q * A (* U), multiplied in that order
      res('drag', st.cd.mean * q * area, st.cd.sem * q * area, 'N');
      res('lift', st.cl.mean * q * area, st.cl.sem * q * area, 'N');
      res('side', st.cy.mean * q * area, st.cy.sem * q * area, 'N');
      res('power', st.cd.mean * q * area * U, st.cd.sem * q * area * U, 'W');
    }
    out.push(',,',
      '# ci95 is from the blocking method (16 blocks) which accounts for,,',
      '# correlation between consecutive timesteps; sigma/sqrt(N) would,,',
      '# understate it badly in a shedding wake.,,',
      ',,', 'sample,Cd,Cl,Cy');
    for (let i = 0; i < series.n; i++) {
      out.push(i + ',' + series.cd[i].toFixed(5) + ',' + series.cl[i].toFixed(5) + ',' + series.cy[i].toFixed(5));
    }
    return out.join('\n');
  }
}

module.exports = { ForceStats, blockStats, HIST, CAP_MAX, BLOCKS, SETTLE, RHO_AIR, NU_AIR };
