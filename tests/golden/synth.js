/* Deterministic synthetic triangle soups, JavaScript side.
 *
 * Mirrors windtunnel_amd/synth.py operation for operation so that both emit
 * the same Float32 bytes: only +, -, *, /, sqrt and floor on doubles, integer
 * hashing with Math.imul, and one rounding to float32 at the end. No
 * transcendental function is used, because libm results differ across
 * platforms while these operations are correctly rounded everywhere.
 */
'use strict';

const M32 = 0x100000000;

/** Counter-based 32-bit generator (golden-ratio increment, murmur3 finaliser). */
function makeRng(seed) {
  let s = seed >>> 0;
  return function () {
    s = (s + 0x9E3779B9) >>> 0;
    let z = s;
    z = Math.imul(z ^ (z >>> 16), 0x85EBCA6B) >>> 0;
    z = Math.imul(z ^ (z >>> 13), 0xC2B2AE35) >>> 0;
    z = (z ^ (z >>> 16)) >>> 0;
    return z / M32;
  };
}

/** Irwin-Hall(12) - 6: zero mean, unit variance, no transcendental. */
function gauss(rng) {
  let s = 0;
  for (let k = 0; k < 12; k++) s += rng();
  return s - 6;
}

function hash3(seed, i, j, k) {
  let h = seed >>> 0;
  h = Math.imul(h ^ (i >>> 0), 0x27D4EB2D) >>> 0; h = (h ^ (h >>> 15)) >>> 0;
  h = Math.imul(h ^ (j >>> 0), 0x165667B1) >>> 0; h = (h ^ (h >>> 15)) >>> 0;
  h = Math.imul(h ^ (k >>> 0), 0x9E3779B1) >>> 0; h = (h ^ (h >>> 16)) >>> 0;
  return h;
}

function lerp(a, b, t) { return a + (b - a) * t; }

/** Value noise in [-1,1] with smoothstep weights. */
function vnoise(seed, x, y, z) {
  const ix = Math.floor(x), iy = Math.floor(y), iz = Math.floor(z);
  const fx = x - ix, fy = y - iy, fz = z - iz;
  const sx = fx * fx * (3 - 2 * fx), sy = fy * fy * (3 - 2 * fy), sz = fz * fz * (3 - 2 * fz);
  const v = (a, b, c) => hash3(seed, ix + a, iy + b, iz + c) / M32 * 2 - 1;
  const x00 = lerp(v(0, 0, 0), v(1, 0, 0), sx), x10 = lerp(v(0, 1, 0), v(1, 1, 0), sx);
  const x01 = lerp(v(0, 0, 1), v(1, 0, 1), sx), x11 = lerp(v(0, 1, 1), v(1, 1, 1), sx);
  return lerp(lerp(x00, x10, sy), lerp(x01, x11, sy), sz);
}

/** Four octaves, base frequency 2, amplitude halving; result in [-1,1]. */
function fbm(seed, x, y, z) {
  let s = 0, a = 1, f = 2, norm = 0;
  for (let o = 0; o < 4; o++) {
    s += a * vnoise(seed + o, x * f, y * f, z * f);
    norm += a; a *= 0.5; f *= 2;
  }
  return s / norm;
}

const ICO_FACES = [
  [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]];

function normalise(x, y, z) {
  const s = Math.sqrt(x * x + y * y + z * z);
  return [x / s, y / s, z / s];
}

/** Unit icosphere: vertices (flat xyz doubles) and faces (flat indices). */
function icosphere(level) {
  const t = (1 + Math.sqrt(5)) / 2;
  const base = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
    [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]];
  const V = [];
  for (const p of base) V.push(...normalise(p[0], p[1], p[2]));
  let F = [];
  for (const f of ICO_FACES) F.push(f[0], f[1], f[2]);
  for (let l = 0; l < level; l++) {
    const cache = new Map();
    const mid = (a, b) => {
      const key = a < b ? a * 2097152 + b : b * 2097152 + a;
      let m = cache.get(key);
      if (m === undefined) {
        const p = normalise((V[3 * a] + V[3 * b]) * 0.5, (V[3 * a + 1] + V[3 * b + 1]) * 0.5,
          (V[3 * a + 2] + V[3 * b + 2]) * 0.5);
        m = V.length / 3; V.push(p[0], p[1], p[2]); cache.set(key, m);
      }
      return m;
    };
    const G = [];
    for (let k = 0; k < F.length; k += 3) {
      const a = F[k], b = F[k + 1], c = F[k + 2];
      const ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
      G.push(a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca);
    }
    F = G;
  }
  return { V, F };
}

function soup(V, F, keep) {
  let nt = 0;
  for (let k = 0; k < F.length / 3; k++) if (!keep || keep[k]) nt++;
  const P = new Float32Array(nt * 9);
  let o = 0;
  for (let k = 0; k < F.length / 3; k++) {
    if (keep && !keep[k]) continue;
    for (let c = 0; c < 3; c++) {
      const v = F[3 * k + c];
      P[o++] = V[3 * v]; P[o++] = V[3 * v + 1]; P[o++] = V[3 * v + 2];
    }
  }
  return P;
}

/** Axis-aligned box surface with unit quads, vertices shared across faces. */
function boxSurface(nxq, nyq, nzq) {
  const V = [], F = [], index = new Map();
  const vid = (i, j, k) => {
    const key = (i * 4096 + j) * 4096 + k;
    let v = index.get(key);
    if (v === undefined) { v = V.length / 3; V.push(i, j, k); index.set(key, v); }
    return v;
  };
  const n = [nxq, nyq, nzq];
  // faces in order -x, +x, -y, +y, -z, +z; (u,v) axes chosen so normals point out
  const faces = [[0, 0, 2, 1], [0, 1, 1, 2], [1, 0, 0, 2], [1, 1, 2, 0], [2, 0, 1, 0], [2, 1, 0, 1]];
  for (const [ax, side, ua, va] of faces) {
    for (let a = 0; a < n[ua]; a++) {
      for (let b = 0; b < n[va]; b++) {
        const corner = (da, db) => {
          const p = [0, 0, 0];
          p[ax] = side ? n[ax] : 0; p[ua] = a + da; p[va] = b + db;
          return vid(p[0], p[1], p[2]);
        };
        const v00 = corner(0, 0), v10 = corner(1, 0), v11 = corner(1, 1), v01 = corner(0, 1);
        F.push(v00, v10, v11, v00, v11, v01);
      }
    }
  }
  return { V, F };
}

/**
 * Bluff body of the TRT+LES workloads, in lattice cells around its own
 * centre: an ellipsoid (semi-axes ex,ey,ez) unioned with a box (size
 * bx,by,bz, centre bc). Triangles of either part whose centroid lies inside
 * the other part are dropped. Every vertex is jittered radially by
 * sigma*N(0,1) cells (Irwin-Hall normal), ellipsoid vertices first.
 */
function ellipsoidBox(opts) {
  const o = Object.assign({ level: 5, ex: 40, ey: 14, ez: 12, bx: 30, by: 16, bz: 10,
    bc: [22, -6, 0], sigma: 0.2, seed: 2, scale: 1 }, opts || {});
  const rng = makeRng(o.seed);
  const sc = o.scale;
  const ico = icosphere(o.level);
  const EV = ico.V.slice();
  const ex = o.ex * sc, ey = o.ey * sc, ez = o.ez * sc;
  for (let v = 0; v < EV.length / 3; v++) {
    const d = o.sigma * sc * gauss(rng);
    const ux = ico.V[3 * v], uy = ico.V[3 * v + 1], uz = ico.V[3 * v + 2];
    EV[3 * v] = ux * ex + ux * d; EV[3 * v + 1] = uy * ey + uy * d; EV[3 * v + 2] = uz * ez + uz * d;
  }
  const bx = o.bx * sc, by = o.by * sc, bz = o.bz * sc;
  const box = boxSurface(o.bx * sc, o.by * sc, o.bz * sc);
  const cx = o.bc[0] * sc, cy = o.bc[1] * sc, cz = o.bc[2] * sc;
  const BV = box.V.slice();
  for (let v = 0; v < BV.length / 3; v++) {
    const px = box.V[3 * v] - bx * 0.5, py = box.V[3 * v + 1] - by * 0.5, pz = box.V[3 * v + 2] - bz * 0.5;
    const u = normalise(px, py, pz);
    const d = o.sigma * sc * gauss(rng);
    BV[3 * v] = cx + px + u[0] * d; BV[3 * v + 1] = cy + py + u[1] * d; BV[3 * v + 2] = cz + pz + u[2] * d;
  }
  const inEll = (x, y, z) => { const a = x / ex, b = y / ey, c = z / ez; return a * a + b * b + c * c < 1; };
  const inBox = (x, y, z) => Math.abs(x - cx) < bx * 0.5 && Math.abs(y - cy) < by * 0.5 && Math.abs(z - cz) < bz * 0.5;
  const centroid = (V, F, k) => {
    const a = F[3 * k], b = F[3 * k + 1], c = F[3 * k + 2];
    return [(V[3 * a] + V[3 * b] + V[3 * c]) / 3, (V[3 * a + 1] + V[3 * b + 1] + V[3 * c + 1]) / 3,
      (V[3 * a + 2] + V[3 * b + 2] + V[3 * c + 2]) / 3];
  };
  const keepE = new Uint8Array(ico.F.length / 3), keepB = new Uint8Array(box.F.length / 3);
  for (let k = 0; k < keepE.length; k++) { const c = centroid(EV, ico.F, k); keepE[k] = inBox(c[0], c[1], c[2]) ? 0 : 1; }
  for (let k = 0; k < keepB.length; k++) { const c = centroid(BV, box.F, k); keepB[k] = inEll(c[0], c[1], c[2]) ? 0 : 1; }
  const A = soup(EV, ico.F, keepE), B = soup(BV, box.F, keepB);
  const P = new Float32Array(A.length + B.length);
  P.set(A, 0); P.set(B, A.length);
  return P;
}

/**
 * Noisy, leaky icosphere of the large voxelizer workload: radius R cells,
 * radius modulated by amp*R*fbm(direction), and each triangle deleted with
 * probability `holes`.
 */
function noisySphere(opts) {
  const o = Object.assign({ level: 7, R: 56, amp: 0.15, holes: 0.005, seed: 5 }, opts || {});
  const ico = icosphere(o.level);
  const V = ico.V.slice();
  for (let v = 0; v < V.length / 3; v++) {
    const ux = ico.V[3 * v], uy = ico.V[3 * v + 1], uz = ico.V[3 * v + 2];
    const r = o.R + o.amp * o.R * fbm(o.seed, ux, uy, uz);
    V[3 * v] = ux * r; V[3 * v + 1] = uy * r; V[3 * v + 2] = uz * r;
  }
  const rng = makeRng(o.seed + 1000);
  const keep = new Uint8Array(ico.F.length / 3);
  for (let k = 0; k < keep.length; k++) keep[k] = rng() < o.holes ? 0 : 1;
  return soup(V, ico.F, keep);
}

module.exports = { makeRng, gauss, hash3, vnoise, fbm, icosphere, boxSurface, ellipsoidBox, noisySphere };
