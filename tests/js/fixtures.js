/* Golden-fixture reading and helpers shared by the Node shim tests. */
'use strict';
const fs = require('fs');
const path = require('path');
const zlib = require('zlib');
const crypto = require('crypto');

const DIR = path.join(__dirname, '..', 'golden', 'fixtures');
const CTOR = { f4: Float32Array, f8: Float64Array, u1: Uint8Array, i4: Int32Array, u4: Uint32Array };

function load(name) { return JSON.parse(fs.readFileSync(path.join(DIR, name + '.json'), 'utf8')); }

function arr(rec) {
  const raw = zlib.gunzipSync(Buffer.from(rec.gz_b64, 'base64'));
  const T = CTOR[rec.dtype];
  const out = new T(raw.length / T.BYTES_PER_ELEMENT);
  new Uint8Array(out.buffer).set(raw);
  return out;
}

function sha(a) {
  return crypto.createHash('sha256').update(Buffer.from(a.buffer, a.byteOffset, a.byteLength)).digest('hex');
}

function assert(c, msg) { if (!c) throw new Error('assertion failed: ' + msg); }

function close(a, b, rtol, atol, msg) {
  for (let i = 0; i < a.length; i++) {
    if (!(Math.abs(a[i] - b[i]) <= atol + rtol * Math.abs(b[i]))) {
      throw new Error(`${msg}: element ${i}: ${a[i]} vs ${b[i]}`);
    }
  }
}

function applyParams(L, p) {
  L.tau = p.tau; L.les = p.les; L.csm = p.csm; L.trt = p.trt; L.lambda = p.lambda; L.u0 = p.u0;
}

function replay(L, rec) {
  L.voxelizeMask(rec.box, arr(rec.mask), arr(rec.wv), rec.ground, rec.roadVel);
}

// MeshIO.meta against the reference's credits for every mesh_io file, under
// each name it was run under (meta needs no device); returns the case count
function checkMeshMeta(MeshIO) {
  const canon = (v) => JSON.stringify(v && typeof v === 'object'
    ? Object.keys(v).sort().reduce((o, k) => { o[k] = v[k]; return o; }, {}) : v);
  let n = 0;
  for (const c of load('mesh_io').cases) {
    const b = fs.readFileSync(path.join(DIR, 'meshio', c.file));
    const got = MeshIO.meta(c.name, b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
    assert(!('meta_error' in c), c.name + ': the reference threw');
    assert(canon(got) === canon(c.meta), `${c.name} meta: ${canon(got)} vs ${canon(c.meta)}`);
    n++;
  }
  return n;
}

module.exports = { load, arr, sha, assert, close, applyParams, replay, checkMeshMeta };
