/* Node drop-in checks that need no GPU: the modules load, construction
 * fails loudly without a device (no CPU fallback), voxelizer argument
 * errors carry the reference's messages, and the pool's failure path
 * rejects so the page can fall back (js/main.js:1414-1418). */
'use strict';
const path = require('path');
const { load, assert, checkMeshMeta } = require('./fixtures.js');
const JS = path.join(__dirname, '..', '..', 'windtunnel_amd', 'js');
const { LBM } = require(path.join(JS, 'lbm.js'));
const { Pool, unavailableReason, suggestedWorkers } = require(path.join(JS, 'pool.js'));
const Voxel = require(path.join(JS, 'voxel.js'));

(async () => {
  assert(LBM.available(), 'addon loads');
  assert(unavailableReason() === null && suggestedWorkers() === 1, 'pool availability');
  assert(LBM.Q === 19 && LBM.OPP === undefined && LBM.W[0] === 1 / 3 && LBM.SOLID_WALL === 2, 'statics');

  let err = null;
  try { new LBM(16, 8, 8); } catch (e) { err = e; }
  assert(err && err.code === 'WT-2', 'no device -> WT_E_DEVICE, got ' + (err && err.code));

  // the addon binds the C ABI one to one
  const addon = require(path.join(JS, '..', 'wt_napi.node'));
  const exported = ('abiVersion create setParams getParams reset voxelize voxelizeGrid voxelizeParts step ' +
    'stepSlabs stepSlab boundaries setParity getState getFields getPops setPops getSolid getWalls timeSteps ' +
    'deviceMemory setBoundaryMode slice sampleMany meshParse setGray adjointForward adjointBackward adjointRelease adjointRecord voxelBounds voxelPlan voxelBuild stepSlabsAsync').split(' ');
  assert(Object.keys(addon).sort().join() === exported.sort().join(), 'addon exports: ' + Object.keys(addon).join(' '));
  err = null;
  try { addon.deviceMemory(); } catch (e) { err = e; }
  assert(err && err.code === 'WT-2', 'deviceMemory without a device -> WT_E_DEVICE');

  const ref = load('vox_errors');
  const msg = f => { try { f(); return null; } catch (e) { return e.message; } };
  assert(msg(() => Voxel.build(new Float32Array(0), 96)) === ref.empty, 'empty');
  assert(msg(() => Voxel.build(new Float32Array(9).fill(1.5), 96)) === ref.degenerate, 'degenerate');
  assert(msg(() => Voxel.build(new Float32Array([0, 0, 0, 1, 0, 0, 0, 1, 1]), 400)) === ref.too_large, 'too large');
  assert(Voxel.bounds(new Float32Array([1, 2, 3, -1, 5, 0])).join() === '-1,2,0,1,5,3', 'bounds');

  // MeshIO.parse needs no device: the reference's results and messages (mesh_io)
  const MeshIO = require(path.join(JS, 'meshio.js'));
  const fs = require('fs'), crypto = require('crypto');
  const dir = path.join(__dirname, '..', 'golden', 'fixtures', 'meshio');
  for (const c of load('mesh_io').cases) {
    const b = fs.readFileSync(path.join(dir, c.file));
    const ab = b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength);
    let P = null, m = null;
    try { P = MeshIO.parse(c.name, ab); } catch (e) { m = e.message; }
    if (!c.ok) { assert(m === c.error, c.name + ' message: ' + m); continue; }
    assert(P && P.length === c.count, c.name + ' count');
    const h = crypto.createHash('sha256').update(Buffer.from(P.buffer, P.byteOffset, P.byteLength)).digest('hex');
    assert(h === c.sha256, c.name + ' soup');
  }
  checkMeshMeta(MeshIO);

  // a worker that fails rejects the pool's promise, leaving the caller to fall back
  const sim = { nx: 8, ny: 8, nz: 8, tau: 0.6, les: true, u0: 0.05, movingWalls: false, parity: 0, steps: 0,
    force: new Float64Array(3), buffers: () => ({ lattice: null }), _dirty() {} };
  const pool = new Pool(sim, 2);
  pool.makeWorker = () => {
    const w = { onmessage: null, onerror: null, terminate() {} };
    w.postMessage = m => Promise.resolve().then(() => {
      if (m.cmd === 'init') w.onmessage({ data: { ok: 'init', id: m.id, z0: 1, z1: 7 } });
      else if (m.cmd === 'step') w.onerror({ message: 'device lost' });
    });
    return w;
  };
  await pool.start();
  let rejected = null;
  try { await pool.step(4); } catch (e) { rejected = e; }
  assert(rejected && /device lost/.test(rejected.message) && pool.busy === false, 'pool rejects on worker error');
  console.log('all passed');
})().catch(e => { console.log('FAIL ' + e.message); process.exit(1); });
