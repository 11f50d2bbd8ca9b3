/* voxel.js -- drop-in for the reference's Voxel module (js/voxel.js:32-235)
 * with the ray-parity voxelizer running on the GPU.
 *
 * Voxel.build(P, res) returns the same object as the reference:
 * {grid, nx, ny, nz, cell, min, inv, ext, centre, bbox, stats:{tris, solid,
 * openness}}, bit-identical, and throws the same Error messages.
 */
'use strict';
const { native } = require('./lbm.js');

function bounds(P) {
  return native().voxelBounds(P instanceof Float32Array ? P : Float32Array.from(P));
}

function build(P, res) {
  res = res || 96;
  const F = P instanceof Float32Array ? P : Float32Array.from(P);
  const plan = native().voxelPlan(F, res);
  const grid = new Uint8Array(plan.nx * plan.ny * plan.nz);
  const m = native().voxelBuild(F, res, grid);
  return {
    grid, nx: m.nx, ny: m.ny, nz: m.nz, cell: m.cell,
    min: m.min, inv: m.inv, ext: m.ext, centre: m.centre, bbox: m.bbox,
    stats: { tris: m.tris, solid: m.solid, openness: m.openness }
  };
}

const Voxel = { build, bounds };
module.exports = Voxel;
if (typeof globalThis !== 'undefined') globalThis.Voxel = Voxel;
