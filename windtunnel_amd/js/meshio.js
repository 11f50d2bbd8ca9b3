/* meshio.js -- drop-in for the reference's MeshIO.parse (js/meshio.js:322-340
 * with its STL, OBJ and binary glTF readers): the file's bytes into a
 * Float32Array triangle soup, 9 floats per triangle, parsed natively by the
 * library (wt_mesh_parse) with the same results and the same Error
 * messages. The attribution helper meta() is page code and is not part of
 * this module.
 */
'use strict';
const { native } = require('./lbm.js');

/** parse(name, ArrayBuffer | typed array) -> Float32Array */
function parse(name, buf) {
  return native().meshParse(String(name), buf);
}

const MeshIO = { parse };
module.exports = MeshIO;
if (typeof globalThis !== 'undefined') globalThis.MeshIO = MeshIO;
