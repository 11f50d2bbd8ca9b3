/* meshio.js -- drop-in for the reference's MeshIO module (js/meshio.js).
 *
 * parse (js/meshio.js:322-340, with its STL, OBJ and binary glTF readers):
 * the file's bytes into a Float32Array triangle soup, 9 floats per
 * triangle, parsed natively by the library (wt_mesh_parse) with the same
 * results and the same Error messages.
 *
 * meta (js/meshio.js:216-314): the credits a model file carries, as the
 * reference's page shows them -- glTF's asset block, the 80-byte header of a
 * binary STL, or the comment lines that open an OBJ or ASCII STL -- with
 * the same fields, the same cleaning and the same null when the file says
 * nothing. It reads a few KiB of text and runs here in JavaScript; the
 * result is untrusted text from the file, to be shown as text.
 */
'use strict';
const { native } = require('./lbm.js');

/** parse(name, ArrayBuffer | typed array) -> Float32Array */
function parse(name, buf) {
  return native().meshParse(String(name), buf);
}

const GLB_MAGIC = 0x46546c67;  // 'glTF'
const GLB_JSON = 0x4e4f534a;   // 'JSON' chunk
const KEEP = 300;              // characters kept per field
const COMMENT_LINES = 13;      // comment lines read from an OBJ / ASCII STL head
const HEAD_BYTES = 4096;

function bytesOf(buf) {
  if (buf instanceof ArrayBuffer) return new Uint8Array(buf);
  if (ArrayBuffer.isView(buf)) return new Uint8Array(buf.buffer, buf.byteOffset, buf.byteLength);
  throw new TypeError('MeshIO.meta: expected an ArrayBuffer or a typed array');
}
const u32 = (b, at) => new DataView(b.buffer, b.byteOffset, b.byteLength).getUint32(at, true);
const text = (b) => new TextDecoder().decode(b);
// runs of characters outside printable ASCII and U+00A0..U+FFFF become one space
const clean = (s) => s.replace(/[^\x20-\x7E\u00A0-\uFFFF]+/g, ' ').trim();

// "key: value" / "key = value" lines, the first spelling that gives a value
const FIELDS = [
  ['title', ['title', 'name']],
  ['author', ['author', 'creator', 'artist', 'by']],
  ['license', ['licen[cs]e']],
  ['source', ['source', 'url']],
];
function fieldsOf(s) {
  const out = {};
  for (const [field, spellings] of FIELDS) {
    let v;
    for (const key of spellings) {
      const hit = new RegExp('^[#\\s]*' + key + '\\s*[:=]\\s*(.+)$', 'im').exec(s);
      v = hit ? hit[1].trim() : undefined;
      if (v) break;
    }
    out[field] = v;
  }
  return out;
}

// glTF credits: asset.extras.{title,author,license,source}, asset.copyright
// and asset.generator; a copyright string that names no author or licence
// of its own is searched for the fields instead
function gltfCredits(doc) {
  const asset = (doc && doc.asset) || {};
  const extras = asset.extras || {};
  const m = { title: extras.title, author: extras.author, license: extras.license, source: extras.source,
    copyright: asset.copyright, generator: asset.generator };
  if (asset.copyright && !m.author && !m.license) Object.assign(m, fieldsOf(asset.copyright));
  return m;
}

// the JSON chunk of a binary glTF, or null (no chunk, or one that does not
// lie inside the bytes or does not parse)
function glbDocument(b) {
  if (b.byteLength < 20 || u32(b, 0) !== GLB_MAGIC) return null;
  const end = Math.min(u32(b, 8), b.byteLength);
  for (let at = 12; at + 8 <= end;) {
    const len = u32(b, at);
    if (u32(b, at + 4) === GLB_JSON) {
      if (at + 8 + len > b.byteLength) return null;
      try { return JSON.parse(text(b.subarray(at + 8, at + 8 + len))); } catch (e) { return null; }
    }
    at += 8 + len + ((4 - (len & 3)) & 3);
  }
  return null;
}

const binarySTL = (b) => b.byteLength >= 84 && 84 + u32(b, 80) * 50 === b.byteLength;

// the comment lines that open the file, up to its first statement
function commentCredits(b) {
  const kept = [];
  for (const raw of text(b.subarray(0, Math.min(HEAD_BYTES, b.byteLength))).split('\n')) {
    const line = raw.trim();
    if (line[0] === '#') kept.push(clean(line.replace(/^#+\s?/, '')));
    else if (line !== '') break;
    if (kept.length >= COMMENT_LINES) break;
  }
  const block = kept.filter((s) => s !== '').join('\n');
  return block ? Object.assign({ notes: block }, fieldsOf(block)) : null;
}

/** meta(name, ArrayBuffer | typed array) -> {title, author, license, source,
 *  copyright, generator, notes} (each only when present) or null */
function meta(name, buf) {
  const b = bytesOf(buf);
  const ext = (name.split('.').pop() || '').toLowerCase();
  let m = null;
  if (ext === 'glb' || (b.byteLength > 12 && u32(b, 0) === GLB_MAGIC)) {
    const doc = glbDocument(b);
    if (doc) m = gltfCredits(doc);
  } else if (ext === 'gltf') {
    try { m = gltfCredits(JSON.parse(text(b))); } catch (e) { m = null; }
  } else if (ext === 'stl' && binarySTL(b)) {
    const head = clean(text(b.subarray(0, 80)));
    if (head) m = Object.assign({ notes: head }, fieldsOf(head));
  } else {
    m = commentCredits(b);
  }
  if (!m) return null;
  for (const k of Object.keys(m)) {
    if (typeof m[k] === 'string') {
      const s = clean(m[k]).slice(0, KEEP);
      if (s) m[k] = s; else delete m[k];
    } else if (m[k] === undefined) {
      delete m[k];
    }
  }
  return Object.keys(m).length ? m : null;
}

const MeshIO = { parse, meta };
module.exports = MeshIO;
if (typeof globalThis !== 'undefined') globalThis.MeshIO = MeshIO;
