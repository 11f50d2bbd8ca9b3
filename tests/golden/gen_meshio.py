"""Write the mesh-reader test files into tests/golden/fixtures/meshio/.

Small seeded STL (binary and ASCII), OBJ and binary glTF files that cover
the readers' paths: formatting variants, polygon fans, negative and
out-of-range indices, glTF node hierarchies with matrices and
translation/rotation/scale, strided and integer accessors, skipped
primitives, extension-less files that are sniffed, and the error cases.
tests/golden/gen_golden.js runs the reference's MeshIO.parse on each file
(case mesh_io) and records what it returns.

    python tests/golden/gen_meshio.py
"""
import json
import os
import struct

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures", "meshio")


def icosa():
    t = (1 + 5 ** 0.5) / 2
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    return V * 0.731, F


def write(name, data):
    with open(os.path.join(OUT, name), "wb") as fh:
        fh.write(data)


def stl_binary(tris, header=b"seeded icosahedron"):
    out = bytearray(header.ljust(80, b" "))
    out += struct.pack("<I", len(tris))
    for t in tris:
        out += struct.pack("<3f", 0, 0, 0) + struct.pack("<9f", *t) + b"\0\0"
    return bytes(out)


def glb(gltf: dict, bin_bytes: bytes, extra_chunks=()):
    js = json.dumps(gltf, separators=(",", ":")).encode()
    js += b" " * ((4 - len(js) % 4) % 4)
    chunks = struct.pack("<II", len(js), 0x4E4F534A) + js
    for typ, data in extra_chunks:
        chunks += struct.pack("<II", len(data), typ) + data
    if bin_bytes is not None:
        binb = bin_bytes + b"\0" * ((4 - len(bin_bytes) % 4) % 4)
        chunks += struct.pack("<II", len(binb), 0x004E4942) + binb
    return struct.pack("<III", 0x46546C67, 2, 12 + len(chunks)) + chunks


def main():
    os.makedirs(OUT, exist_ok=True)
    V, F = icosa()
    rng = np.random.default_rng(21)
    Vj = V + rng.normal(scale=0.01, size=V.shape)
    tris = [Vj[f].reshape(-1).astype(np.float32) for f in F]

    write("ico_binary.stl", stl_binary(tris))
    write("ico_binary_noext", stl_binary(tris[:7], b"solid but binary"))

    # ASCII STL with formatting variants and odd numbers
    lines = ["solid ico", ""]
    fmts = ["{:.6f}", "{:.9e}", "{:+.4f}", "{:.3E}", "{}"]
    for k, t in enumerate(tris):
        lines.append("  facet normal 0 0 0")
        lines.append("\touter loop")
        for v in range(3):
            xyz = [fmts[(k + v + c) % len(fmts)].format(float(t[3 * v + c])) for c in range(3)]
            sep = "   " if (k + v) % 3 == 0 else (" \t " if (k + v) % 3 == 1 else " ")
            lines.append("      vertex " + sep.join(xyz))
        lines.append("\tendloop")
        lines.append("  endfacet")
    lines += ["  facet normal 0 0 0", "    outer loop",
              "      vertex 1. .5 -.25e+1", "      vertex 1e 2 3", "      vertex --1 0x10 1e400",
              "    endloop", "  endfacet",
              "vertexvertex 4 5 6", "vertex 7,8 9", "endsolid ico", ""]
    write("ico_ascii.stl", "\r\n".join(lines).encode())
    write("ico_ascii_noext", ("solid x\n" + "\n".join(
        "vertex %r %r %r" % tuple(float(x) for x in t[3 * v:3 * v + 3]) for t in tris[:5] for v in range(3))).encode())
    write("empty_ascii.stl", b"solid nothing\nendsolid nothing\n")

    # OBJ: comments, vt/vn, v//vn faces, quads and pentagons, negative and
    # out-of-range indices, CRLF, a tab after v (not a vertex line)
    obj = ["# seeded icosahedron", "# author: test", "o ico", ""]
    for v in Vj:
        obj.append("v %.7f %.7f %.7f" % tuple(v))
    obj.append("vt 0.5 0.5")
    obj.append("vn 0 0 1")
    for k, f in enumerate(F):
        a, b, c = (int(i) + 1 for i in f)
        if k % 4 == 0:
            obj.append(f"f {a} {b} {c}")
        elif k % 4 == 1:
            obj.append(f"f {a}/1/1 {b}/1/1 {c}/1/1")
        elif k % 4 == 2:
            obj.append(f"f {a}//1 {b}//1 {c}//1")
        else:
            obj.append(f"f {a - 13} {b - 13} {c - 13}")     # negative: relative to the 12 vertices so far
    obj.append("v\t9 9 9")
    obj.append("v 1 2")
    obj.append("f 1 2 3 4 5")
    obj.append("f -1 -2 -3 -4")
    obj.append("f 1 x 3 4")
    obj.append("f 1 2 99")
    obj.append("f 1 0x5 2")
    obj.append("v 1e-3 Infinity -0")
    obj.append("f -1 1 2")
    write("ico.obj", "\r\n".join(obj).encode())
    write("ico_noext", ("\n".join(obj[:16]) + "\nf 1 2 3\nf 4 5 6\n").encode())
    write("nofaces.obj", b"# only vertices\nv 0 0 0\nv 1 0 0\n")

    # glTF: positions with a stride, uint16 / uint8 / uint32 indices, int16
    # positions, a line primitive, a node tree with a matrix and TRS
    P = Vj.astype(np.float32)
    pos_strided = b"".join(p.tobytes() + b"\xAA\xBB\xCC\xDD" for p in P)          # 16-byte stride
    idx16 = F.astype(np.uint16).reshape(-1).tobytes()
    idx8 = F[:5].astype(np.uint8).reshape(-1).tobytes()
    flat = P[F[5:9].reshape(-1)].astype(np.float32).tobytes()                     # 4 triangles, no indices
    q16 = np.array([[100, -200, 300], [-400, 500, 60], [7, 8, -9]], np.int16).tobytes()
    idx32 = np.array([0, 1, 2, 2, 1, 0], np.uint32).tobytes()
    parts = [pos_strided, idx16, idx8, flat, q16, idx32]
    offs = np.cumsum([0] + [len(p) + (-len(p)) % 4 for p in parts])
    binb = b"".join(p + b"\0" * ((-len(p)) % 4) for p in parts)
    g = {
        "asset": {"version": "2.0", "generator": "gen_meshio"},
        "scene": 0,
        "scenes": [{"nodes": [0, 3]}],
        "nodes": [
            {"children": [1, 2], "translation": [0.5, -1.25, 2.0], "rotation": [0.1, 0.2, 0.3, 0.9273618495495703],
             "scale": [1.5, 0.75, 2.0]},
            {"mesh": 0, "matrix": [0.9, 0.1, 0, 0, -0.1, 0.9, 0.05, 0, 0, -0.05, 1.1, 0, 3, 2, 1, 1]},
            {"mesh": 1, "rotation": [0, 0.7071067811865476, 0, 0.7071067811865476]},
            {"mesh": 2},
        ],
        "meshes": [
            {"primitives": [{"attributes": {"POSITION": 0}, "indices": 1},
                            {"attributes": {"POSITION": 0}, "indices": 2, "mode": 4},
                            {"attributes": {"POSITION": 0}, "indices": 1, "mode": 1}]},
            {"primitives": [{"attributes": {"POSITION": 3}}, {"attributes": {"NORMAL": 3}}]},
            {"primitives": [{"attributes": {"POSITION": 4}, "indices": 5}]},
        ],
        "bufferViews": [
            {"buffer": 0, "byteOffset": int(offs[0]), "byteLength": len(parts[0]), "byteStride": 16},
            {"buffer": 0, "byteOffset": int(offs[1]), "byteLength": len(parts[1])},
            {"buffer": 0, "byteOffset": int(offs[2]), "byteLength": len(parts[2])},
            {"buffer": 0, "byteOffset": int(offs[3]), "byteLength": len(parts[3])},
            {"buffer": 0, "byteOffset": int(offs[4]), "byteLength": len(parts[4])},
            {"buffer": 0, "byteOffset": int(offs[5]), "byteLength": len(parts[5])},
        ],
        "accessors": [
            {"bufferView": 0, "componentType": 5126, "count": 12, "type": "VEC3"},
            {"bufferView": 1, "componentType": 5123, "count": 60, "type": "SCALAR"},
            {"bufferView": 2, "componentType": 5121, "count": 15, "type": "SCALAR"},
            {"bufferView": 3, "componentType": 5126, "count": 12, "type": "VEC3"},
            {"bufferView": 4, "componentType": 5122, "count": 3, "type": "VEC3"},
            {"bufferView": 5, "componentType": 5125, "count": 6, "type": "SCALAR"},
        ],
        "buffers": [{"byteLength": len(binb)}],
    }
    write("scene.glb", glb(g, binb, extra_chunks=[(0x12345678, b"skip")]))
    write("scene_noext", glb(g, binb))
    g2 = dict(g)
    g2.pop("scenes")
    g2.pop("scene")
    write("nodes_only.glb", glb(g2, binb))
    g3 = {k: v for k, v in g.items() if k not in ("scenes", "scene", "nodes")}
    write("meshes_only.glb", glb(g3, binb))
    write("draco.glb", glb(dict(g, extensionsRequired=["KHR_draco_mesh_compression", "EXT_meshopt_compression"]), binb))
    write("nobin.glb", glb(g, None))
    bad = bytearray(glb(g, binb))
    bad[0] = 0x47
    write("badmagic.glb", bytes(bad))
    write("nomesh.glb", glb({"asset": {"version": "2.0"}, "meshes": []}, binb))
    g4 = dict(g, meshes=[{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1, "mode": 0}]}],
              nodes=[{"mesh": 0}], scenes=[{"nodes": [0]}])
    write("notris.glb", glb(g4, binb))
    write("model.gltf", json.dumps(g).encode())
    write("unknown.txt", b"hello world\nthis is not a mesh\n")

    # attribution text for MeshIO.meta: OBJ comment headers (field names in
    # several spellings, an empty value, control characters, more lines and
    # longer lines than are kept, a comment after the first statement), a
    # binary STL header with fields, glTF asset credits (extras with a
    # non-string value, a copyright string that carries the fields itself)
    # and a plain .gltf, which meta reads but parse refuses
    tri = "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"
    write("credits.obj", ("# Title: Seeded Car\n#   Author = A. Person\n## License: CC-BY-4.0\n"
                          "# source: https://example.org/car\n#\n# notes \x01\x02 with controls\n\n" + tri).encode())
    write("credits_long.obj", ("".join(f"# line {i} {'x' * (40 * i)}\n" for i in range(16)) + "# by: later\n"
                               + tri).encode())
    write("credits_url.obj", ("  # name : Wing\n# creator:   \n# artist: B. Other\n# url = http://e.org/w\n"
                              "v 0 0 0\n# author: after\n" + tri).encode())
    write("credits_binary.stl", stl_binary(tris[:2], b"name=Binary Wing\tlicence: MIT"))
    g5 = dict(g, asset={"version": "2.0", "generator": "seeded exporter 1.0", "copyright": "(c) 2024 Someone",
                        "extras": {"title": "Extras Title", "author": "Extras Author", "license": "CC0",
                                   "source": "https://example.org/m", "year": 2024}})
    write("credits.glb", glb(g5, binb))
    g6 = dict(g, asset={"version": "2.0", "copyright": "Title: Packed\nAuthor: C. Third\nLicense = CC-BY\nurl: x",
                        "extras": {"title": "lost"}})
    write("credits_copyright.glb", glb(g6, binb))
    write("credits_many.obj", ("".join(f"#c{i}\n" for i in range(12)) + "# author: thirteenth\n# url: fourteenth\n"
                               + tri).encode())
    g7 = dict(g, asset={"version": "2.0", "copyright": "Author: Z\nTitle: dropped",
                        "extras": {"title": "kept", "license": "MIT"}})
    write("credits_license.glb", glb(g7, binb))
    write("credits.gltf", json.dumps(dict(g, asset={"version": "2.0", "generator": "gltf \u00e9 writer \u2028 sep",
                                                    "extras": {"author": 7}})).encode())


if __name__ == "__main__":
    main()
