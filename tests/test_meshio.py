"""Mesh readers (wt_mesh_parse) against the reference's MeshIO.parse.

The files in tests/golden/fixtures/meshio/ come from tests/golden/gen_meshio.py;
mesh_io.json holds what js/meshio.js returned for each of them under Node
(float count, SHA-256 of the float32 soup, NaN count, first triangle) or the
message it threw, also under names whose extension is not a mesh format so
that the dispatch sniffs the content. The parser is host code: these tests
need no GPU, only the library.
"""
import os

import numpy as np
import pytest

from _util import load
from windtunnel_amd import synth
from windtunnel_amd.lbm import mesh_parse


# every test here also runs in the GPU suite (-m gpu), against the shipped build on the GPU machine
@pytest.fixture(params=[pytest.param("cpu"), pytest.param("gpu", marks=pytest.mark.gpu)], autouse=True)
def suite(request):
    return request.param


DIR = os.path.join(os.path.dirname(__file__), "golden", "fixtures", "meshio")
CASES = load("mesh_io")["cases"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_reference_cases(case):
    with open(os.path.join(DIR, case["file"]), "rb") as fh:
        data = fh.read()
    if not case["ok"]:
        with pytest.raises(ValueError) as e:
            mesh_parse(case["name"], data)
        assert str(e.value) == case["error"]
        return
    P = mesh_parse(case["name"], data)
    assert P.size == case["count"]
    assert int(np.isnan(P).sum()) == case["nan"]
    head = [np.nan if v == "NaN" else v for v in case["head"]]
    np.testing.assert_array_equal(P[:9], np.array(head, dtype=np.float32))
    import hashlib
    assert hashlib.sha256(P.tobytes()).hexdigest() == case["sha256"]


def test_scale_mesh_round_trips():
    """The 640x320x320 body (326 049 triangles) through binary STL and OBJ."""
    P = synth.noisy_sphere(level=7, R=56, holes=0.005, seed=5)
    n = P.size // 9
    stl = bytearray(b"x" * 80) + np.uint32(n).tobytes()
    rec = np.zeros(n, dtype=[("nrm", "<f4", 3), ("v", "<f4", 9), ("attr", "<u2")])
    rec["v"] = P.reshape(n, 9)
    stl += rec.tobytes()
    assert np.array_equal(mesh_parse("body.stl", bytes(stl)), P)
    V = P.reshape(-1, 3).astype(np.float64)
    lines = ["v %r %r %r" % tuple(float(x) for x in v) for v in V]
    lines += ["f %d %d %d" % (3 * t + 1, 3 * t + 2, 3 * t + 3) for t in range(n)]
    got = mesh_parse("body.obj", "\n".join(lines).encode())
    assert np.array_equal(got, P)
