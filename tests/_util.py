"""Shared helpers for the tests: golden-fixture loading and replay."""
from __future__ import annotations

import base64
import gzip
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURES = os.path.join(ROOT, "tests", "golden", "fixtures")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

_DT = {"f4": np.float32, "f8": np.float64, "u1": np.uint8, "i4": np.int32, "u4": np.uint32}


def load(name: str) -> dict:
    with open(os.path.join(FIXTURES, name + ".json")) as fh:
        return json.load(fh)


def arr(rec: dict) -> np.ndarray:
    raw = gzip.decompress(base64.b64decode(rec["gz_b64"]))
    a = np.frombuffer(raw, dtype=_DT[rec["dtype"]]).copy()
    return a.reshape(rec["shape"]) if int(np.prod(rec["shape"])) == a.size else a


def unbits(rec: dict, n: int) -> np.ndarray:
    """LSB-first packed bits back to a 0/1 byte grid of length n."""
    b = arr(rec)
    return np.unpackbits(b, bitorder="little")[:n]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fixture_names(prefix: str) -> list:
    return sorted(f[:-5] for f in os.listdir(FIXTURES) if f.startswith(prefix) and f.endswith(".json"))


def set_params(L, p: dict) -> None:
    """Apply a fixture's solver parameters (movingWalls is set by voxelize)."""
    L.set_params(tau=p["tau"], les=p["les"], csm=p["csm"], trt=p["trt"], **{"lambda": p["lambda"]},
                 u0=p["u0"])


def replay_voxelize(L, rec: dict) -> dict:
    mask = arr(rec["mask"]).reshape(-1)
    wv = arr(rec["wv"]).reshape(-1)
    return L.voxelize_mask(rec["box"], mask, wv, rec["ground"], rec["roadVel"])


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
