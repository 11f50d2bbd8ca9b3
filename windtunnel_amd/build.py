"""Build libwindtunnel.so (HIP, gfx950 only) in-tree.

    python -m windtunnel_amd.build

Compiles every source in windtunnel_amd/csrc with hipcc into one shared
library next to this file. Contraction is off everywhere: the binary64
collision path must round exactly like the reference's IEEE doubles, and
the fast path gains nothing from FMA on a bandwidth-bound kernel. Packed
FP32 instructions are disabled as a code-generation policy.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwindtunnel.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def sources() -> list:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    """Compile the library; `defines` are extra -D switches (tuning builds).
    Each source compiles to an object of its own, in parallel, then one link."""
    import concurrent.futures as cf
    import tempfile
    tmp = out + ".tmp"
    dflags = [f"-D{d}" for d in defines]
    with tempfile.TemporaryDirectory(prefix="wt_build_") as obj:
        def compile_one(src):
            o = os.path.join(obj, os.path.basename(src) + ".o")
            cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], *dflags, "-c", "-o", o, src]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
            return o
        srcs = sources()
        workers = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
        with cf.ThreadPoolExecutor(workers) as ex:
            objs = list(ex.map(compile_one, srcs))
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


ADDON_SRC = os.path.join(HERE, "addon", "wt_napi.c")
ADDON = os.path.join(HERE, "wt_napi.node")
NODE_INCLUDE = os.environ.get("NODE_INCLUDE", "/usr/include/node")


def build_addon(verbose: bool = False, out: str = ADDON):
    """Compile the N-API addon (plain C over the C-ABI) next to the library;
    it finds libwindtunnel.so through its own directory. Returns None when
    the Node.js headers are absent."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        if verbose:
            print(f"node_api.h not found under {NODE_INCLUDE}: N-API addon not built", flush=True)
        return None
    tmp = out + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-shared", "-fPIC", "-Wall", f"-I{NODE_INCLUDE}",
           f"-I{os.path.join(HERE, '..', 'include')}", "-o", tmp, ADDON_SRC,
           f"-L{HERE}", "-lwindtunnel", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
    print(build_addon(verbose="-v" in sys.argv))
