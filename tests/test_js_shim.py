"""The Node.js drop-in (windtunnel_amd/js over the wt_napi addon), run under node."""
import os
import shutil
import subprocess

import pytest

from _util import ROOT

NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "windtunnel_amd", "wt_napi.node")
pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def run(script, env=None, *args):
    e = dict(os.environ)
    e.update(env or {})
    p = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", script), *args], capture_output=True, text=True,
                       timeout=600, env=e)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr


def test_shim_without_gpu():
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a GPU device node is present")
    assert os.path.exists(ADDON), "addon not built"
    run("shim_cpu.js")


@pytest.mark.gpu
def test_shim_on_gpu():
    run("shim_gpu.js")
    run("shim_gpu.js", {"WT_FAULT": "step:3"}, "fault")
