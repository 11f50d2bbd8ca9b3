"""The C-ABI library loads and exports every entry point include/windtunnel.h
declares; without a GPU it refuses to create a lattice with an error, never
a silent fallback. No compute call is made here."""
import ctypes as C
import os
import re

import pytest

from _util import ROOT
from windtunnel_amd import lbm as wl

HEADER = os.path.join(ROOT, "include", "windtunnel.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(wt_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("wt_create", "wt_step", "wt_step_slabs", "wt_voxelize", "wt_voxel_build", "wt_get_fields"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(wl.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding binds exactly the declared set
    assert sorted(wl.SIGNATURES) == declared()


def test_no_gpu_means_an_error_not_a_fallback():
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a GPU device node is present")
    with pytest.raises(wl.WindtunnelError) as e:
        wl.LBM(16, 8, 8)
    assert e.value.code == wl.E_DEVICE
    assert wl.lib().wt_abi_version() == 1


def test_argument_errors_before_device_checks():
    with pytest.raises(ValueError, match="at least 3x3x3"):
        wl.LBM(2, 8, 8)
    with pytest.raises(ValueError, match="unsupported storage"):
        wl.LBM(8, 8, 8, storage="f16", arith="f64")
    # cell indices are 32-bit in the near-wall lists: beyond that, an error
    with pytest.raises(ValueError, match="32-bit cell index"):
        wl.LBM(70000, 60000, 8)
    with pytest.raises(ValueError, match="too large"):
        wl.LBM(1 << 24, 4, 4)
