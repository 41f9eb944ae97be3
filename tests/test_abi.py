"""CPU-side checks of the C ABI: the library loads and exports every
function that include/biqbin_gpu.h declares (no compute calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "biqbin_gpu.h")
LIB = os.path.join(ROOT, "parallel_biqbin_maxcut_amd", "libbiqbin_gpu.so")


def declared_functions():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    # function-pointer typedefs declare no symbol
    txt = re.sub(r"^typedef[^;]*;", "", txt, flags=re.M | re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert "ipm_mc_pk" in names and "bqb_ipm_batch" in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libbiqbin_gpu.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_struct_mirrors_match_the_library():
    """The ctypes mirrors of api.py have the C structures' sizes (a field
    added on one side only shifts every later field)."""
    import ctypes
    import numpy as np
    from parallel_biqbin_maxcut_amd import _native as N, api
    sizes = np.zeros(5, np.int32)
    assert N.lib().bqb_abi_sizes(sizes.ctypes.data_as(N.iptr), 5) == 5
    mirrors = (api.BiqBinParameters, api.SolveOptions, api.SolveStats, api.Trajectory, api.NodeState)
    assert [ctypes.sizeof(m) for m in mirrors] == sizes.tolist()


def test_library_loads_without_gpu():
    from parallel_biqbin_maxcut_amd import _native as N
    L = N.lib()
    assert L.bqb_last_error() is not None
