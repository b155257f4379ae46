"""The C-ABI library loads, exports every symbol include/soltrace_hip.h declares,
and rejects bad arguments without a GPU (no compute calls here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from optixsoltracedemos_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "soltrace_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(st_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = _abi.load()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/soltrace_hip.h but not exported"
        assert n in _abi.SIGNATURES, f"{n} missing from the ctypes bindings"
    assert set(_abi.SIGNATURES) == set(names)


STRUCTS = {"Element": "st_element_f64", "SunDesc": "st_sun_desc", "SunPlane": "st_sun_plane",
           "TraceDesc": "st_trace_desc", "Outputs": "st_outputs", "BatchDesc": "st_batch_desc",
           "ServerConfig": "st_server_config", "JobDesc": "st_job_desc", "JobResult": "st_job_result",
           "ServerStats": "st_server_stats", "TuneResult": "st_tune_result", "LaunchInfo": "st_launch_info"}


def test_abi_version_and_struct_layouts(tmp_path):
    lib = _abi.load()
    assert lib.st_abi_version() == 1
    src = tmp_path / "sz.c"
    body = "".join(f'printf("{py} %zu\\n", sizeof({c}));' for py, c in STRUCTS.items())
    src.write_text('#include <stdio.h>\n#include "soltrace_hip.h"\nint main(void){' + body + "return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout.split()
    sizes = dict(zip(out[0::2], map(int, out[1::2])))
    for py in STRUCTS:
        assert C.sizeof(getattr(_abi, py)) == sizes[py], py


def test_invalid_arguments_without_gpu():
    lib = _abi.load()
    h = C.c_void_p()
    assert lib.st_scene_create(None, None, 0, C.byref(h)) == _abi.ST_ERR_INVALID
    assert b"null" in lib.st_last_error()
    assert lib.st_scene_destroy(None) == _abi.ST_ERR_INVALID
    assert lib.st_trace(None, None, None) == _abi.ST_ERR_INVALID
    assert lib.st_context_create(0, None) == _abi.ST_ERR_INVALID
    assert lib.st_debug_inject_fault(7, 1) == _abi.ST_ERR_INVALID
    assert lib.st_debug_inject_fault(0, 0) == _abi.ST_OK


def test_no_device_is_an_error_not_a_crash():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _abi.load()
    n = C.c_int(-1)
    assert lib.st_device_count(C.byref(n)) == _abi.ST_ERR_DEVICE
    assert n.value == 0
    h = C.c_void_p()
    assert lib.st_context_create(0, C.byref(h)) != _abi.ST_OK
    assert len(lib.st_last_error()) > 0
