"""The C-ABI library loads, exports every symbol include/sray_gpu.h declares,
and its struct layouts match the bindings (no device needed)."""
import ctypes as C
import os
import re

import pytest

import sray_amd as S
from sray_amd import abi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "sray_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sray_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "sray_gpu_trace" in names and "sray_gpu_render" in names
    assert sorted(abi.EXPORTS) == names, "bindings' export list differs from the header"


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(S.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_struct_layouts_match():
    sizes = S.abi_sizes()
    assert sizes == [C.sizeof(s) for s in abi.ABI_STRUCTS]


def test_no_device_is_an_error_not_a_fallback():
    """without a HIP device the context refuses to start (no CPU fallback)"""
    h = C.c_void_p()
    rc = S.lib().sray_gpu_create(0, 0, C.byref(h))
    if rc == abi.SRAY_OK:
        S.lib().sray_gpu_destroy(h)
        pytest.skip("a device is present")
    assert rc in (abi.SRAY_ERR_NODEVICE, abi.SRAY_ERR_HIP)


def test_null_arguments_rejected():
    assert S.lib().sray_gpu_trace(None, 1, None, 0, None, 0, None) == abi.SRAY_ERR_ARG
    assert S.lib().sray_gpu_render(None, 1, None, None, None) == abi.SRAY_ERR_ARG
    assert S.lib().sray_scene_load_xml(None, None, None) == abi.SRAY_ERR_ARG
