"""The C-ABI libraries load and export every entry point their headers
declare (CPU only; no device calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "3d-ose_amd", "lib")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ose[h]?_\w+)\s*\(", text)))


@pytest.mark.parametrize("lib,header", [("libose_hip.so", "ose_api.h"), ("libose_host.so", "ose_host.h")])
def test_exports_every_declared_symbol(lib, header):
    L = ctypes.CDLL(os.path.join(LIB, lib))
    names = declared(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_lists_the_header():
    from ose_amd import _lib
    assert sorted(_lib.hip_symbols()) == declared("ose_api.h")
    from ose_amd import _host_protos
    assert sorted(_host_protos.PROTOS) == declared("ose_host.h")


def test_host_only_entry_points():
    from ose_amd import _lib
    L = _lib.hip()
    assert L.ose_destroy(None) == _lib.OSE_EINVAL
    assert L.ose_last_error(None) == b"null context"
    assert L.ose_extent_is_supported(152) and L.ose_extent_is_supported(864)
    # every extent in [2, 4096] runs: compiled kernels, or the runtime-length path
    assert all(L.ose_extent_is_supported(n) for n in (2, 17, 106, 150, 1018, 2048, 4096))
    assert not L.ose_extent_is_supported(1) and not L.ose_extent_is_supported(4097)
    assert [L.ose_extent_kind(n) for n in (152, 864, 17, 106, 1100, 4097)] == [1, 1, 2, 2, 2, 0]
    assert L.ose_supported_extent(848) == 864
    assert L.ose_supported_extent(272) == 272
    assert L.ose_supported_extent(273) == 280
    assert L.ose_supported_extent(141) == 144
    assert L.ose_add_model(None, None, None, None) == _lib.OSE_EINVAL
    assert L.ose_run(None, 1) == _lib.OSE_EINVAL
    assert L.ose_set_streams(None, 2) == _lib.OSE_EINVAL


def test_create_without_device_reports_edevice():
    """No visible GPU: ose_create fails with OSE_EDEVICE instead of exiting."""
    from ose_amd import _lib
    L = _lib.hip()
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible here")
    ctx = ctypes.c_void_p()
    assert L.ose_create(ctypes.byref(ctx), 0) == _lib.OSE_EDEVICE
    assert not ctx.value
