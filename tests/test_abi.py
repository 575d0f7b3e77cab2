"""The C ABI libraries load and export every function their headers
declare; the Python ctypes tables cover the headers; every component
type exports the make_<TYPE>_factory plugin entry point.  No device
calls: this runs without a GPU."""
import os
import re

import pytest

from conftest import REPO

INC = os.path.join(REPO, "include")
PKG = os.path.join(REPO, "wire-cell-toolkit_amd")


def _declared(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
    return sorted(set(re.findall(r"\b(wcth?_[a-z0-9_]+)\s*\(", text)))


def _exports(path):
    import ctypes
    lib = ctypes.CDLL(path)
    return lib


@pytest.mark.parametrize("header,libname", [("wct_hip.h", "libwct_hip.so"),
                                            ("wct_host.h", "libWireCellHip.so")])
def test_library_exports_every_declared_symbol(header, libname):
    path = os.path.join(PKG, libname)
    assert os.path.exists(path), f"{libname} not built"
    lib = _exports(path)
    names = _declared(header)
    assert len(names) > 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_ctypes_tables_cover_headers():
    from wct_amd import capi, host
    capi.lib()
    host.lib()
    assert capi.MISSING == [] and host.MISSING == []
    assert sorted(capi._SIGS) == _declared("wct_hip.h")
    assert sorted(host._SIGS) == _declared("wct_host.h")


def test_plugin_factories_exported():
    from wct_amd import components as K
    import ctypes
    types = K.known_types()
    for t in ("HipDFT", "HipOmnibusSigProc", "HfFilter", "LfFilter", "ColdElecResponse",
              "FieldResponse", "AnodeChannels"):
        assert t in types
    lib = ctypes.CDLL(os.path.join(PKG, "libWireCellHip.so"))
    for t in types:
        fn = getattr(lib, f"make_{t}_factory")
        fn.restype = ctypes.c_void_p
        assert fn() is not None


def test_abi_version_and_error_string():
    from wct_amd import capi
    L = capi.lib()
    assert L.wct_abi_version() == 5
    assert L.wct_dft_supported(6000) == 1 and L.wct_dft_supported(17) == 1
    assert L.wct_dft_supported(0) == 0 and L.wct_dft_supported((1 << 20) + 1) == 0
    assert L.wct_dft_native_length(6000) == 1 and L.wct_dft_native_length(17) == 0
    assert isinstance(L.wct_last_error(), bytes)


def test_torch_extension_loads():
    """The libtorch extension (autograd Functions over the C ABI) is built
    in-tree and imports without a device."""
    from wct_amd import autograd as A
    mod = A.ext()
    for name in ("Model", "forward_conv", "decon", "apply"):
        assert hasattr(mod, name)
