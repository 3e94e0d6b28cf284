"""The C-ABI library loads and exports every entry point the header declares."""
import ctypes
import re
from pathlib import Path

from seissol_amd import _capi

HEADER = Path(__file__).resolve().parents[1] / "include" / "seissol_amd.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(sdg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = declared_functions()
    assert "sdg_compute_batched_ader" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(declared_functions()) <= set(_capi.SIGNATURES), set(declared_functions()) - set(_capi.SIGNATURES)


def test_errors_are_reported():
    lib = _capi.lib()
    eng = ctypes.c_void_p()
    rc = lib.sdg_engine_create(ctypes.byref(eng), 12, _capi.SDG_FP64)
    assert rc != 0
    assert b"order" in lib.sdg_last_error()
    assert lib.sdg_proxy_kernel_id(b"all") == 0
    assert lib.sdg_proxy_kernel_id(b"nonsense") < 0
