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


def _estimate(order, arithmetic, kernel):
    w = _capi.WorkEstimate()
    _capi.check(_capi.lib().sdg_performance_estimate(order, arithmetic, kernel, ctypes.byref(w)),
                "sdg_performance_estimate")
    return w


def test_work_estimates_match_the_reference_sparsity():
    # non-zero FLOP per element of volume + local flux and of the neighbour
    # flux, counted from the sparsity of the reference's matrices_{20,35,56}.xml
    # and star.xml (24 non-zeros)
    expected = {4: (6516 + 19692, 21372), 5: (16686 + 41400, 44304), 6: (38808 + 80712, 85488)}
    for order, (local, neigh) in expected.items():
        assert _estimate(order, _capi.SDG_FP64, 4).nonzero_flop == local
        assert _estimate(order, _capi.SDG_FP64, 2).nonzero_flop == neigh
        B = order * (order + 1) * (order + 2) // 6
        # bytes: 6 DOF blocks, 3 x 24 star non-zeros, 8 flux solvers (FP64)
        assert _estimate(order, _capi.SDG_FP64, 0).bytes == (6 * 9 * B + 72 + 8 * 81) * 8
        assert _estimate(order, _capi.SDG_FP32, 0).bytes == (6 * 9 * B + 72 + 8 * 81) * 4
    a = _estimate(6, _capi.SDG_FP64, 0)
    parts = [_estimate(6, _capi.SDG_FP64, k) for k in (3, 4, 2)]  # ader, localwoader, neigh
    assert abs(a.nonzero_flop - sum(p.nonzero_flop for p in parts)) < 1e-6
    # split-bf16 forms execute 3 or 6 products per dense contraction
    assert _estimate(6, _capi.SDG_BF16X6, 0).hardware_flop == 6 * _estimate(6, _capi.SDG_FP32, 0).hardware_flop
    assert _capi.lib().sdg_performance_estimate(6, 0, 7, ctypes.byref(_capi.WorkEstimate())) != 0


def test_proxy_cli_rejects_bad_arguments():
    import subprocess
    exe = Path(__file__).resolve().parents[1] / "seissol_amd" / "_lib" / "seissol_amd_proxy"
    r = subprocess.run([str(exe), "10", "1", "bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "Unknown kernel string: bogus" in r.stderr
    r = subprocess.run([str(exe), "10", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "usage" in r.stderr
    r = subprocess.run([str(exe), "-f", "xml", "10", "1", "all"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255
