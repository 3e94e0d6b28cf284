"""The C-ABI library loads and exports every entry point the header declares."""
import ctypes
import re
from pathlib import Path

import numpy as np

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
    # orders 2..8 as the reference's matrices_{4..120}.xml; 9 is rejected before any device call
    assert lib.sdg_engine_create(ctypes.byref(eng), 9, _capi.SDG_FP64) != 0
    assert b"[2, 8]" in lib.sdg_last_error()
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
    assert _capi.lib().sdg_performance_estimate(6, 0, 8, ctypes.byref(_capi.WorkEstimate())) != 0
    # ORDER 8 element kernels are estimated; its DR kernels are not supported
    B8 = 120
    assert _estimate(8, _capi.SDG_FP64, 0).bytes == (6 * 9 * B8 + 72 + 8 * 81) * 8
    assert _estimate(8, _capi.SDG_FP64, 4).nonzero_flop > _estimate(7, _capi.SDG_FP64, 4).nonzero_flop
    assert _capi.lib().sdg_performance_estimate(8, 0, 5, ctypes.byref(_capi.WorkEstimate())) != 0
    assert _capi.lib().sdg_performance_estimate(9, 0, 0, ctypes.byref(_capi.WorkEstimate())) != 0


def test_dr_work_estimates_known_answers():
    # the proxy's DR layer: 4 DR faces per cell, all cell faces DR
    order, B, NP = 6, 56, 49
    taylor = 2 * 9 * (56 + 35 + 20 + 10 + 4 + 1)   # Taylor sum of one side at one time point
    rot = 2 * NP * 81                               # rotation by TinvT / product with the flux solver
    godunov_dense = 4 * order * (2 * taylor + 2 * 2 * NP * B * 9 + 2 * rot)
    nodal_dense = 4 * (rot + 2 * B * NP * 9)
    g = _estimate(order, _capi.SDG_FP64, 6)
    nb = _estimate(order, _capi.SDG_FP64, 5)
    assert g.hardware_flop == godunov_dense == 2860704
    assert nb.hardware_flop == nodal_dense == 229320
    # non-zero counts from the sparsity of the face matrices (entries below
    # 1e-13 of the matrix maximum are zero): plus side h = 0, minus side the
    # mean over h = 1..3, averaged over the four faces
    from seissol_amd.dr import face_matrices
    v, vt = face_matrices(order)
    nz = lambda m: np.count_nonzero(np.abs(m) > 1e-13 * np.abs(m).max())
    nz_v = np.mean([nz(v[i, 0]) for i in range(4)]) + np.mean([nz(v[i, h]) for i in range(4) for h in (1, 2, 3)])
    nz_t = 0.5 * (np.mean([nz(vt[i, 0]) for i in range(4)]) +
                  np.mean([nz(vt[i, h]) for i in range(4) for h in (1, 2, 3)]))
    assert nz_v < 2 * NP * B
    assert abs(g.nonzero_flop - 4 * order * (2 * taylor + 2 * nz_v * 9 + 2 * rot)) < 1e-6
    assert abs(nb.nonzero_flop - 4 * (rot + 2 * nz_t * 9)) < 1e-6
    dsize = 9 * (56 + 35 + 20 + 10 + 4 + 1)
    assert g.bytes == 4 * (2 * dsize * 8 + 81 * 8 + 2 * order * 9 * NP * 8)
    assert nb.bytes == 2 * 9 * B * 8 + 4 * (9 * NP * 8 + 81 * 8)
    # all_dr = ader + volume (no local flux on DR faces) + godunov_dr + neigh_dr
    volume_nz = 38808
    a = _estimate(order, _capi.SDG_FP64, 7)
    ader = _estimate(order, _capi.SDG_FP64, 3)
    assert abs(a.nonzero_flop - (ader.nonzero_flop + volume_nz + g.nonzero_flop + nb.nonzero_flop)) < 1e-6
    # FP32 halves the cell data, not the FP64 face data
    assert _estimate(order, _capi.SDG_FP32, 5).bytes == 2 * 9 * B * 4 + 4 * (9 * NP * 8 + 81 * 8)
    # a matrix-core engine: the DR parts in FP32 (one product each), the local
    # pass with its whole split-product chain; the same non-zero work
    m7 = _estimate(order, _capi.SDG_F16X3, 7)
    assert m7.nonzero_flop == _estimate(order, _capi.SDG_FP32, 7).nonzero_flop
    assert _estimate(order, _capi.SDG_F16X3, 6).hardware_flop == g.hardware_flop
    assert abs(m7.hardware_flop - (_estimate(order, _capi.SDG_F16X3, 1).hardware_flop + g.hardware_flop +
                                   nb.hardware_flop)) < 1e-6
    assert _capi.lib().sdg_performance_estimate(2, _capi.SDG_F16X3, 7, ctypes.byref(_capi.WorkEstimate())) != 0


def test_proxy_kernel_names_and_dr_faces():
    lib = _capi.lib()
    names = ["all", "local", "neigh", "ader", "localwoader", "neigh_dr", "godunov_dr", "all_dr"]
    assert [lib.sdg_proxy_kernel_id(n.encode()) for n in names] == list(range(8))
    assert lib.sdg_proxy_kernel_id(b"dr") == -1
    n = 500
    out = [np.zeros((n, 4), np.int32), np.zeros(4 * n, np.int32), np.zeros(4 * n, np.int32),
           np.zeros((4 * n, 3), np.uint8)]
    _capi.check(lib.sdg_proxy_dr_faces(n, 7, *[a.ctypes.data for a in out]))
    mapping, plus, minus, info = out
    assert mapping.min() >= 0 and mapping.max() < 4 * n
    assert plus.min() >= 0 and plus.max() < n and minus.min() >= 0 and minus.max() < n
    assert set(np.unique(info[:, :2])) == {0, 1, 2, 3} and set(np.unique(info[:, 2])) == {1, 2, 3}
    again = [np.zeros_like(a) for a in out]
    _capi.check(lib.sdg_proxy_dr_faces(n, 7, *[a.ctypes.data for a in again]))
    assert all(np.array_equal(a, b) for a, b in zip(out, again))
    assert lib.sdg_proxy_dr_faces(0, 7, *[a.ctypes.data for a in out]) != 0


def test_proxy_cli_rejects_bad_arguments():
    import subprocess
    exe = Path(__file__).resolve().parents[1] / "seissol_amd" / "_lib" / "seissol_amd_proxy"
    r = subprocess.run([str(exe), "10", "1", "bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "Unknown kernel string: bogus" in r.stderr
    r = subprocess.run([str(exe), "10", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "usage" in r.stderr
    r = subprocess.run([str(exe), "-f", "xml", "10", "1", "all"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255


def test_friction_layer_checks_need_no_device():
    # sizes are checked and an empty fault layer returns before any device call
    lib = _capi.lib()
    prm = _capi.FrictionParams()
    dt = np.full(4, 0.25e-3)
    tw = np.full(4, 0.25e-3)
    layer = _capi.FaultLayer(0, 25, 25, 4, 0)
    assert lib.sdg_friction_evaluate(ctypes.byref(prm), ctypes.byref(layer), 0.0, dt.ctypes.data, tw.ctypes.data,
                                     None, None, None) == 0
    for bad in (_capi.FaultLayer(0, 65, 65, 4, 0), _capi.FaultLayer(0, 16, 25, 4, 0), _capi.FaultLayer(0, 25, 25, 9, 0),
                _capi.FaultLayer(-1, 25, 25, 4, 0)):
        assert lib.sdg_friction_evaluate(ctypes.byref(prm), ctypes.byref(bad), 0.0, dt.ctypes.data, tw.ctypes.data,
                                         None, None, None) != 0
        assert b"invalid layer sizes" in lib.sdg_last_error()
    layer1 = _capi.FaultLayer(1, 25, 25, 4, 0)  # arrays missing
    assert lib.sdg_friction_evaluate(ctypes.byref(prm), ctypes.byref(layer1), 0.0, dt.ctypes.data, tw.ctypes.data,
                                     None, None, None) != 0
    assert b"missing" in lib.sdg_last_error()
