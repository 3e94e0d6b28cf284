"""The reference's own storage layout at the boundary (GPU tests).

The reference's GPU build pads every DOF-like block to [9][ld] (ld = B rounded
up to 128 bytes: 64 at ORDER 6), pads each derivative block likewise, keeps the
star matrices as 24 non-zeros each, and stores them with AplusT inside one
array of LocalIntegrationData structs (AminusT inside NeighboringIntegrationData).
These tests hand exactly such arrays to the C ABI -- every padding entry NaN, so
a kernel that read one would spoil its result -- and compare with the FP64
oracle on the packed data.  Tolerances are those of the packed-layout tests.
"""
import numpy as np
import pytest

import oracle
from helpers import rel_l2
from seissol_amd.dr import DeviceDrFaces, collocation_coefficients, face_matrices, plane_faces
from seissol_amd.kernels import Engine, Spacetime, Time, global_matrices, monomial_integrate
from seissol_amd.lts import DeviceLtsLayer, LocalTimeStepping, make_lts_setup
from seissol_amd.proxy import DeviceLayer, make_setup, run_proxy

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-5, "bf16x3": 2e-5, "f16x3": 5e-6}


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.build()


def _np(t):
    return t.detach().cpu().double().numpy()


def _padding_untouched(layer, t):
    import torch
    B = layer.packed(t).shape[-1]
    return bool(torch.isnan(t[:, :, B:]).all().item()) if t.shape[-1] > B else True


@pytest.mark.parametrize("arith,order,alignment,star_format", [
    ("fp64", 6, 128, "compact"),   # the reference's AMD GPU build: ld 64, 3 x 24 star
    ("fp32", 5, 128, "compact"),   # ld 64 for B = 35
    ("f16x3", 6, 128, "compact"),
    ("f16x3", 5, 16, "compact"),   # ld 36: 16-byte accesses with a partly padded last row group
    ("f16x3", 4, 128, "dense"),
    ("bf16x3", 3, 0, "compact"),   # packed, B = 10: per-value accesses
    ("fp64", 4, 64, "dense"),
])
def test_struct_arrays_proxy_all_match_oracle(gpu, arith, order, alignment, star_format):
    import torch
    steps = 3
    setup = make_setup(order, n=700, mesh="proxy", seed=40 + order)
    print(f"layout proxy all {arith} order={order} alignment={alignment} star={star_format} n={setup.n}",
          flush=True)
    eng = Engine(order, arith)
    eng.set_layout(alignment, star_format)
    layer = DeviceLayer(eng, setup, struct_arrays=True)
    assert layer.ld >= setup.dofs.shape[-1]
    run_proxy(eng, layer, "all", steps, setup.dt)
    torch.cuda.synchronize()
    gm = global_matrices(order)["flat"]
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    q_ref, _ = oracle.proxy_all(order, gm, steps, coeffs, _np(layer.star64), _np(layer.aplus64),
                                _np(layer.aminus64), setup.neighbors, setup.face_relations, setup.dofs)
    err = rel_l2(_np(layer.packed(layer.dofs)), q_ref)
    print(f"rel L2 vs oracle after {steps} steps: {err:.2e}")
    assert err < TOL[arith]
    assert _padding_untouched(layer, layer.dofs)
    assert bool(torch.isnan(layer.local_data[:, 72 + 324 if star_format == "compact" else 243 + 324:]).all())


@pytest.mark.parametrize("arith", ["fp64", "fp32"])
def test_aligned_derivatives_and_taylor_match_oracle(gpu, arith):
    import torch
    order = 6
    setup = make_setup(order, n=301, mesh="proxy", seed=9)
    print(f"aligned derivatives {arith} order={order}", flush=True)
    eng = Engine(order, arith)
    eng.set_layout(128, "compact")
    ld, qreals, dreals = eng.layout_sizes()
    rs = 8 if arith == "fp64" else 4
    # the sizes of tensor::Q and of the dQ family at 128-byte alignment
    lds = [-(-b // (128 // rs)) * (128 // rs) for b in (56, 35, 20, 10, 4, 1)]
    assert ld == 64 and qreals == 576 and dreals == 9 * sum(lds)
    layer = DeviceLayer(eng, setup, with_derivatives=True, struct_arrays=True)
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    Spacetime(eng).computeBatchedAder(coeffs, setup.dt, layer.view)
    out = torch.full_like(layer.time_integrated, float("nan"))
    c2 = monomial_integrate(order, 0.3 * setup.dt, 0.9 * setup.dt)
    Time(eng).evaluateBatched(c2, layer.derivatives, out, setup.n)
    torch.cuda.synchronize()
    gm = global_matrices(order)["flat"]
    I_ref, d_ref = oracle.ader(order, gm, coeffs, setup.dofs, _np(layer.star64), want_derivatives=True)
    tol = TOL[arith]
    assert rel_l2(_np(layer.packed(layer.time_integrated)), I_ref) < tol
    assert rel_l2(_np(layer.derivatives_packed()), d_ref) < tol
    t_ref = oracle.taylor(order, c2, d_ref)
    assert rel_l2(_np(layer.packed(out)), t_ref) < tol
    # the padding rows of every derivative block and of the outputs stay untouched
    for off, bld, m in eng.derivative_blocks():
        blk = layer.derivatives[:, off:off + 9 * bld].reshape(setup.n, 9, bld)
        assert bool(torch.isnan(blk[:, :, m:]).all())
    assert _padding_untouched(layer, out) and _padding_untouched(layer, layer.time_integrated)


@pytest.mark.parametrize("arith,tol", [("fp64", 1e-11), ("fp32", 2e-5)])
def test_lts_on_struct_arrays_matches_oracle(gpu, arith, tol):
    import torch
    order = 5
    lts = make_lts_setup(order, fractions=(0.6, 0.25, 0.1, 0.05), dims=(20, 6, 6), seed=3)
    print(f"lts struct arrays {arith} order={order} n={lts.setup.n}", flush=True)
    eng = Engine(order, arith)
    eng.set_layout(128, "compact")
    layer = DeviceLtsLayer(eng, lts, struct_arrays=True)
    run = LocalTimeStepping(eng, layer)
    run.advance(2)
    torch.cuda.synchronize()
    s = lts.setup
    q_ref = oracle.lts(order, global_matrices(order)["flat"], 2, lts.num_clusters, lts.cluster_ids, lts.dt0,
                       lts.lts_setups, _np(layer.star64), _np(layer.aplus64), _np(layer.aminus64), s.neighbors,
                       s.face_relations, s.dofs)
    err = rel_l2(_np(layer.packed(layer.dofs)), q_ref)
    print(f"rel L2 vs oracle after 2 macro steps: {err:.2e}")
    assert err < tol
    assert _padding_untouched(layer, layer.dofs)


@pytest.mark.parametrize("arith,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_dr_kernels_on_aligned_arrays_match_oracle(gpu, arith, tol):
    import torch
    order = 5
    setup = make_setup(order, mesh="periodic", dims=(8, 6, 6), seed=12, h=100.0)
    faces = plane_faces(setup, plane=4)
    print(f"dr aligned {arith} order={order} faces={faces.num_faces}", flush=True)
    eng = Engine(order, arith)
    eng.set_layout(128, "compact")
    layer = DeviceLayer(eng, setup, with_derivatives=True, struct_arrays=True)
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    Spacetime(eng).computeBatchedAder(coeffs, setup.dt, layer.view)
    dr = DeviceDrFaces(faces)
    qp = torch.full((faces.num_faces, order, 9, faces.points), float("nan"), dtype=torch.float64, device="cuda")
    qm = torch.full_like(qp, float("nan"))
    c = collocation_coefficients(order, setup.dt)
    dr.interpolate(eng, c, layer.derivatives, qp, qm)
    rng = np.random.default_rng(5)
    amp = np.array([1e6] * 6 + [1.0] * 3)[None, :, None]
    gplus = rng.standard_normal((faces.num_faces, 9, faces.points)) * amp
    gminus = rng.standard_normal((faces.num_faces, 9, faces.points)) * amp
    q0 = _np(layer.packed(layer.dofs))
    dr.nodal_flux(eng, layer.view, torch.from_numpy(gplus).cuda(), torch.from_numpy(gminus).cuda())
    torch.cuda.synchronize()
    v3m, v3mtw = face_matrices(order)
    deriv = _np(layer.derivatives_packed())
    rp, rm = oracle.dr_interpolate(order, faces.points, v3m, deriv, faces.plus_element, faces.minus_element,
                                   faces.face_info, faces.tinv_t, c)
    assert rel_l2(_np(qp), rp) < tol and rel_l2(_np(qm), rm) < tol
    ref = oracle.dr_nodal_flux(order, faces.points, v3mtw, faces.dr_mapping, faces.plus_element, faces.face_info,
                               faces.flux_plus, faces.flux_minus, gplus, gminus, q0)
    assert rel_l2(_np(layer.packed(layer.dofs)) - q0, ref - q0) < tol
    assert _padding_untouched(layer, layer.dofs)
