"""Dynamic-rupture face kernels on the GPU against the oracle: space-time
interpolation, nodal flux, and a welded fault through the whole chain."""
import numpy as np
import pytest

import oracle
from helpers import PolyField, project, rel_l2
from seissol_amd import _capi
from seissol_amd.dr import DeviceDrFaces, collocation_coefficients, face_matrices, plane_faces
from seissol_amd.kernels import Engine, Local, Neighbor, Spacetime, global_matrices, monomial_integrate
from seissol_amd.proxy import DeviceLayer, make_setup

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.build()


def _np(t):
    return t.detach().cpu().numpy()


def _case(order, seed, uniform=False, points=None):
    rho = (2500.0, 2500.0) if uniform else (2000.0, 3000.0)
    setup = make_setup(order, mesh="periodic", dims=(8, 6, 6), seed=seed, rho=rho, h=100.0, jitter=0.1)
    faces = plane_faces(setup, plane=4, points=points)
    return setup, faces


@pytest.mark.parametrize("order,arith,tol,extra", [
    (2, "fp64", 1e-12, 0), (3, "fp64", 1e-12, 0), (5, "fp64", 1e-12, 3), (6, "fp64", 1e-12, 0), (7, "fp64", 1e-12, 0),
    (2, "fp32", 2e-5, 2), (4, "fp32", 2e-5, 0), (6, "fp32", 2e-5, 0), (7, "fp32", 2e-5, 0),
])
def test_space_time_interpolation_matches_oracle(gpu, order, arith, tol, extra):
    import torch
    setup, faces = _case(order, seed=order, points=(order + 1) ** 2 + extra)
    print(f"dr interpolate order={order} {arith} faces={faces.num_faces} points={faces.points}", flush=True)
    eng = Engine(order, arith)
    layer = DeviceLayer(eng, setup, with_derivatives=True)
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    Spacetime(eng).computeBatchedAder(coeffs, setup.dt, layer.view)
    dr = DeviceDrFaces(faces)
    qp = torch.full((faces.num_faces, order, 9, faces.points), np.nan, dtype=torch.float64, device="cuda")
    qm = torch.full_like(qp, np.nan)
    c = collocation_coefficients(order, setup.dt)
    dr.interpolate(eng, c, layer.derivatives, qp, qm)
    torch.cuda.synchronize()
    v3m, _ = face_matrices(order)
    deriv = _np(layer.derivatives).astype(np.float64)
    rp, rm = oracle.dr_interpolate(order, faces.points, v3m, deriv, faces.plus_element, faces.minus_element,
                                   faces.face_info, faces.tinv_t, c)
    gp, gm_ = _np(qp), _np(qm)
    assert np.all(np.isfinite(gp)) and np.all(np.isfinite(gm_))
    assert np.all(gp[..., (order + 1) ** 2:] == 0.0)
    # both sides of the oracle from the GPU's own derivatives: the kernel's arithmetic alone
    ep, em = rel_l2(gp, rp), rel_l2(gm_, rm)
    print(f"rel L2 vs oracle: plus {ep:.2e} minus {em:.2e}")
    assert ep < tol and em < tol


@pytest.mark.parametrize("order,arith,tol", [(2, "fp64", 1e-12), (3, "fp64", 1e-12), (6, "fp64", 1e-12),
                                             (7, "fp64", 1e-12), (5, "fp32", 2e-5), (7, "fp32", 2e-5)])
def test_nodal_flux_matches_oracle(gpu, order, arith, tol):
    import torch
    setup, faces = _case(order, seed=10 + order)
    print(f"dr nodal flux order={order} {arith} faces={faces.num_faces}", flush=True)
    rng = np.random.default_rng(order)
    amp = np.array([1e6] * 6 + [1.0] * 3)[None, :, None]
    gplus = rng.standard_normal((faces.num_faces, 9, faces.points)) * amp
    gminus = rng.standard_normal((faces.num_faces, 9, faces.points)) * amp
    eng = Engine(order, arith)
    layer = DeviceLayer(eng, setup)
    q0 = _np(layer.dofs).astype(np.float64)
    dr = DeviceDrFaces(faces)
    tp = torch.from_numpy(gplus).cuda()
    tm = torch.from_numpy(gminus).cuda()
    dr.nodal_flux(eng, layer.view, tp, tm)
    torch.cuda.synchronize()
    _, v3mtw = face_matrices(order)
    ref = oracle.dr_nodal_flux(order, faces.points, v3mtw, faces.dr_mapping, faces.plus_element, faces.face_info,
                               faces.flux_plus, faces.flux_minus, gplus, gminus, q0)
    got = _np(layer.dofs).astype(np.float64)
    err = rel_l2(got - q0, ref - q0)
    print(f"rel L2 of the update vs oracle: {err:.2e}")
    assert err < tol
    untouched = faces.dr_mapping.max(1) < 0
    assert np.array_equal(got[untouched], q0[untouched])


@pytest.mark.parametrize("order", [3, 5])
def test_welded_fault_through_the_gpu_chain_equals_the_regular_update(gpu, order):
    """A fault whose imposed state is the time-integrated trace itself: local
    and neighbour kernels skipping the DR faces, interpolation and nodal flux
    reproduce the oracle's regular update of the continuous field."""
    import torch
    setup, faces = _case(order, seed=20 + order, uniform=True)
    field = PolyField(order, scale=800.0, rng=np.random.default_rng(order),
                      amplitude=np.array([1e6] * 6 + [1.0] * 3))
    setup.dofs = project(order, setup.vertices, field)
    ft_dr = setup.face_types.copy()
    # the oracle's regular update: all faces regular
    gm = global_matrices(order)["flat"]
    ft0 = np.zeros_like(ft_dr)
    star, ap, am = oracle.cell_local_matrices(setup.vertices, setup.material, setup.neighbors, ft0)
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    I = oracle.ader(order, gm, coeffs, setup.dofs, star)
    q_reg = oracle.neighbor(order, gm, I, setup.neighbors, setup.face_relations, am,
                            oracle.local(order, gm, I, star, ap, setup.dofs, ft0), ft0)
    # the GPU chain with DR faces
    print(f"welded fault order={order} faces={faces.num_faces}", flush=True)
    eng = Engine(order, "fp64")
    layer = DeviceLayer(eng, setup, with_derivatives=True)
    Spacetime(eng).computeBatchedAder(coeffs, setup.dt, layer.view)
    Local(eng).computeBatchedIntegral(layer.view)
    Neighbor(eng).computeBatchedNeighborsIntegral(layer.view)
    dr = DeviceDrFaces(faces)
    shape = (faces.num_faces, order, 9, faces.points)
    qp = torch.zeros(shape, dtype=torch.float64, device="cuda")
    qm = torch.zeros_like(qp)
    c = np.zeros((order, order))
    c[0] = coeffs
    dr.interpolate(eng, c, layer.derivatives, qp, qm)
    g = qp[:, 0].contiguous()
    dr.nodal_flux(eng, layer.view, g, g)
    torch.cuda.synchronize()
    got = _np(layer.dofs)
    change = np.linalg.norm(q_reg - setup.dofs)
    err = np.linalg.norm(got - q_reg)
    print(f"|gpu - regular| / |change| = {err / change:.2e}")
    assert err < 1e-10 * change


def test_dr_entry_points_reject_bad_layers(gpu):
    import torch
    setup, faces = _case(3, seed=1)
    eng = Engine(3, "fp64")
    layer = DeviceLayer(eng, setup, with_derivatives=True)
    dr = DeviceDrFaces(faces)
    qp = torch.zeros((faces.num_faces, 3, 9, faces.points), dtype=torch.float64, device="cuda")
    c = collocation_coefficients(3, setup.dt)
    bad = _capi.DrLayer(faces.num_faces, 8, 0, dr.view.plus_element, dr.view.minus_element, dr.view.face_info,
                        dr.view.tinv_t, dr.view.flux_solver_plus, dr.view.flux_solver_minus)
    import ctypes as C
    cc = np.ascontiguousarray(c)
    rc = _capi.lib().sdg_dr_space_time_interpolation(eng.handle, cc.ctypes.data, _capi.ptr(layer.derivatives),
                                                     C.byref(bad), _capi.ptr(qp), _capi.ptr(qp), None)
    assert rc != 0 and b"points" in _capi.lib().sdg_last_error()
    empty = _capi.DrLayer(0, 16, 0, None, None, None, None, None, None)
    assert _capi.lib().sdg_dr_space_time_interpolation(eng.handle, cc.ctypes.data, None, C.byref(empty), None, None,
                                                       None) == 0


@pytest.mark.parametrize("arith,tol", [("fp64", 1e-12), ("fp32", 2e-5), ("f16x3", 2e-5)])
def test_proxy_dr_kernels_match_oracle(gpu, arith, tol):
    """godunov_dr, neigh_dr and all_dr on the proxy's random DR layer (every
    cell face a DR face, 4 n faces between random cells).  A matrix-core
    engine runs ader + local on the matrix cores (writing the derivatives)
    and the DR kernels in FP32."""
    import torch
    from seissol_amd.proxy import DeviceDrProxy, run_proxy
    order = 4
    setup = make_setup(order, n=400, mesh="proxy", seed=5)
    print(f"proxy dr kernels {arith} n={setup.n}", flush=True)
    eng = Engine(order, arith)
    layer = DeviceLayer(eng, setup, with_derivatives=True)
    dr = DeviceDrProxy(eng, layer)
    h = dr.host
    v3m, v3mtw = face_matrices(order)
    c = collocation_coefficients(order, setup.dt)
    ip, im = _np(dr.imposed_plus), _np(dr.imposed_minus)
    # godunov_dr: interpolation of the random derivatives at the Gauss-Legendre points of [0, dt]
    deriv0 = _np(layer.derivatives).astype(np.float64)
    run_proxy(eng, layer, "godunov_dr", 1, setup.dt, dr=dr)
    torch.cuda.synchronize()
    rp, rm = oracle.dr_interpolate(order, dr.points, v3m, deriv0, h["plus"], h["minus"], h["face_info"],
                                   h["tinv_t"], c)
    e1 = max(rel_l2(_np(dr.q_plus), rp), rel_l2(_np(dr.q_minus), rm))
    # neigh_dr: no regular face, the nodal flux of the imposed states
    q0 = _np(layer.dofs).astype(np.float64)
    run_proxy(eng, layer, "neigh_dr", 1, setup.dt, dr=dr)
    torch.cuda.synchronize()
    q1 = _np(layer.dofs).astype(np.float64)
    ref1 = oracle.dr_nodal_flux(order, dr.points, v3mtw, h["dr_mapping"], h["plus"], h["face_info"], h["flux_plus"],
                                h["flux_minus"], ip, im, q0)
    e2 = rel_l2(q1 - q0, ref1 - q0)
    # all_dr: ader + volume (no local flux on DR faces), interpolation of the new derivatives, nodal flux
    run_proxy(eng, layer, "all_dr", 1, setup.dt, dr=dr)
    torch.cuda.synchronize()
    gm = global_matrices(order)["flat"]
    ft = np.full((setup.n, 4), _capi.SDG_FACE_DYNAMIC_RUPTURE, np.uint8)
    star, ap = _np(layer.star64), _np(layer.aplus64)
    I, deriv = oracle.ader(order, gm, monomial_integrate(order, 0.0, setup.dt), q1, star, want_derivatives=True)
    ref2 = oracle.local(order, gm, I, star, ap, q1, ft)
    rp2, rm2 = oracle.dr_interpolate(order, dr.points, v3m, deriv, h["plus"], h["minus"], h["face_info"],
                                     h["tinv_t"], c)
    ref2 = oracle.dr_nodal_flux(order, dr.points, v3mtw, h["dr_mapping"], h["plus"], h["face_info"], h["flux_plus"],
                                h["flux_minus"], ip, im, ref2)
    e3 = rel_l2(_np(layer.dofs).astype(np.float64) - q1, ref2 - q1)
    e4 = max(rel_l2(_np(dr.q_plus), rp2), rel_l2(_np(dr.q_minus), rm2))
    print(f"godunov_dr {e1:.2e} neigh_dr {e2:.2e} all_dr dofs {e3:.2e} all_dr q {e4:.2e}")
    assert max(e1, e2, e3, e4) < tol
    with pytest.raises(ValueError):
        run_proxy(eng, layer, "all_dr", 1, setup.dt)
