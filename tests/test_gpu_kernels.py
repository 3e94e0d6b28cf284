"""Parity of the HIP element-update path with the FP64 oracle (GPU tests).

Every call goes through the C ABI (ctypes).  FP64 results must agree with the
oracle to rel. L2 <= 1e-12 (reassociated sums); FP32 storage/arithmetic to
<= 2e-6 per kernel application (single precision of the data).
"""
import numpy as np
import pytest

import oracle
from helpers import PolyField, coefficient_matrices, evolve, interior_elements, project, rel_l2
from seissol_amd import _capi
from seissol_amd.kernels import (Engine, Local, Neighbor, Spacetime, Time, derivatives_size, global_matrices,
                                 monomial_integrate, num_basis)
from seissol_amd.proxy import DeviceLayer, make_setup, run_proxy

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-6}


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.build()


def _np(t):
    return t.detach().cpu().double().numpy()


def _sync():
    import torch
    torch.cuda.synchronize()


@pytest.mark.parametrize("mesh", ["proxy", "periodic"])
def test_cell_local_matrices_match_oracle(gpu, mesh):
    setup = make_setup(4, n=777, mesh=mesh, seed=21, dims=(4, 4, 6))
    print(f"cell-local n={setup.n} mesh={mesh}", flush=True)
    eng = Engine(4, "fp64")
    layer = DeviceLayer(eng, setup)
    _sync()
    star, ap, am = oracle.cell_local_matrices(setup.vertices, setup.material, setup.neighbors)
    assert rel_l2(_np(layer.star64), star) < 1e-13
    assert rel_l2(_np(layer.aplus64), ap) < 1e-12
    assert rel_l2(_np(layer.aminus64), am) < 1e-12


@pytest.mark.parametrize("arith", ["fp64", "fp32"])
@pytest.mark.parametrize("order", [2, 3, 4, 5, 6, 7, 8])
def test_ader_with_derivatives(gpu, order, arith):
    import torch
    n = 203  # ragged: not a multiple of any block size
    print(f"ader order={order} arith={arith} n={n}", flush=True)
    if order == 8 and arith == "fp32":
        # the proxy mesh's centimetre cells at 6 km/s make the 7th time
        # derivative ~4e43 (beyond FP32) at ORDER 8; a 100 m box keeps it in
        # range (the FP64 case covers the ragged count)
        setup = make_setup(order, mesh="periodic", dims=(4, 4, 4), seed=100 + order)
    else:
        setup = make_setup(order, n=n, mesh="proxy", seed=100 + order)
    eng = Engine(order, arith)
    layer = DeviceLayer(eng, setup, with_derivatives=True)
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    Spacetime(eng).computeBatchedAder(coeffs, setup.dt, layer.view)
    _sync()
    gm = global_matrices(order)["flat"]
    star = _np(layer.star64)
    I_ref, d_ref = oracle.ader(order, gm, coeffs, setup.dofs, star, want_derivatives=True)
    assert rel_l2(_np(layer.time_integrated), I_ref) < TOL[arith]
    assert rel_l2(_np(layer.derivatives), d_ref) < TOL[arith] * 10
    # Time::evaluateBatched on the derivatives reproduces the integral
    out = torch.zeros_like(layer.time_integrated)
    Time(eng).evaluateBatched(coeffs, layer.derivatives, out, setup.n)
    _sync()
    assert rel_l2(_np(out), I_ref) < TOL[arith]
    # DOFs untouched by ader
    assert rel_l2(_np(layer.dofs), setup.dofs) < (0 if arith == "fp64" else 1e-7) + 1e-300


@pytest.mark.parametrize("arith", ["fp64", "fp32"])
@pytest.mark.parametrize("order", [3, 4, 6, 8])
def test_local_integral(gpu, order, arith):
    import torch
    n = 150
    print(f"local order={order} arith={arith}", flush=True)
    setup = make_setup(order, n=n, mesh="proxy", seed=7 * order)
    setup.face_types[::5, 2] = _capi.SDG_FACE_DYNAMIC_RUPTURE
    eng = Engine(order, arith)
    layer = DeviceLayer(eng, setup)
    rng = np.random.default_rng(order)
    I = rng.uniform(-1, 1, setup.dofs.shape)
    layer.time_integrated.copy_(torch.from_numpy(I).to(layer.time_integrated))
    Local(eng).computeBatchedIntegral(layer.view)
    _sync()
    gm = global_matrices(order)["flat"]
    I_used = _np(layer.time_integrated)
    ref = oracle.local(order, gm, I_used, _np(layer.star64), _np(layer.aplus64), setup.dofs, setup.face_types)
    assert rel_l2(_np(layer.dofs) - setup.dofs, ref - setup.dofs) < TOL[arith] * 10


@pytest.mark.parametrize("arith", ["fp64", "fp32"])
@pytest.mark.parametrize("order", [3, 5, 6, 7, 8])
def test_neighbor_integral(gpu, order, arith):
    import torch
    n = 170
    print(f"neighbor order={order} arith={arith}", flush=True)
    setup = make_setup(order, n=n, mesh="proxy", seed=13 * order)
    setup.face_types[::7, 1] = _capi.SDG_FACE_DYNAMIC_RUPTURE
    eng = Engine(order, arith)
    layer = DeviceLayer(eng, setup)
    rng = np.random.default_rng(order + 1)
    I = rng.uniform(-1, 1, setup.dofs.shape)
    layer.time_integrated.copy_(torch.from_numpy(I).to(layer.time_integrated))
    Neighbor(eng).computeBatchedNeighborsIntegral(layer.view)
    _sync()
    gm = global_matrices(order)["flat"]
    ref = oracle.neighbor(order, gm, _np(layer.time_integrated), setup.neighbors, setup.face_relations,
                          _np(layer.aminus64), setup.dofs, setup.face_types)
    assert rel_l2(_np(layer.dofs) - setup.dofs, ref - setup.dofs) < TOL[arith] * 10


@pytest.mark.parametrize("arith", ["fp64", "fp32"])
@pytest.mark.parametrize("order,mesh", [(4, "proxy"), (6, "proxy"), (6, "periodic"), (8, "periodic")])
def test_proxy_all_steps(gpu, order, mesh, arith):
    steps = 3
    setup = make_setup(order, n=1200, mesh=mesh, seed=99, dims=(4, 6, 8))
    print(f"proxy all order={order} mesh={mesh} arith={arith} n={setup.n}", flush=True)
    eng = Engine(order, arith)
    layer = DeviceLayer(eng, setup)
    run_proxy(eng, layer, "all", steps, setup.dt)
    _sync()
    gm = global_matrices(order)["flat"]
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    q_ref, I_ref = oracle.proxy_all(order, gm, steps, coeffs, _np(layer.star64), _np(layer.aplus64),
                                    _np(layer.aminus64), setup.neighbors, setup.face_relations, setup.dofs)
    err = rel_l2(_np(layer.dofs), q_ref)
    assert err < (1e-12 if arith == "fp64" else 1e-5), err
    assert rel_l2(_np(layer.time_integrated), I_ref) < (1e-12 if arith == "fp64" else 1e-5)


@pytest.mark.parametrize("kernel", ["local", "neigh", "ader", "localwoader"])
def test_proxy_kernel_selection(gpu, kernel):
    order = 4
    setup = make_setup(order, n=300, mesh="proxy", seed=3)
    print(f"proxy kernel={kernel}", flush=True)
    eng = Engine(order, "fp64")
    layer = DeviceLayer(eng, setup)
    run_proxy(eng, layer, kernel, 1, setup.dt)
    _sync()
    gm = global_matrices(order)["flat"]
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    star, ap, am = _np(layer.star64), _np(layer.aplus64), _np(layer.aminus64)
    I0 = np.zeros_like(setup.dofs)
    if kernel == "local":
        I = oracle.ader(order, gm, coeffs, setup.dofs, star)
        q = oracle.local(order, gm, I, star, ap, setup.dofs)
    elif kernel == "ader":
        I = oracle.ader(order, gm, coeffs, setup.dofs, star)
        q = setup.dofs
    elif kernel == "neigh":
        I = I0
        q = oracle.neighbor(order, gm, I0, setup.neighbors, setup.face_relations, am, setup.dofs)
    else:
        I = I0
        q = oracle.local(order, gm, I0, star, ap, setup.dofs)
    assert rel_l2(_np(layer.dofs), q) < 1e-12
    assert np.abs(_np(layer.time_integrated) - I).max() <= 1e-12 * max(1.0, np.abs(I).max())


@pytest.mark.parametrize("order", [3, 5, 6, 8])
def test_gpu_step_exact_for_polynomial_states(gpu, order):
    import torch
    dims = (6, 6, 6)
    setup = make_setup(order, mesh="periodic", dims=dims, seed=3, rho=(2500.0, 2500.0), h=100.0, jitter=0.12)
    print(f"polynomial exactness order={order} n={setup.n}", flush=True)
    eng = Engine(order, "fp64")
    layer = DeviceLayer(eng, setup)
    A = coefficient_matrices(*setup.material[0])
    field = PolyField(order, scale=dims[0] * 100.0, rng=np.random.default_rng(3),
                      amplitude=np.array([1e6] * 6 + [1.0] * 3))
    q0 = project(order, setup.vertices, field)
    layer.reset_dofs(q0)
    run_proxy(eng, layer, "all", 1, setup.dt)
    _sync()
    exact = project(order, setup.vertices, evolve(field, A, setup.dt))
    inner = interior_elements(setup)
    q1 = _np(layer.dofs)
    change = np.linalg.norm(exact[inner] - q0[inner])
    assert np.linalg.norm(q1[inner] - exact[inner]) < 1e-9 * change


def test_empty_layer_and_errors(gpu):
    import ctypes
    eng = Engine(4, "fp64")
    empty = _capi.Layer(0, None, None, None, None, None, None, None, None, None)
    coeffs = monomial_integrate(4, 0, 1)
    _capi.check(_capi.lib().sdg_compute_batched_ader_local(eng.handle, coeffs.ctypes.data, 1.0, ctypes.byref(empty),
                                                            None))
    _capi.check(_capi.lib().sdg_compute_batched_neighbors_integral(eng.handle, ctypes.byref(empty), None))
    bad = _capi.Layer(10, None, None, None, None, None, None, None, None, None)
    assert _capi.lib().sdg_compute_batched_neighbors_integral(eng.handle, ctypes.byref(bad), None) != 0
    assert b"missing" in _capi.lib().sdg_last_error()
    # an empty layer is a no-op for every form, the matrix-core kernels included
    for arith in ("fp32", "bf16x3", "f16x3"):
        e2 = Engine(4, arith)
        _capi.check(_capi.lib().sdg_compute_batched_ader_local(e2.handle, coeffs.ctypes.data, 1.0,
                                                                ctypes.byref(empty), None))
        _capi.check(_capi.lib().sdg_compute_batched_neighbors_integral(e2.handle, ctypes.byref(empty), None))
        _capi.check(_capi.lib().sdg_compute_batched_ader(e2.handle, coeffs.ctypes.data, 1.0, ctypes.byref(empty),
                                                         None))


# ---------------------------------------------------------------------------
# matrix-core arithmetic (FP32 data, split bf16/fp16 products, FP32 accumulation)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("arith", ["bf16x3", "bf16x6", "f16x3"])
@pytest.mark.parametrize("order,mesh", [(3, "proxy"), (4, "proxy"), (5, "periodic"), (6, "proxy"), (6, "periodic"),
                                        (7, "periodic")])
def test_split_bf16_proxy_all_matches_oracle(gpu, order, mesh, arith):
    steps = 3
    setup = make_setup(order, n=900, mesh=mesh, seed=77, dims=(4, 6, 8))
    setup.face_types[::11, 3] = _capi.SDG_FACE_DYNAMIC_RUPTURE
    print(f"{arith} proxy all order={order} mesh={mesh} n={setup.n}", flush=True)
    eng = Engine(order, arith)
    layer = DeviceLayer(eng, setup)
    run_proxy(eng, layer, "all", steps, setup.dt)
    _sync()
    gm = global_matrices(order)["flat"]
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    q_ref, _ = oracle.proxy_all(order, gm, steps, coeffs, _np(layer.star64), _np(layer.aplus64),
                                _np(layer.aminus64), setup.neighbors, setup.face_relations, setup.dofs,
                                setup.face_types)
    err = rel_l2(_np(layer.dofs), q_ref)
    print(f"{arith} order {order} {mesh}: rel L2 after {steps} steps {err:.2e}")
    # 2-plane bf16 splits: ~2^-17 per contraction; 2-plane scaled fp16: ~2^-21;
    # 3-plane bf16 splits: FP32 level
    assert err < {"bf16x3": 2e-5, "f16x3": 5e-6, "bf16x6": 2e-6}[arith], err


@pytest.mark.parametrize("arith", ["bf16x3", "bf16x6", "f16x3"])
def test_split_bf16_local_writes_integral_and_traces(gpu, arith):
    import torch
    order = 6
    setup = make_setup(order, n=333, mesh="proxy", seed=5)
    eng = Engine(order, arith)
    layer = DeviceLayer(eng, setup)
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    Spacetime(eng).computeBatchedAderLocal(coeffs, setup.dt, layer.view)
    _sync()
    gm = global_matrices(order)
    I_ref = oracle.ader(order, gm["flat"], coeffs, setup.dofs, _np(layer.star64))
    # split products: about 2^-17 (bf16x3), 2^-21 (f16x3), 2^-24 (bf16x6) per contraction
    tol = {"bf16x3": 1e-5, "f16x3": 3e-6, "bf16x6": 1e-6}[arith]
    err_i = rel_l2(_np(layer.time_integrated), I_ref)
    assert err_i < tol, err_i
    q_ref = oracle.local(order, gm["flat"], I_ref, _np(layer.star64), _np(layer.aplus64), setup.dofs)
    err_q = rel_l2(_np(layer.dofs) - setup.dofs, q_ref - setup.dofs)
    assert err_q < 3 * tol, err_q
    # traces: rT_f I for every face
    tr_ref = np.einsum("fml,nql->nfqm", gm["rT"], I_ref)
    B2 = tr_ref.shape[-1]
    got_t = _np(layer.face_traces)
    assert got_t.shape[-1] % 4 == 0 and np.all(got_t[..., B2:] == 0.0)  # padded rows stay zero
    err_t = rel_l2(got_t[..., :B2], tr_ref)
    assert err_t < tol, err_t
    print(f"{arith} rel errors: I {err_i:.2e}, dQ {err_q:.2e}, traces {err_t:.2e}")


def test_order8_runs_in_fp64_and_fp32_only(gpu):
    # ORDER 8 (120 basis functions) is served by the FP64/FP32 kernels; the
    # matrix-core forms stop at ORDER 7 and say so when the engine is created
    print("order 8 matrix-core rejection", flush=True)
    for arith in ("bf16x3", "bf16x6", "f16x3"):
        with pytest.raises(_capi.SdgError, match="orders 3..7"):
            Engine(8, arith)
    Engine(8, "fp64")
    Engine(8, "fp32")


def test_f16x3_stays_within_1e5_of_the_oracle_over_200_steps(gpu):
    # the chosen form's accuracy bar over the step count of the precision
    # form (200 steps, ORDER 6), on a conforming periodic box of 5120 cells
    order, steps = 6, 200
    setup = make_setup(order, mesh="periodic", dims=(8, 8, 16), seed=21)
    print(f"f16x3 order {order} n={setup.n} {steps} steps", flush=True)
    eng = Engine(order, "f16x3")
    eng.set_layout(0, "compact")
    layer = DeviceLayer(eng, setup)
    run_proxy(eng, layer, "all", steps, setup.dt)
    _sync()
    gm = global_matrices(order)["flat"]
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    q_ref, _ = oracle.proxy_all(order, gm, steps, coeffs, _np(layer.star64), _np(layer.aplus64),
                                _np(layer.aminus64), setup.neighbors, setup.face_relations, setup.dofs)
    q = _np(layer.dofs).astype(np.float64)
    per_q = [rel_l2(q[:, p], q_ref[:, p]) for p in range(9)]
    print(f"f16x3 after {steps} steps: rel L2 per quantity max {max(per_q):.2e}")
    assert max(per_q) < 1e-5, per_q


@pytest.mark.parametrize("order", [3, 4, 5, 6])
def test_fp32_neighbour_rows_with_dr_faces_match_oracle(gpu, order):
    # FP32 at orders 3-6 runs the neighbour pass with one element per 16-lane
    # row; DR faces in the layer take no neighbour flux
    steps = 2
    setup = make_setup(order, n=701, mesh="proxy", seed=43)
    setup.face_types[::5, 1] = _capi.SDG_FACE_DYNAMIC_RUPTURE
    print(f"fp32 neighbour rows order={order} n={setup.n}", flush=True)
    eng = Engine(order, "fp32")
    layer = DeviceLayer(eng, setup)
    gm = global_matrices(order)["flat"]
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    run_proxy(eng, layer, "all", steps, setup.dt)
    _sync()
    q_ref, _ = oracle.proxy_all(order, gm, steps, coeffs, _np(layer.star64), _np(layer.aplus64), _np(layer.aminus64),
                                setup.neighbors, setup.face_relations, setup.dofs, setup.face_types)
    err = rel_l2(_np(layer.dofs).astype(np.float64), q_ref)
    print(f"fp32 order {order}: rel L2 after {steps} steps {err:.2e}")
    assert err < 1e-5, err


@pytest.mark.parametrize("kernels", ["auto", "valu"])
@pytest.mark.parametrize("order", [3, 4, 5, 6, 7])
def test_fp64_matrix_core_and_valu_passes_match_oracle(gpu, order, kernels):
    # "auto": the FP64 ader + local pass on v_mfma_f64_16x16x4 (orders 3-7);
    # "valu": the VALU kernels; both against the oracle at 1e-12, with DR faces
    # in the layer (no local or neighbour flux there) and the derivatives written
    steps = 2
    setup = make_setup(order, n=700, mesh="proxy", seed=41)
    setup.face_types[::7, 2] = _capi.SDG_FACE_DYNAMIC_RUPTURE
    print(f"fp64 {kernels} order={order} n={setup.n}", flush=True)
    eng = Engine(order, "fp64")
    eng.set_kernels(kernels)
    layer = DeviceLayer(eng, setup, with_derivatives=True)
    gm = global_matrices(order)["flat"]
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    star, ap, am = _np(layer.star64), _np(layer.aplus64), _np(layer.aminus64)
    # one fused ader + local pass: I, derivatives, Q
    run_proxy(eng, layer, "local", 1, setup.dt)
    _sync()
    I_ref, d_ref = oracle.ader(order, gm, coeffs, setup.dofs, star, want_derivatives=True)
    q_ref = oracle.local(order, gm, I_ref, star, ap, setup.dofs, setup.face_types)
    assert rel_l2(_np(layer.time_integrated), I_ref) < 1e-12
    assert rel_l2(_np(layer.derivatives_packed()), d_ref) < 1e-12
    assert rel_l2(_np(layer.dofs) - setup.dofs, q_ref - setup.dofs) < 1e-11
    # proxy `all` from the initial state
    layer.reset_dofs(setup.dofs)
    run_proxy(eng, layer, "all", steps, setup.dt)
    _sync()
    q_ref, I_ref = oracle.proxy_all(order, gm, steps, coeffs, star, ap, am, setup.neighbors, setup.face_relations,
                                    setup.dofs, setup.face_types)
    err = rel_l2(_np(layer.dofs), q_ref)
    print(f"fp64 {kernels} order {order}: rel L2 after {steps} steps {err:.2e}")
    assert err < 1e-12, err
    with pytest.raises(_capi.SdgError):
        Engine(order, "f16x3").set_kernels("valu")
