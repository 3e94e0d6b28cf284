"""Local time stepping on the GPU against the oracle, and its hipGraph replay."""
import numpy as np
import pytest

import oracle
from helpers import rel_l2
from seissol_amd import _capi
from seissol_amd.kernels import Engine, global_matrices
from seissol_amd.lts import DeviceLtsLayer, LocalTimeStepping, make_lts_setup

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.build()


def _np(t):
    return t.detach().cpu().numpy()


def _oracle_q(lts, order, steps, layer):
    s = lts.setup
    gm = global_matrices(order)["flat"]
    return oracle.lts(order, gm, steps, lts.num_clusters, lts.cluster_ids, lts.dt0, lts.lts_setups,
                      _np(layer.star64), _np(layer.aplus64), _np(layer.aminus64), s.neighbors, s.face_relations,
                      s.dofs)


@pytest.mark.parametrize("order,fractions,arith,tol", [
    (4, (0.5, 0.3, 0.2), "fp64", 1e-11),
    (5, (0.6, 0.25, 0.1, 0.05), "fp64", 1e-11),
    (3, (0.4, 0.35, 0.25), "fp32", 2e-5),
    (8, (0.5, 0.3, 0.2), "fp64", 1e-11),
])
def test_lts_matches_oracle(gpu, order, fractions, arith, tol):
    import torch
    dims = (20, 6, 6) if len(fractions) == 4 else (10, 6, 6)
    lts = make_lts_setup(order, fractions=fractions, dims=dims, seed=3)
    print(f"lts order={order} clusters={lts.num_clusters} n={lts.setup.n} {arith}", flush=True)
    assert lts.num_clusters == len(fractions) and np.all(np.diff(lts.cluster_start) > 0)
    eng = Engine(order, arith)
    layer = DeviceLtsLayer(eng, lts)
    run = LocalTimeStepping(eng, layer)
    run.advance(2)
    torch.cuda.synchronize()
    q_ref = _oracle_q(lts, order, 2, layer)
    err = rel_l2(_np(layer.dofs).astype(np.float64), q_ref)
    print(f"rel L2 vs oracle after 2 macro steps: {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("order,arith,tol", [
    (2, "fp64", 1e-11), (3, "fp64", 1e-11), (4, "fp32", 2e-5), (6, "fp64", 1e-11), (6, "fp32", 2e-5),
    (7, "fp64", 1e-11), (7, "fp32", 2e-5), (8, "fp32", 2e-5),
])
def test_neighbour_rows_variant_matches_oracle(gpu, order, arith, tol):
    # the neighbour pass with one element per 16-lane row (tile variant 4)
    # at the orders the LTS tile tests above leave out
    import torch
    lts = make_lts_setup(order, fractions=(0.5, 0.3, 0.2), dims=(8, 6, 6), seed=11)
    print(f"neighbour rows order={order} {arith} n={lts.setup.n}", flush=True)
    eng = Engine(order, arith)
    layer = DeviceLtsLayer(eng, lts)
    run = LocalTimeStepping(eng, layer)
    run.set_tile(4, kernel=1)
    run.advance(2)
    torch.cuda.synchronize()
    q_ref = _oracle_q(lts, order, 2, layer)
    err = rel_l2(_np(layer.dofs).astype(np.float64), q_ref)
    print(f"rel L2 vs oracle after 2 macro steps: {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("arith", ["fp64", "fp32"])
def test_every_tile_variant_and_the_autotuner_keep_the_results(gpu, arith):
    import torch
    order = 5
    lts = make_lts_setup(order, fractions=(0.5, 0.3, 0.2), dims=(10, 6, 6), seed=6)
    eng = Engine(order, arith)
    tol = 1e-11 if arith == "fp64" else 2e-5
    q_ref = None
    # FP64 at orders 3-7: local variant 0 is the FP64 matrix-core pass, 1-4 the VALU tiles
    local_variants = 5 if arith == "fp64" else 4
    for variant in range(local_variants):
        layer = DeviceLtsLayer(eng, lts)
        run = LocalTimeStepping(eng, layer)
        run.set_tile(variant, kernel=0)
        run.set_tile(min(variant, 3), kernel=1)
        run.advance(2)
        torch.cuda.synchronize()
        if q_ref is None:
            q_ref = _oracle_q(lts, order, 2, layer)
        err = rel_l2(_np(layer.dofs).astype(np.float64), q_ref)
        print(f"{arith} tile variant {variant}: rel L2 vs oracle {err:.2e}", flush=True)
        assert err < tol
    # the neighbour pass with one element per 16-lane row
    layer = DeviceLtsLayer(eng, lts)
    run = LocalTimeStepping(eng, layer)
    run.set_tile(4, kernel=1)
    run.advance(2)
    torch.cuda.synchronize()
    err = rel_l2(_np(layer.dofs).astype(np.float64), q_ref)
    print(f"{arith} neighbour variant 4: rel L2 vs oracle {err:.2e}", flush=True)
    assert err < tol
    with pytest.raises(_capi.SdgError):
        run.set_tile(5, kernel=1)
    # the autotuner restores the state it ran its trials on
    layer = DeviceLtsLayer(eng, lts)
    run = LocalTimeStepping(eng, layer)
    before = layer.dofs.clone()
    tiles = run.autotune(repeats=2)
    torch.cuda.synchronize()
    assert bool(torch.equal(before, layer.dofs))
    assert set(tiles) == {0, 1, 2}
    assert all(0 <= t["local"] < local_variants and 0 <= t["neigh"] < 5 for t in tiles.values())
    run.advance(2, graph=True)
    torch.cuda.synchronize()
    assert rel_l2(_np(layer.dofs).astype(np.float64), q_ref) < tol


def test_lts_graph_replay_equals_eager(gpu):
    import torch
    order = 4
    lts = make_lts_setup(order, fractions=(0.5, 0.3, 0.2), dims=(10, 6, 6), seed=8)
    eng = Engine(order, "fp64")
    eager_layer = DeviceLtsLayer(eng, lts)
    eager = LocalTimeStepping(eng, eager_layer)
    graph_layer = DeviceLtsLayer(eng, lts)
    graphed = LocalTimeStepping(eng, graph_layer)
    for step in range(3):
        eager.advance(1)
        graphed.advance(1, graph=True)
        torch.cuda.synchronize()
        # the program's own eager work between replays
        scratch = torch.ones(1 << 16, device="cuda").cumsum(0)
        torch.cuda.synchronize()
        assert bool(torch.equal(eager_layer.dofs, graph_layer.dofs)), f"replay {step} differs"
        del scratch


@pytest.mark.parametrize("order,fractions,arith,tol", [
    (5, (0.6, 0.25, 0.1, 0.05), "f16x3", 1e-5),
    (3, (0.4, 0.35, 0.25), "f16x3", 1e-5),
    (4, (0.5, 0.3, 0.2), "bf16x3", 5e-5),
])
def test_lts_on_the_matrix_cores_matches_oracle(gpu, order, fractions, arith, tol):
    # local pass on the matrix cores (buffers and derivatives per the LTS
    # bits), neighbour pass on the FP32 VALU kernel
    import torch
    dims = (20, 6, 6) if len(fractions) == 4 else (10, 6, 6)
    lts = make_lts_setup(order, fractions=fractions, dims=dims, seed=3)
    print(f"lts order={order} clusters={lts.num_clusters} n={lts.setup.n} {arith}", flush=True)
    eng = Engine(order, arith)
    layer = DeviceLtsLayer(eng, lts)
    run = LocalTimeStepping(eng, layer)
    assert all(t["local"] == 0 for t in run.tiles().values())  # the matrix-core pass by default
    run.advance(2)
    torch.cuda.synchronize()
    q_ref = _oracle_q(lts, order, 2, layer)
    q = _np(layer.packed(layer.dofs)).astype(np.float64)
    per_q = [rel_l2(q[:, p], q_ref[:, p]) for p in range(9)]
    print(f"rel L2 per quantity vs oracle after 2 macro steps: max {max(per_q):.2e}", flush=True)
    assert max(per_q) < tol


def test_matrix_core_lts_variants_and_autotuner(gpu):
    import torch
    order = 5
    lts = make_lts_setup(order, fractions=(0.5, 0.3, 0.2), dims=(10, 6, 6), seed=6)
    eng = Engine(order, "f16x3")
    q_ref = None
    for variant in range(5):  # 0: matrix cores, 1..4: the FP32 VALU tiles
        layer = DeviceLtsLayer(eng, lts)
        run = LocalTimeStepping(eng, layer)
        run.set_tile(variant, kernel=0)
        run.advance(2)
        torch.cuda.synchronize()
        if q_ref is None:
            q_ref = _oracle_q(lts, order, 2, layer)
        err = rel_l2(_np(layer.packed(layer.dofs)).astype(np.float64), q_ref)
        print(f"f16x3 local variant {variant}: rel L2 vs oracle {err:.2e}", flush=True)
        assert err < 1e-5
    # the FP32 neighbour pass with one element per 16-lane row
    layer = DeviceLtsLayer(eng, lts)
    run = LocalTimeStepping(eng, layer)
    run.set_tile(4, kernel=1)
    run.advance(2)
    torch.cuda.synchronize()
    err = rel_l2(_np(layer.packed(layer.dofs)).astype(np.float64), q_ref)
    print(f"f16x3 neighbour variant 4: rel L2 vs oracle {err:.2e}", flush=True)
    assert err < 1e-5
    layer = DeviceLtsLayer(eng, lts)
    run = LocalTimeStepping(eng, layer)
    before = layer.dofs.clone()
    tiles = run.autotune(repeats=2)
    torch.cuda.synchronize()
    assert bool(torch.equal(before, layer.dofs))
    assert all(0 <= t["local"] < 5 and 0 <= t["neigh"] < 5 for t in tiles.values())
    with pytest.raises(_capi.SdgError):
        run.set_tile(5, kernel=0)
    run.advance(2, graph=True)
    torch.cuda.synchronize()
    assert rel_l2(_np(layer.packed(layer.dofs)).astype(np.float64), q_ref) < 1e-5


@pytest.mark.parametrize("order,arith,neigh", [
    (2, "fp32", 0), (3, "fp32", 4), (6, "fp32", 4), (7, "fp32", 0), (5, "f16x3", 4),
    (3, "fp64", 0), (4, "fp64", 4), (5, "fp64", 4), (6, "fp64", 0),
])
def test_default_neighbour_tiles(gpu, order, arith, neigh):
    # the untuned neighbour tile: one element per 16-lane row (4) where it was
    # measured faster (FP32 kernel at orders 3-6, FP64 at orders 4-5), else 0
    lts = make_lts_setup(order, fractions=(0.6, 0.4), dims=(6, 4, 4), seed=2)
    eng = Engine(order, arith)
    run = LocalTimeStepping(eng, DeviceLtsLayer(eng, lts))
    assert all(t["neigh"] == neigh for t in run.tiles().values()), run.tiles()


def test_lts_rejects_missing_setup(gpu):
    lts = make_lts_setup(3, fractions=(0.5, 0.5), dims=(6, 4, 4), seed=1)
    eng64 = Engine(3, "fp64")
    layer64 = DeviceLtsLayer(eng64, lts)
    layer64.view.lts_setup = None
    with pytest.raises(_capi.SdgError):
        LocalTimeStepping(eng64, layer64)
