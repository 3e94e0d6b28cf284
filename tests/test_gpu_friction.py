"""GPU friction solver (law 103) against the FP64 oracle, and the warm-start
MLP's kernels and effect (GPU tests)."""
import math

import numpy as np
import pytest

import oracle
from seissol_amd._capi import FrictionOptions, ptr
from seissol_amd.friction import (FaultDriver, FrictionParameters, FrictionSolver, WarmStartTrainer, dr_quadrature,
                                  make_fault_setup, mlp_param_count, pulse_loading, step_times)

pytestmark = pytest.mark.gpu

COMPARED = ["slip_rate_magnitude", "slip_rate1", "slip_rate2", "state_variable", "mu", "slip1", "slip2",
            "accumulated_slip", "traction1", "traction2", "peak_slip_rate", "imposed_state_plus",
            "imposed_state_minus"]


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.build()


def _params_dict(p: FrictionParameters):
    return dict(rs_sr0=p.rs_sr0, rs_slip_rate_tolerance=p.rs_slip_rate_tolerance,
                rs_state_tolerance=p.rs_state_tolerance, rs_max_slip_rate_updates=p.rs_max_slip_rate_updates,
                rs_state_variable_updates=p.rs_state_variable_updates, eta_damp=p.eta_damp,
                eta_damp_end=p.eta_damp_end, nucleation_count=p.nucleation_count, t0=p.t0, s0=p.s0)


def _run_oracle(setup, steps, dt):
    st = oracle.FaultState(setup.arrays, setup.order, setup.points, setup.num_points)
    _, _, R = dr_quadrature(setup.order)
    tp, tw, delta = step_times(setup.order, dt)
    for s in range(steps):
        load = pulse_loading(setup, s * dt + tp)
        st.a["q_interpolated_plus"][:, :, 3, :] = load
        st.a["q_interpolated_minus"][:, :, 3, :] = load
        oracle.friction_evaluate(st, _params_dict(setup.params), s * dt, delta, tw, R)
    return st.a


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("order,faces,padded", [(4, 300, None), (4, 123, 32), (5, 100, None)])
def test_friction_matches_oracle(gpu, order, faces, padded):
    import torch
    steps, dt = 25, 1e-3
    setup = make_fault_setup(num_faces=faces, order=order, seed=11, points=padded, length=4000.0,
                             nucleation_radius=900.0)
    print(f"friction order={order} faces={faces} points={setup.points}", flush=True)
    drv = FaultDriver(setup, FrictionSolver(setup.params, order), dt)
    drv.advance(steps)
    torch.cuda.synchronize()
    ref = _run_oracle(setup, steps, dt)
    np_ = setup.num_points
    assert ref["slip_rate_magnitude"].max() > 0.01, "scenario must rupture"
    for name in COMPARED:
        got = drv.fault.host(name)[..., :np_]
        exp = ref[name][..., :np_]
        err = _rel(got, exp)
        assert err < 1e-9, (name, err)
    for name in ("rupture_time", "dyn_stress_time"):
        np.testing.assert_allclose(drv.fault.host(name)[..., :np_], ref[name][..., :np_], rtol=0, atol=1e-12)
    for name in ("rupture_time_pending", "convergence_inner", "convergence_outer"):
        np.testing.assert_array_equal(drv.fault.host(name)[..., :np_], ref[name][..., :np_])
    # identical face-wide iteration structure: the same number of Newton evaluations per point
    np.testing.assert_array_equal(drv.fault.host("newton_iterations")[..., :np_], ref["newton_iterations"][..., :np_])


def test_mlp_loss_and_gradient_match_oracle(gpu):
    import torch
    rng = np.random.default_rng(1)
    n = 1000
    samples = rng.normal(size=(n, 7)).astype(np.float32)
    tr = WarmStartTrainer(seed=3)
    # identity normalisation so the oracle sees the same numbers
    tr.params[0:6] = 0.0
    tr.params[6:12] = 1.0
    tr.params[12] = 0.0
    tr.params[13] = 1.0
    tr.loss_grad(torch.from_numpy(samples).cuda())
    torch.cuda.synchronize()
    p = tr.params.cpu().double().numpy()
    loss, g = oracle.mlp_loss_grad(samples[:, :6].astype(np.float64), samples[:, 6].astype(np.float64), p[14:])
    assert abs(tr.loss.item() - loss) <= 1e-5 * loss
    gg = tr.grads.cpu().double().numpy()[14:]
    assert _rel(gg, g) < 1e-4, _rel(gg, g)
    assert np.all(tr.grads.cpu().numpy()[:14] == 0.0)  # normalisation constants are not trained


def test_mlp_training_reduces_loss(gpu):
    import torch
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, size=(20000, 6)).astype(np.float32)
    y = (np.sin(2 * x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - x[:, 5]).astype(np.float32)
    samples = torch.from_numpy(np.concatenate([x, y[:, None]], 1)).cuda()
    tr = WarmStartTrainer(seed=0)
    tr.set_normalisation(samples)
    first = float(tr.loss_grad(samples).item())
    tr.fit(samples, steps=1000, batch=1024, lr=3e-3, balance=False)
    last = float(tr.loss_grad(samples).item())
    assert last < 0.1 * first, (first, last)


def test_warm_start_keeps_the_root_and_reports_iterations(gpu):
    import torch
    steps, dt = 40, 1e-3
    setup = make_fault_setup(num_faces=2000, order=4, seed=5, length=6000.0, nucleation_radius=1200.0)
    # plain solver, recording (features, root) samples of its first solves
    solver = FrictionSolver(setup.params, 4)
    drv = FaultDriver(setup, solver, dt)
    cap = 400_000
    buf = torch.zeros((cap, 7), dtype=torch.float32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    opt = FrictionOptions(None, ptr(buf), cap, ptr(cnt), 3, 0)
    drv.advance(steps, opt)
    torch.cuda.synchronize()
    n = min(int(cnt.item()), cap)
    assert n > 50_000
    plain_iters = int(drv.fault.t["newton_iterations"].sum().item())
    perm = torch.randperm(n, generator=torch.Generator(device="cuda").manual_seed(0), device="cuda")
    train, held = buf[:n][perm[: n * 9 // 10]].contiguous(), buf[:n][perm[n * 9 // 10:]].contiguous()
    tr = WarmStartTrainer(seed=0)
    tr.fit(train, steps=3000, batch=1024, lr=3e-3)
    # the network improves on the closed-form creep estimate it corrects, where
    # that estimate is off (elsewhere it is exact to rounding)
    off = held[:, 6].abs() > 0.01
    assert int(off.sum().item()) > 100
    err = (tr.predict(held) - held[:, 6]).abs()[off].median().item()
    err_base = held[:, 6].abs()[off].median().item()
    print(f"held-out median |log10 error| where the creep estimate is off by > 0.01: network {err:.4f}, "
          f"creep estimate {err_base:.4f}", flush=True)
    assert err < 0.5 * err_base
    # warm-started solver from the same initial state
    solver2 = FrictionSolver(setup.params, 4)
    solver2.mlp_params = tr.params
    drv2 = FaultDriver(setup, solver2, dt)
    drv2.advance(steps)
    torch.cuda.synchronize()
    warm_iters = int(drv2.fault.t["newton_iterations"].sum().item())
    print(f"Newton evaluations: plain {plain_iters}, warm-started {warm_iters}, "
          f"ratio {plain_iters / warm_iters:.3f}", flush=True)
    # the warm start adds no evaluation; on this workload the previous slip rate
    # is already within one Newton step for almost every solve, so the count
    # stays about the same (a perfect predictor gives about 1.0x here)
    assert warm_iters <= 1.05 * plain_iters
    # same converged solution (unique root, same tolerance)
    for name in ("slip_rate_magnitude", "state_variable", "slip1"):
        e = _rel(drv2.fault.host(name), drv.fault.host(name))
        assert e < 1e-6, (name, e)
    assert drv2.fault.host("convergence_inner").all()
