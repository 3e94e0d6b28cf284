"""Pins of the friction oracle: the reference's known-answer case, closed
forms of the overflow-safe functions, the steady-state invariant and the
MLP gradient."""
import json
import math
from pathlib import Path

import mpmath
import numpy as np
import pytest

import oracle
from seissol_amd.friction import (FrictionParameters, compute_delta_t, dr_quadrature, gauss_legendre,
                                  make_fault_setup)

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module", autouse=True)
def _build():
    oracle.build()


def test_precompute_and_postcompute_known_answer():
    g = json.loads((GOLDEN / "friction_common.json").read_text())
    O, P, eps = g["order"], g["points"], g["epsilon"]
    qP, qM = np.array(g["q_plus"]), np.array(g["q_minus"])
    n, t1, t2 = oracle.precompute_stress(O, P, g["imp_and_eta"], qP, qM)
    for got, key in ((n, "expected_normal"), (t1, "expected_traction1"), (t2, "expected_traction2")):
        exp = np.array(g[key])
        assert np.all(np.abs(got - exp) <= eps * np.maximum(1.0, np.abs(exp))), key
    plus, minus = oracle.postcompute_imposed(O, P, g["imp_and_eta"], np.zeros((O, P)), np.array(g["traction1"]),
                                             np.array(g["traction2"]), qP, qM, g["time_weights"])
    for got, key in ((plus, "expected_imposed_plus"), (minus, "expected_imposed_minus")):
        exp = np.array(g[key])
        assert np.all(np.abs(got - exp) <= eps * np.maximum(1.0, np.abs(exp))), key


@pytest.mark.parametrize("c", [-60.0, -5.0, 0.0, 3.0, 9.9, 10.1, 40.0, 49.9, 50.0, 120.0, 300.0])
def test_arsinhexp_closed_form(c):
    mpmath.mp.dps = 50
    for x in (1e-30, 1e-16, 1e-8, 1e-3, 0.5, 1.0, 7.0, 1e3, 1e6):
        exact = float(mpmath.asinh(mpmath.mpf(x) * mpmath.e ** mpmath.mpf(c)))
        got = oracle.arsinhexp(x, c)
        # the large-c form c + log(x + sqrt(x^2 + e^-2c)) cancels for x e^c << 1:
        # accurate to a few ulp of max(result, c) (RateAndStateCommon.h:46-79)
        assert abs(got - exact) <= 4e-15 * max(abs(exact), abs(c), 1.0), (x, c, got, exact)
        # derivative d/dx asinh(x e^c) = e^c / sqrt(1 + x^2 e^2c)  (the oracle returns it without the e^c in x)
        ex = mpmath.e ** mpmath.mpf(c)
        dexact = float(ex / mpmath.sqrt(1 + (mpmath.mpf(x) * ex) ** 2))
        dgot = oracle.arsinhexp_derivative(x, c) * float(ex) if False else oracle.arsinhexp_derivative(x, c)
        # the oracle's derivative is d/dx asinh(x e^c) expressed per unit x
        assert abs(dgot - dexact) <= 1e-13 * abs(dexact), (x, c, dgot, dexact)


def test_logsinh_closed_form():
    mpmath.mp.dps = 50
    for x in (1e-10, 2.0, 2e10, 2e22):
        for c in (-30.0, -1.0, 1e-6, 0.7, 25.0, 400.0):
            exact = float(mpmath.log(mpmath.mpf(x) * mpmath.sinh(mpmath.mpf(c)))) if c > 0 else None
            if exact is None:
                continue
            assert abs(oracle.logsinh(x, c) - exact) <= 1e-13 * max(1.0, abs(exact)), (x, c)


def test_initial_state_is_steady_sliding():
    from seissol_amd import _capi
    import ctypes as C
    for a in (0.008, 0.013, 0.02):
        for tau in (60e6, 72e6, 80e6):
            s, m = oracle.rs_initial_state(tau, 0.0, -120e6, a, 1e-6, 1e-16)
            # friction coefficient equals tau / |sigma_n| at V_ini
            assert abs(m * 120e6 - tau) <= 1e-12 * tau
            s2, m2 = C.c_double(), C.c_double()
            _capi.lib().sdg_rs_initial_state(tau, 0.0, -120e6, a, 1e-6, 1e-16, C.byref(s2), C.byref(m2))
            assert abs(s2.value - s) <= 1e-13 * abs(s) and abs(m2.value - m) <= 1e-13 * abs(m)


def test_delta_t_and_time_points():
    tp, tw = gauss_legendre(4, 0.0, 2e-3)
    assert abs(tw.sum() - 2e-3) < 1e-18
    d = compute_delta_t(tp)
    np.testing.assert_allclose(d, oracle.compute_delta_t(tp), rtol=0, atol=0)
    # the sub-steps cover the whole step: sum = dt
    assert abs(d.sum() - 2e-3) < 1e-18


def test_dr_quadrature_matches_reference_data():
    gold = json.loads((GOLDEN / "dr_stroud.json").read_text())
    for order in (4, 5, 6):
        pts, w, R = dr_quadrature(order)
        for name, a in (("quadpoints", pts), ("quadweights", w.reshape(-1, 1)), ("resample", R)):
            c = gold[str(order)][name]
            assert abs(a.sum() - c["sum"]) < 1e-12 * max(1, c["abs_sum"])
            assert abs(np.abs(a).sum() - c["abs_sum"]) < 1e-11 * max(1, c["abs_sum"])
            assert abs(np.sqrt((a * a).sum()) - c["frobenius"]) < 1e-12 * max(1, c["frobenius"])
    d = gold["4"]["data"]
    pts, w, R = dr_quadrature(4)
    np.testing.assert_allclose(R, np.array(d["resample"]), atol=5e-14)
    # resampling is a projection: R R = R
    np.testing.assert_allclose(R @ R, R, atol=1e-13)


def _params_dict(p: FrictionParameters):
    return dict(rs_sr0=p.rs_sr0, rs_slip_rate_tolerance=p.rs_slip_rate_tolerance,
                rs_state_tolerance=p.rs_state_tolerance, rs_max_slip_rate_updates=p.rs_max_slip_rate_updates,
                rs_state_variable_updates=p.rs_state_variable_updates, eta_damp=p.eta_damp,
                eta_damp_end=p.eta_damp_end, nucleation_count=p.nucleation_count, t0=p.t0, s0=p.s0)


def test_steady_sliding_stays_steady():
    # tau = |sigma_n| mu(V, psi_ss(V)) + eta_s V  ->  g(V) = 0 and psi = psi_ss(V)
    prm = FrictionParameters(nucleation_count=0)
    setup = make_fault_setup(num_faces=4, order=4, seed=1, params=prm)
    A = setup.arrays
    V = 0.37
    etaS = A["imp_and_eta"][0, 5]
    a = A["rs_a"]
    low = np.maximum(0, A["rs_f0"] - (A["rs_b"] - a) * np.log(V / prm.rs_sr0))
    mu_ss = A["rs_muw"] + (low - A["rs_muw"]) / (1 + (V / A["rs_srw"]) ** 8) ** 0.125
    psi = a * np.log(2 * prm.rs_sr0 / V * np.sinh(mu_ss / a))
    A["state_variable"][:] = psi
    A["slip_rate1"][:] = V
    A["initial_stress"][:, 3, :] = 120e6 * mu_ss + etaS * V
    st = oracle.FaultState(A, 4, setup.points, setup.num_points)
    _, _, R = dr_quadrature(4)
    tp, tw = gauss_legendre(4, 0.0, 1e-3)
    oracle.friction_evaluate(st, _params_dict(prm), 0.0, compute_delta_t(tp), tw, R)
    np.testing.assert_allclose(st.a["slip_rate_magnitude"], V, rtol=1e-7)
    np.testing.assert_allclose(st.a["state_variable"], psi, rtol=1e-9)
    assert st.a["convergence_inner"].all() and st.a["convergence_outer"].all()


def test_nucleation_starts_slip_and_converges():
    prm = FrictionParameters(t0=0.02)
    setup = make_fault_setup(num_faces=200, order=4, seed=3, params=prm, length=4000.0, nucleation_radius=800.0)
    st = oracle.FaultState(setup.arrays, 4, setup.points, setup.num_points)
    _, _, R = dr_quadrature(4)
    tp, tw = gauss_legendre(4, 0.0, 1e-3)
    d = compute_delta_t(tp)
    for s in range(30):
        oracle.friction_evaluate(st, _params_dict(prm), s * 1e-3, d, tw, R)
    V = st.a["slip_rate_magnitude"]
    assert V.max() > 1e-3  # the patch ruptured
    assert np.isfinite(V).all() and np.isfinite(st.a["state_variable"]).all()
    assert st.a["convergence_outer"].mean() > 0.9


def _drive(setup, steps, dt, closed_form_start):
    from seissol_amd.friction import pulse_loading, step_times
    st = oracle.FaultState(setup.arrays, setup.order, setup.points, setup.num_points)
    _, _, R = dr_quadrature(setup.order)
    tp, tw, delta = step_times(setup.order, dt)
    for s in range(steps):
        load = pulse_loading(setup, s * dt + tp)
        st.a["q_interpolated_plus"][:, :, 3, :] = load
        st.a["q_interpolated_minus"][:, :, 3, :] = load
        oracle.friction_evaluate(st, _params_dict(setup.params), s * dt, delta, tw, R,
                                 closed_form_start=closed_form_start)
    return st.a


def test_closed_form_creep_start_keeps_the_root_and_halves_newton_steps():
    # The creep root V_c = 2 V0 exp(-psi/a) sinh(S/N) solves g(V) + V = 0, so a
    # creeping point (V_c below half the tolerance) passes its first residual
    # test there.  A wrong V_c would fail that test and the counts would not
    # drop; a converged root is unique (g strictly decreasing), so the state
    # must stay the plain one up to the tolerance.
    setup = make_fault_setup(num_faces=300, order=4, seed=11, length=4000.0, nucleation_radius=900.0)
    plain = _drive(setup, 25, 1e-3, False)
    warm = _drive(setup, 25, 1e-3, True)
    np_ = setup.num_points
    assert plain["peak_slip_rate"].max() > 0.01, "scenario must rupture"
    for name in ("slip_rate_magnitude", "state_variable", "slip1", "traction1", "imposed_state_plus"):
        b = plain[name]
        assert np.linalg.norm(warm[name] - b) <= 1e-12 * np.linalg.norm(b), name
    assert warm["convergence_inner"].all() and warm["convergence_outer"][:, :np_].all() == \
        plain["convergence_outer"][:, :np_].all()
    steps_p, steps_w = plain["slip_rate_updates"][:, :np_].sum(), warm["slip_rate_updates"][:, :np_].sum()
    ev_p, ev_w = plain["newton_iterations"][:, :np_].sum(), warm["newton_iterations"][:, :np_].sum()
    # measured on this fault: steps 2,235,750 -> 896,400 (2.49x), evaluations 3,474,700 -> 2,135,350 (1.63x)
    assert steps_p / steps_w >= 2.4, (steps_p, steps_w)
    assert ev_p / ev_w >= 1.55, (ev_p, ev_w)
    # every plain solve evaluates at least once per Newton step plus the passing test
    assert (plain["newton_iterations"] >= plain["slip_rate_updates"]).all()


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n = 16
    x = rng.normal(size=(n, 6))
    t = rng.normal(size=n)
    npar = 32 * 6 + 32 + 32 * 32 + 32 + 32 + 1
    params = rng.normal(scale=0.3, size=npar)
    loss, g = oracle.mlp_loss_grad(x, t, params)
    h = 1e-6
    for idx in rng.choice(npar, 40, replace=False):
        p1, p2 = params.copy(), params.copy()
        p1[idx] += h
        p2[idx] -= h
        fd = (oracle.mlp_loss_grad(x, t, p1)[0] - oracle.mlp_loss_grad(x, t, p2)[0]) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)), idx



def test_dr_face_matrices_match_reference_data():
    # V3mTo2n(i, h): the basis at the Stroud points of face i, h = 0 plus side,
    # h = orientation + 1 minus side; V3mTo2nTWDivM(i, h) = V^T diag(w) / M
    from seissol_amd._capi import check, lib
    gold = json.loads((GOLDEN / "dr_stroud.json").read_text())
    for order in (4, 5, 6):
        n, B = (order + 1) ** 2, order * (order + 1) * (order + 2) // 6
        v = np.zeros((4, 4, n, B))
        vt = np.zeros((4, 4, B, n))
        check(lib().sdg_dr_face_matrices(order, v.ctypes.data, vt.ctypes.data), "sdg_dr_face_matrices")
        for i in range(4):
            for h in range(4):
                for name, a in ((f"V3mTo2n({i},{h})", v[i, h]), (f"V3mTo2nTWDivM({i},{h})", vt[i, h])):
                    c = gold[str(order)][name]
                    scale = max(1.0, c["abs_sum"])
                    assert abs(a.sum() - c["sum"]) < 1e-10 * scale, (order, name)
                    assert abs(np.abs(a).sum() - c["abs_sum"]) < 1e-10 * scale, (order, name)
                    assert abs(float((a * np.arange(1, a.size + 1).reshape(a.shape)).sum()) - c["weighted"]) \
                        < 1e-9 * scale * a.size, (order, name)
