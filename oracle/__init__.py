"""ORACLE -- test infrastructure only.

ctypes binding of the CPU FP64 restatement in ``oracle/oracle.cpp`` and
``oracle/friction_oracle.cpp`` (built by ``oracle/Makefile`` into
``oracle/_build/liboracle.so``).  Only tests, the smoke check and the
CPU-baseline leg of ``bench.py`` use it, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
_LIB = None


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = C.CDLL(str(LIB_PATH))
        P, I64, I = C.c_void_p, C.c_int64, C.c_int
        sig = {
            "oracle_godunov": [P, P, P, P],
            "oracle_cell_local_matrices": [I64, P, P, P, P, P, P, P],
            "oracle_derivatives_size": [I],
            "oracle_ader": [I, P, I64, P, P, P, P, P],
            "oracle_taylor": [I, I64, P, P, P],
            "oracle_local": [I, P, I64, P, P, P, P, P],
            "oracle_neighbor": [I, P, I64, P, P, P, P, P, P],
            "oracle_proxy_all": [I, P, I64, I, P, P, P, P, P, P, P, P, P],
            "oracle_lts": [I, P, I64, I, I, P, C.c_double, P, P, P, P, P, P, P, P],
            "oracle_dr_interpolate": [I, I64, I, P, P, P, P, P, P, P, P, P],
            "oracle_dr_nodal_flux": [I, I64, I, P, P, P, P, P, P, P, P, P],
        }
        for name, args in sig.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = I64 if name == "oracle_derivatives_size" else I
        _LIB = h
    return _LIB


def _p(a):
    """Pointer argument that holds a reference to `a`: a temporary such as
    ``_p(_c(x))`` (a contiguous copy) then lives until the call returns,
    where a bare address would point into freed memory."""
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def godunov(local, neighbor):
    ql = np.zeros(81)
    qn = np.zeros(81)
    lib().oracle_godunov(_p(_c(local)), _p(_c(neighbor)), _p(ql), _p(qn))
    return ql.reshape(9, 9).T.copy(), qn.reshape(9, 9).T.copy()  # [row][col]


def cell_local_matrices(vertices, material, neighbors, face_types=None):
    n = vertices.shape[0]
    star = np.zeros((n, 3, 81))
    ap = np.zeros((n, 4, 81))
    am = np.zeros((n, 4, 81))
    ft = None if face_types is None else _c(face_types, np.uint8)
    lib().oracle_cell_local_matrices(n, _p(_c(vertices)), _p(_c(material)), _p(_c(neighbors, np.int32)),
                                     _p(ft), _p(star), _p(ap), _p(am))
    return star, ap, am


def derivatives_size(order):
    return int(lib().oracle_derivatives_size(order))


def ader(order, gm_flat, coeffs, dofs, star, want_derivatives=False):
    n = dofs.shape[0]
    I = np.zeros_like(dofs)
    d = np.zeros((n, derivatives_size(order))) if want_derivatives else None
    lib().oracle_ader(order, _p(gm_flat), n, _p(_c(coeffs)), _p(_c(dofs)), _p(_c(star)), _p(I), _p(d))
    return (I, d) if want_derivatives else I


def taylor(order, coeffs, derivatives):
    n = derivatives.shape[0]
    B = order * (order + 1) * (order + 2) // 6
    out = np.zeros((n, 9, B))
    lib().oracle_taylor(order, n, _p(_c(coeffs)), _p(_c(derivatives)), _p(out))
    return out


def local(order, gm_flat, I, star, aplus, dofs, face_types=None):
    q = _c(dofs).copy()
    ft = None if face_types is None else _c(face_types, np.uint8)
    lib().oracle_local(order, _p(gm_flat), q.shape[0], _p(_c(I)), _p(_c(star)), _p(_c(aplus)), _p(ft), _p(q))
    return q


def neighbor(order, gm_flat, I, neighbors, face_relations, aminus, dofs, face_types=None):
    q = _c(dofs).copy()
    ft = None if face_types is None else _c(face_types, np.uint8)
    lib().oracle_neighbor(order, _p(gm_flat), q.shape[0], _p(_c(I)), _p(_c(neighbors, np.int32)),
                          _p(_c(face_relations, np.uint8)), _p(_c(aminus)), _p(ft), _p(q))
    return q


def dr_interpolate(order, points, v3m, derivatives, plus, minus, info, tinv_t, coeffs):
    """q_plus, q_minus [f][order][9][points] of the DR faces."""
    nf = len(plus)
    qp = np.zeros((nf, order, 9, points))
    qm = np.zeros_like(qp)
    lib().oracle_dr_interpolate(order, nf, points, _p(_c(v3m)), _p(_c(derivatives)), _p(_c(plus, np.int32)),
                                _p(_c(minus, np.int32)), _p(_c(info, np.uint8)), _p(_c(tinv_t)), _p(_c(coeffs)),
                                _p(qp), _p(qm))
    return qp, qm


def dr_nodal_flux(order, points, v3mtw, dr_mapping, plus, info, flux_plus, flux_minus, imposed_plus, imposed_minus,
                  dofs):
    q = _c(dofs).copy()
    lib().oracle_dr_nodal_flux(order, q.shape[0], points, _p(_c(v3mtw)), _p(_c(dr_mapping, np.int32)),
                               _p(_c(plus, np.int32)), _p(_c(info, np.uint8)), _p(_c(flux_plus)),
                               _p(_c(flux_minus)), _p(_c(imposed_plus)), _p(_c(imposed_minus)), _p(q))
    return q


def lts(order, gm_flat, macro_steps, num_clusters, cluster_ids, dt0, setups, star, aplus, aminus, neighbors,
        face_relations, dofs, face_types=None):
    """Rate-2 local time stepping of `macro_steps` steps of the slowest cluster."""
    q = _c(dofs).copy()
    ft = None if face_types is None else _c(face_types, np.uint8)
    rc = lib().oracle_lts(order, _p(gm_flat), q.shape[0], macro_steps, num_clusters, _p(_c(cluster_ids, np.int32)),
                          float(dt0), _p(_c(setups, np.uint16)), _p(_c(star)), _p(_c(aplus)), _p(_c(aminus)),
                          _p(_c(neighbors, np.int32)), _p(_c(face_relations, np.uint8)), _p(ft), _p(q))
    if rc != 0:
        raise RuntimeError("oracle_lts failed")
    return q


def proxy_all(order, gm_flat, steps, coeffs, star, aplus, aminus, neighbors, face_relations, dofs,
              face_types=None):
    q = _c(dofs).copy()
    I = np.zeros_like(q)
    ft = None if face_types is None else _c(face_types, np.uint8)
    lib().oracle_proxy_all(order, _p(gm_flat), q.shape[0], steps, _p(_c(coeffs)), _p(_c(star)), _p(_c(aplus)),
                           _p(_c(aminus)), _p(_c(neighbors, np.int32)), _p(_c(face_relations, np.uint8)), _p(ft),
                           _p(q), _p(I))
    return q, I


# ---------------------------------------------------------------------------
# friction (oracle/friction_oracle.cpp)
# ---------------------------------------------------------------------------
FAULT_FIELDS = ["imp_and_eta", "q_interpolated_plus", "q_interpolated_minus", "initial_stress", "initial_pressure",
                "nucleation_stress", "nucleation_pressure", "rs_a", "rs_sl0", "rs_f0", "rs_muw", "rs_b", "rs_srw",
                "state_variable", "slip_rate1", "slip_rate2", "slip_rate_magnitude", "mu", "traction1",
                "traction2", "slip1", "slip2", "accumulated_slip", "rupture_time", "peak_slip_rate",
                "dyn_stress_time", "rupture_time_pending", "dyn_stress_time_pending", "convergence_inner",
                "convergence_outer", "imposed_state_plus", "imposed_state_minus", "newton_iterations",
                "slip_rate_updates"]


class OracleFaultLayer(C.Structure):
    _fields_ = [("num_faces", C.c_int64), ("points", C.c_int32), ("num_points", C.c_int32), ("order", C.c_int32),
                ("pad_", C.c_int32)] + [(n, C.c_void_p) for n in FAULT_FIELDS]


class OracleFrictionParams(C.Structure):
    _fields_ = [("rs_sr0", C.c_double), ("rs_slip_rate_tolerance", C.c_double), ("rs_state_tolerance", C.c_double),
                ("rs_max_slip_rate_updates", C.c_int32), ("rs_state_variable_updates", C.c_int32),
                ("eta_damp", C.c_double), ("eta_damp_end", C.c_double), ("nucleation_count", C.c_int32),
                ("closed_form_start", C.c_int32), ("t0", C.c_double), ("s0", C.c_double)]


def _flib():
    h = lib()
    if not hasattr(h, "_friction_ready"):
        P, D, I = C.c_void_p, C.c_double, C.c_int
        h.oracle_friction_evaluate.argtypes = [C.POINTER(OracleFaultLayer), C.POINTER(OracleFrictionParams), D, P, P,
                                               P]
        h.oracle_friction_evaluate.restype = I
        for name in ("oracle_arsinhexp", "oracle_arsinhexp_derivative", "oracle_logsinh"):
            getattr(h, name).argtypes = [D, D]
            getattr(h, name).restype = D
        h.oracle_rs_initial_state.argtypes = [D, D, D, D, D, D, P, P]
        h.oracle_compute_delta_t.argtypes = [I, P, P]
        h.oracle_mlp_loss_grad.argtypes = [C.c_int64, P, P, P, P]
        h.oracle_mlp_loss_grad.restype = D
        h.oracle_precompute_stress.argtypes = [I, I, P, P, P, D, P, P, P]
        h.oracle_postcompute_imposed.argtypes = [I, I, P, P, P, P, P, P, P, P, P]
        h._friction_ready = True
    return h


class FaultState:
    """Host copies of a fault layer (dict of numpy arrays) bound to the
    oracle's struct; arrays are updated in place by `evaluate`."""

    def __init__(self, arrays: dict, order: int, points: int, num_points: int):
        self.a = {k: np.ascontiguousarray(v).copy() for k, v in arrays.items()}
        nf = self.a["rs_a"].shape[0]
        self.view = OracleFaultLayer(nf, points, num_points, order, 0,
                                     *[(self.a[n].ctypes.data if n in self.a else None) for n in FAULT_FIELDS])


def friction_evaluate(state: FaultState, params: dict, full_update_time, delta_t, time_weights, resample,
                      closed_form_start: bool = False):
    """One friction step of every face (FrictionSolver::evaluate, law 103).
    closed_form_start: creeping points start their first slip-rate solve of a
    time point at the closed-form creep root (the library's option of the
    same name); the root is unchanged up to the tolerance."""
    prm = OracleFrictionParams(params["rs_sr0"], params["rs_slip_rate_tolerance"], params["rs_state_tolerance"],
                               params["rs_max_slip_rate_updates"], params["rs_state_variable_updates"],
                               params["eta_damp"], params["eta_damp_end"], params["nucleation_count"],
                               1 if closed_form_start else 0, params["t0"], params["s0"])
    dt = _c(delta_t)
    tw = _c(time_weights)
    rs = _c(resample)
    _flib().oracle_friction_evaluate(C.byref(state.view), C.byref(prm), float(full_update_time), _p(dt), _p(tw),
                                     _p(rs))


def arsinhexp(x, c):
    return _flib().oracle_arsinhexp(float(x), float(c))


def arsinhexp_derivative(x, c):
    return _flib().oracle_arsinhexp_derivative(float(x), float(c))


def logsinh(x, c):
    return _flib().oracle_logsinh(float(x), float(c))


def rs_initial_state(t1, t2, pressure, a, sr0, v_ini):
    s, m = C.c_double(), C.c_double()
    _flib().oracle_rs_initial_state(t1, t2, pressure, a, sr0, v_ini, C.byref(s), C.byref(m))
    return s.value, m.value


def compute_delta_t(time_points):
    tp = _c(time_points)
    out = np.zeros_like(tp)
    _flib().oracle_compute_delta_t(len(tp), _p(tp), _p(out))
    return out


def mlp_loss_grad(x, t, params):
    x = _c(x)
    t = _c(t)
    params = _c(params)
    g = np.zeros_like(params)
    loss = _flib().oracle_mlp_loss_grad(x.shape[0], _p(x), _p(t), _p(params), _p(g))
    return loss, g


def precompute_stress(order, points, imp, q_plus, q_minus, eta_p_damp=1.0):
    n, t1, t2 = (np.zeros((order, points)) for _ in range(3))
    _flib().oracle_precompute_stress(order, points, _p(_c(imp)), _p(_c(q_plus)), _p(_c(q_minus)), eta_p_damp,
                                     _p(n), _p(t1), _p(t2))
    return n, t1, t2


def postcompute_imposed(order, points, imp, normal, t1, t2, q_plus, q_minus, weights):
    plus, minus = np.zeros((9, points)), np.zeros((9, points))
    _flib().oracle_postcompute_imposed(order, points, _p(_c(imp)), _p(_c(normal)), _p(_c(t1)), _p(_c(t2)),
                                       _p(_c(q_plus)), _p(_c(q_minus)), _p(_c(weights)), _p(plus), _p(minus))
    return plus, minus
