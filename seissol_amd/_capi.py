"""ctypes binding of the C ABI declared in ``include/seissol_amd.h``.

The library is loaded from ``seissol_amd/_lib/libseissol_amd.so`` (built by
``seissol_amd.build``).  There is no fallback: if the library is missing,
``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libseissol_amd.so"

SDG_OK = 0
SDG_FP64 = 0
SDG_FP32 = 1
SDG_BF16X3 = 2
SDG_BF16X6 = 3
SDG_F16X3 = 4
SDG_LTS_HAS_BUFFERS = 1 << 8
SDG_LTS_HAS_DERIVATIVES = 1 << 9
SDG_LTS_ACCUMULATE_BUFFERS = 1 << 10
SDG_KERNELS_AUTO = 0
SDG_KERNELS_VALU = 1
SDG_STAR_DENSE = 0
SDG_STAR_COMPACT = 1
SDG_FACE_REGULAR = 0
SDG_FACE_DYNAMIC_RUPTURE = 1

P = C.c_void_p
I64 = C.c_int64
U64 = C.c_uint64
I32 = C.c_int
D = C.c_double
DP = C.POINTER(C.c_double)


class Layer(C.Structure):
    """``sdg_layer``: device pointers of one layer of cells."""

    _fields_ = [
        ("num_elements", I64),
        ("dofs", P),
        ("time_integrated", P),
        ("derivatives", P),
        ("star", P),
        ("aplus", P),
        ("aminus", P),
        ("face_neighbors", P),
        ("face_relations", P),
        ("face_types", P),
        ("face_traces", P),
        ("lts_setup", P),
        ("star_stride", I64),
        ("aplus_stride", I64),
        ("aminus_stride", I64),
    ]


class FrictionParams(C.Structure):
    """``sdg_friction_params``"""

    _fields_ = [("rs_sr0", D), ("rs_slip_rate_tolerance", D), ("rs_state_tolerance", D),
                ("rs_max_slip_rate_updates", C.c_int32), ("rs_state_variable_updates", C.c_int32),
                ("eta_damp", D), ("eta_damp_end", D), ("nucleation_count", C.c_int32), ("pad_", C.c_int32),
                ("t0", D), ("s0", D)]


FAULT_FIELDS = ["imp_and_eta", "q_interpolated_plus", "q_interpolated_minus", "initial_stress", "initial_pressure",
                "nucleation_stress", "nucleation_pressure", "rs_a", "rs_sl0", "rs_f0", "rs_muw", "rs_b", "rs_srw",
                "state_variable", "slip_rate1", "slip_rate2", "slip_rate_magnitude", "mu", "traction1",
                "traction2", "slip1", "slip2", "accumulated_slip", "rupture_time", "peak_slip_rate",
                "dyn_stress_time", "rupture_time_pending", "dyn_stress_time_pending", "convergence_inner",
                "convergence_outer", "imposed_state_plus", "imposed_state_minus", "newton_iterations",
                "slip_rate_updates"]


class FaultLayer(C.Structure):
    """``sdg_fault_layer`` (same layout as the oracle's fault layer)."""

    _fields_ = [("num_faces", I64), ("points", C.c_int32), ("num_points", C.c_int32), ("order", C.c_int32),
                ("pad_", C.c_int32)] + [(name, P) for name in FAULT_FIELDS]


class DrLayer(C.Structure):
    """``sdg_dr_layer``"""

    _fields_ = [("num_faces", I64), ("points", C.c_int32), ("pad_", C.c_int32), ("plus_element", P),
                ("minus_element", P), ("face_info", P), ("tinv_t", P), ("flux_solver_plus", P),
                ("flux_solver_minus", P)]


class ProxyDr(C.Structure):
    """``sdg_proxy_dr``"""

    _fields_ = [("dr_mapping", P), ("faces", DrLayer), ("imposed_plus", P), ("imposed_minus", P), ("q_plus", P),
                ("q_minus", P)]


class WorkEstimate(C.Structure):
    """``sdg_work_estimate``"""

    _fields_ = [("nonzero_flop", C.c_double), ("hardware_flop", C.c_double), ("bytes", C.c_double),
                ("kernel_bytes", C.c_double)]


class FrictionOptions(C.Structure):
    """``sdg_friction_options``"""

    _fields_ = [("mlp_params", P), ("samples", P), ("capacity", I64), ("sample_count", P),
                ("sample_stride", C.c_int32), ("closed_form_start", C.c_int32)]


# name -> (restype, argtypes)
SIGNATURES = {
    "sdg_last_error": (C.c_char_p, []),
    "sdg_version": (I32, []),
    "sdg_device_count": (I32, []),
    "sdg_global_matrices_size": (C.c_size_t, [I32]),
    "sdg_global_matrices": (I32, [I32, P]),
    "sdg_basis_values": (I32, [I32, D, D, D, P]),
    "sdg_basis_gradients": (I32, [I32, D, D, D, P]),
    "sdg_mesh_proxy": (I32, [I64, U64, P, P, P]),
    "sdg_mesh_periodic_size": (I64, [I32, I32, I32]),
    "sdg_mesh_periodic": (I32, [I32, I32, I32, D, D, U64, P, P, P]),
    "sdg_random_materials": (I32, [I64, U64, D, D, D, D, P]),
    "sdg_cfl_timestep": (D, [I64, I32, D, P, P]),
    "sdg_uniform_dofs": (I32, [I64, U64, P]),
    "sdg_monomial_integrate": (I32, [I32, D, D, P]),
    "sdg_cell_local_matrices": (I32, [I64, P, P, P, P, P, P, P, P]),
    "sdg_engine_create": (I32, [C.POINTER(P), I32, I32]),
    "sdg_engine_set_global_matrices": (I32, [P, P]),
    "sdg_engine_destroy": (I32, [P]),
    "sdg_derivatives_size": (I64, [I32]),
    "sdg_engine_set_layout": (I32, [P, I32, I32]),
    "sdg_layout_sizes": (I32, [P, P, P, P]),
    "sdg_compute_batched_ader": (I32, [P, P, D, C.POINTER(Layer), P]),
    "sdg_evaluate_batched": (I32, [P, P, P, P, I64, P]),
    "sdg_compute_batched_integral": (I32, [P, C.POINTER(Layer), P]),
    "sdg_compute_batched_neighbors_integral": (I32, [P, C.POINTER(Layer), P]),
    "sdg_compute_batched_ader_local": (I32, [P, P, D, C.POINTER(Layer), P]),
    "sdg_proxy_kernel_id": (I32, [C.c_char_p]),
    "sdg_engine_set_kernels": (I32, [P, I32]),
    "sdg_proxy_run": (I32, [P, I32, I32, D, C.POINTER(Layer), P]),
    "sdg_proxy_run_static": (I32, [P, I32, I32, D, C.POINTER(Layer), P]),
    "sdg_proxy_dr_faces": (I32, [I64, C.c_uint64, P, P, P, P]),
    "sdg_proxy_run_dr": (I32, [P, I32, I32, D, C.POINTER(Layer), C.POINTER(ProxyDr), P]),
    "sdg_upload_reals": (I32, [P, P, P, I64, P]),
    "sdg_download_reals": (I32, [P, P, P, I64, P]),
    "sdg_convert_fp64_to_real": (I32, [P, P, P, I64, P]),
    "sdg_friction_evaluate": (I32, [C.POINTER(FrictionParams), C.POINTER(FaultLayer), D, P, P, P,
                                    C.POINTER(FrictionOptions), P]),
    "sdg_friction_delta_t": (I32, [I32, P, P]),
    "sdg_gauss_legendre": (I32, [I32, D, D, P, P]),
    "sdg_dr_quadrature": (I32, [I32, P, P, P]),
    "sdg_rs_initial_state": (I32, [D, D, D, D, D, D, P, P]),
    "sdg_mlp_param_count": (I32, []),
    "sdg_mlp_loss_grad": (I32, [P, P, I64, P, P, P]),
    "sdg_mlp_adam": (I32, [P, P, P, P, I32, C.c_float, P]),
    "sdg_mlp_forward": (I32, [P, P, I64, I32, P, P]),
    "sdg_performance_estimate": (I32, [I32, I32, I32, C.POINTER(WorkEstimate)]),
    "sdg_derive_lts_setups": (I32, [I64, P, P, P, P]),
    "sdg_compute_local_integration": (I32, [P, P, D, C.POINTER(Layer), I32, P]),
    "sdg_compute_neighboring_integration": (I32, [P, P, P, C.POINTER(Layer), P]),
    "sdg_lts_create": (I32, [P, P, C.POINTER(Layer), I32, P, D]),
    "sdg_dr_face_matrices": (I32, [I32, P, P]),
    "sdg_dr_face_data": (I32, [I64, P, P, P, P, P, P, P, P, P]),
    "sdg_dr_space_time_interpolation": (I32, [P, P, P, C.POINTER(DrLayer), P, P, P]),
    "sdg_dr_nodal_flux": (I32, [P, C.POINTER(Layer), P, C.POINTER(DrLayer), P, P, P]),
    "sdg_lts_advance": (I32, [P, I32, P]),
    "sdg_lts_graph_advance": (I32, [P, I32, P]),
    "sdg_lts_destroy": (I32, [P]),
    "sdg_lts_autotune": (I32, [P, I32, P]),
    "sdg_lts_tile_variant": (I32, [P, I32, I32]),
    "sdg_lts_set_tile_variant": (I32, [P, I32, I32, I32]),
}

_LIB = None


def lib() -> C.CDLL:
    """Load the library once and attach the signatures."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m seissol_amd.build` "
                "(there is no CPU fallback)")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


class SdgError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> None:
    if rc != SDG_OK:
        msg = lib().sdg_last_error().decode(errors="replace")
        raise SdgError(f"{what} failed ({rc}): {msg}")


def ptr(x) -> int | None:
    """Address of a numpy array or torch tensor (None for None)."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x.ctypes.data
