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
    ]


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
    "sdg_compute_batched_ader": (I32, [P, P, D, C.POINTER(Layer), P]),
    "sdg_evaluate_batched": (I32, [P, P, P, P, I64, P]),
    "sdg_compute_batched_integral": (I32, [P, C.POINTER(Layer), P]),
    "sdg_compute_batched_neighbors_integral": (I32, [P, C.POINTER(Layer), P]),
    "sdg_compute_batched_ader_local": (I32, [P, P, D, C.POINTER(Layer), P]),
    "sdg_proxy_kernel_id": (I32, [C.c_char_p]),
    "sdg_proxy_run": (I32, [P, I32, I32, D, C.POINTER(Layer), P]),
    "sdg_upload_reals": (I32, [P, P, P, I64, P]),
    "sdg_download_reals": (I32, [P, P, P, I64, P]),
    "sdg_convert_fp64_to_real": (I32, [P, P, P, I64, P]),
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
