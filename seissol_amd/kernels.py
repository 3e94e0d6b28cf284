"""Python mirror of the reference kernel call surfaces on top of the C ABI.

Class and method names follow ``seissol::kernels`` (src/Kernels/Spacetime.h:29-70,
Time.h:27-62, Local.h:25-80, Neighbor.h:22-56) for the batched (device) path:

* ``Spacetime.computeBatchedAder``      -> ``sdg_compute_batched_ader``
* ``Time.evaluateBatched``              -> ``sdg_evaluate_batched``
* ``Local.computeBatchedIntegral``      -> ``sdg_compute_batched_integral``
* ``Neighbor.computeBatchedNeighborsIntegral`` -> ``sdg_compute_batched_neighbors_integral``

Device memory comes from PyTorch (plumbing only); all arithmetic runs in the
HIP library.  Errors raise ``SdgError`` with the library's message.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import Layer, check, lib, ptr

ARITHMETIC = {"fp64": _capi.SDG_FP64, "fp32": _capi.SDG_FP32, "bf16x3": _capi.SDG_BF16X3,
              "bf16x6": _capi.SDG_BF16X6, "f16x3": _capi.SDG_F16X3}
MATRIX_CORE = ("bf16x3", "bf16x6", "f16x3")
STAR_FORMAT = {"dense": _capi.SDG_STAR_DENSE, "compact": _capi.SDG_STAR_COMPACT}
# column-major positions q + 9 p of the 24 non-zeros of an elastic star
# matrix, column by column with rows ascending (codegen/matrices/star.xml):
# the order of the compact storage
STAR_NONZEROS = np.array([6, 7, 8, 15, 16, 17, 24, 25, 26, 33, 34, 43, 44, 51, 53, 54, 57, 59, 64, 66, 67, 74, 76,
                          77])


def num_basis(order: int) -> int:
    return order * (order + 1) * (order + 2) // 6


def num_face_basis(order: int) -> int:
    return order * (order + 1) // 2


def padded_face_basis(order: int) -> int:
    """Face rows of the matrix-core face traces: B2 rounded up to a multiple of 4."""
    return (num_face_basis(order) + 3) // 4 * 4


def derivatives_size(order: int) -> int:
    return int(lib().sdg_derivatives_size(order))


def global_matrices(order: int) -> dict:
    """Generated global matrices (FP64, row-major) keyed by family."""
    size = lib().sdg_global_matrices_size(order)
    buf = np.zeros(size, dtype=np.float64)
    check(lib().sdg_global_matrices(order, buf.ctypes.data), "sdg_global_matrices")
    B, B2 = num_basis(order), num_face_basis(order)
    out, off = {}, 0
    for name, cnt, r, c in (("kDivM", 3, B, B), ("kDivMT", 3, B, B), ("rDivM", 4, B, B2),
                            ("fMrT", 4, B2, B), ("rT", 4, B2, B), ("fP", 3, B2, B2)):
        out[name] = buf[off:off + cnt * r * c].reshape(cnt, r, c)
        off += cnt * r * c
    out["mass"] = buf[off:off + B]
    off += B
    out["mass2"] = buf[off:off + B2]
    out["flat"] = buf
    return out


def monomial_integrate(order: int, t0: float, t1: float) -> np.ndarray:
    c = np.zeros(order)
    check(lib().sdg_monomial_integrate(order, t0, t1, c.ctypes.data), "sdg_monomial_integrate")
    return c


def _stream(stream) -> int | None:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream


class Engine:
    """Global matrices resident on the device for one order and arithmetic
    (the role of ``setGlobalData`` in the reference kernels)."""

    def __init__(self, order: int, arithmetic: str = "fp64"):
        self.order = order
        self.arithmetic = arithmetic
        handle = C.c_void_p()
        check(lib().sdg_engine_create(C.byref(handle), order, ARITHMETIC[arithmetic]), "sdg_engine_create")
        self.handle = handle

    def set_kernels(self, family: str = "auto"):
        """Kernel family (``sdg_engine_set_kernels``): "auto" runs the matrix-core
        passes where the engine has them, "valu" the FP64/FP32 VALU kernels only."""
        check(lib().sdg_engine_set_kernels(self.handle, {"auto": 0, "valu": 1}[family]), "sdg_engine_set_kernels")
        self.kernels = family

    def set_layout(self, alignment: int = 0, star_format: str = "dense"):
        """Storage layout of the cells' arrays (``sdg_engine_set_layout``)."""
        check(lib().sdg_engine_set_layout(self.handle, int(alignment), STAR_FORMAT[star_format]),
              "sdg_engine_set_layout")
        self.alignment, self.star_format = int(alignment), star_format

    def layout_sizes(self) -> tuple:
        """(ld of a DOF-like block, reals per cell of it, reals per cell of the derivatives)."""
        ld, q, d = C.c_int32(), C.c_int64(), C.c_int64()
        check(lib().sdg_layout_sizes(self.handle, C.byref(ld), C.byref(q), C.byref(d)), "sdg_layout_sizes")
        return ld.value, q.value, d.value

    def derivative_blocks(self) -> list:
        """(offset, ld, rows) of each derivative block under the engine's layout."""
        out, off = [], 0
        a = max(1, getattr(self, "alignment", 0) // (8 if self.arithmetic == "fp64" else 4))
        for i in range(self.order):
            m = num_basis(self.order - i)
            ld = (m + a - 1) // a * a
            out.append((off, ld, m))
            off += 9 * ld
        return out

    @property
    def real_dtype(self):
        import torch
        return torch.float64 if self.arithmetic == "fp64" else torch.float32

    def close(self):
        if self.handle:
            check(lib().sdg_engine_destroy(self.handle), "sdg_engine_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Spacetime:
    def __init__(self, engine: Engine):
        self.engine = engine

    def computeBatchedAder(self, coeffs, timeStepWidth: float, layer: Layer, stream=None):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        check(lib().sdg_compute_batched_ader(self.engine.handle, c.ctypes.data, float(timeStepWidth),
                                              C.byref(layer), _stream(stream)), "computeBatchedAder")

    def computeBatchedAderLocal(self, coeffs, timeStepWidth: float, layer: Layer, stream=None):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        check(lib().sdg_compute_batched_ader_local(self.engine.handle, c.ctypes.data, float(timeStepWidth),
                                                    C.byref(layer), _stream(stream)), "computeBatchedAderLocal")


class Time:
    def __init__(self, engine: Engine):
        self.engine = engine

    def evaluateBatched(self, coeffs, timeDerivatives, timeEvaluated, numElements: int, stream=None):
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        check(lib().sdg_evaluate_batched(self.engine.handle, c.ctypes.data, ptr(timeDerivatives),
                                          ptr(timeEvaluated), int(numElements), _stream(stream)),
              "evaluateBatched")


class Local:
    def __init__(self, engine: Engine):
        self.engine = engine

    def computeBatchedIntegral(self, layer: Layer, stream=None):
        check(lib().sdg_compute_batched_integral(self.engine.handle, C.byref(layer), _stream(stream)),
              "computeBatchedIntegral")


class Neighbor:
    def __init__(self, engine: Engine):
        self.engine = engine

    def computeBatchedNeighborsIntegral(self, layer: Layer, stream=None):
        check(lib().sdg_compute_batched_neighbors_integral(self.engine.handle, C.byref(layer), _stream(stream)),
              "computeBatchedNeighborsIntegral")
