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
        }
        for name, args in sig.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = I64 if name == "oracle_derivatives_size" else I
        _LIB = h
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data


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


def proxy_all(order, gm_flat, steps, coeffs, star, aplus, aminus, neighbors, face_relations, dofs,
              face_types=None):
    q = _c(dofs).copy()
    I = np.zeros_like(q)
    ft = None if face_types is None else _c(face_types, np.uint8)
    lib().oracle_proxy_all(order, _p(gm_flat), q.shape[0], steps, _p(_c(coeffs)), _p(_c(star)), _p(_c(aplus)),
                           _p(_c(aminus)), _p(_c(neighbors, np.int32)), _p(_c(face_relations, np.uint8)), _p(ft),
                           _p(q), _p(I))
    return q, I
