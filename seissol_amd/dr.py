"""Dynamic-rupture faces on the C ABI: the faces of a plane of a periodic
tetrahedral box turned into a fault, their per-face data
(``sdg_dr_face_data``), the space-time interpolation feeding the friction law
and the nodal flux of its imposed states (DynamicRupture.cpp, Neighbor.cpp)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import DrLayer, check, lib, ptr
from .friction import gauss_legendre
from .proxy import Setup

FACE_NODES = ((0, 2, 1), (0, 1, 3), (0, 3, 2), (1, 2, 3))  # MeshTools.cpp:20


@dataclass
class DrFaces:
    plus_element: np.ndarray   # [f] int32
    minus_element: np.ndarray  # [f] int32
    face_info: np.ndarray      # [f][3] uint8: plus side, minus side, orientation + 1
    dr_mapping: np.ndarray     # [n][4] int32, -1 for non-DR faces
    tinv_t: np.ndarray         # [f][81]
    flux_plus: np.ndarray      # [f][81]
    flux_minus: np.ndarray     # [f][81]
    imp_and_eta: np.ndarray    # [f][11]
    points: int

    @property
    def num_faces(self) -> int:
        return len(self.plus_element)


def face_matrices(order: int):
    n, B = (order + 1) ** 2, order * (order + 1) * (order + 2) // 6
    v = np.zeros((4, 4, n, B))
    vt = np.zeros((4, 4, B, n))
    check(lib().sdg_dr_face_matrices(order, v.ctypes.data, vt.ctypes.data), "sdg_dr_face_matrices")
    return v, vt


def plane_faces(setup: Setup, plane: int, h: float = 100.0, points: int | None = None) -> DrFaces:
    """The faces of the grid plane x-index `plane` of a (jittered) periodic box
    with cube size h become DR faces; the plus side is the cell on the -x side.
    Marks them in setup.face_types (both sides)."""
    order = setup.order
    plus, minus, info = [], [], []
    for e in range(setup.n):
        cx = setup.vertices[e, :, 0].mean()
        if cx >= plane * h:
            continue
        for f in range(4):
            xs = setup.vertices[e, list(FACE_NODES[f]), 0]
            if np.all(np.rint(xs / h) == plane):
                nb = int(setup.neighbors[e, f])
                side, orient = int(setup.face_relations[e, f, 0]), int(setup.face_relations[e, f, 1])
                plus.append(e)
                minus.append(nb)
                info.append((f, side, orient + 1))
    plus = np.array(plus, np.int32)
    minus = np.array(minus, np.int32)
    info = np.array(info, np.uint8).reshape(-1, 3)
    mapping = np.full((setup.n, 4), -1, np.int32)
    for m in range(len(plus)):
        mapping[plus[m], info[m, 0]] = m
        mapping[minus[m], info[m, 1]] = m
        setup.face_types[plus[m], info[m, 0]] = _capi.SDG_FACE_DYNAMIC_RUPTURE
        setup.face_types[minus[m], info[m, 1]] = _capi.SDG_FACE_DYNAMIC_RUPTURE
    nf = len(plus)
    tinv = np.zeros((nf, 81))
    fp = np.zeros((nf, 81))
    fm = np.zeros((nf, 81))
    ie = np.zeros((nf, 11))
    v = np.ascontiguousarray(setup.vertices)
    mat = np.ascontiguousarray(setup.material)
    check(lib().sdg_dr_face_data(nf, v.ctypes.data, mat.ctypes.data, plus.ctypes.data, minus.ctypes.data,
                                 info.ctypes.data, tinv.ctypes.data, fp.ctypes.data, fm.ctypes.data, ie.ctypes.data),
          "sdg_dr_face_data")
    pts = points or (order + 1) ** 2
    return DrFaces(plus, minus, info, mapping, tinv, fp, fm, ie, pts)


def collocation_coefficients(order: int, dt: float) -> np.ndarray:
    """c[o][d] = t_o^d / d! at the Gauss-Legendre points t_o of [0, dt]
    (timeBasis().collocate, KernelDevice.cpp:115-117)."""
    tp, _ = gauss_legendre(order, 0.0, dt)
    c = np.zeros((order, order))
    for o, t in enumerate(tp):
        fact = 1.0
        for d in range(order):
            if d > 0:
                fact *= d
            c[o, d] = t ** d / fact
    return c


class DeviceDrFaces:
    """Device copies of the DR face data and the ``sdg_dr_layer`` view."""

    def __init__(self, faces: DrFaces, device: str = "cuda"):
        import torch
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        self.faces = faces
        self.plus = t(faces.plus_element)
        self.minus = t(faces.minus_element)
        self.info = t(faces.face_info)
        self.mapping = t(faces.dr_mapping)
        self.tinv_t = t(faces.tinv_t)
        self.flux_plus = t(faces.flux_plus)
        self.flux_minus = t(faces.flux_minus)
        self.view = DrLayer(faces.num_faces, faces.points, 0, ptr(self.plus), ptr(self.minus), ptr(self.info),
                            ptr(self.tinv_t), ptr(self.flux_plus), ptr(self.flux_minus))

    def interpolate(self, engine, coeffs: np.ndarray, derivatives, q_plus, q_minus, stream=None):
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        c = np.ascontiguousarray(coeffs, np.float64)
        check(lib().sdg_dr_space_time_interpolation(engine.handle, c.ctypes.data, ptr(derivatives),
                                                    C.byref(self.view), ptr(q_plus), ptr(q_minus), s),
              "sdg_dr_space_time_interpolation")

    def nodal_flux(self, engine, layer_view, imposed_plus, imposed_minus, stream=None):
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        check(lib().sdg_dr_nodal_flux(engine.handle, C.byref(layer_view), ptr(self.mapping), C.byref(self.view),
                                      ptr(imposed_plus), ptr(imposed_minus), s), "sdg_dr_nodal_flux")
