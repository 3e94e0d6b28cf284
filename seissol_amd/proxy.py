"""Proxy harness: synthetic layers and the proxy kernel selection.

Mirrors the reference proxy (src/Proxy/Allocator.cpp:49-134 for the data,
src/Proxy/Common.cpp:22-41 for the kernel names, src/Proxy/Runner.cpp:57-154
for the timing protocol) with physical synthetic data:

* ``mesh="proxy"``: independent random affine tetrahedra (det J in [0.5, 2]),
  random face neighbours and random face relations, as the reference proxy
  fabricates them;
* ``mesh="periodic"``: a conforming periodic jittered tetrahedral box whose
  neighbours and face relations follow from shared vertices.

Materials rho ~ U[rho_min, rho_max], vp, vs fixed; DOFs ~ U[-1, 1]; dt from the
CFL condition (src/Initializer/TimeStepping/GlobalTimestep.cpp:27-47).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import DrLayer, Layer, ProxyDr, check, lib, ptr
from .kernels import MATRIX_CORE, STAR_NONZEROS, Engine, num_basis, padded_face_basis

KERNELS = ("all", "local", "neigh", "ader", "localwoader", "neigh_dr", "godunov_dr", "all_dr")
DR_KERNELS = ("neigh_dr", "godunov_dr", "all_dr")


@dataclass
class Setup:
    """Host-side FP64 description of a layer of cells."""

    order: int
    n: int
    vertices: np.ndarray        # [n][4][3]
    material: np.ndarray        # [n][3] (rho, lambda, mu)
    neighbors: np.ndarray       # [n][4] int32
    face_relations: np.ndarray  # [n][4][2] uint8
    face_types: np.ndarray      # [n][4] uint8
    dofs: np.ndarray            # [n][9][B]
    dt: float
    seed: int
    mesh: str
    dims: tuple = field(default=())


def periodic_dims_for(n_target: int) -> tuple:
    """Even box dimensions (>= 4) whose 5-tet mesh has about n_target cells."""
    cubes = max(64, n_target // 5)
    best = None
    for nx in range(4, 400, 2):
        for ny in range(nx, 400, 2):
            nz = cubes // (nx * ny)
            nz -= nz % 2
            if nz < ny:
                break
            got = 5 * nx * ny * nz
            if best is None or abs(got - n_target) < abs(best[0] - n_target):
                best = (got, (nx, ny, nz))
    return best[1]


def make_setup(order: int, n: int | None = None, mesh: str = "proxy", seed: int = 1234,
               dims: tuple | None = None, rho=(2000.0, 3000.0), vp: float = 6000.0, vs: float = 3464.0,
               cfl: float = 0.5, jitter: float = 0.1, h: float = 100.0) -> Setup:
    L = lib()
    if mesh == "proxy":
        if n is None or n <= 0:
            raise ValueError("proxy mesh needs n > 0")
        vertices = np.zeros((n, 4, 3))
        neighbors = np.zeros((n, 4), dtype=np.int32)
        rel = np.zeros((n, 4, 2), dtype=np.uint8)
        check(L.sdg_mesh_proxy(n, seed, vertices.ctypes.data, neighbors.ctypes.data, rel.ctypes.data),
              "sdg_mesh_proxy")
        dims = ()
    elif mesh == "periodic":
        dims = tuple(dims) if dims else periodic_dims_for(n)
        n = int(L.sdg_mesh_periodic_size(*dims))
        vertices = np.zeros((n, 4, 3))
        neighbors = np.zeros((n, 4), dtype=np.int32)
        rel = np.zeros((n, 4, 2), dtype=np.uint8)
        check(L.sdg_mesh_periodic(*dims, h, jitter, seed, vertices.ctypes.data, neighbors.ctypes.data,
                                  rel.ctypes.data), "sdg_mesh_periodic")
    else:
        raise ValueError(f"unknown mesh kind {mesh!r}")
    material = np.zeros((n, 3))
    check(L.sdg_random_materials(n, seed + 1, rho[0], rho[1], vp, vs, material.ctypes.data),
          "sdg_random_materials")
    B = num_basis(order)
    dofs = np.zeros((n, 9, B))
    check(L.sdg_uniform_dofs(dofs.size, seed + 2, dofs.ctypes.data), "sdg_uniform_dofs")
    dt = float(L.sdg_cfl_timestep(n, order, cfl, vertices.ctypes.data, material.ctypes.data))
    face_types = np.zeros((n, 4), dtype=np.uint8)
    return Setup(order, n, vertices, material, neighbors, rel, face_types, dofs, dt, seed, mesh, dims)


def struct_reals(nreals: int, real_bytes: int, alignment: int) -> int:
    """Reals per element of an array of structs {real data[nreals]; Empty e;}
    aligned to `alignment` bytes (the reference's alignas(Alignment)
    LocalIntegrationData / NeighboringIntegrationData with the empty elastic
    specific data, src/Initializer/Typedefs.h:186-203)."""
    a = max(alignment, real_bytes)
    size = nreals * real_bytes + 1
    return (size + a - 1) // a * a // real_bytes


class DeviceLayer:
    """Device arrays of one layer plus the ``sdg_layer`` view handed to the C ABI.

    The per-element operators are computed on the device in FP64
    (``sdg_cell_local_matrices``) and converted to the engine's real type.

    With ``struct_arrays=True`` the layer is stored as the reference stores it
    under the engine's layout (``Engine.set_layout``): DOF-like blocks
    [n][9][ld] and aligned derivative blocks, the star matrices and AplusT
    inside one array of LocalIntegrationData-like structs and AminusT inside
    one of NeighboringIntegrationData-like structs (passed with their
    strides).  Every padding entry holds NaN, so a kernel that read one would
    spoil its result.  ``packed()`` returns the [n][9][B] view of a DOF-like
    array.
    """

    def __init__(self, engine: Engine, setup: Setup, with_derivatives: bool = False, device: str = "cuda",
                 struct_arrays: bool = False):
        import torch

        self.engine = engine
        self.setup = setup
        n, B = setup.n, num_basis(setup.order)
        rt = engine.real_dtype
        dev = torch.device(device)
        f64 = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.vertices = f64(setup.vertices)
        self.material = f64(setup.material)
        self.neighbors = f64(setup.neighbors)
        self.face_relations = f64(setup.face_relations)
        self.face_types = f64(setup.face_types)
        star64 = torch.zeros((n, 3, 81), dtype=torch.float64, device=dev)
        aplus64 = torch.zeros((n, 4, 81), dtype=torch.float64, device=dev)
        aminus64 = torch.zeros((n, 4, 81), dtype=torch.float64, device=dev)
        stream = torch.cuda.current_stream().cuda_stream
        check(lib().sdg_cell_local_matrices(n, ptr(self.vertices), ptr(self.material), ptr(self.neighbors),
                                            ptr(self.face_types), ptr(star64), ptr(aplus64), ptr(aminus64),
                                            stream), "sdg_cell_local_matrices")
        self.star64, self.aplus64, self.aminus64 = star64, aplus64, aminus64
        ld, qreals, dreals = engine.layout_sizes()
        self.ld = ld
        compact = getattr(engine, "star_format", "dense") == "compact"
        star_src = star64[:, :, torch.from_numpy(STAR_NONZEROS).to(dev)] if compact else star64
        strides = (0, 0, 0)
        nan = float("nan")
        if struct_arrays:
            rs = 8 if rt == torch.float64 else 4
            align = getattr(engine, "alignment", 0)
            sreals = star_src[0].numel()
            lstride = struct_reals(sreals + 324, rs, align)
            nstride = struct_reals(324, rs, align)
            self.local_data = torch.full((n, lstride), nan, dtype=rt, device=dev)
            self.local_data[:, :sreals] = star_src.reshape(n, -1).to(rt)
            self.local_data[:, sreals:sreals + 324] = aplus64.reshape(n, -1).to(rt)
            self.neighbor_data = torch.full((n, nstride), nan, dtype=rt, device=dev)
            self.neighbor_data[:, :324] = aminus64.reshape(n, -1).to(rt)
            self.star = self.local_data
            self.aplus = self.local_data[:, sreals:]
            self.aminus = self.neighbor_data
            strides = (lstride, lstride, nstride)
        elif rt == torch.float64 and not compact:
            self.star, self.aplus, self.aminus = star64, aplus64, aminus64
        else:
            self.star, self.aplus, self.aminus = (star_src.contiguous().to(rt), aplus64.to(rt), aminus64.to(rt))
        fill = nan if struct_arrays else 0.0
        self.dofs = torch.full((n, 9, ld), fill, dtype=rt, device=dev)
        self.dofs[:, :, :B] = torch.from_numpy(setup.dofs).to(dev).to(rt)
        self.time_integrated = torch.full((n, 9, ld), fill, dtype=rt, device=dev)
        self.time_integrated[:, :, :B] = 0.0
        self.derivatives = (torch.full((n, dreals), fill, dtype=rt, device=dev) if with_derivatives else None)
        B2P = padded_face_basis(setup.order)
        if engine.arithmetic in MATRIX_CORE:
            self.face_traces = torch.zeros((n, 4, 9, B2P), dtype=torch.float32, device=dev)
        elif engine.arithmetic == "fp64" and getattr(engine, "kernels", "auto") == "auto" and 3 <= setup.order <= 7:
            # the FP64 matrix-core passes exchange FP64 face traces
            self.face_traces = torch.zeros((n, 4, 9, B2P), dtype=torch.float64, device=dev)
        else:
            self.face_traces = None
        self.view = Layer(n, ptr(self.dofs), ptr(self.time_integrated), ptr(self.derivatives), ptr(self.star),
                          ptr(self.aplus), ptr(self.aminus), ptr(self.neighbors), ptr(self.face_relations),
                          ptr(self.face_types), ptr(self.face_traces), None, *strides)

    def packed(self, t):
        """The [n][9][B] entries of a DOF-like array of this layer."""
        return t[:, :, :num_basis(self.setup.order)]

    def derivatives_packed(self):
        """The derivatives as [n][sum_i 9 B(O-i)] (blocks [9][B(O-i)]), the oracle's layout."""
        import torch
        n = self.setup.n
        parts = [self.derivatives[:, off:off + 9 * ld].reshape(n, 9, ld)[:, :, :m].reshape(n, 9 * m)
                 for off, ld, m in self.engine.derivative_blocks()]
        return torch.cat(parts, 1)

    def reset_dofs(self, dofs: np.ndarray):
        import torch
        self.packed(self.dofs).copy_(torch.from_numpy(dofs).to(self.dofs.device).to(self.dofs.dtype))


class DeviceDrProxy:
    """The proxy's dynamic-rupture data on the device (src/Proxy/Allocator.cpp:186-294):
    every face of every cell is a DR face, 4 n DR faces join random cells with
    random sides and orientations, the layer's derivatives hold random values
    (the proxy's fake derivatives) and the imposed states random values.
    The layer must carry derivatives (``DeviceLayer(..., with_derivatives=True)``)."""

    def __init__(self, engine: Engine, layer: "DeviceLayer", seed: int | None = None, device: str = "cuda"):
        import torch

        setup = layer.setup
        if layer.derivatives is None:
            raise ValueError("the DR proxy kernels need a layer with derivatives")
        n, order = setup.n, setup.order
        seed = setup.seed + 3 if seed is None else seed
        nf = 4 * n
        mapping = np.zeros((n, 4), np.int32)
        plus = np.zeros(nf, np.int32)
        minus = np.zeros(nf, np.int32)
        info = np.zeros((nf, 3), np.uint8)
        check(lib().sdg_proxy_dr_faces(n, seed, mapping.ctypes.data, plus.ctypes.data, minus.ctypes.data,
                                       info.ctypes.data), "sdg_proxy_dr_faces")
        tinv = np.zeros((nf, 81))
        fp = np.zeros((nf, 81))
        fm = np.zeros((nf, 81))
        v = np.ascontiguousarray(setup.vertices)
        m = np.ascontiguousarray(setup.material)
        check(lib().sdg_dr_face_data(nf, v.ctypes.data, m.ctypes.data, plus.ctypes.data, minus.ctypes.data,
                                     info.ctypes.data, tinv.ctypes.data, fp.ctypes.data, fm.ctypes.data, None),
              "sdg_dr_face_data")
        points = (order + 1) ** 2
        dev = torch.device(device)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.mapping, self.plus, self.minus, self.info = t(mapping), t(plus), t(minus), t(info)
        self.tinv_t, self.flux_plus, self.flux_minus = t(tinv), t(fp), t(fm)
        rnd = np.zeros(nf * 9 * points * 2)
        check(lib().sdg_uniform_dofs(rnd.size, seed + 1, rnd.ctypes.data), "sdg_uniform_dofs")
        self.imposed_plus = t(rnd[: nf * 9 * points].reshape(nf, 9, points))
        self.imposed_minus = t(rnd[nf * 9 * points:].reshape(nf, 9, points))
        self.q_plus = torch.zeros((nf, order, 9, points), dtype=torch.float64, device=dev)
        self.q_minus = torch.zeros_like(self.q_plus)
        der = np.zeros(layer.derivatives.numel())
        check(lib().sdg_uniform_dofs(der.size, seed + 2, der.ctypes.data), "sdg_uniform_dofs")
        layer.derivatives.copy_(torch.from_numpy(der).reshape(layer.derivatives.shape).to(layer.derivatives.dtype))
        layer.face_types.fill_(_capi.SDG_FACE_DYNAMIC_RUPTURE)
        self.faces = DrLayer(nf, points, 0, ptr(self.plus), ptr(self.minus), ptr(self.info), ptr(self.tinv_t),
                             ptr(self.flux_plus), ptr(self.flux_minus))
        self.view = ProxyDr(ptr(self.mapping), self.faces, ptr(self.imposed_plus), ptr(self.imposed_minus),
                            ptr(self.q_plus), ptr(self.q_minus))
        self.num_faces, self.points = nf, points
        self.host = {"dr_mapping": mapping, "plus": plus, "minus": minus, "face_info": info, "tinv_t": tinv,
                     "flux_plus": fp, "flux_minus": fm}


def kernel_id(name: str) -> int:
    k = lib().sdg_proxy_kernel_id(name.encode())
    if k < 0:
        raise ValueError(f"unknown proxy kernel {name!r}; expected one of {KERNELS}")
    return k


def run_proxy(engine: Engine, layer: DeviceLayer, kernel: str, steps: int, dt: float, stream=None,
              dr: DeviceDrProxy | None = None):
    """Enqueue `steps` repetitions of proxy kernel `kernel` (no synchronisation);
    the DR kernels take the layer's DR data `dr`."""
    import torch
    s = torch.cuda.current_stream().cuda_stream if stream is None else stream
    if kernel in DR_KERNELS:
        if dr is None:
            raise ValueError(f"proxy kernel {kernel!r} needs the layer's DR data (DeviceDrProxy)")
        check(lib().sdg_proxy_run_dr(engine.handle, kernel_id(kernel), steps, float(dt), C.byref(layer.view),
                                     C.byref(dr.view), s), "sdg_proxy_run_dr")
        return
    check(lib().sdg_proxy_run(engine.handle, kernel_id(kernel), steps, float(dt), C.byref(layer.view), s),
          "sdg_proxy_run")
