"""Local time stepping (rate 2) on the C ABI.

* ``make_lts_setup``: a periodic tetrahedral box whose cells are split into
  clusters by bands along x (the layout keeps face neighbours at most one
  cluster apart, also across the periodic wrap); wave speeds in cluster c are
  scaled by 2^-c so that every cell's CFL step admits its cluster's step
  dt0 * 2^c; cells are sorted by cluster (the reference's per-cluster layers).
* ``LocalTimeStepping``: the clusters' macro step (TimeCluster /
  TimeManager semantics, src/Solver/TimeStepping), enqueued directly or
  replayed from a captured hipGraph.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import check, lib, ptr
from .kernels import Engine
from .proxy import DeviceLayer, Setup, make_setup


@dataclass
class LtsSetup:
    setup: Setup                 # cells sorted by cluster
    cluster_ids: np.ndarray      # [n] int32
    cluster_start: np.ndarray    # [C + 1] int64
    lts_setups: np.ndarray       # [n] uint16 (SDG_LTS_* bits)
    dt0: float

    @property
    def num_clusters(self) -> int:
        return len(self.cluster_start) - 1

    def updates_per_macro_step(self) -> int:
        """Element updates of one macro step: cluster c steps 2^(C-1-c) times."""
        C_ = self.num_clusters
        counts = np.diff(self.cluster_start)
        return int(sum(int(counts[c]) << (C_ - 1 - c) for c in range(C_)))


def band_clusters(x: np.ndarray, fractions) -> np.ndarray:
    """Cluster id of normalised positions x in [0, 1): cluster 0 takes the
    first fractions[0], the others mirrored bands (..., c, c+1, ..., c+1, c, ...)
    so that the ids step by one across every band edge and the periodic wrap."""
    fr = np.asarray(fractions, float)
    fr = fr / fr.sum()
    edges = [(0.0, fr[0], 0)]
    pos = fr[0]
    C_ = len(fr)
    for c in range(1, C_):  # rising half
        w = fr[c] / 2 if c < C_ - 1 else fr[c]
        edges.append((pos, pos + w, c))
        pos += w
    for c in range(C_ - 2, 0, -1):  # falling half
        edges.append((pos, pos + fr[c] / 2, c))
        pos += fr[c] / 2
    out = np.zeros(x.shape, np.int32)
    for lo, hi, c in edges:
        out[(x >= lo) & (x < hi)] = c
    return out


def cubic_dims(n_target: int) -> tuple:
    """Even, near-cubic box dimensions (x the longest) with about n_target cells (5 per cube)."""
    cubes = max(64, n_target // 5)
    c = round(cubes ** (1.0 / 3.0))
    best = None
    for ny in range(max(4, c - 8) & ~1, c + 10, 2):
        for nz in range(max(4, c - 8) & ~1, c + 10, 2):
            nx = int(round(cubes / (ny * nz) / 2.0)) * 2
            if nx < 4:
                continue
            got = 5 * nx * ny * nz
            key = (abs(got - n_target), -nx)
            if best is None or key < best[0]:
                best = (key, (max(nx, ny, nz), *sorted((nx, ny, nz))[:2][::-1]))
    return best[1]


def make_lts_setup(order: int, n: int | None = None, fractions=(0.6, 0.25, 0.1, 0.05), seed: int = 1234,
                   dims: tuple | None = None, scale_speeds: bool = True, cfl: float = 0.5, h: float = 100.0,
                   jitter: float = 0.1, rho=(2000.0, 3000.0)) -> LtsSetup:
    dims = tuple(dims) if dims else cubic_dims(n)
    base = make_setup(order, n=n, mesh="periodic", seed=seed, dims=dims, rho=rho, h=h, jitter=jitter, cfl=cfl)
    nx = base.dims[0]
    cent = base.vertices.mean(1)
    xn = (cent[:, 0] / (nx * h)) % 1.0
    cl = band_clusters(xn, fractions)
    mat = base.material.copy()
    if scale_speeds:
        s2 = 0.25 ** cl  # (2^-c)^2 scales lambda and mu
        mat[:, 1] *= s2
        mat[:, 2] *= s2
    # sort by cluster (stable), remap neighbour indices
    perm = np.argsort(cl, kind="stable")
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    nb = inv[base.neighbors[perm]].astype(np.int32)
    cl = cl[perm]
    setup = Setup(order, base.n, base.vertices[perm].copy(), mat[perm].copy(), nb, base.face_relations[perm].copy(),
                  base.face_types[perm].copy(), base.dofs[perm].copy(), base.dt, seed, "periodic", base.dims)
    C_ = len(fractions)
    counts = np.bincount(cl, minlength=C_)
    start = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    # dt0: every cell's CFL step must admit dt0 * 2^c
    L = lib()
    dt0 = np.inf
    for c in range(C_):
        if counts[c] == 0:
            continue
        sl = slice(start[c], start[c + 1])
        v = np.ascontiguousarray(setup.vertices[sl])
        m = np.ascontiguousarray(setup.material[sl])
        dtc = float(L.sdg_cfl_timestep(int(counts[c]), order, cfl, v.ctypes.data, m.ctypes.data))
        dt0 = min(dt0, dtc / 2 ** c)
    setup.dt = dt0
    setups = derive_lts_setups(cl, setup.neighbors, setup.face_types)
    return LtsSetup(setup, cl.astype(np.int32), start, setups, dt0)


def derive_lts_setups(cluster_ids, neighbors, face_types=None) -> np.ndarray:
    cl = np.ascontiguousarray(cluster_ids, np.int32)
    nb = np.ascontiguousarray(neighbors, np.int32)
    ft = None if face_types is None else np.ascontiguousarray(face_types, np.uint8)
    out = np.zeros(len(cl), np.uint16)
    check(lib().sdg_derive_lts_setups(len(cl), cl.ctypes.data, nb.ctypes.data,
                                      None if ft is None else ft.ctypes.data, out.ctypes.data),
          "sdg_derive_lts_setups")
    return out


class DeviceLtsLayer(DeviceLayer):
    """Device layer with derivatives and LTS bits."""

    def __init__(self, engine: Engine, lts: LtsSetup, device: str = "cuda", struct_arrays: bool = False):
        import torch
        super().__init__(engine, lts.setup, with_derivatives=True, device=device, struct_arrays=struct_arrays)
        self.lts = lts
        self.lts_setups = torch.from_numpy(lts.lts_setups.view(np.int16)).to(device)
        self.view.lts_setup = ptr(self.lts_setups)


class LocalTimeStepping:
    """The clusters of one layer; ``advance`` runs macro steps (the slowest
    cluster's step), eagerly or from a captured hipGraph."""

    def __init__(self, engine: Engine, layer: DeviceLtsLayer):
        self.engine, self.layer = engine, layer
        h = C.c_void_p()
        start = np.ascontiguousarray(layer.lts.cluster_start, np.int64)
        check(lib().sdg_lts_create(C.byref(h), engine.handle, C.byref(layer.view), layer.lts.num_clusters,
                                   start.ctypes.data, float(layer.lts.dt0)), "sdg_lts_create")
        self.handle = h

    def advance(self, macro_steps: int, graph: bool = False, stream=None):
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        fn = lib().sdg_lts_graph_advance if graph else lib().sdg_lts_advance
        check(fn(self.handle, int(macro_steps), s), "sdg_lts_advance")

    def autotune(self, repeats: int = 3, stream=None) -> dict:
        """Pick the fastest tile variant per cluster and kernel; returns the choices."""
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        check(lib().sdg_lts_autotune(self.handle, int(repeats), s), "sdg_lts_autotune")
        return self.tiles()

    def set_tile(self, variant: int, cluster: int | None = None, kernel: int | None = None):
        clusters = range(self.layer.lts.num_clusters) if cluster is None else [cluster]
        kernels = (0, 1) if kernel is None else (kernel,)
        for c in clusters:
            for k in kernels:
                check(lib().sdg_lts_set_tile_variant(self.handle, c, k, int(variant)), "sdg_lts_set_tile_variant")

    def tiles(self) -> dict:
        out = {}
        for c in range(self.layer.lts.num_clusters):
            out[c] = {k: int(lib().sdg_lts_tile_variant(self.handle, c, i)) for i, k in enumerate(("local", "neigh"))}
        return out

    def close(self):
        if self.handle:
            check(lib().sdg_lts_destroy(self.handle), "sdg_lts_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
