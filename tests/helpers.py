"""Test helpers: exact polynomial states, their L2 projection onto the modal
basis of each element, and the exact time evolution of the linear elastic
system for polynomial data (the Taylor series terminates)."""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

from seissol_amd import _capi
from seissol_amd.kernels import num_basis


def tet_quadrature(n: int):
    g, w = np.polynomial.legendre.leggauss(n)
    g = (g + 1) / 2
    w = w / 2
    A, Bm, Cm = np.meshgrid(g, g, g, indexing="ij")
    W = np.einsum("i,j,k->ijk", w, w, w) * (1 - Bm) * (1 - Cm) ** 2
    xi = A * (1 - Bm) * (1 - Cm)
    eta = Bm * (1 - Cm)
    zeta = Cm
    return np.stack([xi.ravel(), eta.ravel(), zeta.ravel()], 1), W.ravel()


@lru_cache(maxsize=None)
def basis_table(order: int, nq: int):
    pts, w = tet_quadrature(nq)
    B = num_basis(order)
    phi = np.zeros((len(w), B))
    lib = _capi.lib()
    for i, (a, b, c) in enumerate(pts):
        lib.sdg_basis_values(order, float(a), float(b), float(c), phi[i].ctypes.data)
    mass = (phi * w[:, None]).T @ phi
    return pts, w, phi, np.diag(mass).copy()


def coefficient_matrices(rho: float, lam: float, mu: float):
    """Physical Jacobians A_d with dQ/dt = -sum_d A_d dQ/dx_d (transpose of the
    reference's transposed coefficient matrices)."""
    A = np.zeros((3, 9, 9))
    l2m = lam + 2 * mu
    # entries (q, p) of the transposed matrices, ElasticSetup.h:27-77
    AT = np.zeros((3, 9, 9))
    AT[0, 6, 0], AT[0, 6, 1], AT[0, 6, 2], AT[0, 7, 3], AT[0, 8, 5] = -l2m, -lam, -lam, -mu, -mu
    AT[0, 0, 6], AT[0, 3, 7], AT[0, 5, 8] = -1 / rho, -1 / rho, -1 / rho
    AT[1, 7, 0], AT[1, 7, 1], AT[1, 7, 2], AT[1, 6, 3], AT[1, 8, 4] = -lam, -l2m, -lam, -mu, -mu
    AT[1, 1, 7], AT[1, 3, 6], AT[1, 4, 8] = -1 / rho, -1 / rho, -1 / rho
    AT[2, 8, 0], AT[2, 8, 1], AT[2, 8, 2], AT[2, 7, 4], AT[2, 6, 5] = -lam, -lam, -l2m, -mu, -mu
    AT[2, 2, 8], AT[2, 5, 6], AT[2, 4, 7] = -1 / rho, -1 / rho, -1 / rho
    for d in range(3):
        A[d] = AT[d].T
    return A


class PolyField:
    """9 polynomials in scaled coordinates s = x / scale, degree < order."""

    def __init__(self, order: int, scale: float, rng: np.random.Generator, amplitude=None):
        self.order, self.scale = order, scale
        self.c = np.zeros((9, order, order, order))
        amp = np.ones(9) if amplitude is None else np.asarray(amplitude)
        for a in range(order):
            for b in range(order - a):
                for cc in range(order - a - b):
                    self.c[:, a, b, cc] = rng.uniform(-1, 1, 9) * amp

    def copy_with(self, c):
        f = PolyField.__new__(PolyField)
        f.order, f.scale, f.c = self.order, self.scale, c
        return f

    def derivative(self, d: int) -> np.ndarray:
        c = np.zeros_like(self.c)
        O = self.order
        for k in range(1, O):
            idx = [slice(None)] * 4
            src = [slice(None)] * 4
            idx[d + 1] = k - 1
            src[d + 1] = k
            c[tuple(idx)] = k * self.c[tuple(src)] / self.scale
        return c

    def apply_operator(self, A) -> "PolyField":
        """L f = -sum_d A_d d_d f."""
        out = np.zeros_like(self.c)
        for d in range(3):
            out -= np.einsum("pq,qabc->pabc", A[d], self.derivative(d))
        return self.copy_with(out)

    def evaluate(self, x: np.ndarray) -> np.ndarray:
        """x: [..., 3] -> [..., 9]"""
        s = x / self.scale
        O = self.order
        px = np.stack([s[..., 0] ** k for k in range(O)], -1)
        py = np.stack([s[..., 1] ** k for k in range(O)], -1)
        pz = np.stack([s[..., 2] ** k for k in range(O)], -1)
        # sum_abc x^a y^b z^c c[p,a,b,c]: the z sum as one matrix product per
        # chunk of points, then a batched product with the x^a y^b weights
        m = px.reshape(-1, O).shape[0]
        px, py, pz = px.reshape(m, O), py.reshape(m, O), pz.reshape(m, O)
        cz = self.c.transpose(3, 0, 1, 2).reshape(O, 9 * O * O)
        out = np.empty((m, 9))
        for i in range(0, m, 32768):
            j = min(m, i + 32768)
            t = (pz[i:j] @ cz).reshape(j - i, 9, O * O)
            w = (px[i:j, :, None] * py[i:j, None, :]).reshape(j - i, O * O, 1)
            out[i:j] = (t @ w)[..., 0]
        return out.reshape(s.shape[:-1] + (9,))


def evolve(field: PolyField, A, t: float, integrate: bool = False) -> PolyField:
    """exp(t L) f (or its integral over [0, t]); exact for polynomial data."""
    total = np.zeros_like(field.c)
    term = field
    for k in range(field.order + 1):
        coef = t ** (k + 1) / math.factorial(k + 1) if integrate else t ** k / math.factorial(k)
        total += coef * term.c
        term = term.apply_operator(A)
    return field.copy_with(total)


def project(order: int, vertices: np.ndarray, field: PolyField, nq: int | None = None) -> np.ndarray:
    """L2 projection of the field onto each element's modal basis: [n][9][B]."""
    nq = nq or order + 2
    pts, w, phi, mass = basis_table(order, nq)
    v0 = vertices[:, 0, :]
    J = np.stack([vertices[:, 1] - v0, vertices[:, 2] - v0, vertices[:, 3] - v0], -1)  # [n][3][3]
    x = v0[:, None, :] + np.einsum("nij,qj->nqi", J, pts)
    f = field.evaluate(x)  # [n][q][9]
    return np.einsum("qk,nqp->npk", phi * w[:, None], f, optimize=True) / mass[None, None, :]


def interior_elements(setup, margin: int = 1) -> np.ndarray:
    """Cells of a periodic box whose cube index is >= margin away from the wrap."""
    nx, ny, nz = setup.dims
    h = (setup.vertices[:, :, 0].max() - setup.vertices[:, :, 0].min()) / nx
    cent = setup.vertices.mean(1)
    idx = np.floor(cent / h + 1e-9).astype(int)
    ok = np.ones(setup.n, bool)
    for d, nd in enumerate((nx, ny, nz)):
        ok &= (idx[:, d] >= margin) & (idx[:, d] <= nd - 1 - margin)
    return np.nonzero(ok)[0]


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_l2_per_quantity(a: np.ndarray, ref: np.ndarray) -> np.ndarray:
    """a, ref: [n][9][B] -> relative L2 error of each quantity."""
    num = np.sqrt(((a - ref) ** 2).sum(axis=(0, 2)))
    den = np.sqrt((ref ** 2).sum(axis=(0, 2)))
    return num / np.maximum(den, 1e-300)
