"""CFL time step (computeCellTimestep, src/Initializer/TimeStepping/GlobalTimestep.cpp:27-47):
dt = cfl * 2 r_in / (vp (2 O - 1)) with r_in the insphere radius, minimised over
the cells.  Pinned by closed forms of the insphere radius, not by restating the
determinant formula: the regular tetrahedron of edge a has r_in = a / (2 sqrt 6),
the unit corner tetrahedron r_in = 1 / (3 + sqrt 3)."""
import math

import numpy as np

from seissol_amd import _capi


def _dt(vertices, material, order, cfl=0.5):
    v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 4, 3)
    m = np.ascontiguousarray(material, np.float64).reshape(-1, 3)
    return float(_capi.lib().sdg_cfl_timestep(len(v), order, cfl, v.ctypes.data, m.ctypes.data))


def _material(rho, vp, vs):
    mu = rho * vs * vs
    return np.array([rho, rho * vp * vp - 2 * mu, mu])


def _rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    return q * np.sign(np.diag(r))


def test_regular_tetrahedron_closed_form():
    a = 37.0
    verts = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float) * a / (2 * math.sqrt(2))
    mat = _material(2670.0, 6000.0, 3464.0)
    for order in (2, 4, 6, 8):
        r_in = a / (2 * math.sqrt(6))
        want = 0.5 * 2 * r_in / (6000.0 * (2 * order - 1))
        assert abs(_dt(verts, mat, order) - want) < 1e-13 * want


def test_unit_tetrahedron_and_the_p_wave_speed():
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    rho, vp = 2200.0, 5000.0
    mat = _material(rho, vp, 2500.0)
    got = _dt(verts, mat, 4, cfl=0.35)
    want = 0.35 * 2 / (3 + math.sqrt(3)) / (vp * 7)
    assert abs(got - want) < 1e-13 * want


def test_rigid_motion_scaling_and_the_minimum_over_cells():
    rng = np.random.default_rng(4)
    mat = _material(2500.0, 6000.0, 3464.0)
    base = rng.standard_normal((4, 3))
    dt0 = _dt(base, mat, 5)
    # invariant under rotation, reflection of the vertex order and translation
    moved = base @ _rotation(rng).T + rng.standard_normal(3) * 1e3
    assert abs(_dt(moved[[1, 0, 2, 3]], mat, 5) - dt0) < 1e-9 * dt0
    # linear in the cell size
    assert abs(_dt(3.5 * base, mat, 5) - 3.5 * dt0) < 1e-12 * dt0
    # the minimum over the cells, with each cell's own wave speed
    slow = _material(2500.0, 3000.0, 1700.0)
    cells = np.stack([base, 3.0 * base, 0.25 * base])
    mats = np.stack([mat, mat, slow])
    # the third cell: a quarter of the radius at half the speed
    assert abs(_dt(cells, mats, 5) - 0.5 * dt0) < 1e-12 * dt0
    assert abs(_dt(cells[:2], mats[:2], 5) - dt0) < 1e-12 * dt0
