"""Dynamic-rupture faces: space-time interpolation and nodal flux of the
oracle, pinned against the regular flux of continuous states (CPU)."""
import numpy as np
import pytest

import oracle
from helpers import PolyField, coefficient_matrices, project, rel_l2
from seissol_amd.dr import collocation_coefficients, face_matrices, plane_faces
from seissol_amd.kernels import global_matrices, monomial_integrate
from seissol_amd.proxy import make_setup


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.build()


def _case(order, seed=3):
    setup = make_setup(order, mesh="periodic", dims=(8, 6, 6), seed=seed, rho=(2500.0, 2500.0), h=100.0,
                       jitter=0.1)
    faces = plane_faces(setup, plane=4)
    field = PolyField(order, scale=800.0, rng=np.random.default_rng(seed),
                      amplitude=np.array([1e6] * 6 + [1.0] * 3))
    setup.dofs = project(order, setup.vertices, field)
    return setup, faces


def test_plane_faces_and_face_data():
    setup, faces = _case(3)
    assert faces.num_faces == 2 * 6 * 6  # two triangles per cube face of the plane
    assert np.all(faces.dr_mapping[faces.plus_element, faces.face_info[:, 0]] == np.arange(faces.num_faces))
    assert np.all(faces.dr_mapping[faces.minus_element, faces.face_info[:, 1]] == np.arange(faces.num_faces))
    assert np.all(faces.face_info[:, 2] >= 1) and np.all(faces.face_info[:, 2] <= 3)
    # flux solvers of the two sides: scales of opposite sign, same material
    ratio = faces.flux_plus / np.where(faces.flux_minus == 0, np.nan, faces.flux_minus)
    assert np.nanmax(ratio) < 0
    # impedances of the friction law: eta = Z/2 for equal materials
    np.testing.assert_allclose(faces.imp_and_eta[:, 5], faces.imp_and_eta[:, 1] / 2, rtol=1e-14)


@pytest.mark.parametrize("order", [3, 4])
def test_interpolation_agrees_on_both_sides_of_a_continuous_field(order):
    setup, faces = _case(order)
    gm = global_matrices(order)["flat"]
    star, ap, am = oracle.cell_local_matrices(setup.vertices, setup.material, setup.neighbors, setup.face_types)
    dt = setup.dt
    _, deriv = oracle.ader(order, gm, monomial_integrate(order, 0.0, dt), setup.dofs, star, want_derivatives=True)
    v3m, _ = face_matrices(order)
    qp, qm = oracle.dr_interpolate(order, faces.points, v3m, deriv, faces.plus_element, faces.minus_element,
                                   faces.face_info, faces.tinv_t, collocation_coefficients(order, dt))
    # the plane is away from the periodic wrap: the field is continuous there
    assert rel_l2(qm, qp) < 1e-11
    assert np.abs(qp).max() > 0


def _updates(order, dr):
    setup, faces = _case(order)
    if not dr:
        setup.face_types[:] = 0
    gm = global_matrices(order)["flat"]
    star, ap, am = oracle.cell_local_matrices(setup.vertices, setup.material, setup.neighbors, setup.face_types)
    coeffs = monomial_integrate(order, 0.0, setup.dt)
    I, deriv = oracle.ader(order, gm, coeffs, setup.dofs, star, want_derivatives=True)
    q = oracle.local(order, gm, I, star, ap, setup.dofs, setup.face_types)
    q = oracle.neighbor(order, gm, I, setup.neighbors, setup.face_relations, am, q, setup.face_types)
    if dr:
        # imposed state of a welded interface: the time-integrated trace itself
        v3m, v3mtw = face_matrices(order)
        c = np.zeros((order, order))
        c[0] = coeffs
        g, _ = oracle.dr_interpolate(order, faces.points, v3m, deriv, faces.plus_element, faces.minus_element,
                                     faces.face_info, faces.tinv_t, c)
        g = g[:, 0]
        q = oracle.dr_nodal_flux(order, faces.points, v3mtw, faces.dr_mapping, faces.plus_element, faces.face_info,
                                 faces.flux_plus, faces.flux_minus, g, g, q)
    return q, setup, faces


@pytest.mark.parametrize("order", [3, 4])
def test_nodal_flux_of_a_welded_fault_equals_the_regular_flux(order):
    q_reg, setup, faces = _updates(order, dr=False)
    q_dr, _, _ = _updates(order, dr=True)
    touched = np.unique(np.concatenate([faces.plus_element, faces.minus_element]))
    change = np.linalg.norm(q_reg[touched] - setup.dofs[touched])
    assert np.linalg.norm(q_dr[touched] - q_reg[touched]) < 1e-10 * change
    rest = np.setdiff1d(np.arange(setup.n), touched)
    assert np.array_equal(q_dr[rest], q_reg[rest])


def test_planted_wrong_minus_orientation_is_caught():
    setup, faces = _case(3)
    gm = global_matrices(3)["flat"]
    star, ap, am = oracle.cell_local_matrices(setup.vertices, setup.material, setup.neighbors, setup.face_types)
    _, deriv = oracle.ader(3, gm, monomial_integrate(3, 0.0, setup.dt), setup.dofs, star, want_derivatives=True)
    v3m, _ = face_matrices(3)
    info = faces.face_info.copy()
    info[:, 2] = 1 + (info[:, 2] % 3)  # orientation + 1 -> next orientation
    qp, qm = oracle.dr_interpolate(3, faces.points, v3m, deriv, faces.plus_element, faces.minus_element, info,
                                   faces.tinv_t, collocation_coefficients(3, setup.dt))
    assert rel_l2(qm, qp) > 1e-3
