"""Local time stepping: the LTS setup rules and the oracle's rate-2 cluster
scheme (CPU)."""
import numpy as np
import pytest

import oracle
from helpers import PolyField, coefficient_matrices, evolve, interior_elements, project, rel_l2
from seissol_amd import _capi
from seissol_amd.kernels import global_matrices, monomial_integrate
from seissol_amd.lts import band_clusters, derive_lts_setups, make_lts_setup

NHD = lambda f: 1 << f          # neighbour provides derivatives
GTS = lambda f: 1 << (4 + f)    # same cluster
BUF, DER, ACC = 1 << 8, 1 << 9, 1 << 10


@pytest.fixture(scope="module", autouse=True)
def _oracle():
    oracle.build()


def test_band_clusters_step_by_one_across_edges_and_wrap():
    x = np.sort(np.random.default_rng(0).uniform(0, 1, 20000))
    c = band_clusters(x, (0.6, 0.25, 0.1, 0.05))
    assert np.all(np.abs(np.diff(c)) <= 1) and abs(int(c[-1]) - int(c[0])) <= 1
    frac = np.bincount(c, minlength=4) / len(c)
    np.testing.assert_allclose(frac, [0.6, 0.25, 0.1, 0.05], atol=0.01)


def test_lts_setup_known_answers():
    # cells 0, 1 in cluster 0; cells 2, 3 in cluster 1
    clusters = np.array([0, 0, 1, 1], np.int32)
    nb = np.array([[1, 1, 1, 1],
                   [0, 2, 0, 0],
                   [1, 3, 3, 3],
                   [2, 2, 2, 2]], np.int32)
    s = derive_lts_setups(clusters, nb)
    # cell 1 sees a larger step on face 1: derivatives from there, accumulated
    # buffer, and (it has same-cluster neighbours) derivatives of its own
    assert s[1] == NHD(1) | ACC | GTS(0) | GTS(2) | GTS(3) | BUF | DER
    # cell 0: same cluster everywhere, but its neighbour accumulates ->
    # "GTS on derivatives" on all four faces (normalisation)
    assert s[0] == GTS(0) | GTS(1) | GTS(2) | GTS(3) | BUF | NHD(0) | NHD(1) | NHD(2) | NHD(3)
    # cell 2: a faster neighbour on face 0 reads its derivatives; buffer for the rest
    assert s[2] == DER | GTS(1) | GTS(2) | GTS(3) | BUF
    assert s[3] == GTS(0) | GTS(1) | GTS(2) | GTS(3) | BUF
    # dynamic-rupture faces: GTS on derivatives
    ft = np.zeros((4, 4), np.uint8)
    ft[3, 0] = _capi.SDG_FACE_DYNAMIC_RUPTURE
    s2 = derive_lts_setups(clusters, nb, ft)
    assert s2[3] == NHD(0) | GTS(0) | DER | GTS(1) | GTS(2) | GTS(3) | BUF
    # neighbours more than one cluster apart are rejected
    with pytest.raises(_capi.SdgError):
        derive_lts_setups(np.array([0, 0, 2, 2], np.int32), nb)


def _oracle_run(lts, order, macro_steps, setups=None):
    s = lts.setup
    gm = global_matrices(order)["flat"]
    star, ap, am = oracle.cell_local_matrices(s.vertices, s.material, s.neighbors)
    q = oracle.lts(order, gm, macro_steps, lts.num_clusters, lts.cluster_ids, lts.dt0,
                   lts.lts_setups if setups is None else setups, star, ap, am, s.neighbors, s.face_relations,
                   s.dofs)
    return q, (gm, star, ap, am)


def test_single_cluster_is_global_time_stepping():
    order = 4
    lts = make_lts_setup(order, fractions=(1.0,), dims=(4, 4, 4), seed=2)
    assert lts.num_clusters == 1 and np.all(lts.lts_setups & (BUF | 0xF0) == (BUF | 0xF0))
    q, (gm, star, ap, am) = _oracle_run(lts, order, 3)
    s = lts.setup
    q_gts, _ = oracle.proxy_all(order, gm, 3, monomial_integrate(order, 0.0, lts.dt0), star, ap, am, s.neighbors,
                                s.face_relations, s.dofs)
    assert rel_l2(q, q_gts) < 1e-14


def _poly_lts(order, setups_override=None):
    lts = make_lts_setup(order, fractions=(0.5, 0.3, 0.2), dims=(16, 12, 12), scale_speeds=False,
                         rho=(2500.0, 2500.0), seed=4, jitter=0.1)
    s = lts.setup
    assert set(np.unique(lts.cluster_ids)) == {0, 1, 2}
    rng = np.random.default_rng(7)
    field = PolyField(order, scale=1600.0, rng=rng, amplitude=np.array([1e6] * 6 + [1.0] * 3))
    q0 = project(order, s.vertices, field)
    s.dofs = q0
    setups = lts.lts_setups if setups_override is None else setups_override(lts)
    q1, _ = _oracle_run(lts, order, 1, setups)
    A = coefficient_matrices(*s.material[0])
    exact = project(order, s.vertices, evolve(field, A, 4 * lts.dt0))
    inner = interior_elements(s, margin=5)
    assert len(inner) > 60 and set(np.unique(lts.cluster_ids[inner])) == {0, 1, 2}
    change = np.linalg.norm(exact[inner] - q0[inner])
    err = np.linalg.norm(q1[inner] - exact[inner])
    return err, change


@pytest.mark.parametrize("order", [3, 4])
def test_lts_macro_step_exact_for_polynomial_states(order):
    err, change = _poly_lts(order)
    assert err < 1e-9 * change, (err, change)


def test_planted_missing_buffer_accumulation_is_caught():
    # a slower cell must see the sum of its faster neighbour's two steps
    err, change = _poly_lts(3, lambda l: l.lts_setups & np.uint16(~ACC & 0xFFFF))
    assert err > 1e-4 * change


def test_planted_buffer_instead_of_derivatives_is_caught():
    # a faster cell must integrate its slower neighbour's derivatives over the sub-interval
    def strip(l):
        s = l.lts_setups.copy()
        for e in range(len(s)):
            for f in range(4):
                nb = l.setup.neighbors[e, f]
                if l.cluster_ids[nb] > l.cluster_ids[e]:
                    s[e] &= np.uint16(~NHD(f) & 0xFFFF)
        return s
    err, change = _poly_lts(3, strip)
    assert err > 1e-4 * change
