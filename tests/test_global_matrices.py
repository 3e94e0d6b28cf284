"""Generated global matrices and basis against the reference's data."""
import json
from pathlib import Path

import numpy as np
import pytest

from seissol_amd import _capi
from seissol_amd.kernels import global_matrices, monomial_integrate, num_basis

GOLDEN = Path(__file__).resolve().parent / "golden"
REF = Path("/root/reference/codegen/matrices")


def checksums(a):
    return {"sum": a.sum(), "abs_sum": np.abs(a).sum(), "frobenius": np.sqrt((a * a).sum()),
            "max_abs": np.abs(a).max(), "weighted": (a * np.arange(1, a.size + 1).reshape(a.shape)).sum()}


@pytest.mark.parametrize("order", [2, 3, 4, 5, 6, 7, 8])
def test_matrix_checksums_match_reference_data(order):
    gold = json.loads((GOLDEN / "global_matrix_checksums.json").read_text())[str(order)]
    gm = global_matrices(order)
    for name, ref in gold.items():
        fam, idx = name[:-1].split("(")
        a = gm[fam][int(idx)]
        assert list(a.shape) == ref["shape"], name
        got = checksums(a)
        scale = ref["abs_sum"]
        for k in ("sum", "abs_sum", "frobenius", "max_abs"):
            assert abs(got[k] - ref[k]) <= 1e-12 * max(scale, 1.0), (name, k, got[k], ref[k])
        assert abs(got["weighted"] - ref["weighted"]) <= 1e-12 * scale * a.size, (name, "weighted")


@pytest.mark.skipif(not REF.exists(), reason="reference data not present")
@pytest.mark.parametrize("order", [4, 6])
def test_matrices_entrywise_against_xml(order):
    import xml.etree.ElementTree as ET
    B = num_basis(order)
    gm = global_matrices(order)
    for m in ET.parse(REF / f"matrices_{B}.xml").getroot():
        name = m.get("name")
        fam = name.split("(")[0]
        if fam not in gm:
            continue
        idx = int(name.split("(")[1][:-1])
        ref = np.zeros((int(m.get("rows")), int(m.get("columns"))))
        for e in m:
            ref[int(e.get("row")) - 1, int(e.get("column")) - 1] = float(e.get("value"))
        np.testing.assert_allclose(gm[fam][idx], ref, atol=1e-14 * max(1, np.abs(ref).max()), rtol=0, err_msg=name)


def test_basis_values_golden():
    gold = json.loads((GOLDEN / "basis_values.json").read_text())
    lib = _capi.lib()
    # BasisFunctions.t.h checks orders 1..6 against prefixes of one list
    for order in range(1, 7):
        B = num_basis(order)
        if B > len(gold["values"]):
            break
        v = np.zeros(B)
        lib.sdg_basis_values(order, 0.3, 0.3, 0.3, v.ctypes.data)
        np.testing.assert_allclose(v, gold["values"][:B], atol=1e-14, rtol=0)
        g = np.zeros((3, B))
        lib.sdg_basis_gradients(order, 0.3, 0.3, 0.3, g.ctypes.data)
        for d in range(3):
            np.testing.assert_allclose(g[d], gold["gradients"][d][:B], atol=1e-13, rtol=0)


def test_mass_matrix_is_diagonal_and_positive():
    from helpers import basis_table
    for order in (3, 6):
        pts, w, phi, mass = basis_table(order, order + 2)
        M = (phi * w[:, None]).T @ phi
        assert np.abs(M - np.diag(np.diag(M))).max() < 1e-14
        np.testing.assert_allclose(global_matrices(order)["mass"], np.diag(M), rtol=1e-13)


def test_monomial_time_basis():
    # tests/Numerical/TimeBasis.t.h: integrate(0, x)[j] = x^(j+1)/(j+1)!
    import math
    for order in range(1, 10):
        for x in (0.0, 1.0, 2.0, 0.123, 15.0):
            c = monomial_integrate(order, 0.0, x)
            exp = [x ** (j + 1) / math.factorial(j + 1) for j in range(order)]
            np.testing.assert_allclose(c, exp, rtol=1e-14, atol=10 * np.finfo(float).eps)


def test_kdivmt_derivative_structure():
    # -kDivMT maps the modal coefficients of f to those of df/dxi_d, which
    # must lower the polynomial degree (derivatives of degree-(O-1) data vanish
    # after O applications).
    for order in (4, 6):
        gm = global_matrices(order)
        B = num_basis(order)
        rng = np.random.default_rng(0)
        v = rng.standard_normal(B)
        for d in range(3):
            x = v.copy()
            bound = np.abs(v).max()
            for _ in range(order):
                x = gm["kDivMT"][d] @ x
                bound *= np.abs(gm["kDivMT"][d]).sum(1).max()
            assert np.abs(x).max() < 1e-13 * bound
