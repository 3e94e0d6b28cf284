"""Write the golden fixtures of this directory from data the reference holds.

Run in a checkout where the SeisSol sources are available (default
``/root/reference``); the fixtures are committed so the tests never read the
reference at run time.  Only numeric data is extracted:

* ``basis_values.json``  -- basis values/gradients at (0.3, 0.3, 0.3), orders 1-6,
  from tests/Numerical/BasisFunctions.t.h:22-300;
* ``godunov_elastic.json`` -- the elastic Godunov-state golden matrices and
  materials of tests/Model/Values.h:24-445, 475-477;
* ``global_matrix_checksums.json`` -- per-matrix sums / norms of
  codegen/matrices/matrices_{4,10,20,35,56,84,120}.xml;
* ``dr_stroud.json`` -- quadrature points, weights, the resampling matrix and
  the face evaluation matrices V3mTo2n / V3mTo2nTWDivM of
  codegen/matrices/dr_stroud_matrices_{4,5,6}.json (checksums + order-4 data);
* ``friction_common.json`` -- the known-answer case of
  tests/DynamicRupture/FrictionLaws/FrictionSolverCommon.t.h:20-170: inputs
  (impedances 10/20/15/25, qPlus = o+q+p, qMinus = 2(o+q+p), traction
  results o+p and 2(o+p), time weights 1..O) and the expected pre- and
  post-computed stresses, evaluated here at O = 4 time points x 25 points.

Usage: python tests/golden/make_fixtures.py [REFERENCE_ROOT]
"""
from __future__ import annotations

import json
import re
import sys
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent
NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def numbers(text: str) -> list[float]:
    return [float(x) for x in re.findall(NUM, text)]


def _arith(node):
    import ast
    if isinstance(node, ast.Expression):
        return _arith(node.body)
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return float(node.value)
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _arith(node.operand)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.Div)):
        a, b = _arith(node.left), _arith(node.right)
        return {ast.Add: a + b, ast.Sub: a - b, ast.Mult: a * b, ast.Div: a / b if b else float("nan")}[type(node.op)]
    raise ValueError("not a plain arithmetic literal")


def entries(text: str) -> list[float]:
    """Comma-separated numeric entries of a brace initialiser; entries may be
    simple arithmetic such as ``1 - 0.25`` (evaluated without eval())."""
    import ast
    flat = text.replace("{", " ").replace("}", " ")
    return [_arith(ast.parse(tok.strip(), mode="eval")) for tok in flat.split(",") if tok.strip()]


def braced_block(text: str, start: int) -> str:
    """Text of the brace-balanced block starting at the first '{' after start."""
    i = text.index("{", start)
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
            if depth == 0:
                return text[i:j + 1]
    raise ValueError("unbalanced braces")


def basis_values(ref: Path) -> dict:
    src = (ref / "tests/Numerical/BasisFunctions.t.h").read_text()
    first = src.index("precomputedValues")
    values = numbers(braced_block(src, first))
    second = src.index("precomputedValues", src.index("Sampled Derivatives Functions"))
    block = braced_block(src, second)
    inner = [braced_block(block, m.start()) for m in re.finditer(r"\{\s*" + NUM, block)]
    # the three innermost vectors (one per direction)
    grads = [numbers(b) for b in inner[:3]] if len(inner) >= 3 else None
    if grads is None or any(len(g) != len(grads[0]) for g in grads):
        # fall back to splitting the flat list into three equal parts
        flat = numbers(block)
        k = len(flat) // 3
        grads = [flat[:k], flat[k:2 * k], flat[2 * k:]]
    return {"point": [0.3, 0.3, 0.3], "values": values, "gradients": grads}


def godunov(ref: Path) -> dict:
    src = (ref / "tests/Model/Values.h").read_text()
    start = src.index("struct ElasticSolutionData")
    end = src.index("struct SolutionData<model::ViscoElasticMaterial>")
    sect = src[start:end]
    out = {}
    for name in ("SolutionHomogeneousLocal", "SolutionHomogeneousNeighbor", "SolutionHeterogeneousLocal",
                 "SolutionHeterogeneousNeighbor"):
        vals = entries(braced_block(sect, sect.index(name)))
        arr = np.array(vals).reshape(9, 9)  # arr[j][i] = Q(i, j)
        out[name] = arr.T.tolist()          # stored as [row][col]
    m = re.search(r"struct SolutionData<model::ElasticMaterial>.*?MaterialVal1\{([^}]*)\}.*?MaterialVal2\{([^}]*)\}",
                  src, re.S)
    # material vectors are (rho, mu, lambda); store as (rho, lambda, mu)
    v1, v2 = numbers(m.group(1)), numbers(m.group(2))
    out["MaterialVal1"] = [v1[0], v1[2], v1[1]]
    out["MaterialVal2"] = [v2[0], v2[2], v2[1]]
    out["tolerance"] = 1e2 * np.finfo(float).eps
    return out


def parse_xml(path: Path) -> dict:
    d = {}
    for m in ET.parse(path).getroot():
        r, c = int(m.get("rows")), int(m.get("columns"))
        a = np.zeros((r, c))
        for e in m:
            a[int(e.get("row")) - 1, int(e.get("column")) - 1] = float(e.get("value"))
        d[m.get("name")] = a
    return d


def checksums(a: np.ndarray) -> dict:
    return {"shape": list(a.shape), "sum": float(a.sum()), "abs_sum": float(np.abs(a).sum()),
            "frobenius": float(np.sqrt((a * a).sum())), "max_abs": float(np.abs(a).max()),
            "weighted": float((a * np.arange(1, a.size + 1).reshape(a.shape)).sum()),
            "nnz": int((np.abs(a) > 1e-13 * max(1.0, np.abs(a).max())).sum())}


def global_matrix_checksums(ref: Path) -> dict:
    out = {}
    for order in range(2, 9):
        B = order * (order + 1) * (order + 2) // 6
        mats = parse_xml(ref / f"codegen/matrices/matrices_{B}.xml")
        out[str(order)] = {name: checksums(a) for name, a in mats.items()
                           if name.split("(")[0] in ("kDivM", "kDivMT", "rDivM", "fMrT", "rT", "fP")}
    return out


def dr_stroud(ref: Path) -> dict:
    out = {}
    for order in (4, 5, 6):
        data = json.loads((ref / f"codegen/matrices/dr_stroud_matrices_{order}.json").read_text())
        mats = {}
        for m in data:
            if m["name"] not in ("quadpoints", "quadweights", "resample") and not m["name"].startswith("V3mTo2n"):
                continue
            a = np.zeros((m["rows"], m["columns"]))
            for r, c, v in m["entries"]:
                a[r - 1, c - 1] = float(v)
            mats[m["name"]] = a
        entry = {k: checksums(v) for k, v in mats.items()}
        if order == 4:
            entry["data"] = {k: v.tolist() for k, v in mats.items() if not k.startswith("V3mTo2n")}
        out[str(order)] = entry
    return out


def friction_common() -> dict:
    """Known-answer data of FrictionSolverCommon.t.h (values only)."""
    O, P = 4, 25
    zp, zs, zpn, zsn = 10.0, 20.0, 15.0, 25.0
    etaP, etaS = zp * zpn / (zp + zpn), zs * zsn / (zs + zsn)
    imp = [zp, zs, zpn, zsn, etaP, etaS, 1 / etaS, 1 / zp, 1 / zs, 1 / zpn, 1 / zsn]
    o, q, p = np.meshgrid(np.arange(O), np.arange(9), np.arange(P), indexing="ij")
    qP = (o + q + p).astype(float)
    qM = 2.0 * (o + q + p)
    oo, pp = np.meshgrid(np.arange(O), np.arange(P), indexing="ij")
    t1 = (oo + pp).astype(float)
    t2 = 2.0 * (oo + pp)
    w = np.arange(1, O + 1, dtype=float)
    normal = etaP * (qM[:, 6] - qP[:, 6] + qP[:, 0] / zp + qM[:, 0] / zpn)
    tr1 = etaS * (qM[:, 7] - qP[:, 7] + qP[:, 3] / zs + qM[:, 3] / zsn)
    tr2 = etaS * (qM[:, 8] - qP[:, 8] + qP[:, 5] / zs + qM[:, 5] / zsn)
    # the post-compute subcase runs with zero fault normal stress
    ns = np.zeros((O, P))
    minus = np.zeros((9, P))
    plus = np.zeros((9, P))
    for k in range(O):
        minus[0] += w[k] * ns[k]; minus[3] += w[k] * t1[k]; minus[5] += w[k] * t2[k]
        minus[6] += w[k] * (qM[k, 6] - (ns[k] - qM[k, 0]) / zpn)
        minus[7] += w[k] * (qM[k, 7] - (t1[k] - qM[k, 3]) / zsn)
        minus[8] += w[k] * (qM[k, 8] - (t2[k] - qM[k, 5]) / zsn)
        plus[0] += w[k] * ns[k]; plus[3] += w[k] * t1[k]; plus[5] += w[k] * t2[k]
        plus[6] += w[k] * (qP[k, 6] + (ns[k] - qP[k, 0]) / zp)
        plus[7] += w[k] * (qP[k, 7] + (t1[k] - qP[k, 3]) / zs)
        plus[8] += w[k] * (qP[k, 8] + (t2[k] - qP[k, 5]) / zs)
    return {"order": O, "points": P, "imp_and_eta": imp, "q_plus": qP.tolist(), "q_minus": qM.tolist(),
            "traction1": t1.tolist(), "traction2": t2.tolist(), "time_weights": w.tolist(),
            "expected_normal": normal.tolist(), "expected_traction1": tr1.tolist(),
            "expected_traction2": tr2.tolist(), "expected_imposed_plus": plus.tolist(),
            "expected_imposed_minus": minus.tolist(), "epsilon": 1e6 * np.finfo(float).eps}


def main(argv):
    ref = Path(argv[1]) if len(argv) > 1 else Path("/root/reference")
    (OUT / "basis_values.json").write_text(json.dumps(basis_values(ref), indent=1))
    (OUT / "godunov_elastic.json").write_text(json.dumps(godunov(ref), indent=1))
    (OUT / "global_matrix_checksums.json").write_text(json.dumps(global_matrix_checksums(ref), indent=1))
    (OUT / "dr_stroud.json").write_text(json.dumps(dr_stroud(ref), indent=1))
    (OUT / "friction_common.json").write_text(json.dumps(friction_common()))
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main(sys.argv)
