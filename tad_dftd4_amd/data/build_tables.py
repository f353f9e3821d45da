"""
Extract the D4 reference-system tables from the reference's Python sources
*as text* and store them as a plain ``.npz`` data file.

Nothing from the reference is imported or executed: each table is located by
its assignment (``name = torch.tensor(``), comments are stripped, and the list
literal is read with :func:`ast.literal_eval`. The resulting arrays are
physics parameters (reference polarisabilities, reference coordination
numbers and charges, quadrature weights), not code.

Sources (paths relative to the reference checkout):

- ``src/tad_dftd4/reference/d4/params.py``: ``refcovcn`` (:40), ``refalpha``
  (:1142), ``refascale`` (:19556), ``refscount`` (:20498), ``refsys``
  (:21440), ``refc`` (:21550), ``secscale`` (:21659), ``secalpha`` (:21683)
- ``src/tad_dftd4/reference/d4/charge_eeq.py``: ``clsq`` (:29), ``clsh`` (:971)
- ``src/tad_dftd4/data/r4r2.py``: ``_r4_over_r2`` (:55-85)
- ``src/tad_dftd4/utils.py``: Casimir-Polder weights of ``trapzd`` (:52-80)
- ``src/tad_dftd4/data/wfpair.py``: ``_wfpair`` (:40), the pairwise Gaussian
  weighting factors of the D4S model, 119 x 119

Usage::

    python -m tad_dftd4_amd.data.build_tables /root/reference
"""

from __future__ import annotations

import ast
import hashlib
import re
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).with_name("d4_ref_tables.npz")


def _strip_comments(text: str) -> str:
    return "\n".join(line.split("#", 1)[0] for line in text.splitlines())


def _balanced(text: str, start: int) -> str:
    """Return ``text[start:end]`` where ``text[start]`` is an opening
    bracket and ``end`` is one past its matching closing bracket."""
    pairs = {"(": ")", "[": "]"}
    stack = []
    for pos in range(start, len(text)):
        ch = text[pos]
        if ch in pairs:
            stack.append(pairs[ch])
        elif ch in (")", "]"):
            if not stack or stack.pop() != ch:
                raise ValueError(f"unbalanced bracket at {pos}")
            if not stack:
                return text[start : pos + 1]
    raise ValueError("unterminated bracket")


def _tensor_literal(src: str, name: str) -> list:
    clean = _strip_comments(src)
    m = re.search(rf"^{name}\s*=\s*torch\.tensor\(", clean, re.M)
    if m is None:
        raise KeyError(name)
    call = _balanced(clean, m.end() - 1)
    first = call.index("[")
    return ast.literal_eval(_balanced(call, first))


def _list_literal(src: str, name: str) -> list:
    clean = _strip_comments(src)
    m = re.search(rf"{name}\s*=\s*\[", clean)
    if m is None:
        raise KeyError(name)
    return ast.literal_eval(_balanced(clean, m.end() - 1))


def _trapzd_weights(src: str) -> list:
    clean = _strip_comments(src)
    m = re.search(r"def trapzd\(", clean)
    m2 = re.search(r"weights\s*=\s*torch\.tensor\(", clean[m.end() :])
    start = m.end() + m2.end() - 1
    call = _balanced(clean, start)
    return ast.literal_eval(_balanced(call, call.index("[")))


def build(reference_root: str | Path) -> dict[str, np.ndarray]:
    root = Path(reference_root) / "src" / "tad_dftd4"
    params = (root / "reference" / "d4" / "params.py").read_text()
    charge = (root / "reference" / "d4" / "charge_eeq.py").read_text()
    r4r2 = (root / "data" / "r4r2.py").read_text()
    utils = (root / "utils.py").read_text()
    wfpair = (root / "data" / "wfpair.py").read_text()

    t: dict[str, np.ndarray] = {}
    for name in ("refcovcn", "refalpha", "refascale", "refscount"):
        t[name] = np.asarray(_tensor_literal(params, name), dtype=np.float64)
    for name in ("refsys", "refc"):
        t[name] = np.asarray(_tensor_literal(params, name), dtype=np.int64)
    t["secscale"] = np.asarray(
        _tensor_literal(params, "secscale"), dtype=np.float64
    ).reshape(-1)
    t["secalpha"] = np.asarray(_tensor_literal(params, "secalpha"), np.float64)
    for name in ("clsq", "clsh"):
        t[name] = np.asarray(_tensor_literal(charge, name), dtype=np.float64)
    t["r4_over_r2"] = np.asarray(_list_literal(r4r2, "_r4_over_r2"), np.float64)
    t["cp_weights"] = np.asarray(_trapzd_weights(utils), dtype=np.float64)
    t["wfpair"] = np.asarray(_list_literal(wfpair, "_wfpair"), np.float64).reshape(119, 119)

    expect = {
        "refcovcn": (104, 7),
        "refalpha": (104, 7, 23),
        "refascale": (104, 7),
        "refscount": (104, 7),
        "refsys": (104, 7),
        "refc": (104, 7),
        "secscale": (18,),
        "secalpha": (18, 23),
        "clsq": (104, 7),
        "clsh": (104, 7),
        "r4_over_r2": (119,),
        "cp_weights": (23,),
        "wfpair": (119, 119),
    }
    for k, shape in expect.items():
        if t[k].shape != shape:
            raise ValueError(f"{k}: shape {t[k].shape} != {shape}")
    return t


def digest(tables: dict[str, np.ndarray]) -> str:
    h = hashlib.sha256()
    for k in sorted(tables):
        h.update(k.encode())
        h.update(np.ascontiguousarray(tables[k]).tobytes())
    return h.hexdigest()


def main(argv: list[str]) -> int:
    ref = argv[1] if len(argv) > 1 else "/root/reference"
    tables = build(ref)
    np.savez(OUT, **tables)
    print(f"wrote {OUT} sha256={digest(tables)}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
