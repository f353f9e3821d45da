"""
Write ``reference_vectors.json``: golden vectors copied *as text* from the
reference checkout (nothing is imported or executed; list literals are read
with ``ast.literal_eval``).

Sources (relative to the reference checkout):

- ``examples/single.py:7-27`` geometry, ``:33-39`` TPSSh parameters,
  ``:52-67`` atom energies (computed by the reference in float32, printed
  with 10 decimals).
- ``examples/batch.py:8-45`` formamide dimer / formamide geometries;
  totals from the docstring of ``src/tad_dftd4/__init__.py:84-87``.
- ``test/test_d4/samples.py`` LiH and SiH4 ``q``, ``disp``, ``disp2``,
  ``disp3`` (dftd4 Fortran test suite values).
- ``test/test_grad/samples_grad.py`` LiH and SiH4 ``grad`` (tblite, TPSS0
  parameters of ``test/test_grad/test_pos.py:55-62``).
- ``test/test_grad/samples_hessian.py`` LiH and SiH4 ``hessian`` (dftd4 3.6.0,
  finite differences of the analytic gradient, TPSS0 parameters of
  ``test/test_grad/test_hessian.py:84-90``; flat, Fortran order of
  ``(nat, 3, nat, 3)``, ``test_hessian.py:92-95``; reference tolerance 1e-7).
- ``test/test_model/samples.py`` LiH and SiH4 ``gw`` (Fortran-ordered
  ``(nref, nat)``) and ``c6``; the D4S ``gw_d4s`` (Fortran-ordered
  ``(nref, nat, nat)``) and ``c6_d4s``.
- ``test/test_d4/samples.py`` LiH and SiH4 ``disp2_d4s`` (two-body energy
  with the D4S model).
- ``test/test_model/samples.py`` MB16_43_01, MB16_43_02, MB16_43_03 ``q``,
  ``gw``, ``c6`` (and the D4S ``gw_d4s``, ``c6_d4s`` where present); these
  molecules' atomic numbers live with their geometries in tad-mctc (absent)
  and are inferred by the tests from the values themselves.

The LiH and SiH4 geometries live in tad-mctc (absent); they are restated
below from tad-mctc's molecule set (``data/molecules.py``: LiH bond along z,
SiH4 tetrahedral). Their correctness is checked by the test suite: the
oracle reproduces the tblite gradient fixtures to ~1e-15 with them.

Usage::

    python tests/golden/make_fixtures.py /root/reference
"""

from __future__ import annotations

import ast
import json
import re
import sys
from pathlib import Path

OUT = Path(__file__).with_name("reference_vectors.json")

SYMBOLS = {"H": 1, "Li": 3, "C": 6, "N": 7, "O": 8, "Si": 14, "S": 16}

_A = 1.61768389755830
RESTATED_GEOMETRY = {
    "LiH": {
        "numbers": [3, 1],
        "positions": [[0.0, 0.0, -1.50796743897235], [0.0, 0.0, 1.50796743897235]],
    },
    "SiH4": {
        "numbers": [14, 1, 1, 1, 1],
        "positions": [
            [0.0, 0.0, 0.0],
            [_A, _A, -_A],
            [-_A, -_A, -_A],
            [_A, -_A, _A],
            [-_A, _A, _A],
        ],
    },
}


# What the MB16-43 ``gw`` fixtures were computed from (reference
# test/test_model/test_weights.py:79-91): 01 with the geometry's CN and q = 0,
# 02 with CN = 0 and the fixture q, 03 with both; ``c6`` of all three is the
# C6 matrix of the molecule at its CN and charges (test_c6.py:32-33 uses 03).
MB16_GW_INPUTS = {"MB16_43_01": "cn, q=0", "MB16_43_02": "cn=0, q", "MB16_43_03": "cn, q"}


def _strip(text: str) -> str:
    return "\n".join(line.split("#", 1)[0] for line in text.splitlines())


def _balanced(text: str, start: int) -> str:
    pairs = {"(": ")", "[": "]"}
    stack = []
    for pos in range(start, len(text)):
        ch = text[pos]
        if ch in pairs:
            stack.append(pairs[ch])
        elif ch in (")", "]"):
            stack.pop()
            if not stack:
                return text[start : pos + 1]
    raise ValueError("unterminated")


def _first_list(text: str, start: int):
    i = text.index("[", start)
    return ast.literal_eval(_balanced(text, i))


def _block(text: str, key: str) -> str:
    """Text of the ``"key": Refs(...)`` block."""
    m = re.search(rf'"{re.escape(key)}":\s*Refs\(', text)
    return _balanced(text, m.end() - 1)


def _entry(block: str, field: str):
    m = re.search(rf'"{field}":\s*', block)
    rest = block[m.end() :]
    if rest.startswith("reshape_fortran("):
        call = _balanced(rest, len("reshape_fortran"))
        values = _first_list(call, 0)
        shape_m = re.search(r"\),\s*(\([0-9,\s]+\))\s*,?\s*\)$", call)
        shape = list(ast.literal_eval(shape_m.group(1)))
        return {"values": values, "fortran_shape": shape}
    return _first_list(rest, 0)


def build(ref: Path) -> dict:
    ex_single = _strip((ref / "examples" / "single.py").read_text())
    ex_batch = _strip((ref / "examples" / "batch.py").read_text())
    init_doc = (ref / "src" / "tad_dftd4" / "__init__.py").read_text()
    d4s = _strip((ref / "test" / "test_d4" / "samples.py").read_text())
    grads = _strip((ref / "test" / "test_grad" / "samples_grad.py").read_text())
    hess = _strip((ref / "test" / "test_grad" / "samples_hessian.py").read_text())
    models = _strip((ref / "test" / "test_model" / "samples.py").read_text())

    out: dict = {"provenance": __doc__.strip().splitlines()[0]}

    sym = re.search(r'symbols="([A-Za-z ]+)"', ex_single).group(1).split()
    pos = _first_list(ex_single, ex_single.index("positions = torch.tensor("))
    ref_e = _first_list(ex_single, ex_single.index("ref = torch.tensor("))
    par = {
        k: float(re.search(rf"{k}=positions\.new_tensor\(([-0-9.e]+)\)", ex_single).group(1))
        for k in ("s6", "s8", "s9", "a1", "a2")
    }
    out["example_single"] = {
        "numbers": [SYMBOLS[s] for s in sym],
        "positions": pos,
        "param": par,
        "energy": ref_e,
        "note": "reference run in float32; printed to 10 decimals",
    }

    syms = re.findall(r'symbol_to_number\("([A-Za-z ]+)"\.split\(\)\)', ex_batch)
    p1 = _first_list(ex_batch, ex_batch.index("torch.tensor("))
    # the second geometry starts after the first tensor block closes
    first_call = _balanced(ex_batch, ex_batch.index("(", ex_batch.index("torch.tensor(")))
    after = ex_batch.index(first_call) + len(first_call)
    p2 = _first_list(ex_batch, ex_batch.index("torch.tensor(", after))
    totals = ast.literal_eval(
        re.search(r"tensor\((\[-0\.0088341432[^\]]*\])\)", init_doc).group(1)
    )
    out["example_batch"] = {
        "numbers": [[SYMBOLS[s] for s in x.split()] for x in syms],
        "positions": [p1, p2],
        "param": par,
        "total_energy": totals,
        "note": "reference run in float32; printed to 10 decimals",
    }

    for name in ("LiH", "SiH4"):
        b = _block(d4s, name)
        g = _block(grads, name)
        m = _block(models, name)
        h = _block(hess, name)
        out[name] = {
            **RESTATED_GEOMETRY[name],
            "q": _entry(b, "q"),
            "disp": _entry(b, "disp"),
            "disp2": _entry(b, "disp2"),
            "disp3": _entry(b, "disp3"),
            "grad": _entry(g, "grad"),
            "hessian": _entry(h, "hessian"),
            "gw": _entry(m, "gw"),
            "c6": _entry(m, "c6"),
            "disp2_d4s": _entry(b, "disp2_d4s"),
            "gw_d4s": _entry(m, "gw_d4s"),
            "c6_d4s": _entry(m, "c6_d4s"),
        }
    for name in ("MB16_43_01", "MB16_43_02", "MB16_43_03"):
        m = _block(models, name)
        out[name] = {
            "q": _entry(m, "q"),
            "gw": _entry(m, "gw"),
            "c6": _entry(m, "c6"),
            "gw_d4s": _entry(m, "gw_d4s"),
            "c6_d4s": _entry(m, "c6_d4s"),
            "gw_inputs": MB16_GW_INPUTS[name],
        }
    out["param_tpssh"] = {
        "s6": 1.0, "s8": 1.85897750, "s9": 1.0, "s10": 0.0, "alp": 16.0,
        "a1": 0.44286966, "a2": 4.60230534,
    }
    out["param_tpss0"] = {
        "s6": 1.0, "s8": 1.62438102, "s9": 1.0, "a1": 0.40329022,
        "a2": 4.80537871,
    }
    return out


def main(argv: list[str]) -> int:
    ref = Path(argv[1] if len(argv) > 1 else "/root/reference")
    data = build(ref)
    OUT.write_text(json.dumps(data, indent=1))
    print(f"wrote {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
