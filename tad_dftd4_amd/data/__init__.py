"""
Element and reference-system data for the D4 model.

``ref_tables()`` returns the reference-system tables text-extracted from the
reference checkout by ``build_tables.py`` (stored in ``d4_ref_tables.npz``);
``elements`` holds the restated dependency tables (covalent radii,
electronegativities, hardnesses, effective charges, EEQ-2019 parameters).
"""

from __future__ import annotations

from functools import lru_cache
from pathlib import Path

import numpy as np

from . import elements

__all__ = ["ref_tables", "elements", "TABLE_FILE"]

TABLE_FILE = Path(__file__).with_name("d4_ref_tables.npz")


@lru_cache(maxsize=1)
def ref_tables() -> dict[str, np.ndarray]:
    """Reference-system tables (read-only numpy arrays)."""
    with np.load(TABLE_FILE, allow_pickle=False) as f:
        out = {k: f[k] for k in f.files}
    for v in out.values():
        v.setflags(write=False)
    return out
