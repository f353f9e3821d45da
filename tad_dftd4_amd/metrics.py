"""
Algorithmic work counts used for roofline fractions.

Convention (fixed for all reports): adds and multiplies count 1 each (an FMA
counts 2); every exp, erf, sqrt, pow, log and divide counts 1; nu_i is the
number of reference systems of atom i's element. For one molecule of n atoms
with two-body cutoff c2 and three-body cutoff c3:

    F_fwd = sum_{i<j, r_ij <= c2} [64 + 2 (2 nu_i nu_j + nu_i + nu_j)]
            + 40 T + 2/3 (n+1)^3 + 2 (n+1)^2 + 40 sum_i nu_i
    F_E+F = 3 F_fwd

where T counts unordered triples with at least one atom whose two incident
edges are within c3. The three-body kernel's share is 40 T (energy) or
120 T (energy + forces). Algorithmic HBM bytes per molecule: inputs
n (8 + 24) + 8 and outputs 8 n (+ 24 n with forces).
"""

from __future__ import annotations

import numpy as np

from .tables import device_tables

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0    # MI355X HBM3E


def triples(positions: np.ndarray, c3: float) -> int:
    """Unordered triples with at least one 'middle' atom (both incident
    edges within c3)."""
    n = len(positions)
    if n < 3:
        return 0
    d2 = ((positions[:, None, :] - positions[None, :, :]) ** 2).sum(-1)
    inside = d2 <= c3 * c3
    np.fill_diagonal(inside, False)
    if inside.sum() == n * (n - 1):
        return n * (n - 1) * (n - 2) // 6
    # sum over middle atoms of C(deg, 2) counts a triple once per middle
    # atom; a triple with two middles has all three edges inside, so it is a
    # triangle and is counted three times
    a = inside.astype(np.float64)
    k = a.sum(1)
    paths = float((k * (k - 1) / 2).sum())
    tri3 = float(np.trace(a @ a @ a)) / 6.0
    deg_pairs = int(round(paths - 2.0 * tri3))
    return deg_pairs


def molecule_flops(numbers: np.ndarray, positions: np.ndarray, c2: float, c3: float,
                   with_forces: bool = True) -> dict:
    t = device_tables()
    z = numbers[numbers != 0]
    p = positions[numbers != 0]
    n = len(z)
    nu = t.nref[z].astype(np.float64)
    d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1)
    iu = np.triu_indices(n, 1)
    keep = d2[iu] <= c2 * c2
    ni, nj = nu[iu[0][keep]], nu[iu[1][keep]]
    pairs = float(np.sum(64.0 + 2.0 * (2.0 * ni * nj + ni + nj)))
    tr = triples(p, c3)
    eeq = 2.0 / 3.0 * (n + 1) ** 3 + 2.0 * (n + 1) ** 2
    fwd = pairs + 40.0 * tr + eeq + 40.0 * nu.sum()
    mult = 3.0 if with_forces else 1.0
    return {
        "total": mult * fwd,
        "atm": mult * 40.0 * tr,
        "triples": tr,
        "bytes": n * 32.0 + 8.0 + n * (32.0 if with_forces else 8.0),
    }


def workload_flops(mols, c2: float, c3: float, with_forces: bool = True) -> dict:
    acc = {"total": 0.0, "atm": 0.0, "triples": 0, "bytes": 0.0}
    for m in mols:
        f = molecule_flops(m.numbers, m.positions, c2, c3, with_forces)
        for k in acc:
            acc[k] += f[k]
    return acc
