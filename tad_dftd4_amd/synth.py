"""
Seeded synthetic molecules for the benchmark workloads and the tests.

Geometry: for a molecule of n atoms take the first n points of an m x m x m
cubic grid (m = smallest integer with m^3 >= n, x slowest), scale by the
spacing and add U(-0.5, 0.5) bohr jitter per coordinate. Elements are drawn
i.i.d. from a mix; total charges from a set. Coordinates are in bohr, FP64.

Workloads (sizes, mixes and charges as the benchmark definitions state):

- ``single_64``: 1 molecule, 64 atoms, H/C/N/O 50/30/10/10, neutral.
- ``batch_256_8_200``: 256 molecules, n ~ U{8..200}, mix ``MIX_Z54``, Q ~ U{-1,0,1}.
- ``batch_1024_20_120``: 1024 molecules, n ~ U{20..120}, ``MIX_Z54``, Q ~ U{-1,0,1}.
- ``cluster_5000``: 5000 atoms, spacing 2.6, H/C/N/O/S 45/30/10/10/5, neutral.
- ``serving_stream``: lognormal sizes (median 40, sigma 0.8, clipped to 4..2000),
  ``MIX_Z54``, Q ~ U{-2..2}, 70 % energy-only / 30 % energy+forces.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

__all__ = [
    "Molecule",
    "grid_positions",
    "draw_elements",
    "molecule",
    "pack",
    "batch",
    "workload",
    "MIX_HCNO",
    "MIX_Z54",
    "MIX_HCNOS",
]

MIX_HCNO = ((1, 0.50), (6, 0.30), (7, 0.10), (8, 0.10))
MIX_HCNOS = ((1, 0.45), (6, 0.30), (7, 0.10), (8, 0.10), (16, 0.05))
_OTHER = tuple(z for z in range(1, 55) if z not in (1, 6, 7, 8))
MIX_Z54 = ((1, 0.45), (6, 0.25), (7, 0.10), (8, 0.10)) + tuple(
    (z, 0.10 / len(_OTHER)) for z in _OTHER
)


@dataclass
class Molecule:
    numbers: np.ndarray  # (n,) int64
    positions: np.ndarray  # (n, 3) float64, bohr
    charge: float


def _grid_edge(n: int) -> int:
    m = 1
    while m * m * m < n:
        m += 1
    return m


def grid_positions(rng: np.random.Generator, n: int, spacing: float) -> np.ndarray:
    m = _grid_edge(n)
    idx = np.arange(n)
    ix, iy, iz = idx // (m * m), (idx // m) % m, idx % m
    grid = np.stack([ix, iy, iz], axis=1).astype(np.float64) * spacing
    return grid + rng.uniform(-0.5, 0.5, size=(n, 3))


def draw_elements(rng: np.random.Generator, n: int, mix) -> np.ndarray:
    z = np.array([a for a, _ in mix], dtype=np.int64)
    p = np.array([b for _, b in mix], dtype=np.float64)
    return rng.choice(z, size=n, p=p / p.sum())


def molecule(rng, n, mix=MIX_HCNO, spacing=2.8, charges=(0,)) -> Molecule:
    numbers = draw_elements(rng, n, mix)
    positions = grid_positions(rng, n, spacing)
    charge = float(rng.choice(np.asarray(charges)))
    return Molecule(numbers, positions, charge)


def pack(mols, pad_to: int | None = None, device=None):
    """Zero-padded batch tensors (numbers, positions, charge) in the
    reference's batch convention (padding atoms have Z = 0 at 0.0)."""
    nmax = max(len(m.numbers) for m in mols)
    if pad_to is not None:
        nmax = max(nmax, pad_to)
    b = len(mols)
    numbers = np.zeros((b, nmax), dtype=np.int64)
    positions = np.zeros((b, nmax, 3), dtype=np.float64)
    charge = np.zeros((b,), dtype=np.float64)
    for k, m in enumerate(mols):
        numbers[k, : len(m.numbers)] = m.numbers
        positions[k, : len(m.numbers)] = m.positions
        charge[k] = m.charge
    t = lambda a: torch.from_numpy(a).to(device) if device else torch.from_numpy(a)
    return t(numbers), t(positions), t(charge)


def batch(seed: int, count: int, nmin: int, nmax: int, mix=MIX_Z54,
          spacing: float = 2.8, charges=(-1, 0, 1)):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=count)
    return [molecule(rng, int(n), mix, spacing, charges) for n in sizes]


def workload(name: str, seed: int | None = None):
    """Molecules of a named workload (see module docstring)."""
    if name == "single_64":
        rng = np.random.default_rng(0 if seed is None else seed)
        return [molecule(rng, 64, MIX_HCNO, 2.8, (0,))]
    if name == "batch_256_8_200":
        return batch(1 if seed is None else seed, 256, 8, 200)
    if name == "batch_1024_20_120":
        return batch(2 if seed is None else seed, 1024, 20, 120)
    if name == "cluster_5000":
        rng = np.random.default_rng(3 if seed is None else seed)
        return [molecule(rng, 5000, MIX_HCNOS, 2.6, (0,))]
    raise KeyError(name)


def serving_sizes(rng: np.random.Generator, count: int) -> np.ndarray:
    n = np.round(rng.lognormal(np.log(40.0), 0.8, size=count)).astype(np.int64)
    return np.clip(n, 4, 2000)
