"""
tad_dftd4_amd: the DFT-D4 dispersion energy and gradient on AMD MI355X.

Drop-in for the reference's ``dftd4(numbers, positions, charge, param)``
entry point (tad-dftd4 ``src/tad_dftd4/disp.py:44-146``), backed by
hand-written FP64 HIP kernels for gfx950 behind a C ABI
(``include/d4amd.h``).

>>> import torch, tad_dftd4_amd as d4
>>> numbers = torch.tensor([6, 1, 1, 1, 1], device="cuda")
>>> energy = d4.dftd4(numbers, positions, 0.0, {"s8": 1.2, "a1": 0.4, "a2": 5.0})
"""

from .__version__ import __version__
from .cutoff import Cutoff
from .damping import Damping, Param, RationalDamping, ZeroDamping
from .disp import dftd4, dftd4_energy_gradient, dftd4_hessian, get_properties
from .model import D4Model, D4SModel
from .ncoord import cn_d4, erf_count

__all__ = [
    "__version__",
    "Cutoff",
    "D4Model",
    "D4SModel",
    "Damping",
    "Param",
    "RationalDamping",
    "ZeroDamping",
    "cn_d4",
    "dftd4",
    "dftd4_energy_gradient",
    "dftd4_hessian",
    "erf_count",
    "get_properties",
]
