"""
Real-space cutoffs (mirrors ``Cutoff``, reference ``src/tad_dftd4/cutoff.py:36-90``).

As in the reference, ``cn`` is stored but the D4 coordination number uses its
own fixed cutoff (the reference never forwards ``Cutoff.cn`` to the CN
function, ``dispersion/base.py:390``).
"""

from __future__ import annotations

import torch

from . import defaults

__all__ = ["Cutoff"]


def _as_tensor(x, device=None, dtype=None) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype) if (device or dtype) else x
    return torch.tensor(x, device=device, dtype=dtype or torch.get_default_dtype())


class Cutoff:
    """Collection of real-space cutoffs (bohr)."""

    __slots__ = ("disp2", "disp3", "cn", "cn_eeq", "device", "dtype")

    def __init__(
        self,
        disp2=defaults.D4_DISP2_CUTOFF,
        disp3=defaults.D4_DISP3_CUTOFF,
        cn=defaults.D4_CN_CUTOFF,
        cn_eeq=defaults.D4_CN_EEQ_CUTOFF,
        device: torch.device | None = None,
        dtype: torch.dtype | None = None,
    ) -> None:
        self.device = device
        self.dtype = dtype
        self.disp2 = _as_tensor(disp2, device, dtype)
        self.disp3 = _as_tensor(disp3, device, dtype)
        self.cn = _as_tensor(cn, device, dtype)
        self.cn_eeq = _as_tensor(cn_eeq, device, dtype)

    def __repr__(self) -> str:
        return (
            f"Cutoff(disp2={float(self.disp2)}, disp3={float(self.disp3)}, "
            f"cn={float(self.cn)}, cn_eeq={float(self.cn_eeq)})"
        )
