"""
Coordination-number functions that ``dftd4`` accepts as its defaults
(the reference takes them from tad-mctc: ``cn_d4`` and ``erf_count``,
``disp.py:30,56-57``).

``cn_d4`` computes the electronegativity-weighted erf-counting D4
coordination number with the HIP kernels; ``erf_count`` is the counting
function 1/2 (1 + erf(-k (r/r0 - 1))) as an elementwise tensor op.
"""

from __future__ import annotations

import torch

from . import defaults

__all__ = ["cn_d4", "erf_count"]


def erf_count(r: torch.Tensor, r0: torch.Tensor, kcn: float = defaults.D4_KCN) -> torch.Tensor:
    """Error-function counting function."""
    return 0.5 * (1.0 + torch.erf(-kcn * (r / r0 - 1.0)))


def cn_d4(numbers: torch.Tensor, positions: torch.Tensor, **kwargs) -> torch.Tensor:
    """D4 coordination numbers of a (batch of) molecule(s) (HIP kernels).
    Keyword arguments: ``counting_function`` (only ``erf_count``) and
    ``cutoff`` (bohr; default 30, pairs beyond it do not count); the other
    keywords of tad-mctc's ``cn_d4`` (radii, electronegativities, constants)
    are not supported."""
    cf = kwargs.pop("counting_function", erf_count)
    cutoff = kwargs.pop("cutoff", None)
    if cf is not erf_count or kwargs:
        raise NotImplementedError("cn_d4 supports only the erf counting function and a cutoff")
    from . import runtime
    from .damping import RationalDamping
    from .disp import _prepare, kernel_params

    model_inst, nums, pos, chg, _, _, lead, n = _prepare(
        numbers, positions.detach(), 0.0, "d4", None, None, None, None, cn_d4,
        erf_count, RationalDamping(),
    )
    kp = kernel_params({"a1": 0.4, "a2": 5.0}, None, model_inst)
    if cutoff is not None:
        kp.cutoff_cn = float(cutoff)
    cn = runtime.properties(nums, pos, chg, kp, cn_only=True)[0]
    return cn.reshape(*lead, n).to(positions.dtype)
