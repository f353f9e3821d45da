"""
Damping-function interface (mirrors reference
``src/tad_dftd4/damping/functions.py`` and ``damping/parameters/base.py``).

The MI355X kernels implement Becke-Johnson rational damping for the two-body
term (``functions.py:262-305``) and zero damping of order 9 for the
three-body term (``functions.py:329-378``) inside their fused pair and
triple passes. The classes here identify the damping in
``dftd4(..., damping_function=...)`` (other damping functions are rejected
there with ``NotImplementedError``) and, called as
``damping(distances, radii, order, only_damping=False, **param)``, evaluate
the damping values elementwise on the tensors' device with the reference's
parameter rules (``functions.py:116-169``):

- an order whose prefactor and radius scale are both ``None`` (s6/rs6,
  s8/rs8, s9/rs9; s10 alone for order 10) gives zeros;
- otherwise the keywords a damping function needs for that order must not be
  ``None`` (``TypeError``, ``functions.py:68-114``).
"""

from __future__ import annotations

from typing import Any, TypedDict

import torch

from . import defaults

__all__ = ["Param", "Damping", "RationalDamping", "ZeroDamping"]


class Param(TypedDict, total=False):
    """Damping parameters (``damping/parameters/base.py:48-85``)."""

    s6: Any
    s8: Any
    s9: Any
    s10: Any
    rs6: Any
    rs8: Any
    rs9: Any
    a1: Any
    a2: Any
    alp: Any
    bet: Any
    doi: Any


class Damping:
    """Base interface; equality by class (``functions.py:171-193,214-238``).
    Subclasses give ``_damp`` and, per order, the keywords it needs in
    ``REQUIRED_BY_ORDER``."""

    radius_type: str = "r4r2"
    doi: str | None = None
    REQUIRED_BY_ORDER: dict[int, tuple[str, ...]] = {}

    def __call__(self, distances: torch.Tensor, radii: torch.Tensor, order: int,
                 only_damping: bool = False, *, a1=None, a2=None, s6=defaults.S6,
                 rs6=None, s8=None, rs8=None, s9=defaults.S9, rs9=defaults.RS9,
                 s10=None, alp=defaults.ALP, bet=None, doi: str | None = None) -> torch.Tensor:
        self.doi = doi
        switch = {6: (s6, rs6), 8: (s8, rs8), 9: (s9, rs9), 10: (s10,)}.get(order)
        if switch is not None and all(v is None for v in switch):
            return torch.zeros_like(distances)
        kw = {"a1": a1, "a2": a2, "rs6": rs6, "rs8": rs8, "rs9": rs9, "alp": alp,
              "bet": bet}
        missing = [k for k in self.REQUIRED_BY_ORDER.get(order, ()) if kw[k] is None]
        if missing:
            raise TypeError(f"{type(self).__name__} (order {order}) requires "
                            f"keyword(s): {', '.join(missing)}")
        return self._damp(distances, radii, order, only_damping, **kw)

    def _damp(self, distances, radii, order, only_damping, **kw) -> torch.Tensor:
        raise NotImplementedError(f"{type(self).__name__} gives no damping values")

    def __eq__(self, other: Any) -> bool:
        return isinstance(other, Damping) and self.__class__ is other.__class__

    def __hash__(self) -> int:
        return hash(self.__class__)


def _divide(x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """x / y with exact zeros of y replaced by the dtype's eps (tad-mctc
    ``storch.divide``, used by ``functions.py:371-378``)."""
    y = torch.as_tensor(y, dtype=x.dtype, device=x.device)
    return x / torch.where(y == 0, torch.finfo(x.dtype).eps, y)


class RationalDamping(Damping):
    """Becke-Johnson rational damping 1 / (r^n + (a1 sqrt(Q) + a2)^n), with
    Q the pair's ``radii`` (3 r4r2_i r4r2_j); ``only_damping`` has no effect,
    as in the reference (``functions.py:241-305``)."""

    radius_type = "r4r2"
    REQUIRED_BY_ORDER = {n: ("a1", "a2") for n in (6, 8, 10)}

    def _damp(self, distances, radii, order, only_damping, *, a1, a2, **_):
        r0 = a1 * torch.sqrt(radii) + a2
        return 1.0 / (distances.pow(order) + r0.pow(order))


class ZeroDamping(Damping):
    """Zero (Chai--Head-Gordon) damping f = 1 / (1 + 6 (rs R / R0)^a) with
    a = alp (orders 6 and 9) or alp + 2 (order 8); f / R^n unless
    ``only_damping`` (``functions.py:308-378``)."""

    radius_type = "rvdw"
    REQUIRED_BY_ORDER = {6: ("rs6", "alp"), 8: ("rs8", "alp"), 9: ("rs9", "alp")}

    def _damp(self, distances, radii, order, only_damping, *, rs6, rs8, rs9, alp, **_):
        if order not in (6, 8, 9):
            raise ValueError("Zero-damping is only implemented for order 6, 8, and 9. "
                             f"Got {order} instead.")
        rs = {6: rs6, 8: rs8, 9: rs9}[order]
        a = alp + 2.0 if order == 8 else alp
        f = 1.0 / (1.0 + 6.0 * (rs * _divide(distances, radii)) ** a)
        return f if only_damping else _divide(f, distances**order)
