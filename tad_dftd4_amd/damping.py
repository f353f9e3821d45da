"""
Damping-function interface (mirrors reference
``src/tad_dftd4/damping/functions.py`` and ``damping/parameters/base.py``).

The MI355X kernels implement Becke-Johnson rational damping for the two-body
term (``functions.py:262-305``) and zero damping of order 9 for the
three-body term (``functions.py:329-378``). The classes here identify the
damping in ``dftd4(..., damping_function=...)``; other damping functions are
not part of this build and are rejected with ``NotImplementedError``.
"""

from __future__ import annotations

from typing import Any, TypedDict

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
    """Base interface; equality by class (``functions.py:171-193``)."""

    radius_type: str = "r4r2"

    def __eq__(self, other: Any) -> bool:
        return isinstance(other, Damping) and self.__class__ is other.__class__

    def __hash__(self) -> int:
        return hash(self.__class__)


class RationalDamping(Damping):
    """Becke-Johnson rational damping 1 / (r^n + (a1 sqrt(Q) + a2)^n)."""

    radius_type = "r4r2"


class ZeroDamping(Damping):
    """Zero damping 1 / (1 + 6 (R0 / r)^alp)."""

    radius_type = "rvdw"
