"""
Host-side construction of the element tables that the kernels keep resident
in HBM (one upload per device and model constants).

The reference rebuilds per-atom tables on every call: reference
polarisabilities with charge scaling (``model/base.py:367-418``) and the
per-atom-pair reference C6 by Casimir-Polder quadrature
(``model/base.py:420-431``, ``utils.py:33-94``). Both depend only on the
element, so they are built here once per element pair:
``RC6[Zi, Zj, a, b]`` for Z < 104 (4.2 MB).

Reference systems of each element are compacted to the front (the reference
marks unused slots with ``refc == 0``; a few actinides have gaps). Sums over
references are invariant under this reordering; the exceptional-weight
fallback compares against the maximum reference CN over all seven original
slots (``model/d4.py:209-217``), kept as ``maxcn``.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache

import numpy as np

from .data import elements as el
from .data import ref_tables

MAX_Z = 104
MAX_REF = 7
GA_DEFAULT = 3.0
GC_DEFAULT = 2.0
WF_DEFAULT = 6.0
_EPS = np.finfo(np.float64).eps


def _zeta(ga: float, gam, qref, qmod):
    """Charge scaling (model/base.py:308-335)."""
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        scale = np.exp(gam * (1.0 - qref / (qmod - _EPS)))
        out = np.exp(ga * (1.0 - scale))
    return np.where(qmod > 0.0, out, np.exp(ga))


def reference_alpha(ga: float = GA_DEFAULT, gc: float = GC_DEFAULT) -> np.ndarray:
    """Reference polarisabilities alpha[Z, ref, w] (original slot order),
    eeq reference charges."""
    t = ref_tables()
    refsys = t["refsys"]
    mask = refsys > 0
    zeff = el.ZEFF[refsys]
    gam = el.GAM[refsys] * gc
    z = np.where(mask, _zeta(ga, gam, zeff, t["clsh"] + zeff), 0.0)
    aiw = t["secscale"][refsys][..., None] * t["secalpha"][refsys] * z[..., None]
    h = t["refalpha"] - t["refscount"][..., None] * aiw
    alpha = t["refascale"][..., None] * h
    return np.where(alpha > 0.0, alpha, 0.0)


@dataclass(frozen=True)
class DeviceTables:
    rcov: np.ndarray
    en: np.ndarray
    r4r2: np.ndarray
    eeq_chi: np.ndarray
    eeq_eta: np.ndarray
    eeq_kcn: np.ndarray
    eeq_rad: np.ndarray
    gam: np.ndarray
    zeff: np.ndarray
    maxcn: np.ndarray
    nref: np.ndarray
    ngw: np.ndarray
    refcn: np.ndarray
    refq: np.ndarray
    alpha0: np.ndarray
    alpha: np.ndarray  # (Z, ref, 23) reference alpha(i w), compacted slots
    rc6: np.ndarray
    wfpair: np.ndarray  # D4S pairwise Gaussian weighting factors (Z, Z)
    perm: np.ndarray  # compacted slot -> original slot (-1: unused)
    refslot: np.ndarray  # perm as int32 (the C ABI's refslot table)


def r4r2_table() -> np.ndarray:
    """sqrt(1/2 <r4>/<r2> sqrt(Z)) for Z = 0..118 (data/r4r2.py:84-88)."""
    raw = ref_tables()["r4_over_r2"]
    return np.sqrt(0.5 * (raw * np.sqrt(np.arange(len(raw), dtype=np.float64))))


@lru_cache(maxsize=8)
def device_tables(ga: float = GA_DEFAULT, gc: float = GC_DEFAULT) -> DeviceTables:
    t = ref_tables()
    refc = t["refc"]
    perm = np.full((MAX_Z, MAX_REF), -1, dtype=np.int64)
    nref = np.zeros(MAX_Z, dtype=np.int32)
    for z in range(MAX_Z):
        used = np.nonzero(refc[z] > 0)[0]
        nref[z] = len(used)
        perm[z, : len(used)] = used

    def compact(a: np.ndarray) -> np.ndarray:
        out = np.zeros((MAX_Z, MAX_REF) + a.shape[2:], dtype=a.dtype)
        for z in range(MAX_Z):
            k = nref[z]
            out[z, :k] = a[z, perm[z, :k]]
        return out

    alpha = compact(reference_alpha(ga, gc))  # (Z, ref, 23)
    w = t["cp_weights"]
    thopi = 3.0 / np.pi
    rc6 = thopi * np.einsum("iaw,jbw->ijab", alpha * w, alpha, optimize=True)

    def c(x, dtype=np.float64):
        return np.ascontiguousarray(x, dtype=dtype)

    return DeviceTables(
        rcov=c(el.COV_D3[:MAX_Z]),
        en=c(el.PAULING_EN[:MAX_Z]),
        r4r2=c(r4r2_table()[:MAX_Z]),
        eeq_chi=c(el.EEQ_CHI[:MAX_Z]),
        eeq_eta=c(el.EEQ_ETA[:MAX_Z]),
        eeq_kcn=c(el.EEQ_KCN[:MAX_Z]),
        eeq_rad=c(el.EEQ_RAD[:MAX_Z]),
        gam=c(el.GAM[:MAX_Z]),
        zeff=c(el.ZEFF[:MAX_Z]),
        maxcn=c(t["refcovcn"].max(axis=1)),
        nref=c(nref, np.int32),
        ngw=c(compact(refc), np.int32),
        refcn=c(compact(t["refcovcn"])),
        refq=c(compact(t["clsq"])),
        alpha0=c(alpha[..., 0]),
        alpha=c(alpha),
        rc6=c(rc6),
        wfpair=c(t["wfpair"][:MAX_Z, :MAX_Z]),
        perm=perm,
        refslot=c(perm, np.int32),
    )
