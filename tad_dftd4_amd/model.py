"""
``D4Model`` and ``D4SModel``: the D4 and D4S dispersion models' constants
(mirror reference ``src/tad_dftd4/model/d4.py``, ``model/d4s.py`` and
``model/base.py:107-151``).

The reference rebuilds per-atom reference polarisabilities and the
per-atom-pair reference C6 tensor on every call. Here the model only carries
its constants (ga, gc, wf, reference charges); the kernels read element
tables that stay resident in HBM (see ``tables.py``). ``rc6`` is still
available as a lazily materialised (..., n, n, 7, 7) tensor in the
reference's slot order.
"""

from __future__ import annotations

import numpy as np
import torch

from .tables import GA_DEFAULT, GC_DEFAULT, MAX_REF, WF_DEFAULT, device_tables

__all__ = ["D4Model", "D4SModel"]


class D4Model:
    """The D4 dispersion model."""

    __slots__ = ("numbers", "ga", "gc", "wf", "ref_charges", "device", "dtype", "_rc6")

    def __init__(
        self,
        numbers: torch.Tensor,
        ga: float = GA_DEFAULT,
        gc: float = GC_DEFAULT,
        wf: torch.Tensor | float | None = None,
        ref_charges: str = "eeq",
        rc6: torch.Tensor | None = None,
        device: torch.device | None = None,
        dtype: torch.dtype | None = None,
    ) -> None:
        if ref_charges not in ("eeq", "gfn2"):
            raise ValueError(f"Unknown reference charges: {ref_charges}")
        if ref_charges != "eeq":
            raise NotImplementedError(
                "only EEQ reference charges are implemented by the MI355X kernels"
            )
        if rc6 is not None:
            raise NotImplementedError(
                "a user-supplied rc6 tensor is not supported: the kernels read the "
                "element-pair reference C6 table resident on the device"
            )
        self.numbers = numbers
        self.ga = float(ga)
        self.gc = float(gc)
        self.wf = WF_DEFAULT if wf is None else float(wf)
        self.ref_charges = ref_charges
        self.device = device if device is not None else numbers.device
        self.dtype = dtype if dtype is not None else torch.get_default_dtype()
        self._rc6 = None

    @property
    def rc6(self) -> torch.Tensor:
        """Reference C6 of all atom pairs, shape (..., n, n, 7, 7), in the
        reference's slot order (unused slots are zero)."""
        if self._rc6 is None:
            t = device_tables(self.ga, self.gc)
            full = np.zeros((t.rc6.shape[0], t.rc6.shape[0], MAX_REF, MAX_REF))
            for zi in range(t.rc6.shape[0]):
                pi = t.perm[zi, : t.nref[zi]]
                for zj in range(t.rc6.shape[0]):
                    pj = t.perm[zj, : t.nref[zj]]
                    full[zi, zj][np.ix_(pi, pj)] = t.rc6[zi, zj, : len(pi), : len(pj)]
            table = torch.as_tensor(full, device=self.device, dtype=self.dtype)
            z = self.numbers.to(self.device)
            self._rc6 = table[z.unsqueeze(-1), z.unsqueeze(-2)]
        return self._rc6

    def __repr__(self) -> str:
        return (f"{type(self).__name__}(ga={self.ga}, gc={self.gc}, wf={self.wf}, "
                f"ref_charges={self.ref_charges!r})")


class D4SModel(D4Model):
    """The D4S dispersion model (``model/d4s.py:56-285``): the Gaussian
    weights of atom A in the pair AB use the element-pair factor
    wf(Z_A, Z_B) of ``data/wfpair.py`` instead of one global factor, so every
    pair has its own weights (and C6). The factors are resident on the device
    with the other tables; the kernels form the pair weights in the pair
    pass."""

    __slots__ = ()

    def __init__(
        self,
        numbers: torch.Tensor,
        ga: float = GA_DEFAULT,
        gc: float = GC_DEFAULT,
        wf: torch.Tensor | float | None = None,
        ref_charges: str = "eeq",
        rc6: torch.Tensor | None = None,
        device: torch.device | None = None,
        dtype: torch.dtype | None = None,
    ) -> None:
        if wf is not None:
            raise NotImplementedError(
                "D4SModel with user-supplied weighting factors is not supported: "
                "the kernels read the element-pair factor table resident on the device"
            )
        super().__init__(numbers, ga, gc, None, ref_charges, rc6, device, dtype)
