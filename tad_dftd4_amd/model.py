"""
``D4Model`` and ``D4SModel``: the D4 and D4S dispersion models (mirror
reference ``src/tad_dftd4/model/d4.py``, ``model/d4s.py`` and
``model/base.py``).

The reference rebuilds per-atom reference polarisabilities and the
per-atom-pair reference C6 tensor on every call. Here the model carries its
constants (ga, gc, wf, reference charges) and the kernels read element
tables that stay resident in HBM (see ``tables.py``). ``rc6`` is still
available as a lazily materialised (..., n, n, 7, 7) tensor in the
reference's slot order.

The model methods of the reference, ``weight_references``,
``get_atomic_c6``, ``get_weighted_pols`` and ``get_polarizabilities``, run as
HIP kernels through the C ABI (``d4_model_*``) on the device of their tensor
arguments, with the reference's shapes, slot order and return conventions,
and are differentiable (hand-written backward kernels through
``torch.autograd``). They compute in FP64 and return the model's dtype.
"""

from __future__ import annotations

import numpy as np
import torch

from . import runtime
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

    # ------------------------------------------------------------ methods

    _pair = False  # D4S: weights per atom pair

    def _device_numbers(self, *tensors) -> torch.Tensor:
        """Atomic numbers as (B, n) int64 on the device of the first tensor
        argument given (else of the model's numbers)."""
        dev = next((t.device for t in tensors if isinstance(t, torch.Tensor)), None)
        dev = dev if dev is not None else torch.device(self.device)
        if dev.type != "cuda":
            raise RuntimeError(
                "tad_dftd4_amd computes on a ROCm device; move the tensors to 'cuda'"
            )
        z = self.numbers.to(device=dev, dtype=torch.int64)
        if z.numel() and (int(z.min()) < 0 or int(z.max()) >= 104):
            raise ValueError("atomic numbers must lie in 0..103 (0 marks padding)")
        n = z.shape[-1]
        return z.reshape(-1, n).contiguous()

    def _lead(self):
        return tuple(self.numbers.shape[:-1]), self.numbers.shape[-1]

    def weight_references(self, cn: torch.Tensor | None = None, q: torch.Tensor | None = None,
                          *, with_dgwdq: bool = False, with_dgwdcn: bool = False):
        """Weights of the reference systems, shape (..., n, 7) (D4S:
        (..., n, n, 7)); with ``with_dgwdcn`` and/or ``with_dgwdq`` a tuple
        (gw, dgwdcn, dgwdq) in that order, as ``model/d4.py:103-266``. As in
        the reference, ``dgwdcn`` is minus the derivative of gw with respect
        to cn (its own tests compare it with the negated autograd derivative,
        ``test/test_model/test_weights.py:245-264``). ``cn`` and ``q``
        default to zero. Differentiable in cn and q."""
        if self._pair and with_dgwdcn:
            raise NotImplementedError(
                "Analytical derivative of Gaussian weights with respect to CN not implemented."
            )
        z = self._device_numbers(cn, q)
        lead, n = self._lead()
        B = z.shape[0]
        cn2 = None if cn is None else cn.reshape(B, n)
        q2 = None if q is None else q.reshape(B, n)
        gw = _Weights.apply(cn2, q2, z, self)
        shape = (*lead, n, n, MAX_REF) if self._pair else (*lead, n, MAX_REF)
        out = [gw.reshape(shape).to(self.dtype)]
        if not (with_dgwdcn or with_dgwdq):
            return out[0]
        _, dc, dq = runtime.model_weights(z, cn2, q2, self.ga, self.gc, self.wf, self._pair,
                                          dgwdcn=with_dgwdcn, dgwdq=with_dgwdq)
        if with_dgwdcn:
            out.append(dc.reshape(shape).to(self.dtype))
        if with_dgwdq:
            out.append(dq.reshape(shape).to(self.dtype))
        return tuple(out)

    def get_atomic_c6(self, gw: torch.Tensor) -> torch.Tensor:
        """C6 coefficients of all atom pairs, shape (..., n, n), from
        weights of shape (..., n, 7) (D4S: (..., n, n, 7); reference
        ``model/d4.py:268-289``, ``model/d4s.py:265-285``). Differentiable
        in gw."""
        z = self._device_numbers(gw)
        lead, n = self._lead()
        B = z.shape[0]
        g = gw.reshape(B, n, n, MAX_REF) if self._pair else gw.reshape(B, n, MAX_REF)
        c6 = _AtomicC6.apply(g, z, self)
        return c6.reshape(*lead, n, n).to(gw.dtype)

    def get_weighted_pols(self, gw: torch.Tensor) -> torch.Tensor:
        """Weighted polarisabilities at the 23 Casimir-Polder frequencies,
        shape (..., n, 23) (reference ``model/d4.py:291-306``; D4S:
        ``model/d4s.py:287-313``, the sum over the pair axis divided by the
        molecule's atom count; the reference's division broadcasts only for a
        single molecule, for a batch it is made per molecule here).
        Differentiable in gw."""
        z = self._device_numbers(gw)
        lead, n = self._lead()
        B = z.shape[0]
        g = gw.reshape(B, n, n, MAX_REF) if self._pair else gw.reshape(B, n, MAX_REF)
        out = _Pols.apply(g, z, self, False)
        return out.reshape(*lead, n, 23).to(gw.dtype)

    def get_polarizabilities(self, weights: torch.Tensor) -> torch.Tensor:
        """Static polarisabilities, shape (..., n) (reference
        ``model/base.py:286-302``; with D4S pair weights (..., n, n): the
        reference's einsum broadcasts over the pair axis, which works for a
        single molecule; for a batch the result is the same per molecule).
        Differentiable in the weights."""
        z = self._device_numbers(weights)
        lead, n = self._lead()
        B = z.shape[0]
        if self._pair:
            g = weights.reshape(B, n, n, MAX_REF)
            shape = (*lead, n, n)
        else:
            g = weights.reshape(B, n, MAX_REF)
            shape = (*lead, n)
        out = _Pols.apply(g, z, self, True)
        return out.reshape(shape).to(weights.dtype)


class _Weights(torch.autograd.Function):
    """gw(cn, q) through d4_model_weights; backward d4_model_weights_backward."""

    @staticmethod
    def forward(ctx, cn, q, z, model):
        gw, _, _ = runtime.model_weights(z, cn, q, model.ga, model.gc, model.wf, model._pair)
        ctx.save_for_backward(cn, q, z)
        ctx.model = model
        return gw

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad):
        cn, q, z = ctx.saved_tensors
        m = ctx.model
        want_cn, want_q = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (want_cn or want_q):
            return None, None, None, None
        gcn, gq = runtime.model_weights_backward(z, cn, q, m.ga, m.gc, m.wf, m._pair, grad,
                                                 want_cn, want_q)
        gcn = gcn.reshape(cn.shape).to(cn.dtype) if want_cn else None
        gq = gq.reshape(q.shape).to(q.dtype) if want_q else None
        return gcn, gq, None, None


class _AtomicC6(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gw, z, model):
        ctx.save_for_backward(gw, z)
        ctx.model = model
        return runtime.model_atomic_c6(z, gw, model.ga, model.gc, model._pair)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad):
        gw, z = ctx.saved_tensors
        m = ctx.model
        g = runtime.model_atomic_c6_backward(z, gw, m.ga, m.gc, m._pair, grad)
        return g.to(gw.dtype), None, None


class _Pols(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gw, z, model, static_only):
        ctx.save_for_backward(z)
        ctx.model = model
        ctx.static_only = static_only
        ctx.gw_dtype = gw.dtype
        return runtime.model_pols(z, gw, model.ga, model.gc, model._pair, static_only)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad):
        (z,) = ctx.saved_tensors
        m = ctx.model
        g = runtime.model_pols_backward(z, m.ga, m.gc, m._pair, ctx.static_only, grad)
        return g.to(ctx.gw_dtype), None, None, None


class D4SModel(D4Model):
    """The D4S dispersion model (``model/d4s.py:56-285``): the Gaussian
    weights of atom A in the pair AB use the element-pair factor
    wf(Z_A, Z_B) of ``data/wfpair.py`` instead of one global factor, so every
    pair has its own weights (and C6). The factors are resident on the device
    with the other tables; the kernels form the pair weights in the pair
    pass."""

    __slots__ = ()
    _pair = True

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
