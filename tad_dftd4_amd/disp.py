"""
Drop-in entry points: ``dftd4`` (reference ``src/tad_dftd4/disp.py:44-146``)
and ``get_properties`` (``disp.py:149-197``), plus the fused
``dftd4_energy_gradient`` for callers that want energies and forces in one
pass.

Arguments, batch convention (zero padding, ``Z = 0``), outputs and error
types follow the reference. The arithmetic runs in hand-written FP64 HIP
kernels on the ROCm device that holds ``positions``; float32 inputs are
computed in FP64 and the result is cast back to the input dtype. Forces come
from hand-written backward kernels through ``torch.autograd``.

Not supported (raise ``NotImplementedError``; there is no fallback):
non-default ``cn_function``/``counting_function``/``damping_function``,
models other than D4 and D4S with EEQ reference charges, gradients with respect to
user-supplied charges, double backward. Gradients with respect to the damping
parameters s6, s8, s10, s9, a1, a2 and alp (0-d tensors with
``requires_grad``) come from the same backward kernels
(``test/test_grad/test_param.py:40-86``).
"""

from __future__ import annotations

import torch

from . import defaults, runtime
from .cutoff import Cutoff
from .damping import Damping, Param, RationalDamping
from .model import D4Model, D4SModel
from .ncoord import cn_d4, erf_count

__all__ = ["dftd4", "get_properties", "dftd4_energy_gradient"]

_MODELS = ("d3", "d4", "d4s", "d5")


def _value(x) -> float:
    if isinstance(x, torch.Tensor):
        return float(x.detach().reshape(()).cpu()) if x.numel() == 1 else _bad(x)
    return float(x)


def _bad(x):
    raise ValueError(f"damping parameters must be scalars, got shape {tuple(x.shape)}")


def kernel_params(param: Param, cutoff: Cutoff | None, model: D4Model | None):
    """Translate the reference's parameter conventions into kernel values.

    - a1 and a2 are required by rational damping (``functions.py:264-268``).
    - The order-8 term exists iff ``s8`` or ``rs8`` is given
      (``functions.py:144-145``); its prefactor defaults to 1.
    - The order-10 term exists iff the key ``s10`` is present
      (``twobody.py:182-197``).
    - The three-body term uses s9 (default 1) and alp (default 16)
      (``threebody.py:232-233``).
    """
    missing = [k for k in ("a1", "a2") if param.get(k) is None]
    if missing:
        raise TypeError(
            f"RationalDamping (order 6) requires keyword(s): {', '.join(missing)}"
        )
    has8 = param.get("s8") is not None or param.get("rs8") is not None
    s10_on = "s10" in param and param["s10"] is not None
    if cutoff is None:
        cutoff = Cutoff()
    kp = runtime.KernelParams(
        s6=_value(param.get("s6", defaults.S6)),
        s8=_value(param.get("s8", defaults.S8)) if has8 else 0.0,
        s10=_value(param["s10"]) if s10_on else 0.0,
        s9=_value(param.get("s9", defaults.S9)),
        a1=_value(param["a1"]),
        a2=_value(param["a2"]),
        alp=_value(param.get("alp", defaults.ALP)),
        cutoff_disp2=_value(cutoff.disp2),
        cutoff_disp3=_value(cutoff.disp3),
        cutoff_cn=defaults.CN_D4_INTERNAL_CUTOFF,
        cutoff_eeq=_value(cutoff.cn_eeq),
        s10_on=s10_on,
    )
    if model is not None:
        kp.ga, kp.gc, kp.wf = model.ga, model.gc, model.wf
        kp.d4s = isinstance(model, D4SModel)
    return kp


class _D4Energy(torch.autograd.Function):
    """Atom energies with a hand-written backward (HIP kernels). The trailing
    inputs are the damping-parameter tensors named by ``pkeys`` (a subset of
    ``runtime.PARAM_ORDER``) whose gradients are requested."""

    @staticmethod
    def forward(ctx, positions, numbers, charge, q, r4r2, kp, pkeys, *ptensors):
        energy, ws = runtime.forward(numbers, positions, charge, kp, q=q, r4r2=r4r2)
        ctx.save_for_backward(positions, numbers, charge, q, r4r2, ws)
        ctx.kp = kp
        ctx.pkeys = pkeys
        ctx.pmeta = [(t.shape, t.dtype) for t in ptensors]
        return energy

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_energy):
        positions, numbers, charge, q, r4r2, ws = ctx.saved_tensors
        want = any(ctx.needs_input_grad[7:])
        out = runtime.backward(numbers, positions, charge, ctx.kp, ws,
                               grad_energy.contiguous(), q=q, r4r2=r4r2, params=want)
        grad, gpar = out if want else (out, None)
        pgrads = []
        for key, (shape, dtype) in zip(ctx.pkeys, ctx.pmeta):
            if gpar is None:
                pgrads.append(None)
            else:
                pgrads.append(gpar[runtime.PARAM_ORDER.index(key)].reshape(shape).to(dtype))
        grad_pos = grad if ctx.needs_input_grad[0] else None
        return (grad_pos, None, None, None, None, None, None, *pgrads)


def _prepare(numbers, positions, charge, model, rcov, r4r2, rvdw, q, cn_function,
             counting_function, damping_function):
    if numbers.shape != positions.shape[:-1]:
        raise ValueError(
            f"Shape of positions ({positions.shape}) is not consistent "
            f"with atomic numbers ({numbers.shape})."
        )
    if isinstance(model, str):
        key = model.casefold()
        if key not in _MODELS:
            raise ValueError(f"Unknown model '{key}'. Please use {', '.join(_MODELS)}.")
        if key not in ("d4", "d4s"):
            raise NotImplementedError(f"model '{key}' is not implemented by the MI355X kernels")
        model_inst = D4SModel(numbers) if key == "d4s" else None
    elif isinstance(model, D4Model):
        model_inst = model
    else:
        raise NotImplementedError(f"model of type {type(model).__name__} is not supported")
    if cn_function is not cn_d4 or counting_function is not erf_count:
        raise NotImplementedError("only the D4 CN with the erf counting function is implemented")
    if not isinstance(damping_function, Damping) or damping_function != RationalDamping():
        raise NotImplementedError("only rational (Becke-Johnson) damping is implemented")
    if rcov is not None and numbers.shape != rcov.shape:
        raise ValueError(
            f"Shape of covalent radii ({rcov.shape}) is not consistent with atomic "
            f"numbers ({numbers.shape})."
        )
    if r4r2 is not None and numbers.shape != r4r2.shape:
        raise ValueError(
            f"Shape of expectation values r4r2 ({r4r2.shape}) is not consistent with "
            f"atomic numbers ({numbers.shape})."
        )
    if rvdw is not None and numbers.shape != rvdw.shape[:-1]:
        raise ValueError(
            f"Shape of van der Waals radii ({rvdw.shape}) is not consistent with "
            f"atomic numbers ({numbers.shape})."
        )
    if q is not None and numbers.shape != q.shape:
        raise ValueError(
            f"Shape of atomic charges ({q.shape}) is not consistent with atomic "
            f"numbers ({numbers.shape})."
        )
    if not positions.is_cuda:
        raise RuntimeError(
            "tad_dftd4_amd computes on a ROCm device; move numbers/positions to 'cuda'"
        )
    dev = positions.device
    lead = numbers.shape[:-1]
    n = numbers.shape[-1]
    nums = numbers.to(device=dev, dtype=torch.int64).reshape(-1, n).contiguous()
    if nums.numel() and (int(nums.min()) < 0 or int(nums.max()) >= 104):
        raise ValueError("atomic numbers must lie in 0..103 (0 marks padding)")
    pos = positions.to(torch.float64).reshape(-1, n, 3)
    if not pos.is_contiguous():
        pos = pos.contiguous()
    chg = torch.as_tensor(charge, dtype=torch.float64, device=dev)
    chg = torch.broadcast_to(chg, lead).reshape(-1).contiguous()
    qq = None if q is None else q.detach().to(dev, torch.float64).reshape(-1, n).contiguous()
    rr = None if r4r2 is None else r4r2.detach().to(dev, torch.float64).reshape(-1, n).contiguous()
    return model_inst, nums, pos, chg, qq, rr, lead, n


def dftd4(
    numbers: torch.Tensor,
    positions: torch.Tensor,
    charge,
    param: Param,
    *,
    model="d4",
    rcov: torch.Tensor | None = None,
    r4r2: torch.Tensor | None = None,
    rvdw: torch.Tensor | None = None,
    q: torch.Tensor | None = None,
    cutoff: Cutoff | None = None,
    cn_function=cn_d4,
    counting_function=erf_count,
    damping_function: Damping = RationalDamping(),
) -> torch.Tensor:
    """Atom-resolved DFT-D4 dispersion energy of a (batch of) molecule(s),
    shape ``numbers.shape`` and dtype ``positions.dtype``. Differentiable
    with respect to ``positions``.

    ``rcov`` and ``rvdw`` are validated and unused, as in the reference
    (``dispersion/base.py:366-387``)."""
    model_inst, nums, pos, chg, qq, rr, lead, n = _prepare(
        numbers, positions, charge, model, rcov, r4r2, rvdw, q, cn_function,
        counting_function, damping_function,
    )
    kp = kernel_params(param, cutoff, model_inst)
    pkeys = tuple(
        k for k in runtime.PARAM_ORDER
        if isinstance(param.get(k), torch.Tensor) and param[k].requires_grad
    )
    if pkeys and (not torch.is_grad_enabled()):
        pkeys = ()
    ptensors = tuple(param[k] for k in pkeys)
    energy = _D4Energy.apply(pos, nums, chg, qq, rr, kp, pkeys, *ptensors)
    return energy.reshape(*lead, n).to(positions.dtype)


def dftd4_energy_gradient(
    numbers: torch.Tensor,
    positions: torch.Tensor,
    charge,
    param: Param,
    *,
    model="d4",
    q: torch.Tensor | None = None,
    r4r2: torch.Tensor | None = None,
    cutoff: Cutoff | None = None,
) -> tuple[torch.Tensor, torch.Tensor]:
    """Atom energies and dE_total/dR in one fused pass (no autograd graph).
    Equivalent to ``dftd4`` followed by
    ``torch.autograd.grad(energy.sum(), positions)``."""
    model_inst, nums, pos, chg, qq, rr, lead, n = _prepare(
        numbers, positions.detach(), charge, model, None, r4r2, None, q, cn_d4,
        erf_count, RationalDamping(),
    )
    kp = kernel_params(param, cutoff, model_inst)
    energy, grad, _ = runtime.energy_gradient(nums, pos, chg, kp, q=qq, r4r2=rr)
    return (energy.reshape(*lead, n).to(positions.dtype),
            grad.reshape(*lead, n, 3).to(positions.dtype))


def get_properties(
    numbers: torch.Tensor,
    positions: torch.Tensor,
    charge=None,
    cutoff: Cutoff | None = None,
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """D4 coordination numbers, EEQ charges, atom-pair C6 coefficients and
    static polarisabilities (reference ``disp.py:149-197``)."""
    if charge is None:
        charge = 0.0
    model_inst, nums, pos, chg, _, _, lead, n = _prepare(
        numbers, positions.detach(), charge, "d4", None, None, None, None, cn_d4,
        erf_count, RationalDamping(),
    )
    kp = kernel_params({"a1": 0.4, "a2": 5.0}, cutoff, model_inst)
    cn, q, c6, alpha = runtime.properties(nums, pos, chg, kp)
    dt = positions.dtype
    return (cn.reshape(*lead, n).to(dt), q.reshape(*lead, n).to(dt),
            c6.reshape(*lead, n, n).to(dt), alpha.reshape(*lead, n).to(dt))
