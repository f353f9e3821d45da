"""
Drop-in entry points: ``dftd4`` (reference ``src/tad_dftd4/disp.py:44-146``)
and ``get_properties`` (``disp.py:149-197``), plus the fused
``dftd4_energy_gradient`` for callers that want energies and forces in one
pass and ``dftd4_hessian`` for the nuclear Hessian (the quantity of
``test/test_grad/test_hessian.py``).

Arguments, batch convention (zero padding, ``Z = 0``), outputs and error
types follow the reference. The arithmetic runs in hand-written FP64 HIP
kernels on the ROCm device that holds ``positions``; float32 inputs are
computed in FP64 and the result is cast back to the input dtype. Forces come
from hand-written backward kernels through ``torch.autograd``.

Not supported (raise ``NotImplementedError``; there is no fallback):
non-default ``cn_function``/``counting_function``/``damping_function``,
models other than D4 and D4S with EEQ reference charges, gradients with
respect to user-supplied ``r4r2`` values, second derivatives involving
user-supplied charges. Gradients with respect to the damping parameters s6,
s8, s10, s9, a1, a2 and alp (0-d tensors with ``requires_grad``) and to
user-supplied charges ``q`` (the reference differentiates through q,
``dispersion/base.py:392-407`` -> ``model/d4.py:219-228``) come from the same
backward kernels (``test/test_grad/test_param.py:40-86``).

Second derivatives with respect to the positions (double backward, e.g.
``torch.autograd.functional.hessian``) are semi-analytic: a Hessian-vector
product is the five-point central difference of the analytic gradient along
the vector (step ``HESSIAN_STEP`` bohr; truncation error O(h^4), about 1e-13
Eh/bohr^2 at h = 2e-4), the four displaced gradients evaluated in one batched
launch. ``dftd4_hessian`` builds the full Hessian the same way, all 12 n
displaced geometries of a molecule in one batch.
"""

from __future__ import annotations

import torch

from . import defaults, runtime
from .cutoff import Cutoff
from .damping import Damping, Param, RationalDamping
from .model import D4Model, D4SModel
from .ncoord import cn_d4, erf_count

__all__ = ["dftd4", "get_properties", "dftd4_energy_gradient", "dftd4_hessian"]

HESSIAN_STEP = 2e-4  # bohr, five-point stencil of the analytic gradient
_STENCIL = (1.0, -1.0, 2.0, -2.0)

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
    def backward(ctx, grad_energy):
        positions, numbers, charge, q, r4r2, ws = ctx.saved_tensors
        want = any(ctx.needs_input_grad[7:])
        want_q = ctx.needs_input_grad[3]
        ge = grad_energy.contiguous()
        gq = None
        if torch.is_grad_enabled() and ctx.needs_input_grad[0]:
            # create_graph: the gradient itself is differentiable (Hessian-
            # vector products through _D4Grad); parameter and charge
            # gradients are first order only
            if want_q:
                raise NotImplementedError(
                    "second derivatives involving user-supplied charges are not supported"
                )
            grad = _D4Grad.apply(positions, ge, numbers, charge, q, r4r2, ctx.kp, ws)
            gpar = None
            if want:
                _, gpar = runtime.backward(numbers, positions.detach(), charge, ctx.kp, ws,
                                           ge.detach(), q=q, r4r2=r4r2, params=True)
        elif want_q:
            grad, gpar, gq = runtime.backward(numbers, positions.detach(), charge, ctx.kp, ws,
                                              ge.detach(), q=q, r4r2=r4r2, params=want,
                                              charges=True)
        else:
            out = runtime.backward(numbers, positions.detach(), charge, ctx.kp, ws,
                                   ge.detach(), q=q, r4r2=r4r2, params=want)
            grad, gpar = out if want else (out, None)
        pgrads = []
        for key, (shape, dtype) in zip(ctx.pkeys, ctx.pmeta):
            if gpar is None:
                pgrads.append(None)
            else:
                pgrads.append(gpar[runtime.PARAM_ORDER.index(key)].reshape(shape).to(dtype))
        grad_pos = grad if ctx.needs_input_grad[0] else None
        return (grad_pos, None, None, gq, None, None, None, *pgrads)


def _stencil(numbers, positions, charge, kp, q, r4r2, disp, grad_energy=None):
    """Energies and gradients at the displaced geometries positions + s h
    disp for s in _STENCIL, one batch of 4 B molecules. Returns (E, G) with
    shapes (4, B, N) and (4, B, N, 3); G is the gradient of
    sum_i grad_energy_i E_i (all ones if None)."""
    h = HESSIAN_STEP
    B, N = numbers.shape
    pos = torch.cat([positions + (s * h) * disp for s in _STENCIL]).contiguous()
    rep = lambda t: None if t is None else t.repeat(4, *([1] * (t.dim() - 1))).contiguous()
    nums, chg, qq, rr = rep(numbers), rep(charge), rep(q), rep(r4r2)
    if grad_energy is None:
        e, g, _ = runtime.energy_gradient(nums, pos, chg, kp, q=qq, r4r2=rr)
    else:
        e, ws = runtime.forward(nums, pos, chg, kp, q=qq, r4r2=rr)
        g = runtime.backward(nums, pos, chg, kp, ws, rep(grad_energy), q=qq, r4r2=rr)
    return e.reshape(4, B, N), g.reshape(4, B, N, 3)


def _five_point(x: torch.Tensor) -> torch.Tensor:
    """d/dh at h = 0 from values at h, -h, 2h, -2h (leading axis)."""
    return (8.0 * (x[0] - x[1]) - (x[2] - x[3])) / (12.0 * HESSIAN_STEP)


class _D4Grad(torch.autograd.Function):
    """Gradient of sum_i grad_energy_i E_i with respect to the positions,
    differentiable once more: the backward is a Hessian-vector product
    (and the matching derivative with respect to grad_energy) by the
    five-point difference of the analytic gradient along the vector."""

    @staticmethod
    def forward(ctx, positions, grad_energy, numbers, charge, q, r4r2, kp, ws):
        grad = runtime.backward(numbers, positions.detach(), charge, kp, ws,
                                grad_energy.detach().contiguous(), q=q, r4r2=r4r2)
        ctx.save_for_backward(positions.detach(), grad_energy.detach(), numbers, charge, q, r4r2)
        ctx.kp = kp
        return grad

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, v):
        positions, ge, numbers, charge, q, r4r2 = ctx.saved_tensors
        v = v.contiguous()
        # unit direction per molecule; the product is linear in v
        norm = v.reshape(v.shape[0], -1).norm(dim=1)
        safe = torch.where(norm > 0, norm, torch.ones_like(norm))
        u = v / safe[:, None, None]
        e, g = _stencil(numbers, positions, charge, ctx.kp, q, r4r2, u, ge.contiguous())
        hv = _five_point(g) * norm[:, None, None]
        de = _five_point(e) * norm[:, None] if ctx.needs_input_grad[1] else None
        return hv, de, None, None, None, None, None, None


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
    # q stays attached to its autograd graph (dL/dq from the backward kernels)
    qq = None if q is None else q.to(dev, torch.float64).reshape(-1, n).contiguous()
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
    (``dispersion/base.py:366-387``). Also differentiable with respect to
    the damping-parameter tensors in ``param`` and a given ``q``."""
    if isinstance(r4r2, torch.Tensor) and r4r2.requires_grad and torch.is_grad_enabled():
        raise NotImplementedError(
            "gradients with respect to user-supplied r4r2 values are not supported"
        )
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


def dftd4_hessian(
    numbers: torch.Tensor,
    positions: torch.Tensor,
    charge,
    param: Param,
    *,
    model="d4",
    q: torch.Tensor | None = None,
    r4r2: torch.Tensor | None = None,
    cutoff: Cutoff | None = None,
    max_batch_bytes: int = 4 << 30,
) -> torch.Tensor:
    """Hessian of the total dispersion energy with respect to the positions,
    shape (..., n, 3, n, 3) (the quantity of reference
    ``test/test_grad/test_hessian.py:73-127``, ``hessian(dftd4, ..., argnums=1)``).

    Column (a, k) is the five-point central difference of the analytic
    gradient under a displacement of atom a along k (step ``HESSIAN_STEP``);
    all 12 n displaced geometries of every molecule go through the fused
    energy + gradient kernels as one batch (split into batches of at most
    ``max_batch_bytes`` of workspace). The result is symmetrised. Padding
    atoms have zero rows and columns."""
    model_inst, nums, pos, chg, qq, rr, lead, n = _prepare(
        numbers, positions.detach(), charge, model, None, r4r2, None, q, cn_d4,
        erf_count, RationalDamping(),
    )
    kp = kernel_params(param, cutoff, model_inst)
    B = nums.shape[0]
    dev = pos.device
    hess = torch.zeros((B, n, 3, n, 3), dtype=torch.float64, device=dev)
    if B == 0 or n == 0:
        return hess.reshape(*lead, n, 3, n, 3).to(positions.dtype)
    # unit displacements of every (atom, direction), one molecule per column
    cols = 3 * n
    eye = torch.eye(cols, dtype=torch.float64, device=dev).reshape(cols, n, 3)
    per = runtime.workspace_bytes(4 * cols, n) / (4 * cols)
    chunk = max(1, min(cols, int(max_batch_bytes // max(per * 4, 1))))
    for b in range(B):
        real = nums[b] != 0
        for c0 in range(0, cols, chunk):
            c1 = min(cols, c0 + chunk)
            m = c1 - c0
            bcast = lambda t: None if t is None else t[b: b + 1].expand(m, *t.shape[1:])
            _, g = _stencil(bcast(nums).contiguous(), bcast(pos).contiguous(),
                            bcast(chg).contiguous(), kp, bcast(qq), bcast(rr), eye[c0:c1])
            hess[b].view(cols, n, 3)[c0:c1] = _five_point(g)
        mask = (real[:, None, None, None] & real[None, None, :, None]).to(torch.float64)
        hess[b] *= mask
    hess = 0.5 * (hess + hess.permute(0, 3, 4, 1, 2))
    return hess.reshape(*lead, n, 3, n, 3).to(positions.dtype)


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
