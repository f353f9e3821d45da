"""
The plain per-molecule GPU path: the reference's dense algorithm expressed
with stock PyTorch tensor ops on the device, one molecule per call, forces
by autograd. This is what running tad-dftd4 itself on a ROCm device does
(per-call model build with the (n, n, 7, 7) reference-C6 tensor, dense
(n, n, n) three-body tensors, ``torch.linalg.solve`` for EEQ; see reference
``dispersion/base.py:285-431``, ``model/base.py:367-431``,
``dispersion/threebody.py:54-163``).

It exists only as the comparison baseline of the benchmark ("batched vs
plain"); ``dftd4()`` never calls it.
"""

from __future__ import annotations

import math

import torch

from . import defaults
from .data import elements as el
from .data import ref_tables
from .tables import r4r2_table, reference_alpha

_CACHE: dict = {}


def _tables(device):
    key = str(device)
    if key not in _CACHE:
        t = ref_tables()
        f = lambda a, dt=torch.float64: torch.as_tensor(a, dtype=dt, device=device)
        _CACHE[key] = dict(
            alpha=f(reference_alpha()), refcn=f(t["refcovcn"]), refc=f(t["refc"], torch.long),
            refq=f(t["clsq"]), w=f(t["cp_weights"]), r4r2=f(r4r2_table()), rcov=f(el.COV_D3),
            en=f(el.PAULING_EN), gam=f(el.GAM), zeff=f(el.ZEFF), chi=f(el.EEQ_CHI),
            eta=f(el.EEQ_ETA), kcn=f(el.EEQ_KCN), rad=f(el.EEQ_RAD),
        )
    return _CACHE[key]


def _zeta(gam, qref, qmod, ga=3.0):
    eps = torch.finfo(qmod.dtype).eps
    scale = torch.exp(gam * (1.0 - qref / (qmod - eps)))
    return torch.where(qmod > 0.0, torch.exp(ga * (1.0 - scale)),
                       torch.exp(torch.tensor(ga, dtype=qmod.dtype, device=qmod.device)))


def energy(numbers, positions, charge, param, cut2=50.0, cut3=25.0, cut_eeq=25.0):
    """Atom energies of one molecule (n,) with dense tensors."""
    t = _tables(positions.device)
    dt, dev = positions.dtype, positions.device
    n = numbers.shape[-1]
    eps = torch.finfo(dt).eps
    eye = torch.eye(n, dtype=torch.bool, device=dev)
    diff = positions.unsqueeze(-2) - positions.unsqueeze(-3)
    r = torch.sqrt(torch.clamp((diff * diff).sum(-1), min=eps))
    r = torch.where(eye, torch.full_like(r, eps), r)
    rc = t["rcov"][numbers].unsqueeze(-1) + t["rcov"][numbers].unsqueeze(-2)
    count = torch.where(eye, torch.zeros_like(r),
                        0.5 * (1.0 + torch.erf(-defaults.D4_KCN * (r / rc - 1.0))))
    en = t["en"][numbers]
    den = defaults.D4_K4 * torch.exp(
        -((torch.abs(en.unsqueeze(-1) - en.unsqueeze(-2)) + defaults.D4_K5) ** 2) / defaults.D4_K6)
    cn = torch.sum(torch.where(r <= defaults.CN_D4_INTERNAL_CUTOFF, den * count, 0.0), -1)
    # EEQ
    cnq = torch.sum(torch.where(r <= cut_eeq, count, 0.0), -1)
    cnq = math.log(1.0 + math.exp(8.0)) - torch.log(1.0 + torch.exp(8.0 - cnq))
    rad = t["rad"][numbers]
    gamma = 1.0 / torch.sqrt(rad.unsqueeze(-1) ** 2 + rad.unsqueeze(-2) ** 2)
    coul = torch.where(eye, torch.diag_embed(t["eta"][numbers] + math.sqrt(2 / math.pi) / rad),
                       torch.erf(r * gamma) / r)
    a = torch.zeros(n + 1, n + 1, dtype=dt, device=dev)
    a[:n, :n] = coul
    a[:n, n] = 1.0
    a[n, :n] = 1.0
    b = torch.cat([-t["chi"][numbers] + t["kcn"][numbers] * torch.sqrt(torch.clamp(cnq, min=eps)),
                   charge.reshape(1)])
    q = torch.linalg.solve(a, b)[:n]
    # model: per-atom-pair reference C6 (rebuilt every call, as the reference)
    alpha = t["alpha"][numbers]
    rc6 = 3.0 / math.pi * torch.einsum("w,iaw,jbw->ijab", t["w"], alpha, alpha)
    refc = t["refc"][numbers]
    mask = refc > 0
    refcn = t["refcn"][numbers]
    zeff = t["zeff"][numbers].unsqueeze(-1)
    gam = t["gam"][numbers].unsqueeze(-1) * 2.0

    def weights(qa):
        d = cn.unsqueeze(-1) - refcn
        tmp = torch.exp(-d * d)
        p1 = torch.pow(tmp, 6.0)
        p3 = p1 + torch.pow(tmp, 12.0) + torch.pow(tmp, 18.0)
        expw = torch.where(mask, torch.where(refc == 3, p3, p1), 0.0)
        norm = torch.where(mask, expw.sum(-1, keepdim=True), 1e-300)
        g = expw / norm
        exc = torch.isnan(g) | (g > torch.finfo(dt).max)
        g = torch.where(exc, (refcn == refcn.max(-1, keepdim=True)[0]).to(dt), g)
        return torch.where(mask, _zeta(gam, t["refq"][numbers] + zeff, qa.unsqueeze(-1) + zeff),
                           0.0) * g

    gw = weights(q)
    c6 = torch.einsum("ijab,ia,jb->ij", rc6, gw, gw)
    r4 = t["r4r2"][numbers]
    qq = 3.0 * r4.unsqueeze(-1) * r4.unsqueeze(-2)
    r0 = param["a1"] * torch.sqrt(qq) + param["a2"]
    inside = (~eye) & (r <= cut2)
    t6 = torch.where(inside, 1.0 / (r**6 + r0**6), 0.0)
    t8 = torch.where(inside, 1.0 / (r**8 + r0**8), 0.0)
    e2 = -0.5 * (param["s6"] * (c6 * t6).sum(-1) + param["s8"] * (c6 * qq * t8).sum(-1))
    # ATM with q = 0 weights, dense (n, n, n)
    c60 = torch.einsum("ijab,ia,jb->ij", rc6, weights(torch.zeros_like(q)), weights(torch.zeros_like(q)))
    c9 = torch.sqrt(torch.clamp(torch.abs(c60.unsqueeze(-1) * c60.unsqueeze(-2) * c60.unsqueeze(-3)),
                                min=eps))
    r2 = torch.where(eye, torch.full_like(r, eps), r * r)
    r2ij, r2ik, r2jk = r2.unsqueeze(-1), r2.unsqueeze(-2), r2.unsqueeze(-3)
    trip = (~eye).unsqueeze(-1) & (~eye).unsqueeze(-2) & (~eye).unsqueeze(-3)
    rr = r2ij * r2ik * r2jk
    r1 = torch.sqrt(rr)
    r3 = torch.where(trip, r1 * rr, eps)
    r5 = torch.where(trip, rr * r3, eps)
    rb0 = r0.unsqueeze(-1) * r0.unsqueeze(-2) * r0.unsqueeze(-3)
    fd = torch.where(trip, 1.0 / (1.0 + 6.0 * (rb0 / torch.where(trip, r1, 1.0)) ** (param["alp"] / 3.0)),
                     0.0)
    s = (r2ij + r2jk - r2ik) * (r2ij - r2jk + r2ik) * (-r2ij + r2jk + r2ik)
    c2 = cut3 * cut3
    ang = torch.where(trip & (r2ij <= c2) & (r2jk <= c2), 0.375 * s / r5 + 1.0 / r3, 0.0)
    e3 = (ang * fd * param["s9"] * c9).sum((-2, -1)) / 6.0
    return e2 + e3


def energy_gradient(numbers, positions, charge, param, **cut):
    """Atom energies and autograd gradient of their sum for one molecule."""
    pos = positions.detach().clone().requires_grad_(True)
    e = energy(numbers, pos, charge, param, **cut)
    (g,) = torch.autograd.grad(e.sum(), pos)
    return e.detach(), g
