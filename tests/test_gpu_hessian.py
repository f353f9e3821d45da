"""
Nuclear Hessian (reference test/test_grad/test_hessian.py): the batched
semi-analytic Hessian (five-point differences of the analytic GPU gradient)
and the double backward of ``dftd4`` against the oracle's exact autograd
Hessian on identical inputs (1e-9 Eh/bohr^2) and against the dftd4 Fortran
Hessians of LiH and SiH4 (reference tolerance 1e-7).
"""

import json
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

import tad_dftd4_amd as d4
from oracle import dense
from tad_dftd4_amd import synth

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())
F64 = torch.float64
H_TOL = 1e-9
TPSS0 = GOLD["param_tpss0"]


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _tp(d):
    return {k: torch.tensor(v, dtype=F64) for k, v in d.items()}


def _fortran(flat, n):
    return torch.tensor(flat, dtype=F64).reshape(3, n, 3, n).permute(3, 2, 1, 0)


def _oracle_hessian(numbers, positions, charge, param, model="d4", cutoff=None):
    def f(p):
        return dense.dftd4(numbers, p, charge, _tp(param), cutoff=cutoff, model=model).sum()
    n = len(numbers)
    h = torch.autograd.functional.hessian(f, positions).reshape(n, 3, n, 3)
    assert torch.isfinite(h).all()
    return h


def _mol(seed, n, mix=synth.MIX_HCNOS, spacing=2.8, charges=(-1, 0, 1)):
    return synth.molecule(np.random.default_rng(seed), n, mix, spacing, charges)


@pytest.mark.parametrize("name", ["LiH", "SiH4"])
def test_hessian_fortran_fixture(name):
    _say(f"Hessian {name} vs dftd4 Fortran and oracle")
    g = GOLD[name]
    numbers = torch.tensor(g["numbers"])
    positions = torch.tensor(g["positions"], dtype=F64)
    h = d4.dftd4_hessian(numbers.cuda(), positions.cuda(), 0.0, TPSS0).cpu()
    ref = _fortran(g["hessian"], len(numbers))
    assert (h - ref).abs().max().item() < 1e-7
    ho = _oracle_hessian(numbers, positions, 0.0, TPSS0)
    assert (h - ho).abs().max().item() < H_TOL


def test_hessian_batch_vs_oracle():
    """Padded batch (12 and 7 atoms, charged) against the exact Hessian."""
    _say("Hessian batch (12, 7) vs oracle autograd")
    mols = [_mol(400, 12), _mol(401, 7)]
    numbers, positions, charge = synth.pack(mols, device="cuda")
    h = d4.dftd4_hessian(numbers, positions, charge, TPSS0).cpu()
    assert h.shape == (2, 12, 3, 12, 3)
    for b, m in enumerate(mols):
        n = len(m.numbers)
        ho = _oracle_hessian(torch.tensor(m.numbers), torch.tensor(m.positions), m.charge, TPSS0)
        assert (h[b, :n, :, :n] - ho).abs().max().item() < H_TOL, b
        assert float(h[b, n:].abs().sum()) == 0.0 and float(h[b, :, :, n:].abs().sum()) == 0.0
    # translation invariance: every row sums to zero over the atoms
    assert h.sum(dim=3).abs().max().item() < 1e-10


def test_hessian_d4s_vs_oracle():
    _say("Hessian D4S vs oracle autograd")
    m = _mol(410, 9)
    numbers, positions, charge = synth.pack([m], device="cuda")
    h = d4.dftd4_hessian(numbers, positions, charge, TPSS0, model="d4s").cpu()[0]
    ho = _oracle_hessian(torch.tensor(m.numbers), torch.tensor(m.positions), m.charge, TPSS0,
                         model="d4s")
    assert (h - ho).abs().max().item() < H_TOL


def test_double_backward_matches():
    """torch.autograd.functional.hessian over dftd4 (double backward through
    the Hessian-vector product) and the derivative of the weighted gradient
    with respect to the weights, against the oracle's autograd."""
    _say("double backward vs batched Hessian and oracle")
    m = _mol(420, 8)
    numbers, positions, charge = synth.pack([m], device="cuda")
    par = {k: torch.tensor(v, dtype=F64, device="cuda") for k, v in TPSS0.items()}

    def f(p):
        return d4.dftd4(numbers, p, charge, par).sum()

    h = torch.autograd.functional.hessian(f, positions).reshape(1, 8, 3, 8, 3)
    hb = d4.dftd4_hessian(numbers, positions, charge, TPSS0)
    assert (h - hb).abs().max().item() < H_TOL
    # d/dw of sum(v * dL/dR), L = sum_i w_i E_i: (dE_i/dR) . v
    w = torch.rand(1, 8, dtype=F64, device="cuda").requires_grad_(True)
    v = torch.rand(1, 8, 3, dtype=F64, device="cuda")
    pos = positions.clone().requires_grad_(True)
    e = d4.dftd4(numbers, pos, charge, par)
    (g,) = torch.autograd.grad((e * w).sum(), pos, create_graph=True)
    gw, gp = torch.autograd.grad((g * v).sum(), (w, pos))
    po = torch.tensor(m.positions).requires_grad_(True)
    wo = w.detach().cpu()[0].clone().requires_grad_(True)
    eo = dense.dftd4(torch.tensor(m.numbers), po, m.charge, _tp(TPSS0))
    (go,) = torch.autograd.grad((eo * wo).sum(), po, create_graph=True)
    gwo, gpo = torch.autograd.grad((go * v.cpu()[0]).sum(), (wo, po))
    assert (gw.cpu()[0] - gwo).abs().max().item() < H_TOL
    assert (gp.cpu()[0] - gpo).abs().max().item() < H_TOL
