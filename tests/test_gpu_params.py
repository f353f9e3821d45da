"""
Gradients with respect to the damping parameters (s6, s8, s10, s9, a1, a2,
alp), the use case of reference test/test_grad/test_param.py:40-186: the HIP
backward against the oracle's autograd on identical inputs, for the one-pass
three-body kernel, the pair-tile kernel and the large-system path, plus a
central finite difference.
"""

import sys

import numpy as np
import pytest
import torch

import tad_dftd4_amd as d4
from oracle import dense
from tad_dftd4_amd import runtime, synth
from tad_dftd4_amd.disp import kernel_params

pytestmark = pytest.mark.gpu

F64 = torch.float64
# TPSS0-D4-ATM parameters of the reference test, s10 present (= 0)
PARAM = {"s6": 1.0, "s8": 0.78981345, "s9": 1.0, "s10": 0.0, "a1": 0.49484001,
         "a2": 5.73083694, "alp": 16.0}
CUT = {"disp2": 50.0, "disp3": 25.0, "cn_eeq": 25.0}


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _mol(seed, n, mix=synth.MIX_Z54, spacing=2.8, charges=(-1, 0, 1)):
    return synth.molecule(np.random.default_rng(seed), n, mix, spacing, charges)


def _params(device):
    return {k: torch.tensor(v, dtype=F64, device=device, requires_grad=True)
            for k, v in PARAM.items()}


def _oracle_grads(mols, weights, cut=CUT):
    out = torch.zeros(len(PARAM), dtype=F64)
    for m, wgt in zip(mols, weights):
        par = _params(None)
        e = dense.dftd4(torch.tensor(m.numbers), torch.tensor(m.positions), m.charge, par,
                        cutoff=cut)
        (e * torch.tensor(wgt[: len(m.numbers)])).sum().backward()
        out += torch.stack([par[k].grad for k in PARAM])
    return out


def _gpu_grads(mols, weights, cut=CUT, **flags):
    numbers, positions, charge = synth.pack(mols, device="cuda")
    par = _params("cuda")
    cutoff = d4.Cutoff(disp2=cut["disp2"], disp3=cut["disp3"], cn=25.0, cn_eeq=cut["cn_eeq"])
    if flags:
        kp = kernel_params(par, cutoff, None)
        for k, v in flags.items():
            setattr(kp, k, v)
        e, ws = runtime.forward(numbers, positions, charge, kp)
        _, gp = runtime.backward(numbers, positions, charge, kp, ws,
                                 grad_energy=torch.tensor(weights, device="cuda"), params=True)
        order = runtime.PARAM_ORDER
        return torch.stack([gp[order.index(k)] for k in PARAM]).cpu()
    e = d4.dftd4(numbers, positions, charge, par, cutoff=cutoff)
    (e * torch.tensor(weights, device="cuda")).sum().backward()
    return torch.stack([par[k].grad for k in PARAM]).cpu()


def _weights(mols, seed):
    nmax = max(len(m.numbers) for m in mols)
    return np.random.default_rng(seed).uniform(0.5, 1.5, size=(len(mols), nmax))


@pytest.mark.parametrize("flags", [{}, {"force_pair_tiles": True},
                                   {"force_pair_tiles": True, "pair_owner": True},
                                   {"force_sort": True, "force_tiled": True}])
def test_param_gradients_vs_oracle(flags):
    _say(f"parameter gradients vs oracle autograd, flags={flags}")
    mols = [_mol(200, 30), _mol(201, 11), _mol(202, 47)]
    wts = _weights(mols, 5)
    ref = _oracle_grads(mols, wts)
    got = _gpu_grads(mols, wts, **flags)
    assert torch.allclose(got, ref, rtol=1e-9, atol=1e-12), (got - ref)


def test_param_gradients_screened_vs_oracle():
    """Short cutoffs: screened pairs and triples drop out of the sums."""
    _say("parameter gradients with cutoffs 20/8/10 vs oracle")
    cut = {"disp2": 20.0, "disp3": 8.0, "cn_eeq": 10.0}
    mols = [_mol(210, 64, spacing=2.6), _mol(211, 25)]
    wts = _weights(mols, 6)
    ref = _oracle_grads(mols, wts, cut)
    for flags in ({}, {"force_pair_tiles": True},
                  {"force_pair_tiles": True, "pair_owner": True}):
        got = _gpu_grads(mols, wts, cut, **flags)
        assert torch.allclose(got, ref, rtol=1e-9, atol=1e-12), (flags, got - ref)


def test_param_gradient_finite_difference():
    """Central differences of the total energy in each parameter."""
    _say("parameter gradients vs central finite differences")
    mols = [_mol(220, 40)]
    numbers, positions, charge = synth.pack(mols, device="cuda")
    wts = np.ones((1, 40))
    got = _gpu_grads(mols, wts)
    for idx, key in enumerate(PARAM):
        h = 1e-5 * max(1.0, abs(PARAM[key]))
        vals = []
        for sgn in (1.0, -1.0):
            par = {k: torch.tensor(v + (sgn * h if k == key else 0.0), dtype=F64)
                   for k, v in PARAM.items()}
            vals.append(d4.dftd4(numbers, positions, charge, par).sum().item())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - got[idx].item()) < 1e-8 + 1e-6 * abs(fd), (key, fd, got[idx].item())
