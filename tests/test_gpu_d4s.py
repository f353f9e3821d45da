"""
The D4S model (reference src/tad_dftd4/model/d4s.py) on the GPU: Gaussian
weights with element-pair factors, formed per pair in the pair kernel. The
HIP path against the FP64 oracle (pinned to the dftd4 Fortran D4S vectors)
on identical inputs: atom energies within 1e-10 Eh, gradients within
1e-9 Eh/bohr, over the small-molecule path, the forced large-system path
(spatial order, tiled EEQ, pair-tile three-body kernel), the fused
energy+gradient entry point and the parameter gradients.
"""

import json
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

import tad_dftd4_amd as d4
from oracle import dense
from tad_dftd4_amd import runtime, synth
from tad_dftd4_amd.disp import kernel_params

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())
F64 = torch.float64
E_TOL = 1e-10
G_TOL = 1e-9
PARAM = {"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.40, "a2": 5.0}
CUT = {"disp2": 50.0, "disp3": 25.0, "cn_eeq": 25.0}


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _tp(d, device=None, grad=False):
    return {k: torch.tensor(v, dtype=F64, device=device, requires_grad=grad)
            for k, v in d.items()}


def _cut(c=CUT):
    return d4.Cutoff(disp2=c["disp2"], disp3=c["disp3"], cn=25.0, cn_eeq=c["cn_eeq"])


def _mol(seed, n, mix=synth.MIX_Z54, spacing=2.8, charges=(-1, 0, 1)):
    return synth.molecule(np.random.default_rng(seed), n, mix, spacing, charges)


def _oracle(mol, param=PARAM, cut=CUT):
    return dense.dftd4_with_grad(torch.tensor(mol.numbers), torch.tensor(mol.positions),
                                 mol.charge, _tp(param), cutoff=cut, model="d4s")


@pytest.mark.parametrize("name", ["LiH", "SiH4"])
def test_d4s_fortran_two_body(name):
    """disp2_d4s of test/test_d4/samples.py (fixture charges, TPSSh, no
    three-body term) and the full energy against the oracle."""
    _say(f"D4S {name} two-body fixture")
    g = GOLD[name]
    numbers = torch.tensor(g["numbers"], device="cuda")
    positions = torch.tensor(g["positions"], dtype=F64, device="cuda")
    q = torch.tensor(g["q"], dtype=F64, device="cuda")
    par = dict(GOLD["param_tpssh"])
    e2 = d4.dftd4(numbers, positions, 0.0, {**par, "s9": 0.0}, q=q, model="d4s").cpu()
    assert (e2 - torch.tensor(g["disp2_d4s"], dtype=F64)).abs().max().item() < 1e-9
    e = d4.dftd4(numbers, positions, 0.0, par, q=q, model=d4.D4SModel(numbers)).cpu()
    eo = dense.dftd4(numbers.cpu(), positions.cpu(), 0.0, _tp(par), q=q.cpu(), model="d4s")
    assert (e - eo).abs().max().item() < E_TOL


def test_d4s_batch_vs_oracle():
    """Padded batch of mixed sizes (2 to 120 atoms, elements up to Xe):
    energies, autograd forces and the fused entry point."""
    sizes = (2, 9, 33, 64, 120)
    _say(f"D4S batch {sizes} vs oracle")
    mols = [_mol(300 + k, n) for k, n in enumerate(sizes)]
    numbers, positions, charge = synth.pack(mols, device="cuda")
    pos = positions.clone().requires_grad_(True)
    e = d4.dftd4(numbers, pos, charge, PARAM, cutoff=_cut(), model="d4s")
    (g,) = torch.autograd.grad(e.sum(), pos)
    ef, gf = d4.dftd4_energy_gradient(numbers, positions, charge, PARAM, cutoff=_cut(),
                                      model="d4s")
    ed = d4.dftd4(numbers, positions, charge, PARAM, cutoff=_cut())
    for b, m in enumerate(mols):
        n = len(m.numbers)
        eo, go = _oracle(m)
        assert (e[b, :n].detach().cpu() - eo).abs().max().item() < E_TOL, b
        assert (g[b, :n].cpu() - go).abs().max().item() < G_TOL, b
        assert (ef[b, :n].cpu() - eo).abs().max().item() < E_TOL, b
        assert (gf[b, :n].cpu() - go).abs().max().item() < G_TOL, b
        assert float(e.detach()[b, n:].abs().sum()) == 0.0
    # the two models differ on this input
    assert (ed - e.detach()).abs().max().item() > 1e-7


@pytest.mark.parametrize("flags", [{"force_sort": True, "force_tiled": True},
                                   {"force_pair_tiles": True}, {"force_lu": True}])
def test_d4s_large_system_paths_vs_oracle(flags):
    """The paths a molecule of 256+ atoms takes, forced at a size the dense
    oracle handles; short cutoffs exercise the screening."""
    _say(f"D4S forced paths {flags} vs oracle")
    cut = {"disp2": 30.0, "disp3": 9.0, "cn_eeq": 12.0}
    mols = [_mol(320, 47, spacing=2.6), _mol(321, 18)]
    numbers, positions, charge = synth.pack(mols, device="cuda")
    kp = kernel_params(PARAM, _cut(cut), d4.D4SModel(numbers))
    for k, v in flags.items():
        setattr(kp, k, v)
    e, g, _ = runtime.energy_gradient(numbers, positions, charge, kp)
    eb, ws = runtime.forward(numbers, positions, charge, kp)
    gb = runtime.backward(numbers, positions, charge, kp, ws)
    for b, m in enumerate(mols):
        n = len(m.numbers)
        eo, go = _oracle(m, cut=cut)
        assert (e[b, :n].cpu() - eo).abs().max().item() < E_TOL, b
        assert (g[b, :n].cpu() - go).abs().max().item() < G_TOL, b
        assert (eb[b, :n].cpu() - eo).abs().max().item() < E_TOL, b
        assert (gb[b, :n].cpu() - go).abs().max().item() < G_TOL, b


def test_d4s_large_molecule_vs_oracle():
    """n = 260 takes the large-system path by itself (spatial order, tiled
    EEQ, pair-tile three-body kernel)."""
    _say("D4S n=260 vs oracle")
    mol = _mol(330, 260, mix=synth.MIX_HCNOS, spacing=2.6, charges=(0,))
    eo, go = _oracle(mol)
    nt, pt, ct = synth.pack([mol], device="cuda")
    ef, gf = d4.dftd4_energy_gradient(nt, pt, ct, PARAM, cutoff=_cut(), model="d4s")
    assert (ef[0].cpu() - eo).abs().max().item() < E_TOL
    assert (gf[0].cpu() - go).abs().max().item() < G_TOL


def test_d4s_param_gradients_vs_oracle():
    """Damping-parameter gradients with the D4S model."""
    _say("D4S parameter gradients vs oracle autograd")
    par = {"s6": 1.0, "s8": 0.79, "s9": 1.0, "s10": 0.0, "a1": 0.49, "a2": 5.73, "alp": 16.0}
    mols = [_mol(340, 30), _mol(341, 11)]
    ref = torch.zeros(len(par), dtype=F64)
    for m in mols:
        pt = _tp(par, grad=True)
        e = dense.dftd4(torch.tensor(m.numbers), torch.tensor(m.positions), m.charge, pt,
                        cutoff=CUT, model="d4s")
        e.sum().backward()
        ref += torch.stack([pt[k].grad for k in par])
    numbers, positions, charge = synth.pack(mols, device="cuda")
    pg = _tp(par, "cuda", grad=True)
    d4.dftd4(numbers, positions, charge, pg, cutoff=_cut(), model="d4s").sum().backward()
    got = torch.stack([pg[k].grad for k in par]).cpu()
    assert torch.allclose(got, ref, rtol=1e-9, atol=1e-12), (got - ref)


def test_d4s_properties_rejected():
    """Properties are evaluated with the D4 model (as disp.py:149-197); the
    C ABI refuses the D4S flag there."""
    _say("D4S flag rejected by d4_properties")
    mol = _mol(350, 6)
    numbers, positions, charge = synth.pack([mol], device="cuda")
    kp = kernel_params(PARAM, None, d4.D4SModel(numbers))
    with pytest.raises(ValueError, match="D4 model only"):
        runtime.properties(numbers, positions, charge, kp)
