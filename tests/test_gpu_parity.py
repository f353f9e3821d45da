"""
GPU parity: the HIP kernels (through the C ABI) against the FP64 CPU oracle
on identical inputs, and against the reference's golden vectors.

Tolerances: atom energies within 1e-10 Eh, gradients
within 1e-9 Eh/bohr of the oracle. Golden vectors printed by the reference
in float32 carry their own, looser bound (stated per test).
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


def _tp(d, device=None):
    return {k: torch.tensor(v, dtype=F64, device=device) for k, v in d.items()}


def _cut():
    return d4.Cutoff(disp2=CUT["disp2"], disp3=CUT["disp3"], cn=25.0, cn_eeq=CUT["cn_eeq"])


def _oracle(mol, param=PARAM, cut=CUT, grad=True):
    numbers = torch.tensor(mol.numbers)
    positions = torch.tensor(mol.positions)
    if grad:
        return dense.dftd4_with_grad(numbers, positions, mol.charge, _tp(param), cutoff=cut)
    return dense.dftd4(numbers, positions, mol.charge, _tp(param), cutoff=cut), None


def _gpu(mols, param=PARAM, pad_to=None):
    numbers, positions, charge = synth.pack(mols, pad_to=pad_to, device="cuda")
    pos = positions.clone().requires_grad_(True)
    e = d4.dftd4(numbers, pos, charge, param, cutoff=_cut())
    (g,) = torch.autograd.grad(e.sum(), pos)
    ef, gf = d4.dftd4_energy_gradient(numbers, positions, charge, param, cutoff=_cut())
    return e.detach().cpu(), g.cpu(), ef.cpu(), gf.cpu()


def _mol(seed, n, mix=synth.MIX_Z54, spacing=2.8, charges=(-1, 0, 1)):
    return synth.molecule(np.random.default_rng(seed), n, mix, spacing, charges)


# ----------------------------------------------------------------- golden


def test_example_single_golden():
    _say("examples/single.py on GPU")
    g = GOLD["example_single"]
    numbers = torch.tensor(g["numbers"], device="cuda")
    positions = torch.tensor(g["positions"], dtype=F64, device="cuda")
    e = d4.dftd4(numbers, positions, 0.0, g["param"]).cpu()
    ref = torch.tensor(g["energy"], dtype=F64)
    assert (e - ref).abs().max().item() < 1e-9  # float32 reference, 10 decimals
    eo = dense.dftd4(numbers.cpu(), positions.cpu(), 0.0, _tp(g["param"]))
    assert (e - eo).abs().max().item() < E_TOL


def test_example_batch_golden():
    _say("examples/batch.py on GPU")
    g = GOLD["example_batch"]
    mols = [synth.Molecule(np.array(n), np.array(p), 0.0)
            for n, p in zip(g["numbers"], g["positions"])]
    numbers, positions, charge = synth.pack(mols, device="cuda")
    e = d4.dftd4(numbers, positions, charge, g["param"]).sum(-1).cpu()
    assert (e - torch.tensor(g["total_energy"], dtype=F64)).abs().max().item() < 2e-9


@pytest.mark.parametrize("name", ["LiH", "SiH4"])
def test_tblite_gradient_golden(name):
    _say(f"tblite gradient {name}")
    g = GOLD[name]
    numbers = torch.tensor(g["numbers"], device="cuda")
    positions = torch.tensor(g["positions"], dtype=F64, device="cuda").requires_grad_(True)
    e = d4.dftd4(numbers, positions, 0.0, GOLD["param_tpss0"])
    (grad,) = torch.autograd.grad(e.sum(), positions)
    ref = torch.tensor(g["grad"], dtype=F64)
    assert (grad.cpu() - ref).abs().max().item() < 1e-12


# ------------------------------------------------------------ vs oracle


@pytest.mark.parametrize("seed,n", [(0, 2), (1, 5), (2, 20), (3, 64), (4, 120)])
def test_single_molecule_vs_oracle(seed, n):
    _say(f"single molecule seed={seed} n={n}")
    mol = _mol(seed, n)
    eo, go = _oracle(mol)
    e, g, ef, gf = _gpu([mol])
    assert (e[0] - eo).abs().max().item() < E_TOL
    assert (g[0] - go).abs().max().item() < G_TOL
    assert (ef[0] - eo).abs().max().item() < E_TOL
    assert (gf[0] - go).abs().max().item() < G_TOL


def test_workload_single_64_vs_oracle():
    """The 64-atom H/C/N/O CPU workload (jittered 2.8 bohr grid)."""
    _say("single_64 workload")
    (mol,) = synth.workload("single_64")
    eo, go = _oracle(mol)
    e, g, _, _ = _gpu([mol])
    assert (e[0] - eo).abs().max().item() < E_TOL
    assert (g[0] - go).abs().max().item() < G_TOL


def test_batch_vs_oracle_and_alone_vs_mix():
    """A padded batch of uneven molecules equals the oracle, and every
    molecule's result is bitwise the same alone and in the batch."""
    _say("batch of 6 uneven molecules")
    mols = [_mol(10 + k, n) for k, n in enumerate([7, 33, 20, 90, 3, 51])]
    e, g, ef, gf = _gpu(mols, pad_to=96)
    for k, mol in enumerate(mols):
        n = len(mol.numbers)
        eo, go = _oracle(mol)
        assert (e[k, :n] - eo).abs().max().item() < E_TOL
        assert (g[k, :n] - go).abs().max().item() < G_TOL
        assert torch.all(e[k, n:] == 0) and torch.all(g[k, n:] == 0)
        ea, ga, efa, gfa = _gpu([mol])
        assert torch.equal(ea[0], e[k, :n])
        assert torch.equal(ga[0], g[k, :n])
        assert torch.equal(efa[0], ef[k, :n])
        assert torch.equal(gfa[0], gf[k, :n])


def test_weighted_backward():
    """Backward with a non-uniform upstream gradient dL/dE_i."""
    _say("weighted backward")
    mol = _mol(21, 30)
    numbers, positions, charge = synth.pack([mol], device="cuda")
    w = torch.linspace(-1.0, 2.0, 30, dtype=F64)
    pos = positions.clone().requires_grad_(True)
    e = d4.dftd4(numbers, pos, charge, PARAM, cutoff=_cut())
    (g,) = torch.autograd.grad((e * w.cuda()).sum(), pos)
    pc = torch.tensor(mol.positions).requires_grad_(True)
    eo = dense.dftd4(torch.tensor(mol.numbers), pc, mol.charge, _tp(PARAM), cutoff=CUT)
    (go,) = torch.autograd.grad((eo * w).sum(), pc)
    assert (g[0].cpu() - go).abs().max().item() < G_TOL


def test_forced_lu_path_matches_cholesky():
    """The pivoted-LU EEQ path (forced from the host) gives the same charges,
    energies and gradients as the Cholesky path."""
    _say("forced LU EEQ path")
    mols = [_mol(30, 40), _mol(31, 17)]
    numbers, positions, charge = synth.pack(mols, device="cuda")
    kp = kernel_params(PARAM, _cut(), None)
    e1, g1, ws1 = runtime.energy_gradient(numbers, positions, charge, kp)
    assert np.all(runtime.molecule_status(ws1, 2) == 0)
    kp.force_lu = True
    e2, g2, ws2 = runtime.energy_gradient(numbers, positions, charge, kp)
    assert np.all(runtime.molecule_status(ws2, 2) == 1)
    assert (e1 - e2).abs().max().item() < 1e-13
    assert (g1 - g2).abs().max().item() < 1e-12
    eo, go = _oracle(mols[0])
    assert (e2[0, :40].cpu() - eo).abs().max().item() < E_TOL
    assert (g2[0, :40].cpu() - go).abs().max().item() < G_TOL


def test_large_molecule_global_eeq():
    """n = 200: the packed factor fills the LDS, the solution vectors live in
    HBM."""
    _say("n=200 molecule")
    mol = _mol(40, 200)
    eo, go = _oracle(mol)
    e, g, _, _ = _gpu([mol])
    assert (e[0] - eo).abs().max().item() < E_TOL
    assert (g[0] - go).abs().max().item() < G_TOL


def test_cutoffs_screening():
    """Short cutoffs exercise the two-body, EEQ-CN and three-body screening
    (including the unscreened r_ik edge)."""
    _say("short cutoffs")
    mol = _mol(50, 60, spacing=2.6)
    cut = {"disp2": 8.0, "disp3": 6.0, "cn_eeq": 7.0}
    eo, go = _oracle(mol, cut=cut)
    numbers, positions, charge = synth.pack([mol], device="cuda")
    c = d4.Cutoff(disp2=8.0, disp3=6.0, cn_eeq=7.0)
    pos = positions.clone().requires_grad_(True)
    e = d4.dftd4(numbers, pos, charge, PARAM, cutoff=c)
    (g,) = torch.autograd.grad(e.sum(), pos)
    assert (e[0].detach().cpu() - eo).abs().max().item() < E_TOL
    assert (g[0].cpu() - go).abs().max().item() < G_TOL


def test_parameter_variants():
    """s10 key present, s8 absent, s9 = 0, given charges, given r4r2."""
    mol = _mol(60, 25)
    numbers, positions, charge = synth.pack([mol], device="cuda")
    nc, pc = torch.tensor(mol.numbers), torch.tensor(mol.positions)
    for name, par in (
        ("s10", {**PARAM, "s10": 0.7}),
        ("no s8", {"s6": 1.0, "a1": 0.4, "a2": 5.0}),
        ("s9=0", {**PARAM, "s9": 0.0}),
        ("alp", {**PARAM, "alp": 14.0}),
    ):
        _say(f"parameter variant {name}")
        e = d4.dftd4(numbers, positions, charge, par, cutoff=_cut())[0].cpu()
        eo = dense.dftd4(nc, pc, mol.charge, _tp(par), cutoff=CUT)
        assert (e - eo).abs().max().item() < E_TOL, name
    _say("given charges and r4r2")
    q = torch.linspace(-0.3, 0.3, 25, dtype=F64)
    e = d4.dftd4(numbers, positions, charge, PARAM, q=q[None].cuda(), cutoff=_cut())[0].cpu()
    eo = dense.dftd4(nc, pc, mol.charge, _tp(PARAM), q=q, cutoff=CUT)
    assert (e - eo).abs().max().item() < E_TOL
    r4 = dense.tables().r4r2[nc] * 1.01
    e = d4.dftd4(numbers, positions, charge, PARAM, r4r2=r4[None].cuda(), cutoff=_cut())[0].cpu()
    parts = dense.dftd4(nc, pc, mol.charge, _tp(PARAM), cutoff=CUT, return_parts=True)
    e2 = dense.dispersion2(nc, pc, _tp(PARAM), parts["c6"], r4, cutoff=CUT["disp2"])
    e3 = dense.atm_term(nc, pc, _tp(PARAM), parts["cn"], dense.get_refc6(nc), r4, CUT["disp3"])
    assert (e - (e2 + e3)).abs().max().item() < E_TOL


def test_inner_padding_and_float32():
    """A Z = 0 atom inside the molecule is a dummy; float32 inputs are
    computed in FP64 and returned as float32."""
    _say("inner padding / float32")
    mol = _mol(70, 12)
    z = mol.numbers.copy()
    z[5] = 0
    m2 = synth.Molecule(z, mol.positions, mol.charge)
    keep = [i for i in range(12) if i != 5]
    m3 = synth.Molecule(mol.numbers[keep], mol.positions[keep], mol.charge)
    e2, g2, _, _ = _gpu([m2])
    e3, g3, _, _ = _gpu([m3])
    assert (e2[0, keep] - e3[0]).abs().max().item() < 1e-15
    assert e2[0, 5] == 0 and torch.all(g2[0, 5] == 0)
    numbers, positions, charge = synth.pack([mol], device="cuda")
    e32 = d4.dftd4(numbers, positions.float(), charge, PARAM, cutoff=_cut())
    assert e32.dtype == torch.float32
    e64 = d4.dftd4(numbers, positions.float().double(), charge, PARAM, cutoff=_cut())
    assert torch.equal(e32, e64.float())


def test_properties_vs_oracle():
    _say("get_properties")
    mol = _mol(80, 33)
    numbers, positions, charge = synth.pack([mol], device="cuda")
    cn, q, c6, alpha = d4.get_properties(numbers, positions, charge, cutoff=_cut())
    nc, pc = torch.tensor(mol.numbers), torch.tensor(mol.positions)
    cno = dense.cn_d4(nc, pc)
    qo = dense.eeq_charges(nc, pc, torch.tensor(mol.charge, dtype=F64))
    gw = dense.weight_references(nc, cno, qo)
    c6o = dense.get_atomic_c6(dense.get_refc6(nc), gw)
    alo = dense.get_polarizabilities(nc, gw)
    assert torch.allclose(cn[0].cpu(), cno, rtol=1e-13, atol=1e-14)
    assert torch.allclose(q[0].cpu(), qo, rtol=0, atol=1e-12)
    assert torch.allclose(c6[0].cpu(), c6o, rtol=1e-12, atol=0)
    assert torch.allclose(alpha[0].cpu(), alo, rtol=1e-12, atol=0)


def test_cn_d4_alone():
    """cn_d4 runs only the coordination-number kernels: the same bits as the
    CN of get_properties (which runs the whole chain) for a ragged batch of
    a 40- and a 300-atom molecule (the second in the spatially sorted
    layout), and the oracle's CN to 1e-13."""
    _say("cn_d4 alone")
    mols = [_mol(81, 40), _mol(82, 300)]
    numbers, positions, _ = synth.pack(mols, device="cuda")
    cn = d4.cn_d4(numbers, positions)
    assert cn.shape == numbers.shape
    assert torch.equal(cn, d4.get_properties(numbers, positions)[0])
    for b, mol in enumerate(mols):
        n = len(mol.numbers)
        cno = dense.cn_d4(torch.tensor(mol.numbers), torch.tensor(mol.positions))
        assert torch.allclose(cn[b, :n].cpu(), cno, rtol=1e-13, atol=1e-14)
        assert torch.all(cn[b, n:] == 0)
    short = d4.cn_d4(numbers, positions, cutoff=9.0)
    for b, mol in enumerate(mols):
        n = len(mol.numbers)
        cno = dense.cn_d4(torch.tensor(mol.numbers), torch.tensor(mol.positions), cutoff=9.0)
        assert torch.allclose(short[b, :n].cpu(), cno, rtol=1e-13, atol=1e-14)
    assert (short - cn).abs().max().item() > 1e-6  # the cutoff removes pairs
    one = d4.cn_d4(numbers[0, :40], positions[0, :40])
    assert one.shape == (40,)
    assert torch.equal(one, d4.get_properties(numbers[0, :40], positions[0, :40])[0])
    assert torch.allclose(one, cn[0, :40], rtol=1e-13, atol=1e-14)


def test_finite_differences_16_atoms():
    """Central differences (h = 1e-5 bohr) of the GPU energy on 16 randomly
    chosen atoms against the GPU gradient."""
    _say("finite differences")
    mol = _mol(90, 48)
    numbers, positions, charge = synth.pack([mol], device="cuda")
    _, g = d4.dftd4_energy_gradient(numbers, positions, charge, PARAM, cutoff=_cut())
    rng = np.random.default_rng(5)
    atoms = rng.choice(48, size=16, replace=False)
    h = 1e-5
    for i in atoms:
        for x in range(3):
            pp, pm = positions.clone(), positions.clone()
            pp[0, i, x] += h
            pm[0, i, x] -= h
            ep = d4.dftd4(numbers, pp, charge, PARAM, cutoff=_cut()).sum()
            em = d4.dftd4(numbers, pm, charge, PARAM, cutoff=_cut()).sum()
            fd = ((ep - em) / (2 * h)).item()
            assert abs(fd - g[0, i, x].item()) < 1e-9


def test_many_elements_vs_oracle():
    """A molecule with more distinct elements than the pair passes' element
    slots (40 elements: reference C6 blocks contracted per pair) batched with
    an ordinary one (contraction vectors per element): energies, fused and
    autograd gradients against the oracle."""
    _say("40 distinct elements + a 30-atom molecule vs oracle")
    rng = np.random.default_rng(470)
    many = synth.Molecule(np.arange(1, 41, dtype=np.int64), synth.grid_positions(rng, 40, 3.0),
                          0.0)
    small = synth.molecule(rng, 30, synth.MIX_Z54, 2.8, (0,))
    numbers, positions, charge = synth.pack([many, small], device="cuda")
    pos = positions.clone().requires_grad_(True)
    e = d4.dftd4(numbers, pos, charge, PARAM, cutoff=_cut())
    (g,) = torch.autograd.grad(e.sum(), pos)
    ef, gf = d4.dftd4_energy_gradient(numbers, positions, charge, PARAM, cutoff=_cut())
    for b, m in enumerate((many, small)):
        n = len(m.numbers)
        eo, go = _oracle(m)
        assert (e[b, :n].detach().cpu() - eo).abs().max().item() < E_TOL, b
        assert (g[b, :n].cpu() - go).abs().max().item() < G_TOL, b
        assert (ef[b, :n].cpu() - eo).abs().max().item() < E_TOL, b
        assert (gf[b, :n].cpu() - go).abs().max().item() < G_TOL, b


@pytest.mark.parametrize("force_lu", [False, True])
def test_paired_eeq_at_padded_length_200(force_lu):
    """At padded length 200 one molecule's EEQ factor fills a CU's LDS; an
    energy-only evaluation on two streams then solves two molecules per
    work-group (largest with smallest). Energies are bitwise those of the
    one-stream evaluation (one molecule per work-group) for an odd batch,
    through the Cholesky and the pivoted-LU paths, and match the oracle."""
    _say(f"paired EEQ at N=200, force_lu={force_lu}")
    sizes = [200, 8, 150, 61, 120, 33, 187]
    mols = [_mol(90 + k, n) for k, n in enumerate(sizes)]
    numbers, positions, charge = synth.pack(mols, pad_to=200, device="cuda")
    kp = kernel_params(PARAM, _cut(), None)
    kp.force_lu = force_lu
    e2, ws2 = runtime.forward(numbers, positions, charge, kp)
    kp.one_stream = True
    e1, ws1 = runtime.forward(numbers, positions, charge, kp)
    st2 = runtime.molecule_status(ws2, len(mols))
    assert np.array_equal(st2, runtime.molecule_status(ws1, len(mols)))
    if force_lu:
        assert np.all(st2 == 1)
    assert torch.equal(e1, e2)
    for k in (1, 4):
        eo, _ = _oracle(mols[k], grad=False)
        assert (e2[k, :sizes[k]].cpu() - eo).abs().max().item() < E_TOL
    # charges given by the caller: both launches copy them per molecule
    qg = torch.where(numbers != 0, 0.01 * (numbers % 5 - 2).to(F64), 0.0)
    kp.one_stream = False
    eq2, _ = runtime.forward(numbers, positions, charge, kp, q=qg)
    kp.one_stream = True
    eq1, _ = runtime.forward(numbers, positions, charge, kp, q=qg)
    assert torch.equal(eq1, eq2) and not torch.equal(eq1, e1)
