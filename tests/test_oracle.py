"""
Pins the CPU oracle (``oracle.dense``) before it is trusted as the parity
checker for the HIP kernels.

Each check is chosen so that a plausible mistake fails it: golden vectors
from the reference (end-to-end energies, tblite gradients, C6 from fixed
weights), an independent loop formulation (``oracle.brute``), closed forms,
invariants and finite differences.
"""

import json
import math
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import brute, dense
from tad_dftd4_amd import synth
from tad_dftd4_amd.data import elements

GOLD = json.loads(
    (Path(__file__).parent / "golden" / "reference_vectors.json").read_text()
)
F64 = torch.float64


def _p(d):
    return {k: torch.tensor(v, dtype=F64) for k, v in d.items()}


def _mol(name):
    g = GOLD[name]
    return torch.tensor(g["numbers"]), torch.tensor(g["positions"], dtype=F64)


# ---------------------------------------------------------------- golden


def test_example_single_energies():
    """examples/single.py:52-67 (C, N, S, H). The reference ran in float32
    and printed 10 decimals, so the bound is 1e-9 Eh."""
    g = GOLD["example_single"]
    e = dense.dftd4(torch.tensor(g["numbers"]), torch.tensor(g["positions"], dtype=F64),
                    0.0, _p(g["param"]))
    ref = torch.tensor(g["energy"], dtype=F64)
    assert (e - ref).abs().max().item() < 1e-9


def test_example_batch_totals():
    """Formamide dimer / formamide totals (src/tad_dftd4/__init__.py:84-87),
    padded batch (C, N, O, H). float32 reference: bound 2e-9 Eh."""
    g = GOLD["example_batch"]
    mols = [synth.Molecule(np.array(n), np.array(p), 0.0)
            for n, p in zip(g["numbers"], g["positions"])]
    numbers, positions, charge = synth.pack(mols)
    e = dense.dftd4(numbers, positions, charge, _p(g["param"])).sum(-1)
    ref = torch.tensor(g["total_energy"], dtype=F64)
    assert (e - ref).abs().max().item() < 2e-9


@pytest.mark.parametrize("name", ["LiH", "SiH4"])
def test_tblite_gradient(name):
    """tblite gradients (test/test_grad/samples_grad.py) with TPSS0-ATM
    parameters; covers CN, EEQ (dq/dR), weights, C6, E2 and ATM."""
    numbers, positions = _mol(name)
    _, g = dense.dftd4_with_grad(numbers, positions, 0.0, _p(GOLD["param_tpss0"]))
    ref = torch.tensor(GOLD[name]["grad"], dtype=F64)
    assert (g - ref).abs().max().item() < 1e-13


@pytest.mark.parametrize("name", ["LiH", "SiH4"])
def test_fortran_energies_with_given_charges(name):
    """test/test_d4/samples.py disp/disp3 with the fixture charges, the same
    call as test/test_d4/test_full.py:58-84 (reference tolerance 1.5e-7)."""
    numbers, positions = _mol(name)
    q = torch.tensor(GOLD[name]["q"], dtype=F64)
    parts = dense.dftd4(numbers, positions, 0.0, _p(GOLD["param_tpssh"]), q=q,
                        return_parts=True)
    disp = torch.tensor(GOLD[name]["disp"], dtype=F64)
    disp3 = torch.tensor(GOLD[name]["disp3"], dtype=F64)
    assert ((parts["e2"] + parts["e3"]) - disp).abs().max().item() < 1e-10
    assert (parts["e3"] - disp3).abs().max().item() < 1e-15


@pytest.mark.parametrize("name", ["LiH", "SiH4"])
def test_eeq_charges_fixture(name):
    """EEQ charges vs the dftd4 test-suite charges. Those were produced by an
    older EEQ build; the reference itself only agrees with them to its
    1.5e-7 test tolerance (test/test_d4/test_twobody.py:154-168)."""
    numbers, positions = _mol(name)
    q = dense.eeq_charges(numbers, positions, torch.tensor(0.0, dtype=F64))
    ref = torch.tensor(GOLD[name]["q"], dtype=F64)
    assert (q - ref).abs().max().item() < 5e-7
    assert abs(q.sum().item()) < 1e-14


def _fixture_gw(name):
    g = GOLD[name]["gw"]
    nref, nat = g["fortran_shape"]
    gw = torch.tensor(g["values"], dtype=F64).reshape(nat, nref)
    return torch.nn.functional.pad(gw, (0, 7 - nref))


@pytest.mark.parametrize("name", ["LiH", "SiH4"])
def test_c6_from_fixture_weights(name):
    """test/test_model/test_c6.py:36-58: C6 from fixed weights pins the
    reference C6 table (alpha, zeta of secondary systems, quadrature)."""
    numbers, _ = _mol(name)
    rc6 = dense.get_refc6(numbers)
    c6 = dense.get_atomic_c6(rc6, _fixture_gw(name))
    n = len(numbers)
    ref = torch.tensor(GOLD[name]["c6"], dtype=F64).reshape(n, n)
    assert torch.allclose(c6, ref, rtol=1e-10, atol=0)


@pytest.mark.parametrize("name", ["LiH", "SiH4"])
def test_weights_fixture_and_normalisation(name):
    """Weights from the fixture geometry and charges agree with the fixture
    weights to the reference's own tolerance (test_model/test_weights.py:45;
    the Fortran fixture differs by ~1.4e-7 relative), and
    sum_r gw_r / zeta_r = 1 (normalisation, model/d4.py:200)."""
    numbers, positions = _mol(name)
    q = torch.tensor(GOLD[name]["q"], dtype=F64)
    cn = dense.cn_d4(numbers, positions)
    gw = dense.weight_references(numbers, cn, q)
    assert torch.allclose(gw, _fixture_gw(name), rtol=1e-6, atol=1e-12)
    t = dense.tables()
    refc = t.refc[numbers]
    zeff = t.zeff[numbers].unsqueeze(-1)
    z = dense.zeta(t.gam[numbers].unsqueeze(-1) * dense.GC,
                   t.clsq[numbers] + zeff, q.unsqueeze(-1) + zeff)
    s = torch.where(refc > 0, gw / z, torch.zeros_like(gw)).sum(-1)
    assert torch.allclose(s, torch.ones_like(s), atol=1e-14)


# ------------------------------------------------------------------ D4S


def _fixture_gw_d4s(name):
    """gw_d4s is Fortran-ordered (nref, nat, nat); test/test_model/test_c6.py:
    63-82 pads it to 7 references and permutes it to (nat, nat, ref)."""
    g = GOLD[name]["gw_d4s"]
    nref, na, nb = g["fortran_shape"]
    src = torch.tensor(g["values"], dtype=F64).reshape(nb, na, nref).permute(2, 1, 0)
    src = torch.nn.functional.pad(src, (0, 0, 0, 0, 0, 7 - nref))
    return src.permute(2, 1, 0)


@pytest.mark.parametrize("name", ["LiH", "SiH4"])
def test_d4s_two_body_fixture(name):
    """test/test_d4/samples.py disp2_d4s (dftd4 Fortran D4S two-body energy),
    same call as test/test_d4/test_twobody.py:68-108 with the fixture
    charges (reference tolerance 1.5e-7); the D4 fixture matches to 1e-10."""
    numbers, positions = _mol(name)
    q = torch.tensor(GOLD[name]["q"], dtype=F64)
    parts = dense.dftd4(numbers, positions, 0.0, _p(GOLD["param_tpssh"]), q=q,
                        return_parts=True, model="d4s")
    ref = torch.tensor(GOLD[name]["disp2_d4s"], dtype=F64)
    assert (parts["e2"] - ref).abs().max().item() < 1e-9
    # the D4 model is measurably different on the same input
    d4 = dense.dftd4(numbers, positions, 0.0, _p(GOLD["param_tpssh"]), q=q,
                     return_parts=True)
    assert (d4["e2"] - ref).abs().max().item() > 1e-6


@pytest.mark.parametrize("name", ["LiH", "SiH4"])
def test_d4s_c6_and_weights_fixture(name):
    """c6_d4s from the fixture weights (test/test_model/test_c6.py:63-82):
    pins the partner/atom index order of the D4S contraction; weights from
    the fixture geometry and charges agree with gw_d4s to the D4 bound."""
    numbers, positions = _mol(name)
    n = len(numbers)
    gw_ref = _fixture_gw_d4s(name)
    c6 = dense.get_atomic_c6_d4s(dense.get_refc6(numbers), gw_ref)
    ref = torch.tensor(GOLD[name]["c6_d4s"], dtype=F64).reshape(n, n)
    assert torch.allclose(c6, ref, rtol=1e-10, atol=0)
    q = torch.tensor(GOLD[name]["q"], dtype=F64)
    gw = dense.weight_references_d4s(numbers, dense.cn_d4(numbers, positions), q)
    assert torch.allclose(gw, gw_ref, rtol=1e-6, atol=1e-12)


def test_d4s_reduces_to_d4_for_constant_factor(monkeypatch):
    """With every pairwise factor equal to the D4 factor (6) the D4S weights
    are the D4 weights for every partner, and the two models give the same
    energies; with the real table the C6 matrix stays symmetric and each
    partner's weights are normalised (sum_r gw / zeta = 1)."""
    m = _rand_small(3, 14)
    numbers, positions = torch.tensor(m.numbers), torch.tensor(m.positions)
    t = dense.tables()
    cn = dense.cn_d4(numbers, positions)
    q = dense.eeq_charges(numbers, positions, torch.tensor(m.charge, dtype=F64))
    gw4 = dense.weight_references(numbers, cn, q)
    gws = dense.weight_references_d4s(numbers, cn, q)
    rc6 = dense.get_refc6(numbers)
    c6s = dense.get_atomic_c6_d4s(rc6, gws)
    assert torch.allclose(c6s, c6s.mT, rtol=1e-14, atol=0)
    refc = t.refc[numbers]
    zeff = t.zeff[numbers].unsqueeze(-1)
    z = dense.zeta(t.gam[numbers].unsqueeze(-1) * dense.GC,
                   t.clsq[numbers] + zeff, q.unsqueeze(-1) + zeff)
    s = torch.where(refc > 0, gws / z, torch.zeros_like(gws)).sum(-1)
    assert torch.allclose(s, torch.ones_like(s), atol=1e-13)
    assert (gws - gw4.unsqueeze(0)).abs().max().item() > 1e-3
    const = dense.Tables.__new__(dense.Tables)
    const.__dict__.update(t.__dict__)
    const.wfpair = torch.full_like(t.wfpair, 6.0)
    monkeypatch.setattr(dense, "tables", lambda dtype=F64: const)
    gwc = dense.weight_references_d4s(numbers, cn, q)
    assert torch.allclose(gwc, gw4.unsqueeze(0).expand_as(gwc), rtol=1e-12, atol=1e-300)
    par = _p(GOLD["param_tpssh"])
    e4 = dense.dftd4(numbers, positions, m.charge, par)
    es = dense.dftd4(numbers, positions, m.charge, par, model="d4s")
    assert torch.allclose(es, e4, rtol=1e-12, atol=1e-18)


def test_d4s_gradient_finite_difference():
    """Analytic (nan-safe autograd) D4S gradient vs central differences."""
    m = _rand_small(5, 10)
    numbers, positions = torch.tensor(m.numbers), torch.tensor(m.positions)
    par = _p(GOLD["param_tpssh"])
    _, g = dense.dftd4_with_grad(numbers, positions, m.charge, par, model="d4s")
    h = 1e-5
    for i, k in [(0, 0), (3, 1), (7, 2)]:
        pp = positions.clone()
        pm = positions.clone()
        pp[i, k] += h
        pm[i, k] -= h
        fd = (dense.dftd4(numbers, pp, m.charge, par, model="d4s").sum()
              - dense.dftd4(numbers, pm, m.charge, par, model="d4s").sum()) / (2 * h)
        assert abs(fd.item() - g[i, k].item()) < 1e-9


# ------------------------------------------------------------- brute force


def _rand_small(seed, n, spacing=2.6, mix=synth.MIX_Z54, charges=(-1, 0, 1)):
    rng = np.random.default_rng(seed)
    return synth.molecule(rng, n, mix, spacing, charges)


@pytest.mark.parametrize("seed", range(4))
def test_brute_cn_eeq_e2_e3(seed):
    m = _rand_small(seed, 7)
    numbers = torch.tensor(m.numbers)
    positions = torch.tensor(m.positions)
    z, p = m.numbers.tolist(), m.positions.tolist()
    rcov, en = elements.COV_D3, elements.PAULING_EN
    cn = dense.cn_d4(numbers, positions)
    assert np.allclose(cn.numpy(), brute.cn_d4(z, p, rcov, en), rtol=1e-13, atol=0)
    q = dense.eeq_charges(numbers, positions, torch.tensor(m.charge, dtype=F64))
    qb = brute.eeq_charges(z, p, m.charge, rcov, elements.EEQ_CHI, elements.EEQ_ETA,
                           elements.EEQ_KCN, elements.EEQ_RAD)
    assert np.allclose(q.numpy(), qb, rtol=0, atol=1e-12)
    par = _p({"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.4, "a2": 5.0})
    # a cutoff that splits the molecule so the screening rules are exercised
    cut = {"disp2": 6.0, "disp3": 5.0, "cn_eeq": 25.0}
    parts = dense.dftd4(numbers, positions, m.charge, par, cutoff=cut,
                        return_parts=True)
    r4r2 = dense.tables().r4r2.tolist()
    e2 = brute.twobody(z, p, parts["c6"].tolist(), r4r2, 1.0, 1.2, 0.4, 5.0, 6.0)
    assert np.allclose(parts["e2"].numpy(), e2, rtol=1e-13, atol=1e-18)
    rc6 = dense.get_refc6(numbers)
    c60 = dense.get_atomic_c6(rc6, dense.weight_references(numbers, cn, None))
    e3 = brute.threebody(z, p, c60.tolist(), r4r2, 1.0, 0.4, 5.0, 16.0, 5.0)
    assert np.allclose(parts["e3"].numpy(), e3, rtol=1e-12, atol=1e-20)


def test_brute_atm_screening_is_asymmetric():
    """A triple whose edge ik alone exceeds the cutoff still contributes to
    atom i (threebody.py:153-160 screens only ij and jk)."""
    numbers = torch.tensor([6, 6, 6])
    positions = torch.tensor([[0.0, 0, 0], [3.0, 0, 0], [6.0, 0, 0]], dtype=F64)
    cn = dense.cn_d4(numbers, positions)
    rc6 = dense.get_refc6(numbers)
    c60 = dense.get_atomic_c6(rc6, dense.weight_references(numbers, cn, None))
    r4r2 = dense.tables().r4r2
    radii = 0.4 * torch.sqrt(3.0 * r4r2[numbers].unsqueeze(-1) * r4r2[numbers]) + 5.0
    c9 = torch.sqrt(torch.abs(c60.unsqueeze(-1) * c60.unsqueeze(-2) * c60.unsqueeze(-3)))
    e = dense.atm_dispersion(numbers, positions, c9, radii, 4.0, 1.0, 16.0)
    # edges 0-1 and 1-2 are inside 4 bohr, 0-2 is not: atoms 0 and 2 each get
    # e/6 from one ordered pair, the middle atom gets nothing.
    assert e[0].item() != 0.0 and e[2].item() != 0.0
    assert e[1].item() == 0.0
    eb = brute.threebody([6, 6, 6], positions.tolist(), c60.tolist(),
                         r4r2.tolist(), 1.0, 0.4, 5.0, 16.0, 4.0)
    assert np.allclose(e.numpy(), eb, rtol=1e-13)


# --------------------------------------------------------- closed forms


def test_two_atom_closed_forms():
    """Two atoms: CN = den * erfcount, EEQ solved by hand, ATM = 0."""
    numbers = torch.tensor([8, 1])
    r = 1.9
    positions = torch.tensor([[0.0, 0, 0], [0.0, 0, r]], dtype=F64)
    t = dense.tables()
    rc = t.cov_d3[8] + t.cov_d3[1]
    den = dense.D4_K4 * math.exp(-((abs(t.en[8] - t.en[1]).item() + dense.D4_K5) ** 2)
                                 / dense.D4_K6)
    cnt = 0.5 * (1 + math.erf(-7.5 * (r / rc.item() - 1)))
    cn = dense.cn_d4(numbers, positions)
    assert math.isclose(cn[0].item(), den * cnt, rel_tol=1e-14)
    # EEQ: [[a b 1][b c 1][1 1 0]] x = [u v Q] -> q0 = (u - v + Q(c - b)) / (a + c - 2b)
    cap = math.log(1 + math.exp(8)) - math.log(1 + math.exp(8 - cnt))
    a = t.eta[8].item() + math.sqrt(2 / math.pi) / t.rad[8].item()
    c = t.eta[1].item() + math.sqrt(2 / math.pi) / t.rad[1].item()
    g = 1 / math.sqrt(t.rad[8].item() ** 2 + t.rad[1].item() ** 2)
    b = math.erf(g * r) / r
    u = -t.chi[8].item() + t.kcn[8].item() * math.sqrt(cap)
    v = -t.chi[1].item() + t.kcn[1].item() * math.sqrt(cap)
    q0 = (u - v) / (a + c - 2 * b)
    q = dense.eeq_charges(numbers, positions, torch.tensor(0.0, dtype=F64))
    assert math.isclose(q[0].item(), q0, rel_tol=1e-13)
    par = _p({"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.4, "a2": 5.0})
    parts = dense.dftd4(numbers, positions, 0.0, par, return_parts=True)
    assert torch.all(parts["e3"] == 0)


def test_trapzd_is_explicit_sum():
    numbers = torch.tensor([6, 1])
    alpha = dense.get_alpha(numbers)
    w = dense.tables().cp_weights
    c = dense.trapzd(alpha)
    manual = 3.0 / math.pi * sum(w[k] * alpha[0, 2, k] * alpha[1, 1, k] for k in range(23))
    assert math.isclose(c[0, 1, 2, 1].item(), manual.item(), rel_tol=1e-14)


def test_c8_and_as_matrix():
    """E = 1/2 sum of the as_matrix form (test_d4/test_twobody.py:113-141)."""
    m = _rand_small(11, 9)
    numbers, positions = torch.tensor(m.numbers), torch.tensor(m.positions)
    par = _p({"s6": 1.0, "s8": 1.2, "a1": 0.4, "a2": 5.0})
    cn = dense.cn_d4(numbers, positions)
    c6 = dense.get_atomic_c6(dense.get_refc6(numbers),
                             dense.weight_references(numbers, cn, None))
    r4r2 = dense.tables().r4r2[numbers]
    e = dense.dispersion2(numbers, positions, par, c6, r4r2)
    em = dense.dispersion2(numbers, positions, par, c6, r4r2, as_matrix=True)
    assert torch.allclose(e, 0.5 * em.sum(-1), rtol=1e-14)


# ------------------------------------------------------------- invariants


def test_invariants_translation_rotation_forces():
    m = _rand_small(5, 10)
    numbers, positions = torch.tensor(m.numbers), torch.tensor(m.positions)
    par = _p({"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.4, "a2": 5.0})
    e, g = dense.dftd4_with_grad(numbers, positions, m.charge, par)
    assert abs(g.sum(0)).max().item() < 1e-15 * 1e3
    rot = torch.tensor(
        [[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]], dtype=F64
    )
    e2 = dense.dftd4(numbers, positions @ rot.T + 3.7, m.charge, par)
    assert torch.allclose(e, e2, rtol=1e-12, atol=1e-18)


def test_charge_conservation_and_padding():
    mols = [_rand_small(s, n) for s, n in ((1, 6), (2, 11), (3, 3))]
    numbers, positions, charge = synth.pack(mols, pad_to=13)
    q = dense.eeq_charges(numbers, positions, charge)
    assert torch.allclose(q.sum(-1), charge, atol=1e-13)
    par = _p({"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.4, "a2": 5.0})
    e = dense.dftd4(numbers, positions, charge, par)
    for k, m in enumerate(mols):
        n = len(m.numbers)
        es = dense.dftd4(torch.tensor(m.numbers), torch.tensor(m.positions), m.charge, par)
        assert torch.allclose(e[k, :n], es, rtol=1e-12, atol=1e-18)
        assert torch.all(e[k, n:] == 0)


def test_term_switches():
    """s9 = 0 leaves the two-body term; s6 = s8 = 0 leaves ATM
    (test_d4/test_twobody.py:144-168, test_d4/test_atm.py:143-169)."""
    m = _rand_small(7, 8)
    numbers, positions = torch.tensor(m.numbers), torch.tensor(m.positions)
    par = {"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.4, "a2": 5.0}
    parts = dense.dftd4(numbers, positions, m.charge, _p(par), return_parts=True)
    e_no3 = dense.dftd4(numbers, positions, m.charge, _p({**par, "s9": 0.0}))
    e_only3 = dense.dftd4(numbers, positions, m.charge, _p({**par, "s6": 0.0, "s8": 0.0}))
    assert torch.allclose(e_no3, parts["e2"], rtol=1e-14)
    assert torch.allclose(e_only3, parts["e3"], rtol=1e-14, atol=1e-25)


def test_finite_difference_gradient():
    """Central differences, h = 1e-5 bohr (examples/forces.py:53-73)."""
    m = _rand_small(9, 8)
    numbers, positions = torch.tensor(m.numbers), torch.tensor(m.positions)
    par = _p({"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.4, "a2": 5.0})
    _, g = dense.dftd4_with_grad(numbers, positions, m.charge, par)
    h = 1e-5
    for i, x in ((0, 0), (3, 1), (7, 2)):
        pp, pm = positions.clone(), positions.clone()
        pp[i, x] += h
        pm[i, x] -= h
        fd = (dense.dftd4(numbers, pp, m.charge, par).sum()
              - dense.dftd4(numbers, pm, m.charge, par).sum()) / (2 * h)
        assert abs(fd.item() - g[i, x].item()) < 1e-10


def test_exceptional_weights_reference_gradient_is_nan():
    """When every Gaussian weight of an atom underflows (its CN is far from
    all reference CNs), the reference falls back to the maximum-CN reference
    (model/d4.py:209-217). Its autograd gradient is then NaN (0/0 in the
    discarded branch of torch.where). The NaN-safe oracle mode keeps the
    energies bit-identical and gives the finite gradient, which matches
    central finite differences; the GPU backward is checked against it."""
    mol = synth.molecule(np.random.default_rng(3), 64, synth.MIX_Z54, 2.8, (-1, 0, 1))
    assert 41 in mol.numbers.tolist()  # Nb, CN ~12 away from its references
    numbers, positions = torch.tensor(mol.numbers), torch.tensor(mol.positions)
    par = _p({"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.4, "a2": 5.0})
    cut = {"disp2": 50.0, "disp3": 25.0, "cn_eeq": 25.0}
    e_safe, g_safe = dense.dftd4_with_grad(numbers, positions, mol.charge, par, cut)
    e_ref, g_ref = dense.dftd4_with_grad(numbers, positions, mol.charge, par, cut,
                                         nan_safe=False)
    assert torch.equal(e_safe, e_ref)
    assert torch.isnan(g_ref).any() and torch.isfinite(g_safe).all()
    i = mol.numbers.tolist().index(41)
    h = 1e-5
    for x in range(3):
        pp, pm = positions.clone(), positions.clone()
        pp[i, x] += h
        pm[i, x] -= h
        fd = (dense.dftd4(numbers, pp, mol.charge, par, cutoff=cut).sum()
              - dense.dftd4(numbers, pm, mol.charge, par, cutoff=cut).sum()) / (2 * h)
        assert abs(fd.item() - g_safe[i, x].item()) < 1e-9


@pytest.mark.parametrize("name", ["LiH", "SiH4"])
def test_hessian_fixture(name):
    """dftd4 3.6.0 Hessians (test/test_grad/samples_hessian.py, TPSS0,
    Fortran order, test_grad/test_hessian.py:73-103; reference tolerance
    1e-7) against the oracle's autograd Hessian, which is symmetric and has
    zero row sums (translation invariance)."""
    numbers, positions = _mol(name)
    n = len(numbers)
    par = _p(GOLD["param_tpss0"])
    h = torch.autograd.functional.hessian(
        lambda p: dense.dftd4(numbers, p, 0.0, par).sum(), positions).reshape(n, 3, n, 3)
    ref = torch.tensor(GOLD[name]["hessian"], dtype=F64).reshape(3, n, 3, n).permute(3, 2, 1, 0)
    assert (h - ref).abs().max().item() < 1e-9
    assert (h - h.permute(2, 3, 0, 1)).abs().max().item() < 1e-18
    assert h.sum(dim=2).abs().max().item() < 1e-15


# ------------------------------------------------------ MB16-43 fixtures
#
# test/test_model/samples.py holds charges, weights and C6 of three
# 16-atom molecules of tad-mctc's MB16-43 set, whose atomic numbers and
# geometries live in tad-mctc (absent). The atomic numbers are inferred from
# the values (test_mb16_atomic_numbers_are_identified checks that each is
# the only element that fits); the fixtures then pin the charge scaling
# (hardness GAM and Z_eff) and the reference C6 of B, F, Na, Mg, Al, Cl as
# well as H, Li, C, N, O, Si, S.

MB16_Z = {
    "MB16_43_01": [11, 1, 8, 1, 9, 1, 1, 8, 7, 1, 1, 17, 5, 5, 7, 13],
    "MB16_43_02": [1, 16, 5, 8, 12, 1, 1, 1, 14, 1, 5, 3, 9, 1, 1, 16],
    "MB16_43_03": [6, 8, 1, 3, 12, 13, 6, 1, 1, 1, 9, 16, 6, 1, 11, 1],
}


def _zeta_ref(numbers, q):
    """zeta(q) of every reference system of every atom (model/base.py:
    308-335), zero on unused slots."""
    t = dense.tables()
    zeff = t.zeff[numbers].unsqueeze(-1)
    z = dense.zeta(t.gam[numbers].unsqueeze(-1) * dense.GC, t.clsq[numbers] + zeff,
                   q.unsqueeze(-1) + zeff)
    return torch.where(t.refc[numbers] > 0, z, torch.zeros_like(z))


def _mb16(name):
    z = torch.tensor(MB16_Z[name])
    q = torch.tensor(GOLD[name]["q"], dtype=F64)
    c6 = torch.tensor(GOLD[name]["c6"], dtype=F64).reshape(16, 16)
    return z, q, _fixture_gw(name), c6


def _mb16_01_weights_at_q(gw, z, q):
    """The 01 weights were made at q = 0 (test_model/test_weights.py:79-81);
    the Gaussian part does not depend on q, so the weights at the charges
    are gw * zeta(q) / zeta(0) (model/d4.py:219-228)."""
    z0 = _zeta_ref(z, torch.zeros_like(q))
    zq = _zeta_ref(z, q)
    return torch.where(z0 > 0, gw * zq / torch.where(z0 > 0, z0, torch.ones_like(z0)),
                       torch.zeros_like(gw))


def _mb16_value(name, zc, i, gw, q, c6):
    """What the fixture pins for atom i if it had atomic number zc: the
    relative C6 diagonal (01, 03) or the largest weight deviation (02)."""
    t = dense.tables()
    zt = torch.tensor([zc])
    if name == "MB16_43_02":
        w = dense.weight_references(zt, None, q[i: i + 1])[0]
        return (w - gw[i]).abs().max().item()
    g = gw[i: i + 1]
    if name == "MB16_43_01":
        g = _mb16_01_weights_at_q(g, zt, q[i: i + 1])
    g = torch.where(t.refc[zt] > 0, g, torch.zeros_like(g))
    v = (g[0] @ dense.get_refc6(zt)[0, 0] @ g[0]).item()
    return abs(v / c6[i, i].item() - 1.0)


@pytest.mark.parametrize("name", sorted(MB16_Z))
def test_mb16_atomic_numbers_are_identified(name):
    _, q, gw, c6 = _mb16(name)
    refc = dense.tables().refc
    for i, zi in enumerate(MB16_Z[name]):
        used = gw[i] != 0
        scores = []
        for zc in range(1, 104):
            if bool((used & (refc[zc] == 0)).any()):
                continue  # weight on a reference system the element lacks
            scores.append((_mb16_value(name, zc, i, gw, q, c6), zc))
        scores.sort()
        assert scores[0][1] == zi, (name, i, scores[:3])
        assert scores[0][0] < 1e-6 and scores[1][0] > 1e-4, (name, i, scores[:3])


def test_mb16_43_03_c6_from_fixture_weights():
    """test/test_model/test_c6.py:32-58 with MB16_43_03 (weights at the
    molecule's CN and charges): pins the reference C6 of every element pair
    of C, O, H, Li, Mg, Al, F, S, Na."""
    z, _, gw, c6 = _mb16("MB16_43_03")
    out = dense.get_atomic_c6(dense.get_refc6(z), gw)
    assert torch.allclose(out, c6, rtol=1e-10, atol=0)


def test_mb16_43_01_c6_from_weights_and_charges():
    """MB16_43_01's weights (q = 0) rescaled to the fixture charges give the
    fixture C6: pins zeta(q) / zeta(0), i.e. GAM and Z_eff, and the
    reference C6 of Na, H, O, F, N, Cl, B, Al (the Fortran weights carry
    ~1e-7 relative, as the LiH/SiH4 weights do)."""
    z, q, gw, c6 = _mb16("MB16_43_01")
    out = dense.get_atomic_c6(dense.get_refc6(z), _mb16_01_weights_at_q(gw, z, q))
    assert torch.allclose(out, c6, rtol=1e-6, atol=0)


def test_mb16_43_02_weights_at_zero_cn():
    """MB16_43_02's weights were made at CN = 0 with the fixture charges
    (test/test_model/test_weights.py:84-86): they follow from the tables
    alone. Pins the reference CNs, Gaussian counts and the charge scaling of
    H, S, B, O, Mg, Si, Li, F to the reference's own tolerance (atol 1e-6,
    test_weights.py:45)."""
    z, q, gw, _ = _mb16("MB16_43_02")
    w = dense.weight_references(z, None, q)
    assert torch.allclose(w, gw, rtol=0, atol=1e-6)
    assert (w - gw).abs().max().item() < 1e-7


@pytest.mark.parametrize("name", sorted(MB16_Z))
def test_mb16_weights_normalised(name):
    """sum_r gw_r / zeta_r = 1 per atom (model/d4.py:200-207), with zeta at
    the charges the weights were made with (q = 0 for 01)."""
    z, q, gw, _ = _mb16(name)
    qq = torch.zeros_like(q) if name == "MB16_43_01" else q
    zz = _zeta_ref(z, qq)
    s = torch.where(zz > 0, gw / torch.where(zz > 0, zz, torch.ones_like(zz)),
                    torch.zeros_like(gw)).sum(-1)
    assert torch.allclose(s, torch.ones_like(s), rtol=0, atol=1e-6)


# ------------------------------------------------------- polarisabilities


def _trapzd_pols(wp):
    """(3/pi) sum_w cp_w a_i(w) a_j(w): utils.trapzd_noref (utils.py:33-94),
    the identity the reference's own test uses
    (test/test_model/test_models.py:60-76)."""
    cp = dense.tables().cp_weights
    return 3.0 / math.pi * torch.einsum("iw,jw,w->ij", wp, wp, cp)


@pytest.mark.parametrize("name", ["LiH", "SiH4", "MB16_43_03", "MB16_43_01"])
def test_weighted_pols_reproduce_fixture_c6(name):
    """get_weighted_pols of the fixture weights through the Casimir-Polder
    quadrature gives the fixture C6 matrix: a wrong weight, slot, element or
    frequency column in the weighted polarisabilities fails this."""
    if name.startswith("MB16"):
        z, q, gw, c6 = _mb16(name)
        if name == "MB16_43_01":
            gw = _mb16_01_weights_at_q(gw, z, q)
        tol = 1e-6 if name == "MB16_43_01" else 1e-10
    else:
        z, _ = _mol(name)
        gw = _fixture_gw(name)
        n = len(z)
        c6 = torch.tensor(GOLD[name]["c6"], dtype=F64).reshape(n, n)
        tol = 1e-10
    wp = dense.get_weighted_pols(z, gw)
    assert wp.shape == (len(z), 23)
    assert torch.allclose(_trapzd_pols(wp), c6, rtol=tol, atol=0)


def test_static_polarizabilities_are_the_zero_frequency_column():
    """The Casimir-Polder weights (utils.py:53-78) are exactly the trapezoid
    weights of dftd4's published imaginary-frequency grid (1e-6, 0.05, 0.1,
    0.2, ..., 2.0, 2.5, 3, 4, 5, 7.5, 10 a.u.; w_0 = (f_1 - f_0)/2,
    w_k = (f_{k+1} - f_{k-1})/2, w_22 = (f_22 - f_21)/2), so column k of
    alpha(i w) is frequency f_k and column 0 (f_0 = 1e-6) is the static
    limit. get_polarizabilities (model/base.py:286-302) is that column of
    the weighted polarisabilities pinned above."""
    f = [1e-6, 0.05, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.2, 1.4, 1.6, 1.8,
         2.0, 2.5, 3.0, 4.0, 5.0, 7.5, 10.0]
    trap = [(f[1] - f[0]) / 2] + [(f[k + 1] - f[k - 1]) / 2 for k in range(1, 22)]
    trap.append((f[22] - f[21]) / 2)
    cp = dense.tables().cp_weights
    assert torch.allclose(torch.tensor(trap, dtype=F64), cp, rtol=0, atol=1e-15)
    for name in ("LiH", "SiH4"):
        z, _ = _mol(name)
        gw = _fixture_gw(name)
        assert torch.equal(dense.get_polarizabilities(z, gw), dense.get_weighted_pols(z, gw)[:, 0])


def test_d4s_weighted_pols_average_over_partners():
    """D4S weighted polarisabilities (model/d4s.py:287-313) sum the pair
    weights over the partner axis n and divide by the atom count: with pair
    weights that do not depend on the partner (gw[m, n] = gw_n) every atom m
    gets the mean of the D4 weighted polarisabilities; a padding atom counts
    in neither."""
    z, _ = _mol("SiH4")
    gw = _fixture_gw("SiH4")
    n = len(z)
    zp = torch.cat([z, torch.zeros(2, dtype=z.dtype)])
    gwp = torch.cat([gw, torch.zeros(2, 7, dtype=F64)])
    pair = gwp.unsqueeze(0).expand(n + 2, n + 2, 7)
    wp = dense.get_weighted_pols(zp, pair, model="d4s")
    mean = dense.get_weighted_pols(z, gw).sum(0) / n
    assert torch.allclose(wp, mean.expand(n + 2, 23), rtol=1e-14, atol=0)
