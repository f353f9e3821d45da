"""
Large-system path: the tiled multi-work-group EEQ solve (and its adjoint)
that molecules beyond the LDS size take. Forced at small sizes from the host
it must agree with the LDS path; above the LDS size it is checked against the
FP64 oracle directly.
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
E_TOL = 1e-10
G_TOL = 1e-9
PARAM = {"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.40, "a2": 5.0}
CUT = {"disp2": 50.0, "disp3": 25.0, "cn_eeq": 25.0}


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _cut():
    return d4.Cutoff(disp2=CUT["disp2"], disp3=CUT["disp3"], cn=25.0, cn_eeq=CUT["cn_eeq"])


def _mol(seed, n, mix=synth.MIX_Z54, spacing=2.8, charges=(-1, 0, 1)):
    return synth.molecule(np.random.default_rng(seed), n, mix, spacing, charges)


def _run(mols, **flags):
    numbers, positions, charge = synth.pack(mols, device="cuda")
    kp = kernel_params(PARAM, _cut(), None)
    for k, v in flags.items():
        setattr(kp, k, v)
    e, g, ws = runtime.energy_gradient(numbers, positions, charge, kp)
    cn, q, c6, alpha = runtime.properties(numbers, positions, charge, kp)
    return e.cpu(), g.cpu(), q.cpu(), runtime.molecule_status(ws, len(mols))


def test_tiled_eeq_matches_lds_path():
    """Sizes across the 64-column block edges, padded together."""
    _say("tiled EEQ vs LDS EEQ, n = 17, 20, 64, 65, 130, 150")
    mols = [_mol(70 + k, n) for k, n in enumerate((17, 20, 64, 65, 130, 150))]
    e1, g1, q1, st1 = _run(mols)
    e2, g2, q2, st2 = _run(mols, force_tiled=True)
    assert np.all(st1 == 0) and np.all(st2 == 0)
    assert (q1 - q2).abs().max().item() < 1e-12
    assert (e1 - e2).abs().max().item() < 1e-12
    assert (g1 - g2).abs().max().item() < 1e-11


def test_tiled_eeq_forced_lu_and_given_charges():
    _say("tiled EEQ with forced LU, and with given charges")
    mols = [_mol(80, 70), _mol(81, 33)]
    e1, g1, q1, _ = _run(mols)
    e2, g2, q2, st2 = _run(mols, force_tiled=True, force_lu=True)
    assert np.all(st2 == 1)
    assert (q1 - q2).abs().max().item() < 1e-12
    assert (e1 - e2).abs().max().item() < 1e-12
    assert (g1 - g2).abs().max().item() < 1e-11
    numbers, positions, charge = synth.pack(mols, device="cuda")
    qg = q1.to("cuda")
    ea = d4.dftd4(numbers, positions, charge, PARAM, q=qg, cutoff=_cut())
    kp = kernel_params(PARAM, _cut(), None)
    kp.force_tiled = True
    eb, _ = runtime.forward(numbers, positions, charge, kp, q=qg)
    assert (ea - eb).abs().max().item() < 1e-13


def test_tiled_eeq_vs_oracle_n260():
    """n = 260 is beyond the LDS factor: the tiled path is taken by itself."""
    _say("n=260 molecule, tiled EEQ, vs oracle")
    mol = _mol(90, 260)
    numbers = torch.tensor(mol.numbers)
    positions = torch.tensor(mol.positions)
    eo, go = dense.dftd4_with_grad(numbers, positions, mol.charge,
                                   {k: torch.tensor(v, dtype=F64) for k, v in PARAM.items()},
                                   cutoff=CUT)
    e, g, q, st = _run([mol])
    assert st[0] == 0
    assert (e[0] - eo).abs().max().item() < E_TOL
    assert (g[0] - go).abs().max().item() < G_TOL
