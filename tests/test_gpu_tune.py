"""
Launch shapes and the shape tuner on the device. Every (lanes per atom row,
three-body pair tile) pair gives the built-in shapes' energies and gradient
to rounding (the shapes change only the order of the row and tile sums) and
the oracle's within the parity bar (1e-10 Eh, 1e-9 Eh/bohr); the row shapes
also on the large-system path; a table from ``tune.tune`` drives later calls,
the autograd path included, with the same results.
"""

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

import tad_dftd4_amd as d4
from oracle import dense
from tad_dftd4_amd import runtime, synth, tune
from tad_dftd4_amd.disp import kernel_params

pytestmark = pytest.mark.gpu

F64 = torch.float64
E_TOL = 1e-10
G_TOL = 1e-9
SAME = 1e-12
PARAM = {"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.40, "a2": 5.0}
CUT = {"disp2": 50.0, "disp3": 25.0, "cn_eeq": 25.0}
SHAPES = [(r, t) for r in tune.ROWS for t in tune.TILES]


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _cut():
    return d4.Cutoff(disp2=CUT["disp2"], disp3=CUT["disp3"], cn=25.0, cn_eeq=CUT["cn_eeq"])


def _kp(rows=0, tile=0):
    kp = kernel_params(PARAM, _cut(), None)
    kp.row_lanes, kp.atm_tile = rows, tile
    return kp


def _mols():
    rng = np.random.default_rng(901)
    return [synth.molecule(rng, n, synth.MIX_Z54, 2.8, (-1, 0, 1)) for n in (9, 47, 90)]


_ORACLE = {}


def _oracle():
    if not _ORACLE:
        tp = {k: torch.tensor(v, dtype=F64) for k, v in PARAM.items()}
        for i, m in enumerate(_mols()):
            e, g = dense.dftd4_with_grad(torch.tensor(m.numbers), torch.tensor(m.positions),
                                         m.charge, tp, cutoff=CUT)
            _ORACLE[i] = (e.numpy(), g.numpy())
    return _ORACLE


def _run(kp, mols, forces=True):
    numbers, positions, charge = synth.pack(mols, device="cuda")
    if forces:
        e, g, _ = runtime.energy_gradient(numbers, positions, charge, kp)
        return e.cpu().numpy(), g.cpu().numpy()
    e, _ = runtime.forward(numbers, positions, charge, kp)
    return e.cpu().numpy(), None


@pytest.mark.parametrize("shape", SHAPES, ids=[f"rows{r}-tile{t}" for r, t in SHAPES])
def test_shapes_match_builtin_and_oracle(shape):
    _say(f"launch shape {shape} vs built-in and oracle, n = 9/47/90")
    mols = _mols()
    e0, g0 = _run(_kp(), mols)
    e, g = _run(_kp(*shape), mols)
    np.testing.assert_allclose(e, e0, rtol=0, atol=SAME)
    np.testing.assert_allclose(g, g0, rtol=0, atol=SAME)
    ee, _ = _run(_kp(*shape), mols, forces=False)
    ee0, _ = _run(_kp(), mols, forces=False)
    np.testing.assert_allclose(ee, ee0, rtol=0, atol=SAME)
    for i, m in enumerate(mols):
        n = len(m.numbers)
        oe, og = _oracle()[i]
        assert np.abs(e[i, :n] - oe).max() <= E_TOL
        assert np.abs(ee[i, :n] - oe).max() <= E_TOL
        assert np.abs(g[i, :n] - og).max() <= G_TOL
        assert np.all(e[i, n:] == 0.0) and np.all(g[i, n:] == 0.0)


@pytest.mark.parametrize("rows", tune.ROWS)
def test_row_shapes_large_system(rows):
    _say(f"row lanes {rows} on the large-system path, n = 300 + 40")
    rng = np.random.default_rng(902)
    mols = [synth.molecule(rng, 300, synth.MIX_Z54, 2.8, (0,)),
            synth.molecule(rng, 40, synth.MIX_Z54, 2.8, (1,))]
    e0, g0 = _run(_kp(), mols)
    e, g = _run(_kp(rows, 0), mols)
    np.testing.assert_allclose(e, e0, rtol=0, atol=SAME)
    np.testing.assert_allclose(g, g0, rtol=0, atol=SAME)


def test_tuned_table_drives_calls():
    _say("tune on buckets 48 and 300, then calls through the activated table")
    table = tune.tune("cuda:0", buckets=(48, 300), count=16, reps=2, warmup=1)
    assert tune.active() is None
    for mode in tune.MODES:
        ent = table.entries[mode]
        assert [n for n, _, _ in ent] == [48, 300]
        for n_max, rows, tile in ent:
            assert (rows, tile) in tune.candidates(n_max)
        assert set(table.timings[mode]["48"]) == {f"{r}x{t}" for r, t in tune.candidates(48)}
    table = tune.ShapeTable.from_json(table.to_json())
    mols = [m for m in _mols() if len(m.numbers) <= 48]
    numbers, positions, charge = synth.pack(mols, device="cuda")
    pos = positions.clone().requires_grad_(True)
    e0 = d4.dftd4(numbers, pos, charge, PARAM, cutoff=_cut())
    (g0,) = torch.autograd.grad(e0.sum(), pos)
    tune.activate(table)
    try:
        assert _kp().shapes(numbers.shape[1], True) == tuple(table.entries["forces"][0][1:])
        pos = positions.clone().requires_grad_(True)
        e = d4.dftd4(numbers, pos, charge, PARAM, cutoff=_cut())
        (g,) = torch.autograd.grad(e.sum(), pos)
        ef, gf = d4.dftd4_energy_gradient(numbers, positions, charge, PARAM, cutoff=_cut())
    finally:
        tune.activate(None)
    for a, b in ((e, e0), (g, g0), (ef, e0), (gf, g0)):
        np.testing.assert_allclose(a.detach().cpu().numpy(), b.detach().cpu().numpy(),
                                   rtol=0, atol=SAME)


def test_shipped_table_drives_calls():
    _say("shipped gfx950 table: energies and gradient of a batch padded to 90")
    table = tune.ShapeTable.load(Path(tune.__file__).with_name("shapes_gfx950.json"))
    mols = _mols()
    e0, _ = _run(_kp(), mols, forces=False)
    ef0, g0 = _run(_kp(), mols)
    tune.activate(table)
    try:
        assert _kp().shapes(90, False) == table.lookup(False, 90)
        e, _ = _run(_kp(), mols, forces=False)
        ef, g = _run(_kp(), mols)
    finally:
        tune.activate(None)
    np.testing.assert_allclose(e, e0, rtol=0, atol=SAME)
    np.testing.assert_allclose(ef, ef0, rtol=0, atol=SAME)
    np.testing.assert_allclose(g, g0, rtol=0, atol=SAME)
    for i, m in enumerate(mols):
        n = len(m.numbers)
        oe, og = _oracle()[i]
        assert np.abs(e[i, :n] - oe).max() <= E_TOL
        assert np.abs(g[i, :n] - og).max() <= G_TOL


@pytest.mark.parametrize("n_pad", [128, 160, 256])
def test_energy_splits_on_rotated_xcds(n_pad):
    """Batches whose 16 x 16 split count is a multiple of 8 (8, 16, 64 at 128,
    160, 256 atoms), where each molecule's splits are rotated over the XCDs by
    its index: every molecule's energies are bitwise those it gets alone at the
    same padded length (the same splits on other XCDs) and, up to 160 atoms,
    alone at its own length (another split count; longer batches take the
    spatial order and the tiled EEQ, whose sums differ in rounding), and the
    oracle's."""
    _say(f"16x16 energies, batch padded to {n_pad}: alone vs batch, oracle")
    rng = np.random.default_rng(903 + n_pad)
    sizes = (12, 40, 75, n_pad - 5, n_pad)
    mols = [synth.molecule(rng, n, synth.MIX_Z54, 2.8, (-1, 0, 1)) for n in sizes]
    kp = _kp(0, 16)
    eb, _ = _run(kp, mols, forces=False)
    assert eb.shape[1] == n_pad
    tp = {k: torch.tensor(v, dtype=F64) for k, v in PARAM.items()}
    for i, m in enumerate(mols):
        n = len(m.numbers)
        numbers, positions, charge = synth.pack([m], pad_to=n_pad, device="cuda")
        ep, _ = runtime.forward(numbers, positions, charge, kp)
        assert np.array_equal(eb[i], ep[0].cpu().numpy())
        if n_pad <= 160:
            ea, _ = _run(kp, [m], forces=False)
            assert np.array_equal(eb[i, :n], ea[0, :n])
        assert np.all(eb[i, n:] == 0.0)
        if n <= 130:
            oe = dense.dftd4(torch.tensor(m.numbers), torch.tensor(m.positions), m.charge, tp,
                             cutoff=CUT).numpy()
            assert np.abs(eb[i, :n] - oe).max() <= E_TOL
