"""
The persistent one-pass 8 x 8 three-body kernel (waves taking (molecule,
split) items from per-XCD work queues) against the one-wave-per-item grid
launch (D4_FLAG_ATM8_GRID): bit for bit, for energies, gradients and
parameter gradients, with 1, 2 and 3 waves per SIMD, on batches with empty
(all-padding) molecules, one-atom molecules and more molecules than waves.
"""

import sys
from dataclasses import replace

import numpy as np
import pytest
import torch

import tad_dftd4_amd as d4
from tad_dftd4_amd import runtime, synth
from tad_dftd4_amd.disp import kernel_params

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason=(
        "the persistent work-queue version of the one-pass 8 x 8 kernel is taken out: the "
        "first GPU run of these tests (default launch = the queue kernel) printed nothing for "
        "180 s and was killed, and reading the code did not find the cause; the batch with "
        "empty and one-atom molecules ran in that same run and stays off the GPU with it")),
]

PARAM = {"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.40, "a2": 5.0}


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _kp(**kw):
    kp = kernel_params(PARAM, d4.Cutoff(disp2=50.0, disp3=25.0, cn=25.0, cn_eeq=25.0), None)
    return replace(kp, **kw)


def _batch(seed, sizes, pad_to=None):
    rng = np.random.default_rng(seed)
    mols = []
    for n in sizes:
        if n == 0:
            mols.append(synth.Molecule(np.zeros(0, dtype=np.int64), np.zeros((0, 3)), 0.0))
        else:
            mols.append(synth.molecule(rng, n, synth.MIX_Z54, 2.8, (-1, 0, 1)))
    numbers, positions, charge = synth.pack([m for m in mols if len(m.numbers)] or mols,
                                             pad_to=pad_to)
    # re-insert the empty molecules as all-padding rows
    B = len(mols)
    N = numbers.shape[1]
    nn = torch.zeros(B, N, dtype=torch.int64)
    pp = torch.zeros(B, N, 3, dtype=torch.float64)
    cc = torch.zeros(B, dtype=torch.float64)
    k = 0
    for b, m in enumerate(mols):
        if len(m.numbers):
            nn[b], pp[b], cc[b] = numbers[k], positions[k], charge[k]
            k += 1
    return nn.cuda(), pp.cuda(), cc.cuda()


CASES = [
    ("mixed sizes with empty and one-atom molecules", 501, [0, 1, 7, 33, 64, 0, 120, 2, 90]),
    ("1024 molecules of 20-120 atoms", 502, None),
]


@pytest.mark.parametrize("label,seed,sizes", CASES)
def test_queue_kernel_matches_grid_launch(label, seed, sizes):
    _say(f"persistent vs grid one-pass kernel: {label}")
    if sizes is None:
        mols = synth.workload("batch_1024_20_120")
        numbers, positions, charge = synth.pack(mols, device="cuda")
    else:
        numbers, positions, charge = _batch(seed, sizes)
    ref_e, ref_g, _ = runtime.energy_gradient(numbers, positions, charge, _kp(atm_grid=True))
    ws = runtime.new_workspace(*numbers.shape, "cuda")
    ref_f, _ = runtime.forward(numbers, positions, charge, _kp(atm_grid=True, atm_tile=8),
                               ws=ws)
    ref_b, ref_p = runtime.backward(numbers, positions, charge, _kp(atm_grid=True), ws,
                                    params=True)
    for wps in (0, 1, 2, 3):
        e, g, _ = runtime.energy_gradient(numbers, positions, charge, _kp(atm_wps=wps))
        assert torch.equal(e, ref_e) and torch.equal(g, ref_g), wps
        ws2 = runtime.new_workspace(*numbers.shape, "cuda")
        f, _ = runtime.forward(numbers, positions, charge, _kp(atm_wps=wps, atm_tile=8),
                               ws=ws2)
        assert torch.equal(f, ref_f), wps
        gb, pb = runtime.backward(numbers, positions, charge, _kp(atm_wps=wps), ws2,
                                  params=True)
        assert torch.equal(gb, ref_b) and torch.equal(pb, ref_p), wps


def test_queue_kernel_repeated_calls_and_graph():
    """The queue counters are reset by every call (replays of a captured
    step, eager calls in between)."""
    _say("persistent one-pass kernel: repeated calls and graph replays")
    numbers, positions, charge = _batch(503, [40, 11, 77, 0, 25])
    kp = _kp()
    e0, g0, _ = runtime.energy_gradient(numbers, positions, charge, kp)
    cap = runtime.CapturedStep(numbers, positions, charge, kp, forces=True)
    for k in range(3):
        e, g = cap.replay()
        torch.cuda.synchronize()
        assert torch.equal(e, e0) and torch.equal(g, g0), k
        e1, g1, _ = runtime.energy_gradient(numbers, positions, charge, kp)
        assert torch.equal(e1, e0) and torch.equal(g1, g0), k
