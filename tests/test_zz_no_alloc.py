"""
d4_energy_gradient_noalloc (no workspace argument, no device allocation: the
library's fixed per-device scratch) against d4_energy_gradient with a
caller workspace, through the C ABI: calls in a row on the same scratch,
each followed by a smaller one, bit for bit (the same kernels); the
large-molecule (spatially sorted) path; a batch above the scratch size is
rejected before any launch.
"""

import sys

import numpy as np
import pytest
import torch

from tad_dftd4_amd import _lib, runtime, synth
from tad_dftd4_amd.cutoff import Cutoff
from tad_dftd4_amd.disp import kernel_params

pytestmark = pytest.mark.gpu

PARAM = {"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.40, "a2": 5.0}


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _batch(seed, sizes, pad_to=None):
    rng = np.random.default_rng(seed)
    mols = [synth.molecule(rng, n, synth.MIX_Z54, 2.8, (-1, 0, 1)) for n in sizes]
    return synth.pack(mols, pad_to=pad_to, device="cuda")


def test_noalloc_matches_workspace_entry_point():
    kp = kernel_params(PARAM, Cutoff(disp2=50.0, disp3=25.0, cn_eeq=25.0), None)
    cases = [
        ("16 molecules padded to 120", _batch(81, list(range(20, 116, 6)), pad_to=120)),
        ("5 molecules padded to 40", _batch(82, (40, 7, 33, 12, 25))),
        ("one 300-atom molecule (sorted path)", _batch(83, (300,))),
        ("2 molecules padded to 24", _batch(84, (24, 9))),
    ]
    for label, (numbers, positions, charge) in cases:
        B, N = numbers.shape
        _say(f"noalloc vs workspace: {label} ({runtime.workspace_bytes(B, N)} bytes)")
        assert runtime.workspace_bytes(B, N) <= _lib.D4_NOALLOC_SCRATCH_BYTES
        e_ref, g_ref, _ = runtime.energy_gradient(numbers, positions, charge, kp)
        energy = torch.full((B, N), float("nan"), dtype=torch.float64, device="cuda")
        grad = torch.full((B, N, 3), float("nan"), dtype=torch.float64, device="cuda")
        e, g = runtime.energy_gradient_noalloc(numbers, positions, charge, kp,
                                               energy=energy, grad=grad)
        torch.cuda.synchronize()
        assert e.data_ptr() == energy.data_ptr() and g.data_ptr() == grad.data_ptr()
        assert torch.equal(e, e_ref), label
        assert torch.equal(g, g_ref), label


def test_noalloc_rejects_batches_above_the_scratch():
    _say("noalloc: batch above the scratch is rejected")
    kp = kernel_params(PARAM, Cutoff(disp2=50.0, disp3=25.0, cn_eeq=25.0), None)
    B, N = 64, 120
    assert runtime.workspace_bytes(B, N) > _lib.D4_NOALLOC_SCRATCH_BYTES
    numbers = torch.zeros((B, N), dtype=torch.int64, device="cuda")
    positions = torch.zeros((B, N, 3), dtype=torch.float64, device="cuda")
    charge = torch.zeros(B, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="scratch"):
        runtime.energy_gradient_noalloc(numbers, positions, charge, kp)
