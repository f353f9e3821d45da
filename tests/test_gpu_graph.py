"""
Replays of a captured evaluation (``runtime.CapturedStep``: the launch
sequence of the caller's stream and its companion stream in one HIP graph)
against eager calls on the same inputs: every replay, with unrelated eager
work and synchronisations between replays, bit for bit; inputs updated in
place between replays; energy-only and large-system (sorted, screened)
captures.
"""

import sys

import numpy as np
import pytest
import torch

import tad_dftd4_amd as d4
from tad_dftd4_amd import runtime, synth
from tad_dftd4_amd.disp import kernel_params

pytestmark = pytest.mark.gpu

PARAM = {"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.40, "a2": 5.0}


def _say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _kp():
    return kernel_params(PARAM, d4.Cutoff(disp2=50.0, disp3=25.0, cn=25.0, cn_eeq=25.0), None)


def _batch(seed, sizes, **kw):
    rng = np.random.default_rng(seed)
    mols = [synth.molecule(rng, n, kw.get("mix", synth.MIX_Z54), kw.get("spacing", 2.8),
                           kw.get("charges", (-1, 0, 1))) for n in sizes]
    return synth.pack(mols, device="cuda")


@pytest.mark.parametrize("forces", [True, False])
def test_replays_match_eager(forces):
    _say(f"captured evaluation, forces={forces}: replays vs eager")
    numbers, positions, charge = _batch(400, (7, 33, 64, 120, 2))
    kp = _kp()
    cap = runtime.CapturedStep(numbers, positions, charge, kp, forces=forces)
    if forces:
        e0, g0, _ = runtime.energy_gradient(numbers, positions, charge, kp)
    else:
        e0, _ = runtime.forward(numbers, positions, charge, kp)
    other = _batch(401, (50, 9))
    for k in range(4):
        out = cap.replay()
        torch.cuda.synchronize()
        e = out[0] if forces else out
        assert torch.equal(e, e0), k
        if forces:
            assert torch.equal(out[1], g0), k
        # unrelated eager work on the same context and stream between replays
        runtime.energy_gradient(*other, kp)
        torch.cuda.synchronize()
    # new coordinates written into the static input
    pos2 = positions + 0.05 * torch.randn_like(positions)
    pos2[numbers == 0] = 0.0
    cap.positions.copy_(pos2)
    out = cap.replay()
    if forces:
        e2, g2, _ = runtime.energy_gradient(numbers, pos2, charge, kp)
        assert torch.equal(out[0], e2) and torch.equal(out[1], g2)
    else:
        e2, _ = runtime.forward(numbers, pos2, charge, kp)
        assert torch.equal(out, e2)
    assert not torch.equal(e2, e0)


def test_large_system_replays_match_eager():
    _say("captured evaluation of a 300-atom cluster (sorted, screened): replays vs eager")
    numbers, positions, charge = _batch(402, (300,), mix=synth.MIX_HCNOS, spacing=2.6,
                                        charges=(0,))
    kp = _kp()
    cap = runtime.CapturedStep(numbers, positions, charge, kp)
    e0, g0, _ = runtime.energy_gradient(numbers, positions, charge, kp)
    for k in range(3):
        e, g = cap.replay()
        torch.cuda.synchronize()
        assert torch.equal(e, e0) and torch.equal(g, g0), k


def test_profile_slot_mask():
    """Kernel timing restricted to one slot brackets only that kernel's
    launches; with every slot enabled all of them are timed."""
    _say("profiling: slot mask vs all slots")
    numbers, positions, charge = _batch(403, (12, 30))
    kp = _kp()
    dev = positions.device
    runtime.energy_gradient(numbers, positions, charge, kp)
    torch.cuda.synchronize()
    runtime.profile_read(dev)  # clear
    runtime.profile_enable(dev, True, slots=("atm",))
    runtime.energy_gradient(numbers, positions, charge, kp)
    runtime.profile_enable(dev, False)
    only = runtime.profile_read(dev)
    assert only["atm"][1] == 1 and only["atm"][0] > 0.0
    assert all(v[1] == 0 for k, v in only.items() if k != "atm")
    runtime.profile_enable(dev, True)
    runtime.energy_gradient(numbers, positions, charge, kp)
    runtime.profile_enable(dev, False)
    every = runtime.profile_read(dev)
    for k in ("cn", "eeq", "weights", "pairs", "atm", "atm_rows", "adjoint", "chain"):
        assert every[k][1] >= 1, k
