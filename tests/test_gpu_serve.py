"""
Continuous batching: requests of mixed size and kind, admitted while other
batches run, give exactly the result each molecule gets alone (a batch of one
at its bucket's padded length), and the result of a direct call at its own
length to rounding, with the built-in launch shapes and with the shipped
tuned shape table.
"""

from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import dense
from tad_dftd4_amd import runtime, synth, tune
from tad_dftd4_amd.cutoff import Cutoff
from tad_dftd4_amd.disp import kernel_params
from tad_dftd4_amd.serve import D4Server, Request, count_class, latency_stats, size_class

pytestmark = pytest.mark.gpu

PARAM = {"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.40, "a2": 5.0}


def _requests(count, seed=7):
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.round(rng.lognormal(np.log(30.0), 0.8, size=count)), 4, 150).astype(int)
    reqs = []
    for k, n in enumerate(sizes):
        m = synth.molecule(rng, int(n), synth.MIX_Z54, 2.8, (-2, -1, 0, 1, 2))
        reqs.append(Request(k, m.numbers, m.positions, m.charge, forces=bool(rng.random() < 0.3)))
    return reqs


MODES = {
    # default: energy-only requests in the energy + forces batches, eager launches
    "default": {},
    "graph replay": {"graphs": True},
    "kinds apart": {"merge_kinds": False},
}


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("shapes", ["builtin", "shipped"])
def test_stream_alone_vs_mix(shapes, mode):
    print(f"[case] serving stream, 120 requests, budget 64 MiB, 3 in flight, {shapes} shapes, "
          f"{mode}", flush=True)
    table = None
    if shapes == "shipped":
        table = tune.ShapeTable.load(Path(tune.__file__).with_name("shapes_gfx950.json"))
    tune.activate(table)
    try:
        _alone_vs_mix(MODES[mode])
    finally:
        tune.activate(None)


def _alone(server, req, kp, bucket=True):
    """The request alone: runtime calls on a one-molecule batch padded to its
    bucket's length as the server pads it (``bucket=False``: at the
    molecule's own length), in the server's way (energy + forces for every
    request when it merges the kinds)."""
    n = len(req.numbers)
    numbers, positions, charge = synth.pack(
        [synth.Molecule(req.numbers, req.positions, req.charge)], device="cuda",
        pad_to=size_class(n) if bucket else None)
    if req.forces or server.merge_kinds:
        e, g, _ = runtime.energy_gradient(numbers, positions, charge, kp)
        return e[0, :n].cpu().numpy(), g[0, :n].cpu().numpy()
    e, _ = runtime.forward(numbers, positions, charge, kp)
    return e[0, :n].cpu().numpy(), None


def _alone_vs_mix(kwargs):
    cut = Cutoff(disp2=50.0, disp3=25.0, cn_eeq=25.0)
    server = D4Server(PARAM, cut, memory_budget=64 << 20, max_inflight=3, max_batch=24,
                      **kwargs)
    reqs = _requests(120)
    arrivals = np.cumsum(np.random.default_rng(3).exponential(2e-4, size=len(reqs)))
    results = server.run_stream(reqs, arrivals)
    assert len(results) == len(reqs)
    assert len({r.batch for r in results}) > 3  # several batches, overlapping
    stats = latency_stats(results)
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
    kp = kernel_params(PARAM, cut, None)
    by_id = {r.rid: r for r in results}
    for req in reqs:
        res = by_id[req.rid]
        e, g = _alone(server, req, kp)
        assert np.array_equal(res.energy, e)
        if req.forces:
            assert np.array_equal(res.gradient, g)
        else:
            assert res.gradient is None
        e1, g1 = _alone(server, req, kp, bucket=False)
        assert np.abs(res.energy - e1).max() < 1e-13
        if req.forces:
            assert np.abs(res.gradient - g1).max() < 1e-12
    # and against the oracle for a few of them
    for req in reqs[:6]:
        eo = dense.dftd4(torch.tensor(req.numbers), torch.tensor(req.positions), req.charge,
                         {k: torch.tensor(v, dtype=torch.float64) for k, v in PARAM.items()},
                         cutoff={"disp2": 50.0, "disp3": 25.0, "cn_eeq": 25.0})
        assert np.abs(by_id[req.rid].energy - eo.numpy()).max() < 1e-10


def test_graph_replays_match_first_batches():
    """The same stream served twice by one server: the second pass replays
    the graphs the first captured (batches of a layout seen before), and
    every result is bitwise the first pass's, batches padded with the filler
    molecule included."""
    print("[case] serving stream twice, graph replay", flush=True)
    cut = Cutoff(disp2=50.0, disp3=25.0, cn_eeq=25.0)
    server = D4Server(PARAM, cut, memory_budget=64 << 20, max_inflight=2, max_batch=24,
                      graphs=True)
    reqs = _requests(80, seed=17)
    arrivals = np.cumsum(np.random.default_rng(6).exponential(2e-4, size=len(reqs)))
    first = server.run_stream(reqs, arrivals)
    captured = sum(b.graph is not None for s in server._free for b in s.layouts.values())
    assert captured > 0
    again = [Request(r.rid, r.numbers, r.positions, r.charge, r.forces) for r in reqs]
    second = server.run_stream(again, arrivals)
    replayed = sum(b.graph is not None for s in server._free for b in s.layouts.values())
    assert replayed >= captured
    a, b = {r.rid: r for r in first}, {r.rid: r for r in second}
    for req in reqs:
        assert a[req.rid].error is None and b[req.rid].error is None
        assert np.array_equal(a[req.rid].energy, b[req.rid].energy)
        if req.forces:
            assert np.array_equal(a[req.rid].gradient, b[req.rid].gradient)
    # some batches ran with filler rows (a count below its count class)
    sizes = {r.rid: len(r.numbers) for r in reqs}
    padded = 0
    for results in (first, second):
        batches = {}
        for r in results:
            batches.setdefault(r.batch, []).append(size_class(sizes[r.rid]))
        for npads in batches.values():
            limit = server._limit(npads[0], True)
            padded += count_class(len(npads), limit) > len(npads)
    assert padded > 0


def test_two_streams_share_a_context():
    """Calls on two torch streams at once (each gets its own companion
    stream for the three-body side) give the results of calls made one after
    the other."""
    print("[case] two streams, one context", flush=True)
    cut = Cutoff(disp2=50.0, disp3=25.0, cn_eeq=25.0)
    kp = kernel_params(PARAM, cut, None)
    reqs = _requests(40, seed=11)
    mols = [synth.Molecule(r.numbers, r.positions, r.charge) for r in reqs]
    a = synth.pack(mols[:20], device="cuda")
    b = synth.pack(mols[20:], device="cuda")
    ea, ga, _ = runtime.energy_gradient(*a, kp)
    eb, _ = runtime.forward(*b, kp)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for _ in range(3):
        with torch.cuda.stream(s1):
            e1, g1, _ = runtime.energy_gradient(*a, kp)
        with torch.cuda.stream(s2):
            e2, _ = runtime.forward(*b, kp)
        outs.append((e1, g1, e2))
    torch.cuda.synchronize()
    for e1, g1, e2 in outs:
        assert torch.equal(e1, ea) and torch.equal(g1, ga) and torch.equal(e2, eb)


def test_bad_requests_fail_alone(monkeypatch):
    """One request per failure kind in a stream of good ones: only those
    fail. A 3000-atom request above a slot's budget is rejected at
    submission; atomic number 120 and a non-finite position are caught when
    their batch is packed; a request with Z = 110 that gets past the host
    checks (switched off for it here) is flagged by the device's status word
    inside its batch."""
    from tad_dftd4_amd import serve

    print("[case] serving stream with bad requests", flush=True)
    cut = Cutoff(disp2=50.0, disp3=25.0, cn_eeq=25.0)
    server = D4Server(PARAM, cut, memory_budget=256 << 20, max_inflight=2, max_batch=16)
    reqs = _requests(30, seed=21)
    bad_z = Request(1000, np.array([6, 120, 1]), np.zeros((3, 3)) + np.arange(3)[:, None] * 2.0,
                    0.0, forces=False)
    rng = np.random.default_rng(2)
    huge = synth.molecule(rng, 3000, synth.MIX_HCNO, 2.8, (0,))
    too_big = Request(1001, huge.numbers, huge.positions, 0.0, forces=True)
    nan_pos = synth.molecule(rng, 12, synth.MIX_HCNO, 2.8, (0,))
    nan_pos.positions[3, 1] = np.nan
    bad_pos = Request(1003, nan_pos.numbers, nan_pos.positions, 0.0, forces=True)
    arrivals = np.cumsum(np.random.default_rng(4).exponential(2e-4, size=len(reqs)))
    server.submit(bad_z)
    server.submit(too_big)
    server.submit(bad_pos)
    results = server.run_stream(reqs, arrivals)
    planted = synth.molecule(rng, 20, synth.MIX_HCNO, 2.8, (0,))
    planted.numbers[5] = 110
    monkeypatch.setattr(serve, "_value_error", lambda r: None)
    server.submit(Request(1002, planted.numbers, planted.positions, 0.0, forces=False))
    server.drain()
    by_id = {r.rid: r for r in server.results}
    assert "memory budget" in by_id[1001].error
    assert "0..103" in by_id[1000].error
    assert "finite" in by_id[1003].error
    assert "out of range" in by_id[1002].error
    for r in (1000, 1001, 1002, 1003):
        assert by_id[r].energy is None and by_id[r].gradient is None
    good = [r for r in results if r.rid < 1000]
    assert len(good) == len(reqs) and all(r.error is None for r in good)
    kp = kernel_params(PARAM, cut, None)
    for req in reqs:
        e, g = _alone(server, req, kp)
        assert np.array_equal(by_id[req.rid].energy, e)
        if req.forces:
            assert np.array_equal(by_id[req.rid].gradient, g)
    assert latency_stats(server.results)["failed"] == 4
