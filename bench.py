#!/usr/bin/env python
"""
Benchmark of the D4 energy + forces path on one MI355X.

Default workload ``batch_1024_20_120``: 1024 synthetic molecules,
n ~ U{20..120}, element mix H/C/N/O 45/25/10/10 % plus 10 % uniform over the
other elements up to Z = 54, total charges in {-1, 0, +1}, jittered 2.8 bohr
grids; s6 = 1, s8 = 1.2, a1 = 0.40, a2 = 5.0 bohr, s9 = 1 (ATM on); cutoffs
50 / 25 / 25 bohr. One step = atom energies and the gradient of the total
energy for the whole batch, inputs resident in HBM (fused kernel sequence).

Prints one JSON line with the throughput (molecules/s) of the step replayed
from a captured HIP graph (the eager launch sequence's figure beside it), a
roofline record for the dominant kernel (three-body term, timed with HIP
events on its launch stream), the largest error of the timed step's output
against the oracle on the CPU-baseline molecules, the plain per-molecule GPU
path and the FP64 CPU restatement timed on the host on bounded samples of the
same workload, the CPU time of the single 64-atom molecule (median of 20), and
the serving stream's p50/p99 latency (``serve``).

Other workloads (``--workload``), one JSON line each:
  single_64        one 64-atom molecule (H/C/N/O), energy + gradient;
  batch_256_8_200  256 molecules, n ~ U{8..200}, energy only, one launch;
  cluster_5000     one 5000-atom cluster (H/C/N/O/S), energy + gradient;
  serving_stream   10,000 requests (n ~ lognormal, median 40, 4..2000 atoms,
                   70 % energy, 30 % energy + forces) served by the
                   continuous-batching loop at 70 % of its measured capacity
                   with Poisson arrivals: p50 / p99 latency; beside it the
                   host's per-request latency on the same mix (oracle).

``--model d4s`` runs the batch workloads with the D4S model (element-pair
Gaussian weighting factors) instead of D4.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--model d4|d4s]
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PARAM = {"s6": 1.0, "s8": 1.2, "s9": 1.0, "a1": 0.40, "a2": 5.0}
CUT = (50.0, 25.0, 25.0)  # disp2, disp3, cn_eeq
WORKLOAD = "batch_1024_20_120"


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--plain-seconds", type=float, default=10.0,
                    help="time budget of the plain per-molecule GPU sample")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="time budget of the CPU baseline sample")
    ap.add_argument("--no-baselines", action="store_true")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving-stream record of the headline workload")
    ap.add_argument("--one-stream", action="store_true",
                    help="profiling: run the three-body side on the caller's stream (no "
                         "companion stream), so a kernel trace shows each kernel's own time")
    ap.add_argument("--requests", type=int, default=10000, help="serving_stream requests")
    ap.add_argument("--model", choices=("d4", "d4s"), default="d4")
    ap.add_argument("--shapes", default="tuned",
                    help="launch shapes: 'tuned' (the table tuned for the device's "
                         "architecture shipped in the package, when there is one), "
                         "'builtin', or the path of a table written by tad_dftd4_amd.tune")
    return ap.parse_args()


def use_shapes(choice: str, dev):
    """Activate the launch-shape table; returns its label for the config."""
    import torch

    from tad_dftd4_amd import tune

    if choice == "builtin":
        tune.activate(None)
        return "built-in rules"
    if choice == "tuned":
        arch = torch.cuda.get_device_properties(dev).gcnArchName.split(":")[0]
        path = ROOT / "tad_dftd4_amd" / f"shapes_{arch}.json"
        if not path.exists():
            tune.activate(None)
            return f"built-in rules (no tuned table for {arch})"
    else:
        path = Path(choice)
    tune.activate(tune.ShapeTable.load(path))
    try:
        return f"tuned table {path.relative_to(ROOT)}"
    except ValueError:
        return f"tuned table {path}"


def traffic_from_profiles(workload: str):
    """HBM bytes per launch of the three-body kernel from the committed
    rocprofv3 PMC summary, if present."""
    f = ROOT / "profiles" / f"atm_pmc_{workload}.txt"
    if not f.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read_text(), re.M)
    return float(m.group(1)) if m else None


METRICS = {
    "batch_1024_20_120": "mol/s E+forces (20-120 atoms) batched vs plain; p50/p99 serve latency",
    "batch_256_8_200": "mol/s energy (8-200 atoms), one launch",
    "cluster_5000": "5000-atom cluster energy + forces evaluations/s",
    "single_64": "single 64-atom molecule energy + forces, molecules/s",
}
FORCES = {"batch_1024_20_120": True, "batch_256_8_200": False, "cluster_5000": True,
          "single_64": True}


def plain_gpu(mols, device, budget_s: float, workload: str = WORKLOAD) -> dict:
    import torch

    from tad_dftd4_amd import plain

    par = {k: torch.tensor(v, dtype=torch.float64, device=device)
           for k, v in {**PARAM, "alp": 16.0}.items()}
    data = [(torch.tensor(m.numbers, device=device), torch.tensor(m.positions, device=device),
             torch.tensor(m.charge, dtype=torch.float64, device=device)) for m in mols]
    # warm-up (kernel selection, allocator)
    for n, p, c in data[:2]:
        plain.energy_gradient(n, p, c, par, cut2=CUT[0], cut3=CUT[1], cut_eeq=CUT[2])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    for n, p, c in data:
        plain.energy_gradient(n, p, c, par, cut2=CUT[0], cut3=CUT[1], cut_eeq=CUT[2])
        done += 1
        if done % 8 == 0:
            torch.cuda.synchronize()
            if time.perf_counter() - t0 > budget_s:
                break
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "molecules/s", "sample": f"first {done} molecules of "
            f"{workload}, one molecule per call, energy + autograd forces, dense torch ops on "
            "the GPU"}


def host_cpu() -> dict:
    """Host CPU model and the threads the CPU legs use (the box's CPU share:
    OMP_NUM_THREADS where the environment sets it, else every core)."""
    import torch

    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    torch.set_num_threads(threads)
    return {"model": model, "threads": threads, "os_cpu_count": os.cpu_count()}


def cpu_baseline(mols, budget_s: float, workload: str = WORKLOAD, forces: bool = True,
                 model: str = "d4") -> tuple[dict, list]:
    """The oracle (dense FP64 restatement) on the host, one molecule per call,
    over the first molecules of the workload until ``budget_s`` is spent.
    Returns the record and the oracle's (energy, gradient) per molecule done."""
    import torch

    from oracle import dense

    cpu = host_cpu()
    par = {k: torch.tensor(v, dtype=torch.float64) for k, v in PARAM.items()}
    cut = {"disp2": CUT[0], "disp3": CUT[1], "cn_eeq": CUT[2]}
    outs = []
    t0 = time.perf_counter()
    for m in mols:
        if forces:
            e, g = dense.dftd4_with_grad(torch.tensor(m.numbers), torch.tensor(m.positions),
                                         m.charge, par, cutoff=cut, model=model)
        else:
            e = dense.dftd4(torch.tensor(m.numbers), torch.tensor(m.positions), m.charge, par,
                            cutoff=cut, model=model)
            g = None
        outs.append((e, g))
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    done = len(outs)
    what = "energy + autograd forces" if forces else "energy"
    rec = {"value": done / dt, "unit": "molecules/s", "cores": cpu["threads"],
           "kind": "port", "cpu_model": cpu["model"], "os_cpu_count": cpu["os_cpu_count"],
           "sample": f"first {done} molecules of {workload}, {what}, FP64 dense restatement "
                     f"on the host ({cpu['threads']} threads), one molecule per call"}
    return rec, outs


def cpu_single_64(runs: int = 20) -> dict:
    """The single 64-atom molecule (H/C/N/O, neutral) on the host: energies and
    the autograd gradient, 3 warm-up runs then the median of ``runs``."""
    import statistics

    import torch

    from oracle import dense
    from tad_dftd4_amd import synth

    cpu = host_cpu()
    (m,) = synth.workload("single_64")
    par = {k: torch.tensor(v, dtype=torch.float64) for k, v in PARAM.items()}
    cut = {"disp2": CUT[0], "disp3": CUT[1], "cn_eeq": CUT[2]}
    nt, pt = torch.tensor(m.numbers), torch.tensor(m.positions)
    times = []
    for k in range(3 + runs):
        t0 = time.perf_counter()
        dense.dftd4_with_grad(nt, pt, m.charge, par, cutoff=cut)
        if k >= 3:
            times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"workload": "single_64", "median_ms": round(med * 1e3, 3),
            "molecules_per_s": round(1.0 / med, 2), "runs": runs, "cores": cpu["threads"],
            "cpu_model": cpu["model"], "kind": "port",
            "what": "1 x 64 atoms, energy + autograd gradient, FP64 dense restatement"}


def cpu_serving(reqs, budget_s: float, max_atoms: int = 300) -> dict:
    """Per-request latency of the serving mix on the host: the oracle (dense
    FP64 restatement) on the requests in arrival order, energy or energy +
    autograd forces as each request asks, until ``budget_s`` is spent;
    requests above ``max_atoms`` are skipped (their n^3 tensors take seconds
    and gigabytes each)."""
    import numpy as np
    import torch

    from oracle import dense

    cpu = host_cpu()
    par = {k: torch.tensor(v, dtype=torch.float64) for k, v in PARAM.items()}
    cut = {"disp2": CUT[0], "disp3": CUT[1], "cn_eeq": CUT[2]}
    lat, skipped = [], 0
    t0 = time.perf_counter()
    for r in reqs:
        if len(r.numbers) > max_atoms:
            skipped += 1
            continue
        nt, pt = torch.tensor(r.numbers), torch.tensor(r.positions)
        t1 = time.perf_counter()
        if r.forces:
            dense.dftd4_with_grad(nt, pt, r.charge, par, cutoff=cut)
        else:
            dense.dftd4(nt, pt, r.charge, par, cutoff=cut)
        lat.append(time.perf_counter() - t1)
        if time.perf_counter() - t0 > budget_s:
            break
    lat = np.array(lat) * 1e3
    return {"value": round(float(np.median(lat)), 3), "unit": "ms per request (median)",
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "molecules_per_s": round(float(len(lat) / (lat.sum() / 1e3)), 1),
            "cores": cpu["threads"],
            "kind": "port", "cpu_model": cpu["model"],
            "sample": f"the first {len(lat)} requests of the stream in arrival order, one at a "
                      f"time, as each asks (energy or energy + autograd forces), FP64 dense "
                      f"restatement; {skipped} requests above {max_atoms} atoms skipped"}


def oracle_errors(outs, energy, grad) -> dict:
    """Largest |E - E_oracle| and |g - g_oracle| over the molecules the CPU
    baseline computed, from the timed step's own outputs."""
    import torch

    de = dg = 0.0
    e = energy.cpu()
    g = grad.cpu() if grad is not None else None
    for b, (eo, go) in enumerate(outs):
        n = eo.shape[0]
        de = max(de, (e[b, :n] - eo).abs().max().item())
        if go is not None and g is not None:
            dg = max(dg, (g[b, :n] - go).abs().max().item())
    rec = {"max_abs_dE": de, "molecules": len(outs), "tol_E": 1e-10}
    if g is not None:
        rec.update({"max_abs_dg": dg, "tol_g": 1e-9})
    return rec


def serve_bench(args) -> dict:
    """Continuous batching of a 10,000-request stream (see module docstring)."""
    import numpy as np
    import torch

    from tad_dftd4_amd import synth
    from tad_dftd4_amd.cutoff import Cutoff
    from tad_dftd4_amd.serve import D4Server, Request, latency_stats

    host_cpu()
    rng = np.random.default_rng(4)
    sizes = synth.serving_sizes(rng, args.requests)
    reqs = []
    for k, n in enumerate(sizes):
        m = synth.molecule(rng, int(n), synth.MIX_Z54, 2.8, (-2, -1, 0, 1, 2))
        reqs.append(Request(k, m.numbers, m.positions, m.charge, forces=bool(rng.random() < 0.3)))
    cut = Cutoff(disp2=CUT[0], disp3=CUT[1], cn_eeq=CUT[2])

    def fresh():
        return [Request(r.rid, r.numbers, r.positions, r.charge, r.forces) for r in reqs]

    # one server throughout: an untimed closed-loop pass sizes its slot
    # buffers, then the closed-loop capacity (every request waiting at t = 0);
    # an untimed open pass at 70 % of it (other arrival times) makes the
    # buffer views of the smaller batches an open stream forms, then the
    # measured open pass
    server = D4Server(PARAM, cut)
    server.run_stream(fresh(), np.zeros(len(reqs)))
    torch.cuda.synchronize()
    cap = latency_stats(server.run_stream(fresh(), np.zeros(len(reqs))))
    rate = 0.7 * cap["molecules_per_s"]
    warm = np.cumsum(np.random.default_rng(6).exponential(1.0 / rate, size=len(reqs)))
    server.run_stream(fresh(), warm)
    arrivals = np.cumsum(np.random.default_rng(5).exponential(1.0 / rate, size=len(reqs)))
    res = server.run_stream(fresh(), arrivals)
    st = latency_stats(res)
    extra = {}
    if args.workload == "serving_stream" and not args.no_baselines:
        order = np.argsort(arrivals, kind="stable")
        extra["cpu_baseline"] = cpu_serving([reqs[k] for k in order], args.cpu_seconds)
    return {
        "metric": "p50/p99 serve latency, 10k requests (4-2000 atoms, 70% E / 30% E+F), "
                  "Poisson at 70% of capacity",
        "value": round(st["p99_ms"], 3),
        "unit": "ms (p99)",
        "n_gpus": 1,
        "higher_is_better": False,
        "p50_ms": round(st["p50_ms"], 3),
        "p99_ms": round(st["p99_ms"], 3),
        "mean_ms": round(st["mean_ms"], 3),
        "molecules_per_s": round(st["molecules_per_s"], 1),
        "capacity_molecules_per_s": round(cap["molecules_per_s"], 1),
        "offered_rate_per_s": round(rate, 1),
        "batches": st["batches"],
        "failed": st["failed"],
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "serving_stream", "requests": len(reqs),
                   "atoms_min_max": [int(sizes.min()), int(sizes.max())],
                   "atoms_median": float(np.median(sizes)), "cutoffs_bohr": list(CUT),
                   "damping": PARAM, "parallelism": "single GPU",
                   "server": {"max_inflight": server.max_inflight,
                              "merge_kinds": server.merge_kinds, "graphs": server.graphs}},
        **extra,
    }


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(molecules shard across GPUs without communication)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from tad_dftd4_amd import metrics, runtime, synth
    from tad_dftd4_amd.cutoff import Cutoff
    from tad_dftd4_amd.disp import kernel_params
    from tad_dftd4_amd.model import D4SModel

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    shapes = use_shapes(args.shapes, dev)
    if args.workload == "serving_stream":
        out = serve_bench(args)
        out["config"]["launch_shapes"] = shapes
        print(json.dumps(out), flush=True)
        return 0
    forces = FORCES[args.workload]
    mols = synth.workload(args.workload)
    numbers, positions, charge = synth.pack(mols, device=dev)
    B, N = numbers.shape
    model = D4SModel(numbers) if args.model == "d4s" else None
    kp = kernel_params(PARAM, Cutoff(disp2=CUT[0], disp3=CUT[1], cn_eeq=CUT[2]), model)
    kp.one_stream = args.one_stream
    rows, tile = kp.shapes(N, forces)
    rows = rows or (16 if N <= 160 else 32 if N <= 384 else 64)
    tile = tile or (8 if forces else 16)
    ws = runtime.new_workspace(B, N, dev)
    energy = torch.empty((B, N), dtype=torch.float64, device=dev)
    grad = torch.empty((B, N, 3), dtype=torch.float64, device=dev)

    def step():
        if forces:
            runtime.energy_gradient(numbers, positions, charge, kp, ws=ws, energy=energy,
                                    grad=grad)
        else:
            runtime.forward(numbers, positions, charge, kp, ws=ws, energy=energy)

    def timed(fn, steps: int) -> float:
        """Milliseconds of ``steps`` calls, bracketed by synchronisations and
        HIP events on the current stream (where every launch sequence here
        starts and, via the companion stream's join, ends)."""
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
        for _ in range(steps):
            fn()
        stop.record()
        torch.cuda.synchronize()
        return start.elapsed_time(stop)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # eager launch sequence (host launches every kernel of every step)
    ms_eager = timed(step, args.steps)
    # the same steps with HIP events on the three-body kernel's launch stream
    # bracketing that kernel alone (events around every kernel add ~5 % here)
    runtime.profile_read(dev)  # clear
    runtime.profile_enable(dev, True, slots=("atm",))
    timed(step, args.steps)
    runtime.profile_enable(dev, False)
    prof_step = runtime.profile_read(dev)
    # the roofline figure: the same steps on one stream, so that the
    # three-body kernel runs with no other kernel beside it (in the two-stream
    # step the other side's kernels share its CUs by design, which stretches
    # its own duration)
    one = args.one_stream
    kp.one_stream = True
    runtime.profile_enable(dev, True, slots=("atm",))
    timed(step, args.steps)
    runtime.profile_enable(dev, False)
    prof = runtime.profile_read(dev)
    kp.one_stream = one
    # per-kernel averages from a fully instrumented pass of the same steps
    runtime.profile_enable(dev, True)
    for _ in range(args.steps):
        step()
    runtime.profile_enable(dev, False)
    prof_all = runtime.profile_read(dev)
    # the headline: the step captured once in a HIP graph (both streams'
    # launch sequences) and replayed; K replays timed after W untimed ones
    cap = runtime.CapturedStep(numbers, positions, charge, kp, forces=forces)
    # W untimed replays, and more until 0.2 s of replays have run: the first
    # few dozen replays after the capture run ~5 % slower (clocks settling)
    t_warm = time.perf_counter()
    warm_replays = 0
    while warm_replays < args.warmup or time.perf_counter() - t_warm < 0.2:
        cap.replay()
        warm_replays += 1
        if warm_replays % 10 == 0:
            torch.cuda.synchronize()
    ms_total = timed(cap.replay, args.steps)
    ms_step = ms_total / args.steps
    value = B * args.steps / (ms_total / 1e3)
    # end to end: the inputs copied in from pinned host memory before every
    # replay and the outputs copied back after it, on the same stream
    h_in = [t.cpu().pin_memory() for t in (cap.numbers, cap.positions, cap.charge)]
    h_out = [torch.empty(t.shape, dtype=t.dtype).pin_memory()
             for t in ((cap.energy, cap.grad) if forces else (cap.energy,))]

    def e2e():
        for d, h in zip((cap.numbers, cap.positions, cap.charge), h_in):
            d.copy_(h, non_blocking=True)
        outs = cap.replay()
        outs = outs if forces else (outs,)
        for h, d in zip(h_out, outs):
            h.copy_(d, non_blocking=True)

    for _ in range(3):
        e2e()
    ms_e2e = timed(e2e, args.steps)
    g_energy, g_grad = (cap.energy, cap.grad) if forces else (cap.energy, None)
    same = bool(torch.equal(g_energy, energy) and (not forces or torch.equal(g_grad, grad)))

    work = metrics.workload_flops(mols, CUT[0], CUT[1], with_forces=forces)
    atm_ms, atm_n = prof["atm"]
    atm_avg_s = atm_ms / max(atm_n, 1) / 1e3
    atm_step_ms = prof_step["atm"][0] / max(prof_step["atm"][1], 1)
    achieved = work["atm"] / atm_avg_s / 1e12
    kernel_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in prof_all.items() if v[1]}

    out = {
        "metric": METRICS[args.workload],
        "value": round(value, 1),
        "unit": "molecules/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "model": args.model,
            "molecules": B,
            "atoms_min_max": [int(min(len(m.numbers) for m in mols)),
                              int(max(len(m.numbers) for m in mols))],
            "padded_atoms": N,
            "outputs": "atom energies + gradient" if forces else "atom energies",
            "cutoffs_bohr": list(CUT),
            "damping": PARAM,
            "parallelism": "single GPU",
            "launch_shapes": {"source": shapes, "row_lanes": rows,
                              "atm_tile": None if N > 256 else tile},
            "streams": "one (profiling)" if args.one_stream else "caller + companion",
        },
        "roofline": {
            "kernel": ("k_atm8s (three-body, each triple once, 8x8 pair tiles, group screening, "
                       "energy + gradient)" if N > 256 else
                       f"{'k_atm8' if tile == 8 else 'k_atm1'} (three-body, each triple once, "
                       f"{tile}x{tile} pair tiles, "
                       f"{'energy + gradient' if forces else 'energy'})"),
            "bound": "fp64_valu",
            "note": "FP64 vector-ALU kernel (MFMA unused: rocprof shows no MFMA instructions); "
                    "MI355X FP64 peak is 78.6 TFLOP/s",
            "achieved": round(achieved, 3),
            "peak": metrics.PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / metrics.PEAK_FP64_TFLOPS, 4),
            "traffic": traffic_from_profiles(args.workload),
            "algorithmic_flops_per_launch": work["atm"],
            "avg_launch_ms": round(atm_avg_s * 1e3, 4),
            "timing": "HIP events around each launch of the kernel on its stream, in eager "
                      "steps on one stream (the kernel alone on the GPU)",
            "in_step_avg_launch_ms": round(atm_step_ms, 4),
            "in_step_note": "the same kernel in the two-stream step, sharing the CUs with the "
                            "two-body side",
        },
        "step_flops": {
            "algorithmic_tflops": round(work["total"] / (ms_step / 1e3) / 1e12, 3),
            "frac_of_fp64_peak": round(work["total"] / (ms_step / 1e3) / 1e12
                                       / metrics.PEAK_FP64_TFLOPS, 4),
        },
        "eager": {"ms_per_step": round(ms_eager / args.steps, 4),
                  "value": round(B * args.steps / (ms_eager / 1e3), 1),
                  "note": "the same step launched kernel by kernel from the host"},
        "timed_path": "HIP graph replay of the captured step (both streams)",
        "end_to_end": {"ms_per_step": round(ms_e2e / args.steps, 4),
                       "value": round(B * args.steps / (ms_e2e / 1e3), 1),
                       "note": "inputs copied in from pinned host memory before each replay "
                               "and the outputs copied back after it (PCIe both ways)"},
        "graph_warmup_replays": warm_replays,
        "graph_equals_eager": same,
        "kernel_avg_ms": kernel_ms,
        "kernel_avg_note": "per-kernel HIP-event averages from an eager pass of the same steps "
                           "with every kernel bracketed; the roofline pass brackets only the "
                           "three-body kernel",
    }
    if args.workload == "cluster_5000":
        out["unit"] = "evaluations/s"
        out["cpu_baseline"] = {"value": None, "unit": "evaluations/s", "cores": None,
                               "kind": "port", "sample": "not run: the dense restatement needs "
                               "n^3 = 1.25e11-element tensors (1 TB each) for 5000 atoms"}
    elif not args.no_baselines:
        if forces and args.model == "d4":
            pl = plain_gpu(mols, dev, args.plain_seconds, args.workload)
            out["plain_gpu"] = pl
            out["speedup_vs_plain"] = round(value / pl["value"], 1)
        out["cpu_baseline"], outs = cpu_baseline(mols, args.cpu_seconds, args.workload, forces,
                                                 args.model)
        out["oracle_check"] = oracle_errors(outs, g_energy, g_grad)
        if args.workload in (WORKLOAD, "single_64"):
            out["cpu_single_64"] = cpu_single_64()
    if args.workload == WORKLOAD and args.model == "d4" and not args.no_serve:
        sv = serve_bench(args)
        out["serve"] = {k: sv[k] for k in (
            "p50_ms", "p99_ms", "mean_ms", "molecules_per_s", "capacity_molecules_per_s",
            "offered_rate_per_s", "batches", "failed")}
        out["serve"]["requests"] = sv["config"]["requests"]
        out["serve"]["atoms_min_max"] = sv["config"]["atoms_min_max"]
        out["serve"]["atoms_median"] = sv["config"]["atoms_median"]
        out["serve"]["workload"] = ("serving_stream: lognormal sizes (median 40, sigma 0.8, "
                                    "4..2000 atoms), 70 % energy / 30 % energy + forces, "
                                    "Poisson arrivals at 70 % of the closed-loop capacity")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
