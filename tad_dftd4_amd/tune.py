"""
Launch-shape tuner.

Two launch shapes of the batched path change speed but not results (beyond
rounding): the lanes per atom row of the CN, pair and chain passes (16, 32
or 64: a quarter, half or full wave walks one atom's partners) and the pair
tile of the one-pass three-body kernel (8 x 8 or 16 x 16 pairs per
work-group, used up to 256 atoms). Built in, both follow fixed rules of the
padded length and the mode. ``tune`` times every candidate on the device for
a list of padded-length buckets and both modes (atom energies; energies and
gradient), on synthetic molecules of the batch workloads' kind (jittered
2.8 bohr grid, Z <= 54 mix, charges -1/0/+1), with HIP events on the calling
stream, and keeps the fastest per bucket in a ``ShapeTable``. The table
saves to and loads from JSON; ``activate(table)`` makes every later call
whose ``KernelParams`` leave both shapes at 0 take the table's choice for
its padded length (the first bucket at least as long; longer batches keep
the built-in rules). Up to 192 atoms a molecule's results are bitwise the
same alone and in any batch whose padded length falls in the same bucket (the
buckets are the serving loop's padded lengths); otherwise they agree to
rounding (longer batches also take the spatial order and the tiled EEQ).

    from tad_dftd4_amd import tune
    table = tune.tune(torch.device("cuda:0"))
    table.save("shapes.json")
    tune.activate(tune.ShapeTable.load("shapes.json"))

The reference has no counterpart (its dense torch ops have no launch shapes).
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field, replace
from pathlib import Path

import numpy as np
import torch

from . import runtime, synth

ROWS = (16, 32, 64)
TILES = (8, 16)
ONE_PASS_MAX = 256  # longest padded length the one-pass three-body kernels take
# the padded lengths of the serving loop's buckets up to 512 atoms: a request
# alone and the serving batch it joins then take the same table entry (so the
# same shapes and the same bits)
BUCKETS = (32, 64, 96, 128, 160, 192, 256, 384, 512)
MODES = ("energy", "forces")


@dataclass
class ShapeTable:
    """Chosen (row_lanes, atm_tile) per mode and padded-length bucket.

    ``entries[mode]`` is a list of ``[n_max, row_lanes, atm_tile]`` sorted by
    ``n_max``; ``timings`` keeps the measured milliseconds per candidate."""

    device: str = ""
    entries: dict = field(default_factory=dict)
    timings: dict = field(default_factory=dict)

    def lookup(self, grad: bool, natoms: int) -> tuple[int, int]:
        for n_max, rows, tile in self.entries.get("forces" if grad else "energy", ()):
            if natoms <= n_max:
                return int(rows), int(tile)
        return 0, 0

    def to_json(self) -> str:
        return json.dumps({"device": self.device, "entries": self.entries,
                           "timings": self.timings}, indent=1, sort_keys=True)

    @classmethod
    def from_json(cls, text: str) -> "ShapeTable":
        d = json.loads(text)
        entries = {}
        for mode, lst in d.get("entries", {}).items():
            if mode not in MODES:
                raise ValueError(f"unknown mode {mode!r} in shape table")
            rows_ = []
            for n_max, rows, tile in lst:
                if rows not in (0,) + ROWS or tile not in (0,) + TILES:
                    raise ValueError(f"invalid shape ({rows}, {tile}) in shape table")
                if int(n_max) < 1:
                    raise ValueError(f"invalid bucket length {n_max} in shape table")
                rows_.append([int(n_max), int(rows), int(tile)])
            entries[mode] = sorted(rows_)
        return cls(d.get("device", ""), entries, d.get("timings", {}))

    def save(self, path) -> None:
        Path(path).write_text(self.to_json() + "\n")

    @classmethod
    def load(cls, path) -> "ShapeTable":
        return cls.from_json(Path(path).read_text())


def activate(table: ShapeTable | None) -> None:
    """Use ``table`` for calls that leave the shapes at 0 (None: built-in rules)."""
    if table is not None and not isinstance(table, ShapeTable):
        raise TypeError(f"expected a ShapeTable or None, got {type(table).__name__}")
    runtime._SHAPES = table


def active() -> ShapeTable | None:
    return runtime._SHAPES


def candidates(n_max: int) -> list[tuple[int, int]]:
    """(row_lanes, atm_tile) pairs worth timing for a bucket; (0, 0) is the
    built-in rule. Past the one-pass limit the tile has no effect."""
    tiles = TILES if n_max <= ONE_PASS_MAX else (0,)
    return [(0, 0)] + [(r, t) for r in ROWS for t in tiles]


def sample_batch(n_max: int, count: int, seed: int, device):
    """``count`` molecules with sizes in (n_max / 2, n_max], padded to n_max."""
    rng = np.random.default_rng(seed)
    lo = max(4, n_max // 2 + 1)
    mols = [synth.molecule(rng, int(n), synth.MIX_Z54, 2.8, (-1, 0, 1))
            for n in rng.integers(lo, n_max + 1, size=count)]
    return synth.pack(mols, pad_to=n_max, device=device)


def time_calls(fns: dict, reps: int, warmup: int, device) -> dict:
    """Median milliseconds per entry of ``fns`` (name -> callable) over
    ``reps`` rounds, HIP events on the current stream (each call joins its
    companion stream back to it). Every round times every candidate once, so
    clock and cache state drift alike over all of them."""
    for _ in range(warmup):
        for fn in fns.values():
            fn()
    stream = torch.cuda.current_stream(device)
    ms = {k: [] for k in fns}
    for _ in range(reps):
        for k, fn in fns.items():
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            ms[k].append(a.elapsed_time(b))
    return {k: float(np.median(v)) for k, v in ms.items()}


def tune(device=None, buckets=BUCKETS, modes=MODES, count: int = 512, reps: int = 5,
         warmup: int = 2, seed: int = 0, kp: runtime.KernelParams | None = None,
         log=None) -> ShapeTable:
    """Time every candidate shape per bucket and mode; return the fastest.

    ``kp`` gives the damping parameters and cutoffs of the calls being tuned
    (default: s8 = 1.2, a1 = 0.4, a2 = 5, s9 = 1, cutoffs 50/25/25); its own
    shapes are ignored. ``count`` molecules per bucket (fewer for long
    buckets, so that each bucket holds about the same number of atoms)."""
    bad = [m for m in modes if m not in MODES]
    if bad:
        raise ValueError(f"unknown mode(s) {bad}; expected some of {MODES}")
    device = torch.device(device or "cuda:0")
    if kp is None:
        kp = runtime.KernelParams(s8=1.2, cutoff_disp2=50.0, cutoff_disp3=25.0)
    prev = active()
    activate(None)
    arch = torch.cuda.get_device_properties(device).gcnArchName.split(":")[0]
    table = ShapeTable(device=f"{torch.cuda.get_device_name(device)} ({arch})")
    try:
        for mode in modes:
            chosen = []
            for i, n_max in enumerate(sorted(buckets)):
                cnt = max(8, min(count, count * 120 // n_max))
                numbers, positions, charge = sample_batch(n_max, cnt, seed + i, device)
                ws = runtime.new_workspace(cnt, n_max, device)
                fns = {}
                for rows, tile in candidates(n_max):
                    k = replace(kp, row_lanes=rows, atm_tile=tile)
                    if mode == "energy":
                        def fn(k=k):
                            runtime.forward(numbers, positions, charge, k, ws=ws)
                    else:
                        def fn(k=k):
                            runtime.energy_gradient(numbers, positions, charge, k, ws=ws)
                    fns[(rows, tile)] = fn
                ms = time_calls(fns, reps, warmup, device)
                times = {f"{r}x{t}": round(v, 4) for (r, t), v in ms.items()}
                t_best, (r_best, t_tile) = min((v, k) for k, v in ms.items())
                best = (t_best, r_best, t_tile)
                chosen.append([n_max, best[1], best[2]])
                table.timings.setdefault(mode, {})[str(n_max)] = times
                if log is not None:
                    log(f"[tune] {mode} n<={n_max} ({cnt} molecules): best rows={best[1]} "
                        f"tile={best[2]} {best[0]:.4f} ms; built-in {times['0x0']:.4f} ms")
                del ws
            table.entries[mode] = chosen
    finally:
        activate(prev)
    return table


def main(argv=None) -> None:
    import argparse

    ap = argparse.ArgumentParser(description="Tune the launch shapes on cuda:0 and "
                                 "write the table as JSON.")
    ap.add_argument("--out", required=True)
    ap.add_argument("--count", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=9)
    ap.add_argument("--warmup", type=int, default=3)
    a = ap.parse_args(argv)
    table = tune("cuda:0", count=a.count, reps=a.reps, warmup=a.warmup,
                 log=lambda m: print(m, flush=True))
    table.save(a.out)
    print(json.dumps(table.entries), flush=True)


if __name__ == "__main__":
    main()
