"""
Continuous-batching request loop for D4 energies (and forces) on one GPU.

Requests (one molecule each: atomic numbers, positions in bohr, total charge,
energy-only or energy + forces) arrive on the host at any time. The loop

- keeps the element tables resident on the device (one context per device);
- groups waiting requests by kind and padded size bucket (multiples of 32
  atoms up to 192, the largest size the single-work-group EEQ solver takes,
  then 256, then multiples of 128), so a batch holds many requests while
  padding stays moderate (padding rows exit at once in every kernel);
- runs every batch in one of ``max_inflight`` slots, each with its own HIP
  stream and its own device arena and pinned host staging buffers, kept
  between batches (no allocation in steady state); a slot's share of the
  memory budget bounds the batch size (``d4_workspace_bytes`` + inputs +
  outputs);
- packs a batch with vectorised copies into the pinned buffer and moves it
  with one host-to-device copy, and the results with one copy back;
- reports each result when its batch's completion event has fired;
- fails a bad request alone: a request with no atoms, atomic numbers
  outside 0..103, positions of the wrong shape or not finite, or a footprint
  above one slot's memory budget is answered at submission with a failed
  ``Result`` (``error`` set); after a batch, a molecule whose status word
  reports a singular EEQ system or whose energies are not finite gets a
  failed result while the rest of its batch succeeds.

The kernels give every molecule a result that does not depend on its batch
neighbours, nor on the padded length within one of the buckets up to 192
atoms, so such a request's result is bitwise the same alone and in the mix;
from 193 atoms on the padded batch takes the spatial order and the tiled EEQ
and the results agree with the request alone to rounding.
"""

from __future__ import annotations

import collections
import functools
import gc
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import runtime
from .cutoff import Cutoff
from .disp import kernel_params

__all__ = ["Request", "Result", "D4Server", "size_class"]

_ALIGN = 256
# molecule status bits of the workspace (include/d4amd.h, d4_molecule_status)
STATUS_LU = 1
STATUS_SINGULAR = 2
STATUS_BAD_Z = 4


def size_class(n: int) -> int:
    """Padded length of the batches a molecule of n atoms joins."""
    if n <= 192:
        return max(32, (n + 31) // 32 * 32)
    if n <= 256:
        return 256
    return (n + 127) // 128 * 128


def _up(x: int) -> int:
    return (x + _ALIGN - 1) // _ALIGN * _ALIGN


@dataclass
class Request:
    rid: int
    numbers: np.ndarray
    positions: np.ndarray
    charge: float
    forces: bool
    arrival: float = 0.0  # host clock (time.perf_counter) when it arrived


@dataclass
class Result:
    rid: int
    energy: np.ndarray | None
    gradient: np.ndarray | None
    arrival: float
    done: float
    batch: int
    error: str | None = None  # set for a failed request (energy and gradient None)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class _Slot:
    """One in-flight batch position: a stream, a device arena (inputs,
    outputs, workspace) and a pinned host buffer (inputs, then outputs),
    grown on demand and reused."""

    def __init__(self, device: torch.device):
        self.device = device
        self.stream = torch.cuda.Stream(device)
        self.dev = torch.empty(0, dtype=torch.uint8, device=device)
        self.host = torch.empty(0, dtype=torch.uint8).pin_memory()
        self.event = None
        self.reqs = None
        self.bid = 0
        self.shape = None

    def reserve(self, dev_bytes: int, host_bytes: int) -> None:
        if self.dev.numel() < dev_bytes:
            self.dev = torch.empty(dev_bytes, dtype=torch.uint8, device=self.device)
        if self.host.numel() < host_bytes:
            self.host = torch.empty(host_bytes, dtype=torch.uint8).pin_memory()


@functools.lru_cache(maxsize=8192)
def _layout(nbatch: int, natoms: int, forces: bool) -> dict:
    """Byte offsets of one batch in a slot: inputs (numbers, positions,
    charge) and outputs (energy, gradient) in both buffers, the workspace on
    the device only."""
    bn = nbatch * natoms
    off = {}
    o = 0
    for name, size in (("numbers", 8 * bn), ("positions", 24 * bn), ("charge", 8 * nbatch)):
        off[name] = o
        o = _up(o + size)
    off["in_end"] = o
    for name, size in (("energy", 8 * bn), ("grad", 24 * bn if forces else 0)):
        off[name] = o
        o = _up(o + size)
    off["out_end"] = o
    # host only: the molecules' status words, copied back with the outputs
    off["status"] = o
    off["host_end"] = _up(o + 4 * nbatch)
    off["ws"] = o
    off["ws_bytes"] = runtime.workspace_bytes(nbatch, natoms)
    off["dev_end"] = o + _up(off["ws_bytes"])
    return off


def footprint(nbatch: int, natoms: int, forces: bool) -> int:
    """Device bytes of a batch: inputs + outputs + workspace."""
    return _layout(nbatch, natoms, forces)["dev_end"]


def _shape_error(req: Request, slot_budget: int) -> str | None:
    """The checks that need no pass over the request's values: atom count,
    array types and shapes, footprint (cached per padded length)."""
    numbers, positions = req.numbers, req.positions
    if not isinstance(numbers, np.ndarray) or not isinstance(positions, np.ndarray):
        return "numbers and positions must be numpy arrays"
    n = len(numbers)
    if n < 1:
        return "a request needs at least one atom"
    if numbers.ndim != 1 or numbers.dtype.kind not in "iu":
        return "atomic numbers must be a 1-d integer array"
    if positions.shape != (n, 3):
        return f"positions must have shape ({n}, 3), got {positions.shape}"
    if footprint(1, size_class(n), req.forces) > slot_budget:
        return f"a {n}-atom request exceeds the per-slot memory budget ({slot_budget} bytes)"
    return None


def _value_error(req: Request) -> str | None:
    """The checks on the request's values: atomic-number range, finiteness."""
    if req.numbers.min() < 0 or req.numbers.max() >= 104:
        return "atomic numbers must lie in 0..103 (0 marks padding)"
    if not np.isfinite(req.positions).all() or not np.isfinite(req.charge):
        return "positions and charge must be finite"
    return None


def request_error(req: Request, slot_budget: int) -> str | None:
    """Why ``req`` cannot be served by a slot of ``slot_budget`` device
    bytes (None if it can): no atoms, atomic numbers outside 0..103,
    positions of the wrong shape, non-finite inputs, or a footprint above the
    budget even alone. (The serving loop makes the shape checks at
    submission and the value checks once per batch on the packed arrays.)"""
    return _shape_error(req, slot_budget) or _value_error(req)


def _view(buf: torch.Tensor, off: int, dtype, shape) -> torch.Tensor:
    count = int(np.prod(shape))
    item = torch.empty(0, dtype=dtype).element_size()
    return buf[off: off + count * item].view(dtype).view(shape)


class D4Server:
    """Continuous batching of D4 requests on one device."""

    def __init__(self, param: dict, cutoff: Cutoff | None = None, device="cuda:0",
                 memory_budget: int = 16 << 30, max_inflight: int = 4,
                 max_batch: int = 1024):
        self.device = torch.device(device)
        self.kp = kernel_params(param, cutoff, None)
        runtime.context(self.device, self.kp.ga, self.kp.gc)  # tables resident now
        self.budget = int(memory_budget)
        self.max_inflight = int(max_inflight)
        self.slot_budget = self.budget // self.max_inflight
        self.max_batch = int(max_batch)
        self.queues: dict[tuple, collections.deque] = collections.defaultdict(collections.deque)
        self.inflight: list[_Slot] = []
        self.results: list[Result] = []
        self._bid = 0
        self._free = [_Slot(self.device) for _ in range(self.max_inflight)]

    # ------------------------------------------------------------ admission

    footprint = staticmethod(footprint)

    def _admissible(self, npad: int, forces: bool, limit: int) -> int:
        lo, hi = 0, limit
        while lo < hi:
            mid = (lo + hi + 1) // 2
            if self.footprint(mid, npad, forces) <= self.slot_budget:
                lo = mid
            else:
                hi = mid - 1
        return lo

    def validate(self, req: Request) -> str | None:
        """Why ``req`` cannot be served (None if it can)."""
        return request_error(req, self.slot_budget)

    def _fail(self, req: Request, why: str, batch: int = 0) -> None:
        self.results.append(Result(req.rid, None, None, req.arrival, time.perf_counter(),
                                   batch, why))

    def submit(self, req: Request) -> None:
        """Queue ``req``; a request that cannot be served gets a failed
        ``Result`` (at once for shape and size problems, when its batch is
        packed for bad values) and does not enter a batch."""
        why = _shape_error(req, self.slot_budget)
        if why is not None:
            self._fail(req, why)
            return
        self.queues[(req.forces, size_class(len(req.numbers)))].append(req)

    def pending(self) -> int:
        return sum(len(q) for q in self.queues.values())

    def _pick_queue(self):
        """Queue whose oldest request waited longest."""
        best, key = None, None
        for k, q in self.queues.items():
            if q and (best is None or q[0].arrival < best):
                best, key = q[0].arrival, k
        return key

    def _launch(self, key) -> bool:
        forces, npad = key
        q = self.queues[key]
        count = self._admissible(npad, forces, min(len(q), self.max_batch))
        if count == 0:  # submit() rejects such requests; kept as a guard
            raise MemoryError(f"a {npad}-atom request exceeds the per-slot memory budget "
                              f"({self.slot_budget} bytes)")
        reqs = [q.popleft() for _ in range(count)]
        zc = np.concatenate([r.numbers for r in reqs])
        pc = np.concatenate([r.positions for r in reqs])
        charges = np.fromiter((r.charge for r in reqs), dtype=np.float64, count=count)
        if (zc.min() < 0 or zc.max() >= 104 or not np.isfinite(pc).all()
                or not np.isfinite(charges).all()):
            # rare: find the bad requests, fail them alone, pack the rest
            good = []
            for r in reqs:
                why = _value_error(r)
                if why is None:
                    good.append(r)
                else:
                    self._fail(r, why)
            reqs = good
            count = len(reqs)
            if count == 0:
                return False
            zc = np.concatenate([r.numbers for r in reqs])
            pc = np.concatenate([r.positions for r in reqs])
            charges = np.fromiter((r.charge for r in reqs), dtype=np.float64, count=count)
        slot = self._free.pop()
        lay = _layout(count, npad, forces)
        slot.reserve(lay["dev_end"], lay["host_end"])
        # pack into the pinned buffer: zero padding, then scatter every
        # request's atoms to (row, column) in one vectorised copy
        lens = np.fromiter((len(r.numbers) for r in reqs), dtype=np.int64, count=count)
        # flat destination of every atom: row k starts at k * npad
        start = np.arange(0, count * npad, npad, dtype=np.int64)
        flat = np.arange(int(lens.sum()), dtype=np.int64) + np.repeat(
            start - (np.cumsum(lens) - lens), lens)
        h = slot.host.numpy()
        hn = h[lay["numbers"]: lay["numbers"] + 8 * count * npad].view(np.int64)
        hp = h[lay["positions"]: lay["positions"] + 24 * count * npad].view(np.float64)
        hc = h[lay["charge"]: lay["charge"] + 8 * count].view(np.float64)
        hn.fill(0)  # padding atoms: Z = 0 at the origin
        hp.fill(0.0)
        hn[flat] = zc
        hp.reshape(-1, 3)[flat] = pc
        hc[:] = charges
        with torch.cuda.stream(slot.stream):
            slot.dev[: lay["in_end"]].copy_(slot.host[: lay["in_end"]], non_blocking=True)
            nt = _view(slot.dev, lay["numbers"], torch.int64, (count, npad))
            pt = _view(slot.dev, lay["positions"], torch.float64, (count, npad, 3))
            ct = _view(slot.dev, lay["charge"], torch.float64, (count,))
            ws = slot.dev[lay["ws"]: lay["ws"] + lay["ws_bytes"]]
            energy = _view(slot.dev, lay["energy"], torch.float64, (count, npad))
            if forces:
                grad = _view(slot.dev, lay["grad"], torch.float64, (count, npad, 3))
                runtime.energy_gradient(nt, pt, ct, self.kp, ws=ws, energy=energy, grad=grad)
            else:
                runtime.forward(nt, pt, ct, self.kp, ws=ws, energy=energy)
            lo, hi = lay["energy"], lay["out_end"]
            slot.host[lo:hi].copy_(slot.dev[lo:hi], non_blocking=True)
            # status words: the first array of the workspace
            slot.host[lay["status"]: lay["status"] + 4 * count].copy_(
                ws[: 4 * count], non_blocking=True)
            slot.event = torch.cuda.Event()
            slot.event.record(slot.stream)
        self._bid += 1
        slot.bid = self._bid
        slot.reqs = reqs
        slot.shape = (count, npad, forces, lay)
        self.inflight.append(slot)
        return True

    # ------------------------------------------------------------------ loop

    def poll(self) -> int:
        """Collect finished batches; returns the number of results added."""
        added = 0
        still = []
        for slot in self.inflight:
            if not slot.event.query():
                still.append(slot)
                continue
            now = time.perf_counter()
            count, npad, forces, lay = slot.shape
            h = slot.host.numpy()
            e = h[lay["energy"]: lay["energy"] + 8 * count * npad].view(np.float64)
            e = e.reshape(count, npad)
            g = None
            if forces:
                g = h[lay["grad"]: lay["grad"] + 24 * count * npad].view(np.float64)
                g = g.reshape(count, npad, 3)
            status = h[lay["status"]: lay["status"] + 4 * count].view(np.int32).copy()
            # one copy of the batch's outputs out of the pinned buffer (reused by
            # the slot's next batch); each result is a view of its rows
            e = e.copy()
            g = None if g is None else g.copy()
            bad = status & (STATUS_SINGULAR | STATUS_BAD_Z)
            finite = np.isfinite(e).all(axis=1)
            if g is not None:
                finite &= np.isfinite(g).all(axis=(1, 2))
            ok = (bad == 0) & finite
            results = self.results
            bid = slot.bid
            for k, r in enumerate(slot.reqs):
                n = len(r.numbers)
                if ok[k]:
                    results.append(Result(r.rid, e[k, :n], None if g is None else g[k, :n],
                                          r.arrival, now, bid))
                    continue
                if status[k] & STATUS_SINGULAR:
                    why = "the EEQ system of this molecule is singular"
                elif status[k] & STATUS_BAD_Z:
                    why = "atomic numbers out of range"
                else:
                    why = "non-finite result"
                results.append(Result(r.rid, None, None, r.arrival, now, bid, why))
            added += len(slot.reqs)
            slot.reqs = None
            self._free.append(slot)
        self.inflight = still
        return added

    def step(self) -> None:
        """Poll completions, then admit as many batches as free slots allow."""
        self.poll()
        while self._free and self.pending():
            self._launch(self._pick_queue())

    def drain(self) -> None:
        while self.pending() or self.inflight:
            self.step()

    def run_stream(self, requests: list[Request], arrivals: np.ndarray) -> list[Result]:
        """Replay requests at the given arrival offsets (seconds from now),
        serve them to completion and return their results. The cyclic
        garbage collector is paused meanwhile (a full collection over the
        request objects takes tens of milliseconds, longer than most requests
        take to serve)."""
        first = len(self.results)
        order = np.argsort(arrivals, kind="stable")
        was_enabled = gc.isenabled()
        gc.disable()
        try:
            t0 = time.perf_counter()
            i = 0
            while i < len(order) or self.pending() or self.inflight:
                now = time.perf_counter() - t0
                while i < len(order) and arrivals[order[i]] <= now:
                    r = requests[order[i]]
                    r.arrival = t0 + arrivals[order[i]]
                    self.submit(r)
                    i += 1
                self.step()
        finally:
            if was_enabled:
                gc.enable()
        return self.results[first:]


def latency_stats(results: list[Result]) -> dict:
    """Latency percentiles and throughput of the successful results."""
    failed = sum(r.error is not None for r in results)
    results = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in results]) * 1e3
    span = max(r.done for r in results) - min(r.arrival for r in results)
    return {
        "failed": failed,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "molecules_per_s": len(results) / span,
        "batches": len({r.batch for r in results}),
    }
