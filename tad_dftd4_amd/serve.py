"""
Continuous-batching request loop for D4 energies (and forces) on one GPU.

Requests (one molecule each: atomic numbers, positions in bohr, total charge,
energy-only or energy + forces) arrive on the host at any time. The loop

- keeps the element tables resident on the device (one context per device);
- groups waiting requests by kind and padded size class (multiples of 16
  atoms, the three-body tile edge) so padding stays small;
- admits a batch whenever an in-flight slot is free and the batch's device
  footprint (workspace + inputs + outputs, ``d4_workspace_bytes``) fits in the
  memory budget next to the batches already on the device;
- runs every batch on its own HIP stream, so a batch admitted while others
  are still running overlaps with them on the device;
- reports each result when its batch's completion event has fired.

The kernels give every molecule a result that does not depend on its batch
neighbours or the padded length, so a request's result is the same alone and
in the mix.
"""

from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import runtime
from .cutoff import Cutoff
from .disp import kernel_params

__all__ = ["Request", "Result", "D4Server", "size_class"]

_CLASS_STEP = 16


def size_class(n: int) -> int:
    """Padded length of a molecule of n atoms."""
    return max(_CLASS_STEP, (n + _CLASS_STEP - 1) // _CLASS_STEP * _CLASS_STEP)


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
    energy: np.ndarray
    gradient: np.ndarray | None
    arrival: float
    done: float
    batch: int

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Batch:
    bid: int
    reqs: list
    stream: torch.cuda.Stream
    event: torch.cuda.Event
    energy: torch.Tensor
    grad: torch.Tensor | None
    nbytes: int
    keep: list = field(default_factory=list)


class D4Server:
    """Continuous batching of D4 requests on one device."""

    def __init__(self, param: dict, cutoff: Cutoff | None = None, device="cuda:0",
                 memory_budget: int = 8 << 30, max_inflight: int = 4,
                 max_batch: int = 1024):
        self.device = torch.device(device)
        self.kp = kernel_params(param, cutoff, None)
        runtime.context(self.device, self.kp.ga, self.kp.gc)  # tables resident now
        self.budget = int(memory_budget)
        self.max_inflight = int(max_inflight)
        self.max_batch = int(max_batch)
        self.queues: dict[tuple, collections.deque] = collections.defaultdict(collections.deque)
        self.inflight: list[_Batch] = []
        self.in_use = 0
        self.results: list[Result] = []
        self._bid = 0
        self._streams = [torch.cuda.Stream(self.device) for _ in range(self.max_inflight)]
        self._free_streams = list(self._streams)

    # ------------------------------------------------------------ admission

    @staticmethod
    def footprint(nbatch: int, natoms: int, forces: bool) -> int:
        """Device bytes of a batch: workspace + inputs + outputs."""
        io = natoms * (8 + 24 + 8 + (24 if forces else 0)) + 8
        return runtime.workspace_bytes(nbatch, natoms) + nbatch * io

    def _admissible(self, npad: int, forces: bool, free: int, limit: int) -> int:
        lo, hi = 0, limit
        while lo < hi:
            mid = (lo + hi + 1) // 2
            if self.footprint(mid, npad, forces) <= free:
                lo = mid
            else:
                hi = mid - 1
        return lo

    def submit(self, req: Request) -> None:
        if len(req.numbers) < 1:
            raise ValueError("a request needs at least one atom")
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
        free = self.budget - self.in_use
        count = self._admissible(npad, forces, free, min(len(q), self.max_batch))
        if count == 0:
            if not self.inflight and self.footprint(1, npad, forces) > self.budget:
                raise MemoryError(f"a {npad}-atom request exceeds the memory budget")
            return False
        reqs = [q.popleft() for _ in range(count)]
        stream = self._free_streams.pop()
        numbers = np.zeros((count, npad), dtype=np.int64)
        positions = np.zeros((count, npad, 3), dtype=np.float64)
        charge = np.zeros((count,), dtype=np.float64)
        for k, r in enumerate(reqs):
            numbers[k, : len(r.numbers)] = r.numbers
            positions[k, : len(r.numbers)] = r.positions
            charge[k] = r.charge
        with torch.cuda.stream(stream):
            nt = torch.from_numpy(numbers).pin_memory().to(self.device, non_blocking=True)
            pt = torch.from_numpy(positions).pin_memory().to(self.device, non_blocking=True)
            ct = torch.from_numpy(charge).pin_memory().to(self.device, non_blocking=True)
            if forces:
                energy, grad, ws = runtime.energy_gradient(nt, pt, ct, self.kp)
            else:
                energy, ws = runtime.forward(nt, pt, ct, self.kp)
                grad = None
            e_host = energy.to("cpu", non_blocking=True)
            g_host = grad.to("cpu", non_blocking=True) if forces else None
            event = torch.cuda.Event()
            event.record(stream)
        nbytes = self.footprint(count, npad, forces)
        self.in_use += nbytes
        self._bid += 1
        self.inflight.append(_Batch(self._bid, reqs, stream, event, e_host, g_host, nbytes,
                                    keep=[nt, pt, ct, energy, grad, ws]))
        return True

    # ------------------------------------------------------------------ loop

    def poll(self) -> int:
        """Collect finished batches; returns the number of results added."""
        added = 0
        still = []
        for b in self.inflight:
            if b.event.query():
                now = time.perf_counter()
                e = b.energy.numpy()
                g = b.grad.numpy() if b.grad is not None else None
                for k, r in enumerate(b.reqs):
                    n = len(r.numbers)
                    self.results.append(Result(r.rid, e[k, :n].copy(),
                                               None if g is None else g[k, :n].copy(),
                                               r.arrival, now, b.bid))
                    added += 1
                self.in_use -= b.nbytes
                self._free_streams.append(b.stream)
            else:
                still.append(b)
        self.inflight = still
        return added

    def step(self) -> None:
        """Poll completions, then admit as many batches as slots and memory
        allow."""
        self.poll()
        while self._free_streams and self.pending():
            key = self._pick_queue()
            if not self._launch(key):
                break

    def drain(self) -> None:
        while self.pending() or self.inflight:
            self.step()

    def run_stream(self, requests: list[Request], arrivals: np.ndarray) -> list[Result]:
        """Replay requests at the given arrival offsets (seconds from now)
        and serve them to completion."""
        order = np.argsort(arrivals, kind="stable")
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
        return self.results


def latency_stats(results: list[Result]) -> dict:
    lat = np.array([r.latency for r in results]) * 1e3
    span = max(r.done for r in results) - min(r.arrival for r in results)
    return {
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "molecules_per_s": len(results) / span,
        "batches": len({r.batch for r in results}),
    }
