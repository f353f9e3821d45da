"""
Continuous-batching request loop for D4 energies (and forces) on one GPU.

Requests (one molecule each: atomic numbers, positions in bohr, total charge,
energy-only or energy + forces) arrive on the host at any time. The loop

- keeps the element tables resident on the device (one context per device);
- groups waiting requests by padded size bucket (64, 128 and 192 atoms, 192
  being the largest size the single-work-group EEQ solver takes, then 256,
  then multiples of 128): few queues, so that each comes round often, while
  padding stays moderate (padding rows exit at once in every kernel); by default
  (``merge_kinds``) energy-only requests share the energy + forces batches of
  their bucket and their forces are dropped, which halves the number of
  queues a request waits behind;
- runs every batch in one of ``max_inflight`` slots, each with its own HIP
  stream and its own device arena and pinned host staging buffers, kept
  between batches (no allocation in steady state); a slot's share of the
  memory budget bounds the batch size (``d4_workspace_bytes`` + inputs +
  outputs);
- pads a batch of c requests to a count class (1, 2, 3, 4, 6, 8, 12, ...:
  at most 1.5 c) with copies of a 4-atom filler molecule whose results are
  dropped, so that every batch of one (count class, bucket, kind) in a slot
  has the same layout, whose buffer views are made once; with ``graphs``
  the first such batch runs eagerly and is then captured into a HIP graph
  (both streams of the evaluation) that later ones replay. Graph replay is
  off by default: on a Poisson stream new layouts keep appearing, each
  capture synchronises the device, and the measured latency was several
  times that of the eager launches (which suits a loop whose batch shapes
  repeat);
- packs a batch with one copy per request into the pinned buffer and moves
  it with one host-to-device copy, and the results with one copy back;
- when a slot frees up, launches the queue whose oldest request's wait times
  the square root of its length is largest;
- reports each result when its batch's completion event has fired;
- fails a bad request alone: a request with no atoms, atomic numbers
  outside 0..103 or all zero, positions of the wrong shape or not finite, or
  a footprint above one slot's memory budget is answered at submission (or
  when its batch is packed, for the value checks) with a failed ``Result``
  (``error`` set); after a batch, a molecule whose status word reports a
  singular EEQ system or whose energies are not finite gets a failed result
  while the rest of its batch succeeds.

The kernels give every molecule a result that does not depend on its batch
neighbours nor on the batch size, so a request's result is bitwise the same
alone (a batch of one at its bucket's padded length) and in the mix. Against
a direct call at the molecule's own length it is bitwise the same where both
lengths take the same launch shapes (up to 192 atoms with the shipped shape
table for energy + forces), else equal to rounding.
"""

from __future__ import annotations

import collections
import functools
import gc
import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import runtime
from .cutoff import Cutoff
from .disp import kernel_params

__all__ = ["Request", "Result", "D4Server", "size_class", "count_class"]

_ALIGN = 256
# molecule status bits of the workspace (include/d4amd.h, d4_molecule_status)
STATUS_LU = 1
STATUS_SINGULAR = 2
STATUS_BAD_Z = 4


def size_class(n: int) -> int:
    """Padded length of the batches a molecule of n atoms joins."""
    if n <= 192:
        return max(64, (n + 63) // 64 * 64)
    if n <= 256:
        return 256
    return (n + 127) // 128 * 128


# a batch of c requests runs padded to the smallest class >= c (at most 1.5 c)
COUNT_CLASSES = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024)

# the filler molecule of the padding rows (ammonia, bohr): an ordinary small
# molecule, so a padded batch holds no case the single requests do not
_FILL_Z = np.array([7, 1, 1, 1], dtype=np.int64)
_FILL_R = np.array([[0.0, 0.0, 0.0], [1.91, 0.0, -0.68], [-0.96, 1.65, -0.68],
                    [-0.96, -1.65, -0.68]])


def count_class(count: int, limit: int) -> int:
    """Padded batch size of a batch of ``count`` requests whose bucket admits
    at most ``limit`` molecules per slot."""
    for c in COUNT_CLASSES:
        if c >= count:
            return min(c, limit)
    return min(count, limit)


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
    grown on demand and reused, and the batch layouts (with their captured
    graphs) used in it so far."""

    def __init__(self, device: torch.device):
        self.device = device
        self.stream = torch.cuda.Stream(device)
        self.dev = torch.empty(0, dtype=torch.uint8, device=device)
        self.host = torch.empty(0, dtype=torch.uint8).pin_memory()
        self.event = None
        self.reqs = None
        self.bid = 0
        self.batch = None
        self.layouts: dict[tuple, _Batch] = {}

    def reserve(self, dev_bytes: int, host_bytes: int) -> None:
        grown = False
        if self.dev.numel() < dev_bytes:
            self.dev = None  # (free the old arena first)
            self.dev = torch.empty(dev_bytes, dtype=torch.uint8, device=self.device)
            grown = True
        if self.host.numel() < host_bytes:
            self.host = torch.empty(host_bytes, dtype=torch.uint8).pin_memory()
            grown = True
        if grown:  # views and graphs point into the old buffers
            self.layouts.clear()


class _Batch:
    """One (count class, padded length, kind) layout in a slot: numpy views
    of the pinned buffer, tensor views of the device arena, and the HIP graph
    of the evaluation once captured."""

    def __init__(self, slot: _Slot, nbatch: int, npad: int, forces: bool):
        lay = _layout(nbatch, npad, forces)
        self.nbatch, self.npad, self.forces = nbatch, npad, forces
        self.in_bytes, self.out = lay["in_end"], (lay["energy"], lay["out_end"])
        self.status = lay["status"]
        h = slot.host.numpy()
        bn = nbatch * npad

        def hv(name, nbytes, dtype, shape):
            return h[lay[name]: lay[name] + nbytes].view(dtype).reshape(shape)

        self.hn = hv("numbers", 8 * bn, np.int64, (nbatch, npad))
        self.hp = hv("positions", 24 * bn, np.float64, (nbatch, npad, 3))
        self.hc = hv("charge", 8 * nbatch, np.float64, (nbatch,))
        self.he = hv("energy", 8 * bn, np.float64, (nbatch, npad))
        self.hg = hv("grad", 24 * bn, np.float64, (nbatch, npad, 3)) if forces else None
        self.hs = hv("status", 4 * nbatch, np.int32, (nbatch,))
        d = slot.dev
        self.nt = _view(d, lay["numbers"], torch.int64, (nbatch, npad))
        self.pt = _view(d, lay["positions"], torch.float64, (nbatch, npad, 3))
        self.ct = _view(d, lay["charge"], torch.float64, (nbatch,))
        self.ws = d[lay["ws"]: lay["ws"] + lay["ws_bytes"]]
        self.et = _view(d, lay["energy"], torch.float64, (nbatch, npad))
        self.gt = _view(d, lay["grad"], torch.float64, (nbatch, npad, 3)) if forces else None
        self.graph = None

    def pack(self, reqs: list[Request]) -> None:
        """Requests into rows 0.. of the pinned input buffer (zero padding),
        the filler molecule into the rest (every time: the layouts of a slot
        share its buffer)."""
        hn, hp, hc = self.hn, self.hp, self.hc
        count = len(reqs)
        hn.fill(0)
        hp.fill(0.0)
        for k, r in enumerate(reqs):
            n = len(r.numbers)
            hn[k, :n] = r.numbers
            hp[k, :n] = r.positions
            hc[k] = r.charge
        if count < self.nbatch:
            hn[count:, :4] = _FILL_Z
            hp[count:, :4] = _FILL_R
            hc[count:] = 0.0

    def _call(self, kp) -> None:
        if self.forces:
            runtime.energy_gradient(self.nt, self.pt, self.ct, kp, ws=self.ws, energy=self.et,
                                    grad=self.gt)
        else:
            runtime.forward(self.nt, self.pt, self.ct, kp, ws=self.ws, energy=self.et)

    def run(self, kp, stream, graphs: bool) -> None:
        """The evaluation on ``stream``: graph replay once captured; else
        eager, then (with ``graphs``) captured for the next batch of this
        layout."""
        if self.graph is not None:
            with torch.cuda.stream(stream):
                self.graph.replay()
            return
        with torch.cuda.stream(stream):
            self._call(kp)
        if graphs:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=stream):
                self._call(kp)
            self.graph = g


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
    if not req.numbers.any():
        return "a request needs at least one atom (all atomic numbers are 0)"
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
                 memory_budget: int = 16 << 30, max_inflight: int = 3,
                 max_batch: int = 1024, merge_kinds: bool = True, graphs: bool = False):
        self.device = torch.device(device)
        self.kp = kernel_params(param, cutoff, None)
        runtime.context(self.device, self.kp.ga, self.kp.gc)  # tables resident now
        self.budget = int(memory_budget)
        self.max_inflight = int(max_inflight)
        self.slot_budget = self.budget // self.max_inflight
        self.max_batch = int(max_batch)
        self.merge_kinds = bool(merge_kinds)
        self.graphs = bool(graphs)
        self._limits: dict[tuple, int] = {}
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
        self.queues[(req.forces or self.merge_kinds, size_class(len(req.numbers)))].append(req)

    def pending(self) -> int:
        return sum(len(q) for q in self.queues.values())

    def _pick_queue(self):
        """Queue with the largest (wait of its oldest request) x sqrt(requests
        waiting): a full queue goes before an equally old short one, whose
        requests would mostly wait for the next round anyway, and a short
        queue's score still grows with its age, so it is not passed over for
        long (measured against oldest-first on a 10k-request Poisson stream:
        p50 1.4-1.7 against 2.4-2.8 ms at 120k requests/s, p99 4.2-5.0 against
        4.9-5.9 ms)."""
        now = time.perf_counter()
        best, key = -1.0, None
        for k, q in self.queues.items():
            if q:
                score = (now - q[0].arrival) * math.sqrt(len(q))
                if score > best:
                    best, key = score, k
        return key

    def _limit(self, npad: int, forces: bool) -> int:
        """Most molecules of padded length ``npad`` one slot admits."""
        key = (npad, forces)
        if key not in self._limits:
            self._limits[key] = self._admissible(npad, forces, self.max_batch)
        return self._limits[key]

    def _packed(self, slot: _Slot, reqs: list[Request], npad: int, forces: bool,
                limit: int) -> _Batch:
        """The slot's layout for these requests (padded to their count
        class), with the requests packed into its pinned buffer."""
        nbatch = count_class(len(reqs), limit)
        lay = _layout(nbatch, npad, forces)
        slot.reserve(lay["dev_end"], lay["host_end"])
        b = slot.layouts.get((nbatch, npad, forces))
        if b is None:
            b = slot.layouts[(nbatch, npad, forces)] = _Batch(slot, nbatch, npad, forces)
        b.pack(reqs)
        return b

    def _checked(self, slot, reqs: list[Request], npad: int, forces: bool, limit: int):
        """Pack the requests and check their values on the packed arrays;
        when that fails (rare), fail the bad requests alone (``_value_error``)
        and pack the rest once more. Returns (layout, requests), layout None
        when no request is left."""
        b = self._packed(slot, reqs, npad, forces, limit)
        c = len(reqs)
        hn = b.hn[:c]
        if (hn.min() >= 0 and hn.max() < 104 and hn.any(axis=1).all()
                and np.isfinite(b.hp[:c]).all() and np.isfinite(b.hc[:c]).all()):
            return b, reqs
        good = []
        for r in reqs:
            why = _value_error(r)
            if why is None:
                good.append(r)
            else:
                self._fail(r, why)
        if not good:
            return None, good
        return self._packed(slot, good, npad, forces, limit), good

    def _launch(self, key) -> bool:
        forces, npad = key
        q = self.queues[key]
        limit = self._limit(npad, forces)
        if limit == 0:  # submit() rejects such requests; kept as a guard
            raise MemoryError(f"a {npad}-atom request exceeds the per-slot memory budget "
                              f"({self.slot_budget} bytes)")
        reqs = [q.popleft() for _ in range(min(len(q), limit))]
        slot = self._free.pop()
        b, reqs = self._checked(slot, reqs, npad, forces, limit)
        if b is None:
            self._free.append(slot)
            return False
        nbatch = b.nbatch
        st = slot.stream
        with torch.cuda.stream(st):
            slot.dev[: b.in_bytes].copy_(slot.host[: b.in_bytes], non_blocking=True)
        b.run(self.kp, st, self.graphs)
        with torch.cuda.stream(st):
            lo, hi = b.out
            slot.host[lo:hi].copy_(slot.dev[lo:hi], non_blocking=True)
            # status words: the first array of the workspace
            slot.host[b.status: b.status + 4 * nbatch].copy_(b.ws[: 4 * nbatch],
                                                             non_blocking=True)
            slot.event = torch.cuda.Event()
            slot.event.record(st)
        self._bid += 1
        slot.bid = self._bid
        slot.reqs = reqs
        slot.batch = b
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
            b = slot.batch
            count = len(slot.reqs)
            # one copy of the batch's outputs out of the pinned buffer (reused by
            # the slot's next batch); each result is a view of its rows
            e = b.he[:count].copy()
            g = None if b.hg is None else b.hg[:count].copy()
            status = b.hs[:count].copy()
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
                    results.append(Result(r.rid, e[k, :n], g[k, :n] if r.forces else None,
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
