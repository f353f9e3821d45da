"""
Device runtime: per-device contexts (element tables resident in HBM),
workspace sizing, and thin wrappers that call the C ABI on torch tensors
with torch's current HIP stream. No host synchronisation happens here.
"""

from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .tables import GA_DEFAULT, GC_DEFAULT, WF_DEFAULT, device_tables

_CTX: dict[tuple, ctypes.c_void_p] = {}
# Launch-shape table in use (``tune.activate``): consulted by calls whose
# KernelParams leave both shapes at 0.
_SHAPES = None
_ROW_FLAGS = {0: 0, 16: _lib.D4_FLAG_ROWS_16, 32: _lib.D4_FLAG_ROWS_32, 64: _lib.D4_FLAG_ROWS_64}
_CTX_LOCK = threading.Lock()
_KEEPALIVE: list = []


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def context(device: torch.device, ga: float = GA_DEFAULT, gc: float = GC_DEFAULT):
    """Context for ``device`` and the model constants (ga, gc), created once
    and kept for the life of the process."""
    lib = _lib.load()
    index = device.index if device.index is not None else torch.cuda.current_device()
    key = (index, float(ga), float(gc))
    with _CTX_LOCK:
        if key in _CTX:
            return _CTX[key]
        t = device_tables(float(ga), float(gc))
        arrays = [t.rcov, t.en, t.r4r2, t.eeq_chi, t.eeq_eta, t.eeq_kcn, t.eeq_rad,
                  t.gam, t.zeff, t.maxcn, t.refcn, t.refq, t.alpha0, t.rc6, t.nref, t.ngw,
                  t.wfpair, t.alpha, t.refslot]
        _KEEPALIVE.append(arrays)
        tb = _lib.Tables(
            rcov=_ptr(t.rcov, ctypes.c_double),
            en=_ptr(t.en, ctypes.c_double),
            r4r2=_ptr(t.r4r2, ctypes.c_double),
            eeq_chi=_ptr(t.eeq_chi, ctypes.c_double),
            eeq_eta=_ptr(t.eeq_eta, ctypes.c_double),
            eeq_kcn=_ptr(t.eeq_kcn, ctypes.c_double),
            eeq_rad=_ptr(t.eeq_rad, ctypes.c_double),
            gam=_ptr(t.gam, ctypes.c_double),
            zeff=_ptr(t.zeff, ctypes.c_double),
            maxcn=_ptr(t.maxcn, ctypes.c_double),
            nref=_ptr(t.nref, ctypes.c_int32),
            ngw=_ptr(t.ngw, ctypes.c_int32),
            refcn=_ptr(t.refcn, ctypes.c_double),
            refq=_ptr(t.refq, ctypes.c_double),
            alpha0=_ptr(t.alpha0, ctypes.c_double),
            rc6=_ptr(t.rc6, ctypes.c_double),
            wfpair=_ptr(t.wfpair, ctypes.c_double),
            alpha=_ptr(t.alpha, ctypes.c_double),
            refslot=_ptr(t.refslot, ctypes.c_int32),
        )
        out = ctypes.c_void_p()
        _lib.check(lib.d4_context_create(index, ctypes.byref(tb), ctypes.byref(out)),
                   "d4_context_create")
        _CTX[key] = out
        return out


def workspace_bytes(nbatch: int, natoms: int) -> int:
    n = ctypes.c_size_t()
    _lib.check(_lib.load().d4_workspace_bytes(nbatch, natoms, ctypes.byref(n)),
               "d4_workspace_bytes")
    return int(n.value)


def new_workspace(nbatch: int, natoms: int, device) -> torch.Tensor:
    return torch.empty(workspace_bytes(nbatch, natoms), dtype=torch.uint8, device=device)


@dataclass
class KernelParams:
    """Damping parameters, model constants and cutoffs for one call."""

    s6: float = 1.0
    s8: float = 0.0
    s10: float = 0.0
    s9: float = 1.0
    a1: float = 0.4
    a2: float = 5.0
    alp: float = 16.0
    ga: float = GA_DEFAULT
    gc: float = GC_DEFAULT
    wf: float = WF_DEFAULT
    cutoff_disp2: float = 60.0
    cutoff_disp3: float = 40.0
    cutoff_cn: float = 30.0
    cutoff_eeq: float = 25.0
    s10_on: bool = False
    force_lu: bool = False
    force_tiled: bool = False
    force_sort: bool = False
    force_pair_tiles: bool = False
    d4s: bool = False  # D4S model (pairwise Gaussian weighting factors)
    atm16: bool = False  # testing: one-pass three-body kernel with 16 x 16 pair tiles
    pair_owner: bool = False  # testing: large systems take the pair-owner three-body kernel
    # launch shapes (results agree to rounding): lanes per atom row of the CN,
    # pair and chain passes (16, 32, 64; 0: by padded length) and the pair tile
    # of the one-pass three-body kernel (8, 16; 0: 16 for energies, 8 with
    # gradients). With both at 0 an activated shape table decides.
    row_lanes: int = 0
    atm_tile: int = 0
    one_stream: bool = False  # profiling: no companion stream for the three-body side

    def shapes(self, natoms: int | None = None, grad: bool = False) -> tuple[int, int]:
        """(row_lanes, atm_tile) for a call on a batch padded to ``natoms``."""
        if self.row_lanes == 0 and self.atm_tile == 0 and _SHAPES is not None and natoms:
            return _SHAPES.lookup(grad, natoms)
        return self.row_lanes, self.atm_tile

    def c(self, natoms: int | None = None, grad: bool = False) -> _lib.Params:
        rows, tile = self.shapes(natoms, grad)
        if rows not in _ROW_FLAGS:
            raise ValueError(f"row_lanes must be 0, 16, 32 or 64, got {rows}")
        if tile not in (0, 8, 16):
            raise ValueError(f"atm_tile must be 0, 8 or 16, got {tile}")
        flags = (
            (_lib.D4_FLAG_S10 if self.s10_on else 0)
            | (_lib.D4_FLAG_FORCE_LU if self.force_lu else 0)
            | (_lib.D4_FLAG_FORCE_TILED if self.force_tiled else 0)
            | (_lib.D4_FLAG_FORCE_SORT if self.force_sort else 0)
            | (_lib.D4_FLAG_FORCE_PAIR_TILES if self.force_pair_tiles else 0)
            | (_lib.D4_FLAG_D4S if self.d4s else 0)
            | (_lib.D4_FLAG_ATM16 if self.atm16 or tile == 16 else 0)
            | (_lib.D4_FLAG_PAIR_OWNER if self.pair_owner else 0)
            | _ROW_FLAGS[rows]
            | (_lib.D4_FLAG_ATM8 if tile == 8 else 0)
            | (_lib.D4_FLAG_ONE_STREAM if self.one_stream else 0)
        )
        return _lib.Params(
            self.s6, self.s8, self.s10, self.s9, self.a1, self.a2, self.alp,
            self.ga, self.gc, self.wf, self.cutoff_disp2, self.cutoff_disp3,
            self.cutoff_cn, self.cutoff_eeq, flags, 0,
        )


def _vp(t: torch.Tensor | None):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _check_inputs(numbers, positions, charge):
    if not positions.is_cuda:
        raise RuntimeError("tad_dftd4_amd runs on a ROCm device; got CPU tensors")
    assert numbers.dtype == torch.int64 and numbers.is_contiguous()
    assert positions.dtype == torch.float64 and positions.is_contiguous()
    assert charge.dtype == torch.float64 and charge.is_contiguous()
    assert numbers.device == positions.device == charge.device


def forward(numbers, positions, charge, kp: KernelParams, q=None, r4r2=None,
            ws=None, q_out=None, energy=None):
    """Atom energies (B, N); returns (energy, workspace)."""
    _check_inputs(numbers, positions, charge)
    B, N = numbers.shape
    dev = positions.device
    ctx = context(dev, kp.ga, kp.gc)
    if ws is None:
        ws = new_workspace(B, N, dev)
    if energy is None:
        energy = torch.empty((B, N), dtype=torch.float64, device=dev)
    p = kp.c(N, grad=False)
    _lib.check(
        _lib.load().d4_forward(
            ctx, _stream(dev), B, N, _vp(numbers), _vp(positions), _vp(charge), _vp(q),
            _vp(r4r2), ctypes.byref(p), _vp(energy), _vp(q_out), _vp(ws), ws.numel()),
        "d4_forward",
    )
    return energy, ws


PARAM_ORDER = ("s6", "s8", "s10", "s9", "a1", "a2", "alp")


def backward(numbers, positions, charge, kp: KernelParams, ws, grad_energy=None,
             q=None, r4r2=None, params: bool = False, charges: bool = False):
    """Gradient of L = sum(grad_energy * E) w.r.t. positions, (B, N, 3); with
    ``params`` also dL/d(s6, s8, s10, s9, a1, a2, alp) as a (7,) tensor
    (order ``PARAM_ORDER``), returned as a second value; with ``charges``
    (given ``q``) also dL/dq (B, N): then the result is (grad, gpar or None,
    grad_q)."""
    _check_inputs(numbers, positions, charge)
    B, N = numbers.shape
    dev = positions.device
    ctx = context(dev, kp.ga, kp.gc)
    grad = torch.empty((B, N, 3), dtype=torch.float64, device=dev)
    gpar = torch.empty(7, dtype=torch.float64, device=dev) if params else None
    gq = torch.empty((B, N), dtype=torch.float64, device=dev) if charges else None
    if grad_energy is not None:
        grad_energy = grad_energy.to(torch.float64).contiguous()
    p = kp.c(N, grad=True)
    _lib.check(
        _lib.load().d4_backward(
            ctx, _stream(dev), B, N, _vp(numbers), _vp(positions), _vp(charge), _vp(q),
            _vp(r4r2), ctypes.byref(p), _vp(grad_energy), _vp(grad), _vp(gpar), _vp(gq),
            _vp(ws), ws.numel()),
        "d4_backward",
    )
    if charges:
        return grad, gpar, gq
    return (grad, gpar) if params else grad


def energy_gradient(numbers, positions, charge, kp: KernelParams, q=None, r4r2=None,
                    ws=None, energy=None, grad=None):
    """Fused atom energies and gradient of the total energy."""
    _check_inputs(numbers, positions, charge)
    B, N = numbers.shape
    dev = positions.device
    ctx = context(dev, kp.ga, kp.gc)
    if ws is None:
        ws = new_workspace(B, N, dev)
    if energy is None:
        energy = torch.empty((B, N), dtype=torch.float64, device=dev)
    if grad is None:
        grad = torch.empty((B, N, 3), dtype=torch.float64, device=dev)
    p = kp.c(N, grad=True)
    _lib.check(
        _lib.load().d4_energy_gradient(
            ctx, _stream(dev), B, N, _vp(numbers), _vp(positions), _vp(charge), _vp(q),
            _vp(r4r2), ctypes.byref(p), _vp(energy), _vp(grad), _vp(ws), ws.numel()),
        "d4_energy_gradient",
    )
    return energy, grad, ws


def energy_gradient_noalloc(numbers, positions, charge, kp: KernelParams, q=None, r4r2=None,
                            energy=None, grad=None):
    """``energy_gradient`` through d4_energy_gradient_noalloc: the library's
    fixed per-device scratch (``_lib.D4_NOALLOC_SCRATCH_BYTES``) instead of a
    workspace; no device allocation when ``energy`` and ``grad`` are given.
    Calls that overlap on one GPU race on the scratch."""
    _check_inputs(numbers, positions, charge)
    B, N = numbers.shape
    dev = positions.device
    ctx = context(dev, kp.ga, kp.gc)
    if energy is None:
        energy = torch.empty((B, N), dtype=torch.float64, device=dev)
    if grad is None:
        grad = torch.empty((B, N, 3), dtype=torch.float64, device=dev)
    p = kp.c(N, grad=True)
    _lib.check(
        _lib.load().d4_energy_gradient_noalloc(
            ctx, _stream(dev), B, N, _vp(numbers), _vp(positions), _vp(charge), _vp(q),
            _vp(r4r2), ctypes.byref(p), _vp(energy), _vp(grad)),
        "d4_energy_gradient_noalloc",
    )
    return energy, grad


def properties(numbers, positions, charge, kp: KernelParams, ws=None, cn_only: bool = False):
    """(cn, q, c6, alpha) as in get_properties; with ``cn_only`` just the
    coordination numbers are computed and (cn, None, None, None) returned."""
    _check_inputs(numbers, positions, charge)
    B, N = numbers.shape
    dev = positions.device
    ctx = context(dev, kp.ga, kp.gc)
    if ws is None:
        ws = new_workspace(B, N, dev)
    f64 = dict(dtype=torch.float64, device=dev)
    cn = torch.empty((B, N), **f64)
    q = None if cn_only else torch.empty((B, N), **f64)
    c6 = None if cn_only else torch.empty((B, N, N), **f64)
    alpha = None if cn_only else torch.empty((B, N), **f64)
    p = kp.c(N, grad=False)
    _lib.check(
        _lib.load().d4_properties(
            ctx, _stream(dev), B, N, _vp(numbers), _vp(positions), _vp(charge),
            ctypes.byref(p), _vp(cn), _vp(q), _vp(c6), _vp(alpha), _vp(ws), ws.numel()),
        "d4_properties",
    )
    return cn, q, c6, alpha


# ------------------------------------------------------------ model methods


def _model_params(ga: float, gc: float, wf: float, d4s: bool) -> _lib.Params:
    return _lib.Params(0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, ga, gc, wf, 0.0, 0.0, 0.0, 0.0,
                       _lib.D4_FLAG_D4S if d4s else 0, 0)


def _f64(t: torch.Tensor | None, device) -> torch.Tensor | None:
    if t is None:
        return None
    return t.detach().to(device=device, dtype=torch.float64).contiguous()


def model_weights(numbers, cn, q, ga: float, gc: float, wf: float, d4s: bool,
                  dgwdcn: bool = False, dgwdq: bool = False):
    """Reference weights (B, N, 7) (D4S: (B, N, N, 7)) for ``numbers`` (B, N)
    int64 on the device and optional cn, q (B, N); with ``dgwdcn``/``dgwdq``
    also the reference's analytic derivatives (-d gw/d cn, d gw/d q).
    Returns (gw, dgwdcn or None, dgwdq or None)."""
    dev = numbers.device
    B, N = numbers.shape
    ctx = context(dev, ga, gc)
    shape = (B, N, N, 7) if d4s else (B, N, 7)
    f64 = dict(dtype=torch.float64, device=dev)
    gw = torch.empty(shape, **f64)
    dc = torch.empty(shape, **f64) if dgwdcn else None
    dq = torch.empty(shape, **f64) if dgwdq else None
    cn, q = _f64(cn, dev), _f64(q, dev)
    p = _model_params(ga, gc, wf, d4s)
    _lib.check(
        _lib.load().d4_model_weights(ctx, _stream(dev), B, N, _vp(numbers), _vp(cn), _vp(q),
                                     ctypes.byref(p), _vp(gw), _vp(dc), _vp(dq)),
        "d4_model_weights",
    )
    return gw, dc, dq


def model_weights_backward(numbers, cn, q, ga, gc, wf, d4s, grad_gw, want_cn=True,
                           want_q=True):
    dev = numbers.device
    B, N = numbers.shape
    ctx = context(dev, ga, gc)
    f64 = dict(dtype=torch.float64, device=dev)
    gcn = torch.empty((B, N), **f64) if want_cn else None
    gq = torch.empty((B, N), **f64) if want_q else None
    cn, q, grad_gw = _f64(cn, dev), _f64(q, dev), _f64(grad_gw, dev)
    p = _model_params(ga, gc, wf, d4s)
    _lib.check(
        _lib.load().d4_model_weights_backward(ctx, _stream(dev), B, N, _vp(numbers), _vp(cn),
                                              _vp(q), ctypes.byref(p), _vp(grad_gw), _vp(gcn),
                                              _vp(gq)),
        "d4_model_weights_backward",
    )
    return gcn, gq


def model_atomic_c6(numbers, gw, ga, gc, d4s):
    dev = numbers.device
    B, N = numbers.shape
    ctx = context(dev, ga, gc)
    c6 = torch.empty((B, N, N), dtype=torch.float64, device=dev)
    gw = _f64(gw, dev)
    p = _model_params(ga, gc, 0.0, d4s)
    _lib.check(
        _lib.load().d4_model_atomic_c6(ctx, _stream(dev), B, N, _vp(numbers), _vp(gw),
                                       ctypes.byref(p), _vp(c6)),
        "d4_model_atomic_c6",
    )
    return c6


def model_atomic_c6_backward(numbers, gw, ga, gc, d4s, grad_c6):
    dev = numbers.device
    B, N = numbers.shape
    ctx = context(dev, ga, gc)
    gw = _f64(gw, dev)
    out = torch.empty_like(gw)
    grad_c6 = _f64(grad_c6, dev)
    p = _model_params(ga, gc, 0.0, d4s)
    _lib.check(
        _lib.load().d4_model_atomic_c6_backward(ctx, _stream(dev), B, N, _vp(numbers), _vp(gw),
                                                ctypes.byref(p), _vp(grad_c6), _vp(out)),
        "d4_model_atomic_c6_backward",
    )
    return out


def _pols_shape(B, N, d4s, static_only):
    if static_only:
        return (B, N, N) if d4s else (B, N)
    return (B, N, 23)


def model_pols(numbers, gw, ga, gc, d4s, static_only):
    dev = numbers.device
    B, N = numbers.shape
    ctx = context(dev, ga, gc)
    out = torch.empty(_pols_shape(B, N, d4s, static_only), dtype=torch.float64, device=dev)
    gw = _f64(gw, dev)
    p = _model_params(ga, gc, 0.0, d4s)
    _lib.check(
        _lib.load().d4_model_pols(ctx, _stream(dev), B, N, _vp(numbers), _vp(gw),
                                  ctypes.byref(p), int(static_only), _vp(out)),
        "d4_model_pols",
    )
    return out


def model_pols_backward(numbers, ga, gc, d4s, static_only, grad_out):
    dev = numbers.device
    B, N = numbers.shape
    ctx = context(dev, ga, gc)
    shape = (B, N, N, 7) if d4s else (B, N, 7)
    out = torch.empty(shape, dtype=torch.float64, device=dev)
    grad_out = _f64(grad_out, dev)
    p = _model_params(ga, gc, 0.0, d4s)
    _lib.check(
        _lib.load().d4_model_pols_backward(ctx, _stream(dev), B, N, _vp(numbers),
                                           ctypes.byref(p), int(static_only), _vp(grad_out),
                                           _vp(out)),
        "d4_model_pols_backward",
    )
    return out


class CapturedStep:
    """One evaluation of a fixed-shape batch captured into a HIP graph and
    replayed: the launch sequence of the caller's stream and of its companion
    stream (forked and joined by events) without the host-side launch gaps.

    The inputs are copied into static device buffers (``numbers``,
    ``positions``, ``charge``, and ``q``/``r4r2`` when given); write new values
    into them in place between replays. Damping parameters and cutoffs are
    fixed at capture. ``replay()`` enqueues the graph on the current stream
    and returns the static outputs: the atom energies, with ``forces`` also
    the gradient (the same values as ``energy_gradient``/``forward``)."""

    def __init__(self, numbers, positions, charge, kp: KernelParams, forces: bool = True,
                 q=None, r4r2=None):
        _check_inputs(numbers, positions, charge)
        dev = positions.device
        B, N = numbers.shape
        self.kp = kp
        self.forces = forces
        self.numbers = numbers.clone()
        self.positions = positions.clone()
        self.charge = charge.clone()
        self.q = None if q is None else q.clone()
        self.r4r2 = None if r4r2 is None else r4r2.clone()
        self.ws = new_workspace(B, N, dev)
        self.energy = torch.empty((B, N), dtype=torch.float64, device=dev)
        self.grad = torch.empty((B, N, 3), dtype=torch.float64, device=dev) if forces else None
        # capture on a stream of our own; one eager call on it first creates
        # its companion stream and sets kernel attributes outside the capture
        self.stream = torch.cuda.Stream(device=dev)
        cur = torch.cuda.current_stream(dev)
        self.stream.wait_stream(cur)
        with torch.cuda.stream(self.stream):
            self._launch()
        cur.wait_stream(self.stream)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            self._launch()

    def _launch(self):
        if self.forces:
            energy_gradient(self.numbers, self.positions, self.charge, self.kp, q=self.q,
                            r4r2=self.r4r2, ws=self.ws, energy=self.energy, grad=self.grad)
        else:
            forward(self.numbers, self.positions, self.charge, self.kp, q=self.q,
                    r4r2=self.r4r2, ws=self.ws, energy=self.energy)

    def replay(self):
        self.graph.replay()
        return (self.energy, self.grad) if self.forces else self.energy


def molecule_status(ws: torch.Tensor, nbatch: int) -> np.ndarray:
    """Per-molecule status bits of the last call on ``ws`` (synchronises)."""
    out = np.zeros(nbatch, dtype=np.int32)
    _lib.check(
        _lib.load().d4_molecule_status(_stream(ws.device), nbatch, _vp(ws),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
        "d4_molecule_status",
    )
    return out


def profile_enable(device, on: bool, ga: float = GA_DEFAULT, gc: float = GC_DEFAULT,
                   slots=None) -> None:
    """Per-kernel HIP-event timing of calls through the context of ``device``:
    every kernel, or only the kernel slots named in ``slots`` (names of
    ``profile_read``)."""
    code = int(bool(on))
    if on and slots is not None:
        code = _lib.D4_PROFILE_SLOTS
        for name in slots:
            code |= 1 << _lib.SLOTS.index(name)
    _lib.check(_lib.load().d4_profile_enable(context(device, ga, gc), code),
               "d4_profile_enable")


def profile_read(device, ga: float = GA_DEFAULT, gc: float = GC_DEFAULT) -> dict:
    """{kernel: (total_ms, launches)} since the last read (synchronises)."""
    n = len(_lib.SLOTS)
    ms = (ctypes.c_double * n)()
    cnt = (ctypes.c_int64 * n)()
    _lib.check(_lib.load().d4_profile_read(context(device, ga, gc), ms, cnt, n),
               "d4_profile_read")
    return {name: (ms[k], cnt[k]) for k, name in enumerate(_lib.SLOTS)}
