"""
CPU tests of the host side: the C-ABI library loads and exports every
symbol declared in include/d4amd.h (no compute calls), the device tables
built on the host match the oracle, and the Python boundary follows the
reference's argument conventions and error types.
"""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import torch

import tad_dftd4_amd as d4
from oracle import dense
from tad_dftd4_amd import _lib, tables
from tad_dftd4_amd.disp import kernel_params

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    text = (ROOT / "include" / "d4amd.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(d4_\w+)\s*\(", text, re.M)))


def test_library_exports_header_symbols():
    lib = _lib.load()
    names = _declared_symbols()
    assert len(names) >= 9
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.SIGNATURES)


def test_workspace_bytes_and_errors():
    lib = _lib.load()
    n = ctypes.c_size_t()
    assert lib.d4_workspace_bytes(4, 10, ctypes.byref(n)) == 0
    small = n.value
    assert lib.d4_workspace_bytes(4, 20, ctypes.byref(n)) == 0
    assert n.value > small
    assert lib.d4_workspace_bytes(-1, 3, ctypes.byref(n)) == _lib.D4_ERR_INVALID
    assert b"negative" in lib.d4_last_error()
    # argument validation happens before any device work
    p = _lib.Params()
    rc = lib.d4_forward(None, None, 1, 1, None, None, None, None, None, ctypes.byref(p),
                        None, None, None, 0)
    assert rc == _lib.D4_ERR_INVALID and b"context" in lib.d4_last_error()


def test_context_rejects_null_tables():
    lib = _lib.load()
    out = ctypes.c_void_p()
    tb = _lib.Tables()
    assert lib.d4_context_create(0, ctypes.byref(tb), ctypes.byref(out)) == _lib.D4_ERR_INVALID


def test_device_tables_match_oracle_rc6():
    """Compacted element-pair C6 table equals the oracle's per-atom rc6."""
    t = tables.device_tables()
    numbers = torch.tensor([1, 6, 7, 8, 16, 3, 14, 26, 53, 92])
    rc6 = dense.get_refc6(numbers).numpy()
    for a, za in enumerate(numbers.tolist()):
        for b, zb in enumerate(numbers.tolist()):
            pa = t.perm[za, : t.nref[za]]
            pb = t.perm[zb, : t.nref[zb]]
            want = rc6[a, b][np.ix_(pa, pb)]
            got = t.rc6[za, zb, : len(pa), : len(pb)]
            assert np.allclose(got, want, rtol=1e-13, atol=0)
            assert np.all(t.rc6[za, zb, len(pa):, :] == 0)


def test_device_tables_reference_data():
    t = tables.device_tables()
    assert t.nref[1] == 2 and t.nref[6] == 7 and t.nref[92] == 6
    assert np.all(t.ngw[1, :2] == [3, 1])
    alpha = dense.get_alpha(torch.tensor([6])).numpy()[0]
    assert np.allclose(t.alpha0[6, :7], alpha[t.perm[6, :7], 0], rtol=1e-14)
    assert np.allclose(t.r4r2, dense.tables().r4r2.numpy()[:104], rtol=1e-15)


def test_kernel_params_conventions():
    kp = kernel_params({"a1": 0.4, "a2": 5.0}, None, None)
    assert kp.s8 == 0.0 and not kp.s10_on and kp.s6 == 1.0 and kp.s9 == 1.0
    assert kp.alp == 16.0 and kp.cutoff_disp2 == 60.0 and kp.cutoff_disp3 == 40.0
    kp = kernel_params({"a1": 0.4, "a2": 5.0, "s8": torch.tensor(1.2, dtype=torch.float64), "s10": 0.0},
                       d4.Cutoff(disp2=50, disp3=25, cn=3, cn_eeq=20), None)
    assert kp.s8 == 1.2 and kp.s10_on and kp.cutoff_eeq == 20.0
    assert kp.cutoff_cn == 30.0  # Cutoff.cn is ignored, as in the reference
    with pytest.raises(TypeError):
        kernel_params({"a1": 0.4}, None, None)


def test_boundary_errors_before_device_work():
    numbers = torch.tensor([6, 1])
    positions = torch.zeros(2, 3, dtype=torch.float64)
    par = {"a1": 0.4, "a2": 5.0}
    with pytest.raises(ValueError):
        d4.dftd4(numbers, torch.zeros(3, 3), 0.0, par)
    with pytest.raises(ValueError):
        d4.dftd4(numbers, positions, 0.0, par, model="d9")
    with pytest.raises(NotImplementedError):
        d4.dftd4(numbers, positions, 0.0, par, model="d3")
    with pytest.raises(NotImplementedError):
        d4.dftd4(numbers, positions, 0.0, par, model=d4.D4SModel(numbers, wf=4.0))
    with pytest.raises(ValueError):
        d4.dftd4(numbers, positions, 0.0, par, r4r2=torch.ones(3))
    with pytest.raises(ValueError):
        d4.dftd4(numbers, positions, 0.0, par, q=torch.ones(3))
    with pytest.raises(NotImplementedError):
        d4.dftd4(numbers, positions, 0.0, par, damping_function=d4.ZeroDamping())
    # CPU tensors are rejected: there is no CPU fallback
    with pytest.raises(RuntimeError):
        d4.dftd4(numbers, positions, 0.0, par)
    # cn_d4: only erf counting and a cutoff; shapes checked; no CPU fallback
    with pytest.raises(NotImplementedError):
        d4.cn_d4(numbers, positions, counting_function=lambda r, r0: r)
    with pytest.raises(NotImplementedError):
        d4.cn_d4(numbers, positions, kcn=7.0)
    with pytest.raises(ValueError):
        d4.cn_d4(numbers, torch.zeros(3, 3))
    with pytest.raises(RuntimeError):
        d4.cn_d4(numbers, positions, cutoff=10.0)


def test_cutoff_and_model_objects():
    c = d4.Cutoff()
    assert float(c.disp2) == 60.0 and float(c.cn_eeq) == 25.0
    m = d4.D4Model(torch.tensor([6, 1]))
    assert m.wf == 6.0 and m.ga == 3.0 and m.gc == 2.0
    rc6 = m.rc6
    ref = dense.get_refc6(torch.tensor([6, 1])).to(rc6.dtype)
    assert torch.allclose(rc6.double(), ref.double(), rtol=1e-6)
    with pytest.raises(NotImplementedError):
        d4.D4Model(torch.tensor([6]), ref_charges="gfn2")
    s = d4.D4SModel(torch.tensor([6, 1]))
    assert isinstance(s, d4.D4Model) and repr(s).startswith("D4SModel(")
    kp = d4.disp.kernel_params({"a1": 0.4, "a2": 5.0}, None, s)
    assert kp.d4s and kp.c().flags & 32
    assert not d4.disp.kernel_params({"a1": 0.4, "a2": 5.0}, None, m).d4s
    t = d4.tables.device_tables()
    assert t.wfpair.shape == (104, 104) and t.wfpair[6, 1] == t.wfpair[1, 6]


def test_five_point_stencil():
    """The Hessian stencil: (8 (f(h) - f(-h)) - (f(2h) - f(-2h))) / 12h is
    exact for polynomials up to degree 4 and O(h^4) otherwise."""
    from tad_dftd4_amd import disp

    h = disp.HESSIAN_STEP
    x = torch.tensor([s * h for s in disp._STENCIL], dtype=torch.float64)
    for c in ([1.0, 2.0, 3.0, 4.0, 5.0], [0.5, -1.0, 0.0, 7.0, -2.0]):
        f = sum(ci * x**i for i, ci in enumerate(c))
        assert abs(disp._five_point(f).item() - c[1]) < 1e-9
    assert abs(disp._five_point(torch.sin(1.0 + x)).item() - np.cos(1.0)) < 1e-12


def test_serving_request_validation():
    """Requests the serving loop answers with a failed result at submission
    (no GPU needed: the checks are on host arrays and workspace sizes)."""
    from tad_dftd4_amd.serve import Request, footprint, request_error, size_class

    pos = np.zeros((3, 3))
    ok = Request(0, np.array([1, 6, 8]), pos, 0.0, forces=True)
    budget = footprint(1, size_class(3), True)
    assert request_error(ok, budget) is None
    assert "at least one atom" in request_error(
        Request(1, np.zeros(0, dtype=np.int64), np.zeros((0, 3)), 0.0, False), budget)
    assert "0..103" in request_error(Request(2, np.array([1, 104, 8]), pos, 0.0, False), budget)
    assert "0..103" in request_error(Request(3, np.array([-1, 6, 8]), pos, 0.0, False), budget)
    assert "shape" in request_error(Request(4, np.array([1, 6, 8]), pos[:2], 0.0, False), budget)
    bad = pos.copy()
    bad[1, 2] = np.nan
    assert "finite" in request_error(Request(5, np.array([1, 6, 8]), bad, 0.0, False), budget)
    assert "finite" in request_error(Request(6, np.array([1, 6, 8]), pos, np.inf, False), budget)
    assert "memory budget" in request_error(ok, budget - 1)
    assert "integer" in request_error(Request(7, np.array([1.0, 6.0, 8.0]), pos, 0.0, False),
                                      budget)
    assert "all atomic numbers are 0" in request_error(
        Request(8, np.zeros(3, dtype=np.int64), pos, 0.0, False), budget)


class _HostSlot:
    """A slot with plain host memory for both buffers (no device)."""

    def __init__(self):
        import torch

        self.dev = torch.empty(0, dtype=torch.uint8)
        self.host = torch.empty(0, dtype=torch.uint8)
        self.layouts = {}

    def reserve(self, dev_bytes, host_bytes):
        import torch

        if self.dev.numel() < dev_bytes or self.host.numel() < host_bytes:
            self.dev = torch.empty(max(dev_bytes, self.dev.numel()), dtype=torch.uint8)
            self.host = torch.empty(max(host_bytes, self.host.numel()), dtype=torch.uint8)
            self.layouts.clear()


def _host_server():
    from tad_dftd4_amd.serve import D4Server

    server = object.__new__(D4Server)
    server.results = []
    return server


def test_serving_pack_and_value_checks(monkeypatch):
    """Packing a batch into its count class: each request in its own row
    with zero padding, the filler molecule in the rows after the last
    request; a request with bad values fails alone and the rest are packed
    again; when the per-request check lets a value through that the batch
    check rejects (as a planted Z = 110 does, for the device's status word
    to catch), the batch goes ahead after one repacking."""
    from tad_dftd4_amd import serve
    from tad_dftd4_amd.serve import Request

    rng = np.random.default_rng(3)
    reqs = [Request(k, rng.integers(1, 20, size=n), rng.normal(size=(n, 3)) * 3.0, 0.0, True)
            for k, n in enumerate((5, 17, 9, 30, 12))]
    server = _host_server()
    slot = _HostSlot()
    b, kept = server._checked(slot, reqs, 32, True, 1024)
    assert kept == reqs and b.nbatch == 6  # five requests: count class 6
    for k, r in enumerate(reqs):
        n = len(r.numbers)
        assert np.array_equal(b.hn[k, :n], r.numbers) and not b.hn[k, n:].any()
        assert np.array_equal(b.hp[k, :n], r.positions) and not b.hp[k, n:].any()
    assert np.array_equal(b.hn[5, :4], [7, 1, 1, 1]) and not b.hn[5, 4:].any()
    assert b.hc[5] == 0.0
    bad = Request(9, np.array([6, 120, 1]), np.zeros((3, 3)), 0.0, True)
    b, kept = server._checked(slot, reqs[:3] + [bad], 32, True, 1024)
    assert kept == reqs[:3] and b.nbatch == 3
    assert server.results[-1].rid == 9 and "0..103" in server.results[-1].error
    planted = Request(10, np.array([6, 110, 1]), np.zeros((3, 3)), 0.0, True)
    monkeypatch.setattr(serve, "_value_error", lambda r: None)
    b, kept = server._checked(slot, [planted] + reqs[:1], 32, True, 1024)
    assert kept == [planted] + reqs[:1] and b.hn[0, 1] == 110
    b, kept = server._checked(slot, [bad], 32, True, 1024)  # nothing left to fail
    assert kept == [bad]


def test_serving_count_classes():
    """A batch of c requests is padded to the smallest count class >= c:
    never more than 1.5 c, never above what a slot admits, and the same
    class for every c between two classes (one layout and graph each)."""
    from tad_dftd4_amd.serve import COUNT_CLASSES, count_class

    assert list(COUNT_CLASSES) == sorted(COUNT_CLASSES)
    for c in range(1, 1025):
        k = count_class(c, 1024)
        assert c <= k <= max(c, int(1.5 * c)) and k in COUNT_CLASSES
        assert count_class(c, 40) == min(k, 40)
    assert count_class(2000, 4096) == 2000  # beyond the table: no padding
    assert {count_class(c, 1024) for c in range(97, 129)} == {128}


def test_r4r2_gradient_is_refused():
    """dftd4 raises before any launch when r4r2 requires grad (the backward
    kernels do not form dE/dr4r2)."""
    numbers = torch.tensor([6, 1])
    positions = torch.zeros(2, 3, dtype=torch.float64)
    r4r2 = torch.ones(2, dtype=torch.float64, requires_grad=True)
    with pytest.raises(NotImplementedError):
        d4.dftd4(numbers, positions, 0.0, {"a1": 0.4, "a2": 5.0}, r4r2=r4r2)
