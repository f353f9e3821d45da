"""
The callable damping functions (reference ``damping/functions.py:116-169,
241-378``) on the CPU: closed-form values, the parameter rules of
``Damping.__call__`` and the two-body energy they assemble against the
oracle's ``dispersion2``.
"""

import numpy as np
import pytest
import torch

from oracle import dense
from tad_dftd4_amd import synth
from tad_dftd4_amd.damping import Damping, RationalDamping, ZeroDamping

F64 = torch.float64


def test_rational_closed_forms():
    """At r = a1 sqrt(Q) + a2 the value is 1 / (2 r^n); a1 = a2 = 0 gives
    r^-n; ``only_damping`` changes nothing (functions.py:262-305)."""
    q = torch.tensor([4.0, 9.0, 30.0], dtype=F64)
    r0 = 0.4 * torch.sqrt(q) + 5.0
    for n in (6, 8, 10):
        v = RationalDamping()(r0, q, n, a1=0.4, a2=5.0, s8=1.0, s10=1.0)
        assert torch.allclose(v, 0.5 / r0**n, rtol=1e-15, atol=0)
        r = torch.tensor([1.5, 2.0, 7.0], dtype=F64)
        assert torch.allclose(RationalDamping()(r, q, n, True, a1=0.0, a2=0.0, s8=1.0, s10=1.0),
                              r**-n, rtol=1e-15, atol=0)


def test_zero_closed_forms():
    """At rs R = R0 the damping is 1/7 whatever the exponent; order 8 uses
    alp + 2; without ``only_damping`` the value is divided by R^n; a zero
    radius is replaced by eps (functions.py:329-378)."""
    r = torch.tensor([2.0, 3.5, 6.0], dtype=F64)
    for n, kw in ((6, {"rs6": 1.0}), (8, {"rs8": 1.0, "s8": 1.0}), (9, {"rs9": 1.0})):
        f = ZeroDamping()(r, r, n, True, **kw)
        assert torch.allclose(f, torch.full_like(r, 1 / 7), rtol=1e-15, atol=0)
        assert torch.allclose(ZeroDamping()(r, r, n, **kw), f / r**n, rtol=1e-15, atol=0)
    r0 = torch.tensor([4.0, 7.0, 12.0], dtype=F64)
    f6 = ZeroDamping()(r, r0, 6, True, rs6=1.0, alp=14.0)
    f8 = ZeroDamping()(r, r0, 8, True, rs8=1.0, s8=1.0, alp=12.0)
    assert torch.allclose(f6, f8, rtol=1e-15, atol=0)
    assert torch.allclose(f6, 1 / (1 + 6 * (r / r0) ** 14), rtol=1e-15, atol=0)
    fz = ZeroDamping()(r, torch.zeros(3, dtype=F64), 9, True, alp=1.0)
    eps = torch.finfo(F64).eps
    assert torch.allclose(fz, 1 / (1 + 6 * r / eps), rtol=1e-15, atol=0)


def test_parameter_rules():
    """Key-presence rules of Damping.__call__ (functions.py:141-152) and the
    per-order requirements (functions.py:68-114,255-259,319-324)."""
    r = torch.tensor([3.0, 4.0], dtype=F64)
    q = torch.tensor([10.0, 20.0], dtype=F64)
    zero = torch.zeros(2, dtype=F64)
    assert torch.equal(RationalDamping()(r, q, 8, a1=0.4, a2=5.0), zero)  # no s8, rs8
    assert torch.equal(RationalDamping()(r, q, 10, a1=0.4, a2=5.0), zero)  # no s10
    assert torch.equal(RationalDamping()(r, q, 6, s6=None, rs6=None), zero)
    assert torch.equal(ZeroDamping()(r, q, 9, s9=None, rs9=None), zero)
    with pytest.raises(TypeError, match=r"RationalDamping \(order 6\) requires keyword\(s\): a1, a2"):
        RationalDamping()(r, q, 6)
    with pytest.raises(TypeError, match=r"requires keyword\(s\): a2"):
        RationalDamping()(r, q, 8, a1=0.4, rs8=1.0)
    with pytest.raises(TypeError, match=r"ZeroDamping \(order 6\) requires keyword\(s\): rs6"):
        ZeroDamping()(r, q, 6)
    with pytest.raises(TypeError, match=r"requires keyword\(s\): alp"):
        ZeroDamping()(r, q, 9, alp=None)
    with pytest.raises(ValueError, match="order 6, 8, and 9"):
        ZeroDamping()(r, q, 10, s10=1.0)
    with pytest.raises(NotImplementedError):
        Damping()(r, q, 6)
    d = RationalDamping()
    d(r, q, 6, a1=0.4, a2=5.0, doi="10.1063/1.5090222")
    assert d.doi == "10.1063/1.5090222"
    assert RationalDamping() == RationalDamping() and RationalDamping() != ZeroDamping()


def test_two_body_energy_from_the_callable():
    """-1/2 sum_j (s6 C6 f6 + s8 C8 f8 + s10 C10 f10) over real pairs inside the
    cutoff, with f from RationalDamping and Q = 3 r4r2_i r4r2_j, equals the
    oracle's dispersion2 (twobody.py:89-201)."""
    m = synth.molecule(np.random.default_rng(7), 14, synth.MIX_Z54, 2.6, (0,))
    z, pos = torch.tensor(m.numbers), torch.tensor(m.positions)
    cn = dense.cn_d4(z, pos)
    c6 = dense.get_atomic_c6(dense.get_refc6(z), dense.weight_references(z, cn, None))
    r4r2 = dense.tables().r4r2[z]
    par = {"s6": 1.0, "s8": 1.2, "s10": 0.3, "a1": 0.4, "a2": 5.0}
    cutoff = 9.0
    r = dense.cdist(pos)
    qq = 3 * r4r2.unsqueeze(-1) * r4r2.unsqueeze(-2)
    inside = (r <= cutoff) & ~torch.eye(len(z), dtype=torch.bool)
    damp = RationalDamping()
    e = (par["s6"] * c6 * damp(r, qq, 6, **par)
         + par["s8"] * c6 * qq * damp(r, qq, 8, **par)
         + par["s10"] * c6 * qq**2 * 49 / 40 * damp(r, qq, 10, **par))
    e = -0.5 * torch.where(inside, e, 0.0).sum(-1)
    eo = dense.dispersion2(z, pos, {k: torch.tensor(v, dtype=F64) for k, v in par.items()},
                           c6, r4r2, cutoff=cutoff)
    assert torch.allclose(e, eo, rtol=1e-13, atol=1e-16)
