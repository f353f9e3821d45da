"""
ctypes binding of the C ABI declared in ``include/d4amd.h``.

The shared library is built in-tree by ``tad_dftd4_amd.build`` into
``tad_dftd4_amd/lib/libd4amd.so``. Loading fails loudly when it is missing:
there is no CPU or PyTorch fallback for the kernels.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libd4amd.so"

D4_OK = 0
D4_ERR_INVALID = -1
D4_ERR_DEVICE = -2
D4_ERR_NOMEM = -3
D4_ERR_WORKSPACE = -4
D4_MAX_Z = 104
D4_MAX_REF = 7
D4_FLAG_S10 = 1
D4_FLAG_FORCE_LU = 2
D4_FLAG_FORCE_TILED = 4
D4_FLAG_FORCE_SORT = 8
D4_FLAG_FORCE_PAIR_TILES = 16
D4_FLAG_D4S = 32
D4_FLAG_ATM16 = 64
D4_FLAG_PAIR_OWNER = 128
D4_FLAG_ROWS_16 = 256
D4_FLAG_ROWS_32 = 512
D4_FLAG_ROWS_64 = 768
D4_FLAG_ROWS_MASK = 768
D4_FLAG_ATM8 = 1024
D4_FLAG_ONE_STREAM = 2048
D4_NOALLOC_SCRATCH_BYTES = 64 << 20

SLOTS = ("prep", "cn", "eeq", "weights", "pairs", "atm", "atm_rows", "adjoint", "chain")
D4_PROFILE_SLOTS = 1 << 16

STATUS_LU = 1
STATUS_SINGULAR = 2
STATUS_BAD_Z = 4

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class Tables(ctypes.Structure):
    _fields_ = [
        ("rcov", _dp),
        ("en", _dp),
        ("r4r2", _dp),
        ("eeq_chi", _dp),
        ("eeq_eta", _dp),
        ("eeq_kcn", _dp),
        ("eeq_rad", _dp),
        ("gam", _dp),
        ("zeff", _dp),
        ("maxcn", _dp),
        ("nref", _ip),
        ("ngw", _ip),
        ("refcn", _dp),
        ("refq", _dp),
        ("alpha0", _dp),
        ("rc6", _dp),
        ("wfpair", _dp),
        ("alpha", _dp),
        ("refslot", _ip),
    ]


class Params(ctypes.Structure):
    _fields_ = [
        ("s6", ctypes.c_double),
        ("s8", ctypes.c_double),
        ("s10", ctypes.c_double),
        ("s9", ctypes.c_double),
        ("a1", ctypes.c_double),
        ("a2", ctypes.c_double),
        ("alp", ctypes.c_double),
        ("ga", ctypes.c_double),
        ("gc", ctypes.c_double),
        ("wf", ctypes.c_double),
        ("cutoff_disp2", ctypes.c_double),
        ("cutoff_disp3", ctypes.c_double),
        ("cutoff_cn", ctypes.c_double),
        ("cutoff_eeq", ctypes.c_double),
        ("flags", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


# exported symbol -> (restype, argtypes)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
SIGNATURES = {
    "d4_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Tables), ctypes.POINTER(_vp)]),
    "d4_context_destroy": (ctypes.c_int, [_vp]),
    "d4_workspace_bytes": (ctypes.c_int, [_i64, _i64, ctypes.POINTER(ctypes.c_size_t)]),
    "d4_forward": (
        ctypes.c_int,
        [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Params), _vp, _vp, _vp,
         ctypes.c_size_t],
    ),
    "d4_backward": (
        ctypes.c_int,
        [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Params), _vp, _vp, _vp,
         _vp, _vp, ctypes.c_size_t],
    ),
    "d4_energy_gradient": (
        ctypes.c_int,
        [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Params), _vp, _vp, _vp,
         ctypes.c_size_t],
    ),
    "d4_energy_gradient_noalloc": (
        ctypes.c_int,
        [_vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Params), _vp, _vp],
    ),
    "d4_properties": (
        ctypes.c_int,
        [_vp, _vp, _i64, _i64, _vp, _vp, _vp, ctypes.POINTER(Params), _vp, _vp, _vp, _vp, _vp,
         ctypes.c_size_t],
    ),
    "d4_model_weights": (
        ctypes.c_int,
        [_vp, _vp, _i64, _i64, _vp, _vp, _vp, ctypes.POINTER(Params), _vp, _vp, _vp],
    ),
    "d4_model_weights_backward": (
        ctypes.c_int,
        [_vp, _vp, _i64, _i64, _vp, _vp, _vp, ctypes.POINTER(Params), _vp, _vp, _vp],
    ),
    "d4_model_atomic_c6": (
        ctypes.c_int, [_vp, _vp, _i64, _i64, _vp, _vp, ctypes.POINTER(Params), _vp],
    ),
    "d4_model_atomic_c6_backward": (
        ctypes.c_int, [_vp, _vp, _i64, _i64, _vp, _vp, ctypes.POINTER(Params), _vp, _vp],
    ),
    "d4_model_pols": (
        ctypes.c_int,
        [_vp, _vp, _i64, _i64, _vp, _vp, ctypes.POINTER(Params), ctypes.c_int, _vp],
    ),
    "d4_model_pols_backward": (
        ctypes.c_int,
        [_vp, _vp, _i64, _i64, _vp, ctypes.POINTER(Params), ctypes.c_int, _vp, _vp],
    ),
    "d4_molecule_status": (ctypes.c_int, [_vp, _i64, _vp, _ip]),
    "d4_profile_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "d4_profile_read": (
        ctypes.c_int,
        [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64), ctypes.c_int],
    ),
    "d4_last_error": (ctypes.c_char_p, []),
}

_LIB = None


class D4Error(RuntimeError):
    """Error reported by the HIP library."""


def load(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load the library (once). torch is imported first so that the library
    binds to torch's HIP runtime."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    import torch  # noqa: F401  (binds libamdhip64 before our library)

    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"HIP library {p} is missing; build it with `python -m tad_dftd4_amd.build`"
        )
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc == D4_OK:
        return
    msg = load().d4_last_error().decode(errors="replace")
    if rc in (D4_ERR_INVALID, D4_ERR_WORKSPACE):
        raise ValueError(f"{what}: {msg}")
    raise D4Error(f"{what} failed ({rc}): {msg}")
