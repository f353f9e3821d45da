"""
Build the HIP library ``tad_dftd4_amd/lib/libd4amd.so`` for gfx950 in-tree.

The library is linked against the HIP runtime that PyTorch ships, so that
torch's streams and device pointers are valid inside it (one runtime per
process). Without torch it links against ROCm's runtime.

Usage::

    python -m tad_dftd4_amd.build [--arch gfx950] [--debug]
"""

from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libd4amd.so"
SOURCES = [CSRC / "d4_kernels.hip", CSRC / "d4_eeq.hip", CSRC / "d4_sort.hip",
           CSRC / "d4_model.hip", CSRC / "d4_capi.hip"]
HEADERS = [CSRC / "d4_internal.h", PKG.parent / "include" / "d4amd.h"]


def _hip_runtime_dir() -> str:
    try:
        import torch  # noqa: F401

        d = Path(torch.__file__).resolve().parent / "lib"
        if (d / "libamdhip64.so").exists():
            return str(d)
    except ImportError:
        pass
    return os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")


def hipcc() -> str:
    exe = shutil.which("hipcc") or os.path.join(
        os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"
    )
    return exe


def command(arch: str = "gfx950", debug: bool = False, out: Path = LIB) -> list[str]:
    rt = _hip_runtime_dir()
    opt = ["-O1", "-g"] if debug else ["-O3"]
    return [
        hipcc(),
        *opt,
        "-std=c++17",
        "-fPIC",
        "-shared",
        f"--offload-arch={arch}",
        # no packed FP32 instructions anywhere
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-no-hip-rt",
        "-Wall",
        "-Wno-unused-function",
        *[str(s) for s in SOURCES],
        f"-L{rt}",
        f"-Wl,-rpath,{rt}",
        "-lamdhip64",
        "-o",
        str(out),
    ]


def build(arch: str = "gfx950", debug: bool = False, verbose: bool = False) -> Path:
    LIBDIR.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = command(arch, debug, tmp)
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(
        line for line in res.stderr.splitlines() if "packed-fp32-ops" not in line
    )
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{' '.join(cmd)}\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--debug", action="store_true")
    args = ap.parse_args(argv)
    path = build(args.arch, args.debug, verbose=True)
    print(path)
    return 0


if __name__ == "__main__":
    sys.exit(main())
