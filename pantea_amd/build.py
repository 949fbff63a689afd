"""Build libpantea_hip.so (gfx950) in-tree with hipcc.

    python -m pantea_amd.build

The library lands next to this file (pantea_amd/libpantea_hip.so) so that it
travels with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "libpantea_hip.so"

ARCH = os.environ.get("PANTEA_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 instructions (v_pk_{add,mul,fma}_f32) are generated
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]

SOURCES = ["md.hip", "nbr.hip", "acsf.hip", "mlp.hip", "scaler.hip", "kalman.hip"]

# the Kalman filter update calls rocBLAS (GEMM, SYRK, TRMM, TRSM, GEMV) and rocSOLVER (dpotrs)
LINK_FLAGS = ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found")
    return exe


def build(verbose: bool = False) -> Path:
    srcs = [str(CSRC / s) for s in SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, *srcs, *LINK_FLAGS, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        sys.stderr.write(res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
