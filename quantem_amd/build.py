"""Builds the gfx950 HIP library ``quantem_amd/_lib/libqem_ptycho.so`` in-tree with hipcc.

The library is a plain C-ABI shared object (see ``include/qem_ptycho.h``); it is loaded with
ctypes, so no torch extension machinery is involved.  ``python -m quantem_amd.build`` rebuilds it.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_DIR = os.path.join(PKG_DIR, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libqem_ptycho.so")
SOURCES = [os.path.join(CSRC, "ptycho.hip"), os.path.join(CSRC, "ptycho256.hip"), os.path.join(CSRC, "nn.hip"),
           os.path.join(CSRC, "optim.hip"), os.path.join(CSRC, "constraints.hip"),
           os.path.join(CSRC, "preprocess.hip")]
HEADERS = [
    os.path.join(CSRC, "fft2.h"),
    os.path.join(CSRC, "twiddles.h"),
    os.path.join(CSRC, "common.h"),
    os.path.join(CSRC, "ptycho_common.h"),
    os.path.join(REPO_DIR, "include", "qem_ptycho.h"),
    os.path.join(REPO_DIR, "include", "qem_warmstart.h"),
]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (expected /opt/rocm/bin/hipcc)")


def compile_command(out: str = LIB_PATH) -> list[str]:
    return [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-shared",
        "-fPIC",
        # packed FP32 (v_pk_*_f32) is disabled for all device code of this library
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-I",
        os.path.join(REPO_DIR, "include"),
        "-o",
        out,
        *SOURCES,
    ]


def build(verbose: bool = False) -> str:
    """Compiles the library (always) and returns its path."""
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = compile_command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-4000:]}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    path = build(verbose=True)
    print(path)
    sys.exit(0)
