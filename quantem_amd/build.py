"""Builds the gfx950 HIP library ``quantem_amd/_lib/libqem_ptycho.so`` in-tree with hipcc.

The library is a plain C-ABI shared object (see ``include/qem_ptycho.h``); it is loaded with
ctypes, so no torch extension machinery is involved.  ``python -m quantem_amd.build`` rebuilds it.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_DIR = os.path.join(PKG_DIR, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libqem_ptycho.so")
SOURCES = [os.path.join(CSRC, "ptycho.hip"), os.path.join(CSRC, "ptycho256.hip"), os.path.join(CSRC, "nn.hip"),
           os.path.join(CSRC, "optim.hip"), os.path.join(CSRC, "constraints.hip"),
           os.path.join(CSRC, "preprocess.hip"), os.path.join(CSRC, "probe.hip"),
           os.path.join(CSRC, "ptycho_gen.hip")]
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


def _flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        # packed FP32 (v_pk_*_f32) is disabled for all device code of this library
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-I",
        os.path.join(REPO_DIR, "include"),
    ]


def compile_command(out: str = LIB_PATH) -> list[str]:
    """One hipcc line that compiles and links every source (the serial form of :func:`build`)."""
    return [hipcc(), *_flags(), "-shared", "-o", out, *SOURCES]


def build(verbose: bool = False) -> str:
    """Compiles the library (always) and returns its path.

    Each translation unit is compiled to an object of its own, in parallel, and the objects are linked
    into the shared library; no device code crosses translation units.
    """
    os.makedirs(LIB_DIR, exist_ok=True)
    obj_dir = os.path.join(LIB_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, os.path.splitext(os.path.basename(s))[0] + ".o") for s in SOURCES]
    cmds = [[hipcc(), *_flags(), "-c", "-o", o, s] for s, o in zip(SOURCES, objs)]
    if verbose:
        for c in cmds:
            print(" ".join(c), flush=True)
    with ThreadPoolExecutor(max_workers=min(len(cmds), max(1, (os.cpu_count() or 1)))) as ex:
        results = list(ex.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds))
    for c, res in zip(cmds, results):
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed ({res.returncode}) on {c[-1]}:\n{res.stderr[-4000:]}")
    tmp = LIB_PATH + ".tmp"
    link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(link), flush=True)
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc link failed ({res.returncode}):\n{res.stderr[-4000:]}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    path = build(verbose=True)
    print(path)
    sys.exit(0)
