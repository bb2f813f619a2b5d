"""The C ABI refuses to run when the translation units that share qem::Params were compiled with
different layouts of it (a build that mixed headers of two versions).

The check is exercised on the CPU: the library's sources are compiled once more, unoptimised, with
ptycho256.hip or ptycho_gen.hip given an extra leading Params field (QEM_PARAMS_LAYOUT_TEST); the
workspace query of that library must fail with QEM_ERR_LAUNCH before it reads the descriptor, while
the regular build passes the same query."""

import ctypes
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from quantem_amd import _native as nat
from quantem_amd import build as qbuild

pytestmark = pytest.mark.slow


def _desc():
    return nat.QemDesc(num_slices=2, num_modes=1, batch=1, roi=32, obj_rows=64, obj_cols=64, num_gpts=1,
                       loss_type=0, mean_intensity=1.0, flags=nat.QEM_FLAG_STORE_WAVES, segment_slices=0)


def test_regular_build_layouts_agree():
    lib = nat.load()
    out = ctypes.c_size_t(0)
    assert lib.qem_ptycho_workspace_bytes(ctypes.byref(_desc()), ctypes.byref(out)) == nat.QEM_OK
    assert out.value > 0


UNITS = ("ptycho.hip", "ptycho256.hip", "ptycho_gen.hip", "preprocess.hip")
ODD_UNITS = ("ptycho256.hip", "ptycho_gen.hip")


@pytest.fixture(scope="module")
def objects(tmp_path_factory):
    """Every unit compiled normally, and the two kernel units once more with the extra field (one
    parallel batch of compiles)."""
    tmp = str(tmp_path_factory.mktemp("layout"))
    flags = [f"--offload-arch={qbuild.ARCH}", "-O0", "-std=c++17", "-fPIC", "-I",
             os.path.join(qbuild.REPO_DIR, "include")]
    jobs = {(u, False): [] for u in UNITS}
    jobs.update({(u, True): ["-DQEM_PARAMS_LAYOUT_TEST"] for u in ODD_UNITS})
    cmds = {}
    for (u, odd), extra in jobs.items():
        out = os.path.join(tmp, u + (".odd" if odd else "") + ".o")
        cmds[(u, odd)] = (out, [qbuild.hipcc(), *flags, *extra, "-c", "-o", out, os.path.join(qbuild.CSRC, u)])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        res = dict(zip(cmds, ex.map(lambda c: subprocess.run(c[1], capture_output=True, text=True),
                                    cmds.values())))
    for r in res.values():
        assert r.returncode == 0, r.stderr[-2000:]
    return tmp, {k: v[0] for k, v in cmds.items()}


@pytest.mark.parametrize("odd_unit", ODD_UNITS)
def test_mismatched_params_layout_is_rejected(objects, odd_unit):
    tmp, objs = objects
    so = os.path.join(tmp, f"libmismatch_{odd_unit}.so")
    parts = [objs[(u, u == odd_unit)] for u in UNITS]
    r = subprocess.run([qbuild.hipcc(), f"--offload-arch={qbuild.ARCH}", "-shared", "-fPIC", "-o", so, *parts],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    lib = ctypes.CDLL(so)
    lib.qem_ptycho_workspace_bytes.argtypes = [ctypes.POINTER(nat.QemDesc), ctypes.POINTER(ctypes.c_size_t)]
    lib.qem_last_error.restype = ctypes.c_char_p
    out = ctypes.c_size_t(0)
    rc = lib.qem_ptycho_workspace_bytes(ctypes.byref(_desc()), ctypes.byref(out))
    assert rc == nat.QEM_ERR_LAUNCH
    assert b"layout" in lib.qem_last_error()
    # the descriptor is not even looked at: a NULL one gives the same answer
    assert lib.qem_ptycho_workspace_bytes(None, ctypes.byref(out)) == nat.QEM_ERR_LAUNCH
