"""C-ABI checks that need no GPU: the library loads, exports every function include/*.h declares,
and rejects bad arguments on the host before touching the device."""

import ctypes
import glob
import os
import re

import pytest

from quantem_amd import _native as nat
from quantem_amd import ptychography as qp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(qem_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_expected_functions():
    assert declared_functions() == set(nat.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = nat.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.qem_abi_version() == 104


def _desc(**kw):
    d = dict(num_slices=8, num_modes=4, batch=256, roi=128, obj_rows=384, obj_cols=384, num_gpts=4096,
             loss_type=0, mean_intensity=1e5, flags=nat.QEM_FLAG_STORE_WAVES)
    d.update(kw)
    return nat.QemDesc(**d)


def test_workspace_size_formula():
    # stored waves + far-field waves dominate: B*M*(S+1)*N^2*8 bytes
    ws = qp.workspace_bytes(_desc())
    assert ws >= 256 * 4 * 9 * 128 * 128 * 8
    assert ws < 256 * 4 * 9 * 128 * 128 * 8 * 1.1
    ws_nostore = qp.workspace_bytes(_desc(flags=0))
    assert ws - ws_nostore == 256 * 4 * 8 * 128 * 128 * 8


def test_checkpoint_workspace_formula():
    # segment_slices k in [2, S): ceil(S/k) checkpoints + one k-slice segment instead of S stored waves
    full = qp.workspace_bytes(_desc(num_slices=20, roi=256, num_modes=2, batch=64, obj_rows=512, obj_cols=512))
    ck = qp.workspace_bytes(_desc(num_slices=20, roi=256, num_modes=2, batch=64, obj_rows=512, obj_cols=512,
                                  segment_slices=5))
    wave = 64 * 2 * 256 * 256 * 8
    assert full - ck == (20 - (4 + 5)) * wave
    # 0, 1 and >= S all mean "store every slice"
    for k in (0, 1, 20, 40):
        assert qp.workspace_bytes(_desc(num_slices=20, segment_slices=k)) == qp.workspace_bytes(_desc(num_slices=20))
    # without stored waves the setting is irrelevant
    assert qp.workspace_bytes(_desc(flags=0, segment_slices=3)) == qp.workspace_bytes(_desc(flags=0))


@pytest.mark.parametrize("field,value,code", [
    ("segment_slices", -1, nat.QEM_ERR_ARG),
    ("roi", 97, nat.QEM_ERR_UNSUPPORTED), ("roi", 1024, nat.QEM_ERR_UNSUPPORTED), ("roi", 1, nat.QEM_ERR_UNSUPPORTED),
    ("roi_cols", 11, nat.QEM_ERR_UNSUPPORTED), ("roi_cols", -1, nat.QEM_ERR_ARG), ("obj_cols", 120, nat.QEM_ERR_ARG),
    ("num_slices", 0, nat.QEM_ERR_ARG), ("batch", 0, nat.QEM_ERR_ARG),
    ("obj_rows", 64, nat.QEM_ERR_ARG), ("loss_type", 9, nat.QEM_ERR_ARG), ("mean_intensity", 0.0, nat.QEM_ERR_ARG),
    ("num_gpts", 0, nat.QEM_ERR_ARG), ("num_modes", 0, nat.QEM_ERR_ARG),
    ("batch", 32768, nat.QEM_ERR_ARG),  # 32768 x 4 modes x 128^2 = 2^31 far-field elements
])
def test_descriptor_validation(field, value, code):
    lib = nat.load()
    out = ctypes.c_size_t(0)
    rc = lib.qem_ptycho_workspace_bytes(ctypes.byref(_desc(**{field: value})), ctypes.byref(out))
    assert rc == code
    assert nat.last_error()


def test_null_and_misaligned_pointers_rejected_on_host():
    lib = nat.load()
    d = _desc(batch=2, num_slices=1)
    rc = lib.qem_ptycho_forward_loss(ctypes.byref(d), None, None, None, None, None, None, None, None, None,
                                     None, 0, None)
    assert rc == nat.QEM_ERR_NULL and "obj" in nat.last_error()
    # Made-up addresses only where a host check fails whatever the order of the checks: every call
    # below also claims a 16-byte workspace, far below the requirement, so none can reach a launch.
    rc = lib.qem_ptycho_forward_loss(ctypes.byref(d), 4100, 4096, 4096, None, 4096, 4096, None, 4096, None,
                                     4096, 16, None)
    assert rc == nat.QEM_ERR_ALIGN
    rc = lib.qem_ptycho_forward_loss(ctypes.byref(d), 4096, 4096, 4096, None, 4096, 4096, None, 4096, None,
                                     4096, 16, None)
    assert rc == nat.QEM_ERR_WORKSPACE
    d2 = _desc(batch=2, num_slices=1, flags=0)
    rc = lib.qem_ptycho_adjoint(ctypes.byref(d2), 4096, 4096, 4096, None, None, 4096, 4096, 4096, None, None,
                                4096, 16, None)
    assert rc == nat.QEM_ERR_ARG and "STORE_WAVES" in nat.last_error()
    rc = nat.load().qem_fft2_c64(48, 1, 0, 4096, 4096, None)
    assert rc == nat.QEM_ERR_UNSUPPORTED


def test_check_maps_codes_to_reference_exception_classes():
    with pytest.raises(ValueError):
        nat.check(nat.QEM_ERR_ARG, "x")
    with pytest.raises(RuntimeError):
        nat.check(nat.QEM_ERR_LAUNCH, "x")


def test_python_boundary_rejects_cpu_tensors():
    import torch

    obj = torch.ones((1, 64, 64), dtype=torch.complex64)
    probe = torch.ones((1, 32, 32), dtype=torch.complex64)
    with pytest.raises(ValueError):
        qp.multislice_loss(obj, probe, torch.zeros((1, 2)), None, torch.zeros((1, 32, 32)), num_gpts=1,
                           mean_intensity=1.0)


def test_generic_shapes_are_accepted_and_sized():
    """Nr x Nc patterns with prime factors 2, 3, 5, 7 run the generic kernels: stored waves, far field and
    the per-item row-transformed wave are Nr * Nc complex64 each, the shift ramps Nr + Nc per pattern."""
    for nr, nc in ((96, 96), (160, 192), (48, 80), (100, 120), (16, 16), (512, 7), (2, 2)):
        assert nat.roi_supported(nr, nc)
        d = _desc(roi=nr, roi_cols=nc, obj_rows=600, obj_cols=600)
        ws = qp.workspace_bytes(d)
        per = 256 * 4 * nr * nc * 8
        assert ws >= per * (8 + 1 + 1)  # 8 stored slices, far field, row-transformed wave
        assert ws - qp.workspace_bytes(_desc(roi=nr, roi_cols=nc, obj_rows=600, obj_cols=600, flags=0)) == 8 * per
    for nr, nc in ((97, 96), (96, 11), (1024, 64), (1, 8)):
        assert not nat.roi_supported(nr, nc)
    # roi_cols 0 is the square pattern
    assert qp.workspace_bytes(_desc(roi_cols=0)) == qp.workspace_bytes(_desc(roi_cols=128))
