"""The allocation-free forward (qem_ptycho_forward_loss_static) against the workspace forward
(qem_ptycho_forward_loss): same kernels, so loss and intensities are bitwise equal; two calls in a row
on the shared static workspace, the second with a smaller input, at 128 x 128, at 256 x 256 then 64 x 64
(the 256 x 256 propagators once entry by entry, once as row x column factors at 0.1 A sampling), and at
the generic shapes 96 x 80 then 30 x 42."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from quantem_amd import _native as nat
from quantem_amd import ptychography as qp
from tests._cases import attach_targets, make_case


def _ptr(t):
    return None if t is None else t.data_ptr()


def _inputs(n, s, m, b, seed, dev, sampling=None):
    case = make_case(N=n, S=s, M=m, B=b, seed=seed, sampling=sampling)
    I = po.forward(case["obj"], case["probe"], case["positions"], case["propagators"])["I"]
    case = attach_targets(case, I)
    t = {k: torch.from_numpy(np.ascontiguousarray(case[k])).to(dev)
         for k in ("obj", "probe", "positions", "propagators", "targets", "mask")}
    nr, nc = (n, n) if isinstance(n, int) else n
    d = nat.QemDesc(num_slices=s, num_modes=m, batch=b, roi=nr, obj_rows=case["obj"].shape[1],
                    obj_cols=case["obj"].shape[2], num_gpts=case["num_gpts"], loss_type=0,
                    mean_intensity=case["mean_intensity"], flags=0, segment_slices=0, roi_cols=nc)
    return d, t


def _forward_workspace(d, t, dev):
    lib = nat.load()
    ws = torch.empty(qp.workspace_bytes(d), dtype=torch.uint8, device=dev)
    loss = torch.empty((), dtype=torch.float32, device=dev)
    I = torch.empty(t["targets"].shape, dtype=torch.float32, device=dev)
    st = torch.cuda.current_stream(dev).cuda_stream
    nat.check(lib.qem_ptycho_forward_loss(ctypes.byref(d), _ptr(t["obj"]), _ptr(t["probe"]), _ptr(t["positions"]),
                                          _ptr(t["propagators"]), _ptr(t["targets"]), _ptr(t["mask"]), None,
                                          _ptr(loss), _ptr(I), _ptr(ws), ws.numel(), st), "forward_loss")
    return loss, I


def _forward_static(d, t, dev):
    lib = nat.load()
    loss = torch.full((), float("nan"), dtype=torch.float32, device=dev)
    I = torch.full(t["targets"].shape, float("nan"), dtype=torch.float32, device=dev)
    st = torch.cuda.current_stream(dev).cuda_stream
    nat.check(lib.qem_ptycho_forward_loss_static(ctypes.byref(d), _ptr(t["obj"]), _ptr(t["probe"]),
                                                 _ptr(t["positions"]), _ptr(t["propagators"]), _ptr(t["targets"]),
                                                 _ptr(t["mask"]), None, _ptr(loss), _ptr(I), st),
              "forward_loss_static")
    return loss, I


@pytest.mark.gpu
@pytest.mark.parametrize("first,second", [((128, 8, 4, 40), (128, 3, 2, 9)), ((256, 3, 2, 12), (64, 2, 1, 5)),
                                          ((256, 3, 2, 12, 0.1), (64, 2, 1, 5)),
                                          (((96, 80), 3, 2, 20), ((30, 42), 2, 1, 6))])
def test_static_forward_matches_workspace_forward(cuda, first, second):
    for k, (n, s, m, b, *samp) in enumerate((first, second)):
        print(f"static forward call {k}: shape={n} s={s} m={m} b={b} sampling={samp}", flush=True)
        d, t = _inputs(n, s, m, b, 40 + k + (n if isinstance(n, int) else sum(n)), cuda, *samp)
        ref_loss, ref_I = _forward_workspace(d, t, cuda)
        loss, I = _forward_static(d, t, cuda)
        torch.cuda.synchronize()
        assert torch.equal(loss, ref_loss)
        assert torch.equal(I, ref_I)


def test_static_forward_rejects_before_touching_the_device():
    """Host-side refusals (no pointer is dereferenced, no device call is made): gradient flags, and a
    batch whose workspace exceeds the static one."""
    lib = nat.load()
    d = nat.QemDesc(num_slices=8, num_modes=4, batch=110, roi=128, obj_rows=384, obj_cols=384, num_gpts=4096,
                    loss_type=0, mean_intensity=1.0, flags=0, segment_slices=0)
    args = [None] * 10
    assert lib.qem_ptycho_forward_loss_static(ctypes.byref(d), *args) == nat.QEM_ERR_WORKSPACE
    assert "static workspace" in nat.last_error()
    d.batch = 109
    assert lib.qem_ptycho_forward_loss_static(ctypes.byref(d), *args) == nat.QEM_ERR_NULL
    d.flags = nat.QEM_FLAG_STORE_WAVES
    assert lib.qem_ptycho_forward_loss_static(ctypes.byref(d), *args) == nat.QEM_ERR_ARG
