"""Object hard constraints kernels against torch autograd of the reference formula
(object_models.py:289-301, as restated in quantem_amd.reconstruct.object_hard_constraints)."""

import ctypes

import pytest
import torch

from quantem_amd import _native as nat
from quantem_amd import constraints as qc
from quantem_amd import reconstruct as rc


def test_constraints_reject_bad_arguments_on_host():
    lib = nat.load()
    n = ctypes.c_size_t(0)
    assert lib.qem_object_constraints_workspace_bytes(0, ctypes.byref(n)) == nat.QEM_ERR_ARG
    assert lib.qem_object_constraints_workspace_bytes(10, ctypes.byref(n)) == nat.QEM_OK and n.value >= 8 * 1026
    assert lib.qem_object_constraints_forward(10, 0, None, 8, 8, 1 << 20, None) == nat.QEM_ERR_NULL
    assert lib.qem_object_constraints_forward(10, 0, 12, 8, 8, 1 << 20, None) == nat.QEM_ERR_ALIGN
    assert lib.qem_object_constraints_forward(10, 0, 8, 8, 8, 16, None) == nat.QEM_ERR_WORKSPACE
    assert lib.qem_object_constraints_backward(10, 0, 8, 8, None, 8, 1 << 20, None) == nat.QEM_ERR_NULL
    with pytest.raises(ValueError):
        qc.object_constraints(torch.ones(2, 2, dtype=torch.complex64), "potential")


@pytest.mark.gpu
@pytest.mark.parametrize("obj_type,shape", [("complex", (3, 70, 90)), ("pure_phase", (2, 64, 64)),
                                            ("complex", (8, 384, 384))])
def test_constraints_match_torch_autograd(cuda, obj_type, shape):
    g = torch.Generator(device=cuda).manual_seed(5)
    amp = 0.5 + torch.rand(shape, device=cuda, generator=g)  # 0.5 .. 1.5: the clamp is active on half
    ph = 3.0 * torch.rand(shape, device=cuda, generator=g) - 1.0
    obj = (amp * torch.exp(1j * ph)).to(torch.complex64)
    up = torch.randn(shape, dtype=torch.complex64, device=cuda, generator=g)
    a = obj.clone().requires_grad_(True)
    out = qc.object_constraints(a, obj_type)
    out.backward(up)
    b = obj.clone().requires_grad_(True)
    ref = rc.object_hard_constraints(b, obj_type)
    ref.backward(up)
    rel = lambda x, y: float((x - y).abs().norm() / y.abs().norm())
    assert rel(out.detach(), ref.detach()) < 2e-6
    assert rel(a.grad, b.grad) < 2e-5
