"""Object hard constraints on the GPU, differentiable (the reconstruction loop's per-batch constraints).

``ObjectConstraints.apply_hard_constraints`` for obj_type "complex" and "pure_phase"
(``object_models.py:289-301``, no FOV mask): ``clamp(|O|, 0, 1) * exp(i (angle(O) - mean angle(O)))``
(amplitude 1 for "pure_phase").  Forward and backward are two kernels each (``include/qem_constraints.h``)
instead of a dozen torch elementwise and reduction kernels; the reductions run in a fixed order.
"""

from __future__ import annotations

import ctypes

import torch

from . import _native as nat


class ObjectConstraints(torch.autograd.Function):
    @staticmethod
    def forward(ctx, obj: torch.Tensor, pure_phase: bool):
        if obj.dtype != torch.complex64:
            raise TypeError(f"obj must be complex64, got {obj.dtype}")
        if obj.device.type != "cuda":
            raise ValueError(f"obj must be on a GPU device, got {obj.device}")
        o = obj.contiguous()
        n = o.numel()
        lib = nat.load()
        nb = ctypes.c_size_t(0)
        nat.check(lib.qem_object_constraints_workspace_bytes(n, ctypes.byref(nb)), "constraints workspace")
        ws = torch.empty(int(nb.value), dtype=torch.uint8, device=o.device)
        out = torch.empty_like(o)
        stream = torch.cuda.current_stream(o.device).cuda_stream
        nat.check(lib.qem_object_constraints_forward(n, int(pure_phase), o.data_ptr(), out.data_ptr(), ws.data_ptr(),
                                                     ws.numel(), stream), "qem_object_constraints_forward")
        ctx.save_for_backward(o, ws)
        ctx.pure_phase = bool(pure_phase)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        o, ws = ctx.saved_tensors
        g = grad_out.contiguous()
        gin = torch.empty_like(o)
        lib = nat.load()
        stream = torch.cuda.current_stream(o.device).cuda_stream
        nat.check(lib.qem_object_constraints_backward(o.numel(), int(ctx.pure_phase), o.data_ptr(), g.data_ptr(),
                                                      gin.data_ptr(), ws.data_ptr(), ws.numel(), stream),
                  "qem_object_constraints_backward")
        return gin, None


def object_constraints(obj: torch.Tensor, obj_type: str = "complex") -> torch.Tensor:
    """Constrained object for obj_type "complex" or "pure_phase" (GPU, differentiable)."""
    if obj_type not in ("complex", "pure_phase"):
        raise ValueError(f"obj_type must be 'complex' or 'pure_phase', got {obj_type}")
    return ObjectConstraints.apply(obj, obj_type == "pure_phase")
