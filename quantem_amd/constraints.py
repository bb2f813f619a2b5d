"""Hard constraints of the reconstruction loop on the GPU, differentiable (applied every batch).

* ``ObjectConstraints.apply_hard_constraints`` for obj_type "complex" and "pure_phase"
  (``object_models.py:289-301``, no FOV mask): ``clamp(|O|, 0, 1) * exp(i (angle(O) - mean angle(O)))``
  (amplitude 1 for "pure_phase").  Forward and backward are two kernels each
  (``include/qem_constraints.h``) instead of a dozen torch elementwise and reduction kernels.
* ``ProbeConstraints._probe_orthogonalization_constraint`` (``probe_models.py:484-525``): the probe
  modes orthogonalised, rescaled to their norms and sorted by intensity, forward and backward three
  kernels each (``include/qem_probe.h``) instead of the O(M^2) torch kernels of Gram-Schmidt.

The reductions run in a fixed order.
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


class ProbeOrthogonalization(torch.autograd.Function):
    @staticmethod
    def forward(ctx, probe: torch.Tensor):
        if probe.dtype != torch.complex64 or probe.dim() != 3:
            raise TypeError(f"probe must be complex64 (M, N, N), got {probe.dtype} {tuple(probe.shape)}")
        if probe.device.type != "cuda":
            raise ValueError(f"probe must be on a GPU device, got {probe.device}")
        p = probe.contiguous()
        M = p.shape[0]
        pixels = p.shape[1] * p.shape[2]
        lib = nat.load()
        nb = ctypes.c_size_t(0)
        nat.check(lib.qem_probe_orthogonalize_workspace_bytes(M, pixels, ctypes.byref(nb)), "orthogonalize workspace")
        ws = torch.empty(int(nb.value), dtype=torch.uint8, device=p.device)
        out = torch.empty_like(p)
        stream = torch.cuda.current_stream(p.device).cuda_stream
        nat.check(lib.qem_probe_orthogonalize_forward(M, pixels, p.data_ptr(), out.data_ptr(), ws.data_ptr(),
                                                      ws.numel(), stream), "qem_probe_orthogonalize_forward")
        ctx.save_for_backward(p, ws)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        p, ws = ctx.saved_tensors
        g = grad_out.contiguous()
        gin = torch.empty_like(p)
        lib = nat.load()
        stream = torch.cuda.current_stream(p.device).cuda_stream
        nat.check(lib.qem_probe_orthogonalize_backward(p.shape[0], p.shape[1] * p.shape[2], p.data_ptr(), g.data_ptr(),
                                                       gin.data_ptr(), ws.data_ptr(), ws.numel(), stream),
                  "qem_probe_orthogonalize_backward")
        return gin


def orthogonalize_probe(probe: torch.Tensor) -> torch.Tensor:
    """Orthogonalised probe modes, each rescaled to its input norm, sorted strongest first (GPU,
    differentiable; ``probe_models.py:484-525``)."""
    return ProbeOrthogonalization.apply(probe)
