"""The reconstruction loop's optimizer step on the GPU: object-gradient rescale + Adam in one kernel.

The reference optimises the object and probe with ``torch.optim.Adam`` (``core/ml/optimizer_mixin.py:
84-126``) after rescaling the object gradient by ``upsample_factor**2 / 2`` (``ptychography.py:330-332``).
:class:`Adam` is that optimizer (same update as ``torch.optim.Adam`` without amsgrad) with a per-group
``grad_scale`` that folds the rescale into the same pass; each parameter is one launch of
``qem_adam_step_f32`` (``include/qem_optim.h``).  Complex parameters are updated through their
interleaved float32 view, as torch's Adam does.
"""

from __future__ import annotations

import torch

from . import _native as nat


class Adam(torch.optim.Optimizer):
    """torch.optim.Adam (no amsgrad) for float32 / complex64 GPU parameters, one HIP kernel per tensor.

    Args:
        params: parameters or parameter groups; a group may set ``grad_scale`` (default 1), which
            multiplies its gradients inside the step (the gradient tensors are left untouched).
        lr, betas, eps, weight_decay: as ``torch.optim.Adam``.
    """

    def __init__(self, params, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8,
                 weight_decay: float = 0.0, grad_scale: float = 1.0):
        if lr < 0.0 or eps < 0.0 or not (0.0 <= betas[0] < 1.0) or not (0.0 <= betas[1] < 1.0):
            raise ValueError(f"invalid Adam hyper-parameters lr={lr}, betas={betas}, eps={eps}")
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay,
                                      grad_scale=grad_scale))

    @staticmethod
    def _flat(t: torch.Tensor) -> torch.Tensor:
        if t.is_complex():
            t = torch.view_as_real(t)
        if t.dtype != torch.float32:
            raise TypeError(f"Adam needs float32 or complex64 tensors, got {t.dtype}")
        if t.device.type != "cuda":
            raise ValueError(f"Adam needs GPU tensors, got {t.device}")
        return t

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        lib = nat.load()
        for group in self.param_groups:
            beta1, beta2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                pv = self._flat(p)
                if not pv.is_contiguous():
                    raise ValueError("Adam needs contiguous parameters")
                gv = self._flat(p.grad).contiguous()
                state = self.state[p]
                if not state:
                    state["step"] = 0
                    state["exp_avg"] = torch.zeros_like(pv)
                    state["exp_avg_sq"] = torch.zeros_like(pv)
                state["step"] += 1
                rc = lib.qem_adam_step_f32(pv.numel(), pv.data_ptr(), gv.data_ptr(), state["exp_avg"].data_ptr(),
                                           state["exp_avg_sq"].data_ptr(), float(group["lr"]), float(beta1),
                                           float(beta2), float(group["eps"]), float(group["weight_decay"]),
                                           float(group["grad_scale"]), int(state["step"]),
                                           torch.cuda.current_stream(pv.device).cuda_stream)
                nat.check(rc, "qem_adam_step_f32")
        return loss
