"""The reconstruction loop's optimizer step on the GPU: object-gradient rescale + the step in one kernel.

The reference optimises each model with ``torch.optim.Adam`` (default), ``AdamW`` or ``SGD``
(``core/ml/optimizer_mixin.py:84-126``) after rescaling the object gradient by ``upsample_factor**2 / 2``
(``ptychography.py:330-332``).  :class:`Adam`, :class:`AdamW` and :class:`SGD` are those optimizers (same
updates as torch's, without amsgrad / maximize) with a per-group ``grad_scale`` that folds the rescale
into the same pass; each parameter is one launch of ``qem_adam_step_f32``, ``qem_adamw_step_f32`` or
``qem_sgd_step_f32`` (``include/qem_optim.h``).  Complex parameters are updated through their
interleaved float32 view, as torch's optimizers do.
"""

from __future__ import annotations

import torch

from . import _native as nat


def _flat(t: torch.Tensor, what: str) -> torch.Tensor:
    if t.is_complex():
        t = torch.view_as_real(t)
    if t.dtype != torch.float32:
        raise TypeError(f"{what} needs float32 or complex64 tensors, got {t.dtype}")
    if t.device.type != "cuda":
        raise ValueError(f"{what} needs GPU tensors, got {t.device}")
    return t


class Adam(torch.optim.Optimizer):
    """torch.optim.Adam (no amsgrad) for float32 / complex64 GPU parameters, one HIP kernel per tensor.

    Args:
        params: parameters or parameter groups; a group may set ``grad_scale`` (default 1), which
            multiplies its gradients inside the step (the gradient tensors are left untouched).
        lr, betas, eps, weight_decay: as ``torch.optim.Adam``.
    """

    _entry = "qem_adam_step_f32"
    _name = "Adam"

    def __init__(self, params, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8,
                 weight_decay: float = 0.0, grad_scale: float = 1.0):
        if lr < 0.0 or eps < 0.0 or not (0.0 <= betas[0] < 1.0) or not (0.0 <= betas[1] < 1.0) or weight_decay < 0:
            raise ValueError(f"invalid {self._name} hyper-parameters lr={lr}, betas={betas}, eps={eps}, "
                             f"weight_decay={weight_decay}")
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay,
                                      grad_scale=grad_scale))

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        fn = getattr(nat.load(), self._entry)
        for group in self.param_groups:
            beta1, beta2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                pv = _flat(p, self._name)
                if not pv.is_contiguous():
                    raise ValueError(f"{self._name} needs contiguous parameters")
                gv = _flat(p.grad, self._name).contiguous()
                state = self.state[p]
                if not state:
                    state["step"] = 0
                    state["exp_avg"] = torch.zeros_like(pv)
                    state["exp_avg_sq"] = torch.zeros_like(pv)
                state["step"] += 1
                rc = fn(pv.numel(), pv.data_ptr(), gv.data_ptr(), state["exp_avg"].data_ptr(),
                        state["exp_avg_sq"].data_ptr(), float(group["lr"]), float(beta1), float(beta2),
                        float(group["eps"]), float(group["weight_decay"]), float(group["grad_scale"]),
                        int(state["step"]), torch.cuda.current_stream(pv.device).cuda_stream)
                nat.check(rc, self._entry)
        return loss


class AdamW(Adam):
    """torch.optim.AdamW (decoupled weight decay, default 1e-2; no amsgrad), one HIP kernel per tensor,
    with the per-group ``grad_scale`` of :class:`Adam`."""

    _entry = "qem_adamw_step_f32"
    _name = "AdamW"

    def __init__(self, params, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8,
                 weight_decay: float = 1e-2, grad_scale: float = 1.0):
        super().__init__(params, lr=lr, betas=betas, eps=eps, weight_decay=weight_decay, grad_scale=grad_scale)


class SGD(torch.optim.Optimizer):
    """torch.optim.SGD (momentum, dampening, weight decay, Nesterov; no maximize) for float32 /
    complex64 GPU parameters, one HIP kernel per tensor, with a per-group ``grad_scale``."""

    def __init__(self, params, lr: float = 1e-3, momentum: float = 0.0, dampening: float = 0.0,
                 weight_decay: float = 0.0, nesterov: bool = False, grad_scale: float = 1.0):
        if lr < 0.0 or momentum < 0.0 or weight_decay < 0.0:
            raise ValueError(f"invalid SGD hyper-parameters lr={lr}, momentum={momentum}, weight_decay={weight_decay}")
        if nesterov and (momentum <= 0 or dampening != 0):
            raise ValueError("Nesterov momentum requires a momentum and zero dampening")
        super().__init__(params, dict(lr=lr, momentum=momentum, dampening=dampening, weight_decay=weight_decay,
                                      nesterov=nesterov, grad_scale=grad_scale))

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        lib = nat.load()
        for group in self.param_groups:
            for p in group["params"]:
                if p.grad is None:
                    continue
                pv = _flat(p, "SGD")
                if not pv.is_contiguous():
                    raise ValueError("SGD needs contiguous parameters")
                gv = _flat(p.grad, "SGD").contiguous()
                state = self.state[p]
                first = 0
                buf = None
                if group["momentum"] != 0:
                    buf = state.get("momentum_buffer")
                    if buf is None:
                        buf = torch.empty_like(pv)
                        state["momentum_buffer"] = buf
                        first = 1
                rc = lib.qem_sgd_step_f32(pv.numel(), pv.data_ptr(), gv.data_ptr(),
                                          None if buf is None else buf.data_ptr(), float(group["lr"]),
                                          float(group["momentum"]), float(group["dampening"]),
                                          float(group["weight_decay"]), int(bool(group["nesterov"])),
                                          float(group["grad_scale"]), first,
                                          torch.cuda.current_stream(pv.device).cuda_stream)
                nat.check(rc, "qem_sgd_step_f32")
        return loss


OPTIMIZER_TYPES = {"adam": Adam, "adamw": AdamW, "sgd": SGD}


def make_optimizer(params, opt_params: dict | None, default_lr: float):
    """The optimizer of one model as the reference builds it (``optimizer_mixin.py:84-126``): type
    "adam" (default), "adamw" or "sgd"; the remaining keys are the optimizer's keyword arguments, lr
    defaulting to the model's rate.  ``params`` may carry parameter groups (e.g. with ``grad_scale``)."""
    opt_params = dict(opt_params or {})
    opt_type = str(opt_params.pop("type", "adam")).lower()
    if opt_type not in OPTIMIZER_TYPES:
        raise NotImplementedError(f"Unknown optimizer type: {opt_type}")
    opt_params.setdefault("lr", default_lr)
    return OPTIMIZER_TYPES[opt_type](params, **opt_params)


SCHEDULER_TYPES = ("cyclic", "plateau", "exp", "gamma", "exponential", "linear", "none")


def make_scheduler(optimizer: torch.optim.Optimizer, params: dict | None, num_iter: int | None = None):
    """The learning-rate schedule of one model's optimizer, as the reference builds it
    (``core/ml/optimizer_mixin.py:128-183``), or None.

    ``params["type"]`` (case-insensitive): "none"; "cyclic" (CyclicLR between base_lr = lr/4 and
    max_lr = 4 lr, step_size_up 100, mode "triangular2", no momentum cycling); "plateau" (any type
    starting with "plat" or "reducelronplat": ReduceLROnPlateau, factor 0.5, patience 10, threshold
    1e-3, min_lr = lr/20, cooldown 50); "exp" / "gamma" / "exponential" (ExponentialLR with ``gamma``,
    else factor ** (1 / num_iter) with factor 0.01, else 0.9); "linear" (LinearLR from 0.1 to 1.0 over
    num_iter).  lr is the optimizer's first group's rate; every default can be overridden by a key of
    the same name (``momentum`` for cycle_momentum).  Step it once per epoch with :func:`step_scheduler`.
    """
    if not params:
        return None
    sched_type = str(params.get("type", "none")).lower()
    base = optimizer.param_groups[0]["lr"]
    sch = torch.optim.lr_scheduler
    if sched_type == "none":
        return None
    if sched_type == "cyclic":
        up = params.get("step_size_up", 100)
        return sch.CyclicLR(optimizer, base_lr=params.get("base_lr", base / 4), max_lr=params.get("max_lr", base * 4),
                            step_size_up=up, step_size_down=params.get("step_size_down", up),
                            mode=params.get("mode", "triangular2"), cycle_momentum=params.get("momentum", False))
    if sched_type.startswith(("plat", "reducelronplat")):
        return sch.ReduceLROnPlateau(optimizer, mode="min", factor=params.get("factor", 0.5),
                                     patience=params.get("patience", 10), threshold=params.get("threshold", 1e-3),
                                     min_lr=params.get("min_lr", base / 20), cooldown=params.get("cooldown", 50))
    if sched_type in ("exp", "gamma", "exponential"):
        if "gamma" in params:
            gamma = params["gamma"]
        elif num_iter is not None:
            gamma = params.get("factor", 0.01) ** (1.0 / num_iter)
        else:
            gamma = 0.9
        return sch.ExponentialLR(optimizer, gamma=gamma)
    if sched_type == "linear":
        return sch.LinearLR(optimizer, start_factor=params.get("start_factor", 0.1),
                            end_factor=params.get("end_factor", 1.0), total_iters=params.get("total_iters", num_iter))
    raise ValueError(f"Unknown scheduler type: {sched_type}, expected one of {list(SCHEDULER_TYPES)}")


def step_scheduler(scheduler, loss: float | None) -> None:
    """One epoch's step (``optimizer_mixin.py:195-202``): ReduceLROnPlateau takes the epoch loss (and
    is skipped without one), the others step unconditionally."""
    if scheduler is None:
        return
    if isinstance(scheduler, torch.optim.lr_scheduler.ReduceLROnPlateau):
        if loss is not None:
            scheduler.step(loss)
    else:
        scheduler.step()
