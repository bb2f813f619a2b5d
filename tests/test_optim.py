"""The optimizer step (object-gradient rescale + Adam in one kernel) against torch.optim.Adam."""

import pytest
import torch

from quantem_amd import _native as nat
from quantem_amd import optim


def test_adam_rejects_bad_arguments_on_host():
    lib = nat.load()
    assert lib.qem_adam_step_f32(-1, None, None, None, None, 1e-3, 0.9, 0.999, 1e-8, 0.0, 1.0, 1, None) == nat.QEM_ERR_ARG
    assert lib.qem_adam_step_f32(8, None, None, None, None, 1e-3, 0.9, 0.999, 1e-8, 0.0, 1.0, 0, None) == nat.QEM_ERR_ARG
    assert lib.qem_adam_step_f32(8, None, None, None, None, 1e-3, 1.5, 0.999, 1e-8, 0.0, 1.0, 1, None) == nat.QEM_ERR_ARG
    assert lib.qem_adam_step_f32(8, None, 64, 64, 64, 1e-3, 0.9, 0.999, 1e-8, 0.0, 1.0, 1, None) == nat.QEM_ERR_NULL
    assert lib.qem_adam_step_f32(8, 72, 64, 64, 64, 1e-3, 0.9, 0.999, 1e-8, 0.0, 1.0, 1, None) == nat.QEM_ERR_ALIGN
    assert lib.qem_adam_step_f32(0, None, None, None, None, 1e-3, 0.9, 0.999, 1e-8, 0.0, 1.0, 1, None) == nat.QEM_OK
    with pytest.raises(ValueError):
        optim.Adam([torch.zeros(3)], lr=-1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,dtype,wd,gscale", [((8, 33, 41), torch.complex64, 0.0, 18.0), ((4, 16, 16), torch.complex64, 0.0, 1.0),
                                                   ((1001,), torch.float32, 0.01, 0.5)])
def test_adam_matches_torch(cuda, shape, dtype, wd, gscale):
    """Five steps on random gradients: parameters and moments within float32 rounding of torch.optim.Adam
    (foreach=False) fed the same gradients pre-multiplied by grad_scale; odd sizes exercise the tail."""
    g = torch.Generator(device=cuda).manual_seed(3)
    p0 = torch.randn(shape, dtype=dtype, device=cuda, generator=g)
    a = p0.clone().requires_grad_(True)
    b = p0.clone().requires_grad_(True)
    ours = optim.Adam([{"params": [a], "lr": 5e-3, "grad_scale": gscale}], weight_decay=wd)
    ref = torch.optim.Adam([b], lr=5e-3, weight_decay=wd, foreach=False)
    for _ in range(5):
        gr = torch.randn(shape, dtype=dtype, device=cuda, generator=g)
        a.grad = gr.clone()
        b.grad = gr * gscale
        ours.step()
        ref.step()
        assert torch.equal(a.grad, gr)  # the scale is applied inside the step, not to .grad
    scale = b.detach().abs().max()
    assert (a.detach() - b.detach()).abs().max() <= 1e-6 * scale
    for k in ("exp_avg", "exp_avg_sq"):
        ra = ours.state[a][k]
        rb = ref.state[b][k]
        rb = torch.view_as_real(rb) if rb.is_complex() else rb
        assert (ra - rb).abs().max() <= 1e-5 * rb.abs().max(), k


def test_adamw_sgd_reject_bad_arguments_on_host():
    lib = nat.load()
    assert lib.qem_adamw_step_f32(8, None, None, None, None, 1e-3, 0.9, 0.999, 1e-8, -0.1, 1.0, 1, None) == nat.QEM_ERR_ARG
    assert lib.qem_adamw_step_f32(8, 72, 64, 64, 64, 1e-3, 0.9, 0.999, 1e-8, 0.01, 1.0, 1, None) == nat.QEM_ERR_ALIGN
    assert lib.qem_sgd_step_f32(-1, None, None, None, 1e-3, 0.0, 0.0, 0.0, 0, 1.0, 1, None) == nat.QEM_ERR_ARG
    assert lib.qem_sgd_step_f32(8, 64, 64, None, 1e-3, 0.0, 0.0, 0.0, 1, 1.0, 1, None) == nat.QEM_ERR_ARG  # nesterov
    assert lib.qem_sgd_step_f32(8, 64, 64, None, 1e-3, 0.9, 0.0, 0.0, 0, 1.0, 1, None) == nat.QEM_ERR_NULL
    assert lib.qem_sgd_step_f32(8, 64, 72, None, 1e-3, 0.0, 0.0, 0.0, 0, 1.0, 1, None) == nat.QEM_ERR_ALIGN
    assert lib.qem_sgd_step_f32(0, None, None, None, 1e-3, 0.0, 0.0, 0.0, 0, 1.0, 1, None) == nat.QEM_OK
    with pytest.raises(ValueError):
        optim.SGD([torch.zeros(3)], lr=0.1, nesterov=True)
    with pytest.raises(NotImplementedError):
        optim.make_optimizer([torch.zeros(3)], {"type": "lbfgs"}, 1e-3)
    assert isinstance(optim.make_optimizer([torch.zeros(3)], None, 5e-3), optim.Adam)
    o = optim.make_optimizer([torch.zeros(3)], {"type": "AdamW", "weight_decay": 0.1}, 5e-3)
    assert isinstance(o, optim.AdamW) and o.param_groups[0]["lr"] == 5e-3 and o.param_groups[0]["weight_decay"] == 0.1


def _steps(ours, ref, a, b, shape, dtype, gscale, dev, n=5):
    g = torch.Generator(device=dev).manual_seed(5)
    for _ in range(n):
        gr = torch.randn(shape, dtype=dtype, device=dev, generator=g)
        a.grad = gr.clone()
        b.grad = gr * gscale
        ours.step()
        ref.step()
        assert torch.equal(a.grad, gr)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,dtype,wd,gscale", [((8, 33, 41), torch.complex64, 0.01, 7.75),
                                                   ((1001,), torch.float32, 0.1, 1.0)])
def test_adamw_matches_torch(cuda, shape, dtype, wd, gscale):
    """AdamW (decoupled decay) against torch.optim.AdamW (foreach=False) over five steps."""
    p0 = torch.randn(shape, dtype=dtype, device=cuda, generator=torch.Generator(device=cuda).manual_seed(4))
    a = p0.clone().requires_grad_(True)
    b = p0.clone().requires_grad_(True)
    ours = optim.AdamW([{"params": [a], "lr": 5e-3, "grad_scale": gscale}], weight_decay=wd)
    ref = torch.optim.AdamW([b], lr=5e-3, weight_decay=wd, foreach=False)
    _steps(ours, ref, a, b, shape, dtype, gscale, cuda)
    assert (a.detach() - b.detach()).abs().max() <= 1e-6 * b.detach().abs().max()
    for k in ("exp_avg", "exp_avg_sq"):
        rb = ref.state[b][k]
        rb = torch.view_as_real(rb) if rb.is_complex() else rb
        assert (ours.state[a][k] - rb).abs().max() <= 1e-5 * rb.abs().max(), k


@pytest.mark.gpu
@pytest.mark.parametrize("momentum,dampening,wd,nesterov,gscale", [(0.0, 0.0, 0.0, False, 7.75), (0.9, 0.0, 0.0, False, 1.0),
                                                                   (0.9, 0.1, 0.01, False, 2.0),
                                                                   (0.8, 0.0, 0.001, True, 1.0)])
def test_sgd_matches_torch(cuda, momentum, dampening, wd, nesterov, gscale):
    """SGD (plain, momentum, dampening, weight decay, Nesterov) against torch.optim.SGD (foreach=False)."""
    shape, dtype = (8, 33, 41), torch.complex64
    p0 = torch.randn(shape, dtype=dtype, device=cuda, generator=torch.Generator(device=cuda).manual_seed(6))
    a = p0.clone().requires_grad_(True)
    b = p0.clone().requires_grad_(True)
    kw = dict(momentum=momentum, dampening=dampening, weight_decay=wd, nesterov=nesterov)
    ours = optim.SGD([{"params": [a], "lr": 1e-2, "grad_scale": gscale}], **kw)
    ref = torch.optim.SGD([b], lr=1e-2, foreach=False, **kw)
    _steps(ours, ref, a, b, shape, dtype, gscale, cuda)
    assert (a.detach() - b.detach()).abs().max() <= 1e-6 * b.detach().abs().max()
    if momentum:
        rb = torch.view_as_real(ref.state[b]["momentum_buffer"])
        assert (ours.state[a]["momentum_buffer"] - rb).abs().max() <= 1e-5 * rb.abs().max()


def test_scheduler_factory_matches_reference_defaults():
    """make_scheduler builds the reference's schedulers (optimizer_mixin.py:128-183) with its defaults."""
    from quantem_amd.optim import make_scheduler, step_scheduler

    def opt(lr=0.04):
        return torch.optim.SGD([torch.nn.Parameter(torch.zeros(3))], lr=lr)

    assert make_scheduler(opt(), None) is None and make_scheduler(opt(), {"type": "none"}) is None
    o = opt()
    c = make_scheduler(o, {"type": "cyclic"})
    assert isinstance(c, torch.optim.lr_scheduler.CyclicLR)
    assert c.base_lrs == [0.01] and c.max_lrs == [0.16] and c.total_size == 200 and c.mode == "triangular2"
    o = opt()
    p = make_scheduler(o, {"type": "Plateau"})
    assert isinstance(p, torch.optim.lr_scheduler.ReduceLROnPlateau)
    assert (p.factor, p.patience, p.threshold, p.cooldown, p.min_lrs) == (0.5, 10, 1e-3, 50, [0.002])
    step_scheduler(p, None)  # no loss: skipped
    assert p.last_epoch == 0
    e = make_scheduler(opt(), {"type": "exp"}, num_iter=50)
    assert e.gamma == pytest.approx(0.01 ** (1 / 50))
    assert make_scheduler(opt(), {"type": "gamma"}).gamma == 0.9
    assert make_scheduler(opt(), {"type": "exponential", "gamma": 0.5}).gamma == 0.5
    o = opt()
    lin = make_scheduler(o, {"type": "linear"}, num_iter=10)
    assert isinstance(lin, torch.optim.lr_scheduler.LinearLR) and lin.total_iters == 10
    assert o.param_groups[0]["lr"] == pytest.approx(0.004)  # starts at 0.1 x
    for _ in range(10):
        o.step()  # no gradients: a no-op, in the order torch expects
        step_scheduler(lin, 1.0)
    assert o.param_groups[0]["lr"] == pytest.approx(0.04)
    with pytest.raises(ValueError):
        make_scheduler(opt(), {"type": "cosine"})


def test_plateau_schedule_halves_after_patience():
    from quantem_amd.optim import make_scheduler, step_scheduler

    o = torch.optim.SGD([torch.nn.Parameter(torch.zeros(1))], lr=1.0)
    p = make_scheduler(o, {"type": "plateau", "patience": 2, "cooldown": 0})
    for loss in (1.0, 1.0, 1.0, 1.0):
        o.step()
        step_scheduler(p, loss)
    assert o.param_groups[0]["lr"] == pytest.approx(0.5)  # 3 epochs without improvement after the best
