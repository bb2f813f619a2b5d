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
