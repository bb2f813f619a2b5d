"""Torch restatement of quantem's per-batch forward/loss, differentiated by torch autograd.

TEST INFRASTRUCTURE ONLY (checker and CPU baseline; never the measured product path).

It replays the reference's op sequence on CPU tensors (quantem, src/quantem/diffractive_imaging/):
  dset.forward: frac = pos - round(pos)                    dataset_models.py:1482-1499
  patch indices: (round(pos) + fftfreq offsets) % shape     dataset_models.py:514-539
  probe_model.forward: fourier_shift_expand(...).swapaxes   probe_models.py:718-720, ptycho_utils.py:128-176
  obj_model.forward: real/imag advanced-index gather        object_models.py:252-265
  forward_operator / overlap_projection / _propagate_array  ptychography_base.py:892-989 (with the torch.stack)
  DetectorPixelated.forward                                 detector_models.py:22-30
  error_estimate                                            ptychography_base.py:908-932
  loss.backward()                                           ptychography.py:328
In complex64 it is the CPU baseline of the benchmark; in complex128 it is an autograd gradient
oracle independent of the hand-derived adjoint in ptycho_oracle.py.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch


def _fftfreq(n: int, d: float, like: torch.Tensor) -> torch.Tensor:
    return torch.from_numpy(np.fft.fftfreq(n, d=d).astype(np.float32)).to(like.device)


def patch_indices(positions: torch.Tensor, n, obj_shape) -> torch.Tensor:
    nr, nc = (n, n) if isinstance(n, int) else n
    r0 = torch.round(positions[:, 0]).to(torch.int32)
    c0 = torch.round(positions[:, 1]).to(torch.int32)
    x_ind = torch.fft.fftfreq(nr, d=1 / nr).to(positions.device)
    y_ind = torch.fft.fftfreq(nc, d=1 / nc).to(positions.device)
    row = (r0[:, None, None] + x_ind[None, :, None]) % obj_shape[-2]
    col = (c0[:, None, None] + y_ind[None, None, :]) % obj_shape[-1]
    return (row * obj_shape[-1] + col).to(torch.int32)


def fourier_translation_operator(positions: torch.Tensor, shape, dtype) -> torch.Tensor:
    nr, nc = shape[-2:]
    r = positions[..., 0][:, None, None]
    c = positions[..., 1][:, None, None]
    kr = _fftfreq(nr, 1.0, positions).to(positions.dtype)
    kc = _fftfreq(nc, 1.0, positions).to(positions.dtype)
    ramp = torch.exp(-2.0j * np.pi * kr[None, :, None] * r) * torch.exp(-2.0j * np.pi * kc[None, None, :] * c)
    for _ in range(len(shape) - 2):
        ramp = ramp[:, None, ...]
    return ramp.to(dtype)


def fourier_shift_expand(array: torch.Tensor, positions: torch.Tensor) -> torch.Tensor:
    phase = fourier_translation_operator(positions, array.shape, array.dtype)
    return torch.fft.ifft2(torch.fft.fft2(array) * phase)


def batch_loss(obj, probe, positions, propagators, targets, mask, num_gpts, mean_intensity,
               loss_type="l2_amplitude", descan: Optional[torch.Tensor] = None):
    """Returns (loss, pred_intensities) following the reference op order."""
    n = tuple(probe.shape[-2:])  # roi_shape (Nr, Nc)
    frac = positions - torch.round(positions)
    with torch.no_grad():
        idx = patch_indices(positions.detach(), n, obj.shape)
    shifted = fourier_shift_expand(probe, frac).swapaxes(0, 1)  # (M, B, n, n)
    obj_flat = obj.reshape(obj.shape[0], -1)
    patches = torch.complex(obj_flat.real[:, idx], obj_flat.imag[:, idx])  # (S, B, n, n)
    propagated = [shifted]
    overlap = patches[0] * shifted
    for s in range(1, obj.shape[0]):
        p = torch.fft.ifft2(torch.fft.fft2(overlap) * propagators[s - 1])
        overlap = patches[s] * p
        propagated.append(p)
    torch.stack(propagated, dim=0)  # the reference materialises this stack every batch
    if descan is not None:
        overlap = overlap * fourier_translation_operator(descan, n, overlap.dtype)[None]
    exit_fft = torch.fft.fft2(overlap, norm="ortho")
    intens = torch.fft.fftshift(torch.sum(torch.abs(exit_fft) ** 2, dim=0), dim=(-2, -1))
    if "amplitude" in loss_type:
        preds = torch.sqrt(intens + 1e-9)
    else:
        preds = intens
    m = torch.ones(n, dtype=preds.dtype) if mask is None else mask
    diff = preds * m - targets * m
    B = diff.shape[0]
    if "l1" in loss_type:
        err = torch.sum(torch.abs(diff)) / (B / num_gpts)
    elif "l2" in loss_type:
        err = torch.sum(torch.abs(diff) ** 2) / (B / num_gpts)
    elif loss_type == "poisson":
        err = torch.sum(preds - targets * torch.log(preds + 1e-6))
    else:
        raise ValueError(f"Unknown loss type {loss_type}")
    return err / mean_intensity, intens


def loss_and_grads(obj, probe, positions, propagators, targets, mask, num_gpts, mean_intensity,
                   loss_type="l2_amplitude", descan=None, dtype=torch.complex128, grad_positions=True):
    """Autograd gradients (torch convention) of the batch loss; numpy in, numpy out."""
    rdt = torch.float64 if dtype == torch.complex128 else torch.float32
    o = torch.tensor(np.asarray(obj), dtype=dtype, requires_grad=True)
    p = torch.tensor(np.asarray(probe), dtype=dtype, requires_grad=True)
    pos = torch.tensor(np.asarray(positions), dtype=rdt, requires_grad=grad_positions)
    props = None if propagators is None or len(propagators) == 0 else \
        torch.tensor(np.asarray(propagators), dtype=dtype, requires_grad=True)
    t = torch.tensor(np.asarray(targets), dtype=rdt)
    m = None if mask is None else torch.tensor(np.asarray(mask), dtype=rdt)
    d = None if descan is None else torch.tensor(np.asarray(descan), dtype=rdt, requires_grad=True)
    loss, intens = batch_loss(o, p, pos, props, t, m, num_gpts, mean_intensity, loss_type, d)
    loss.backward()
    out = {"loss": loss.item(), "I": intens.detach().numpy(), "obj": o.grad.numpy(), "probe": p.grad.numpy()}
    if grad_positions:
        out["positions"] = pos.grad.numpy()
    if d is not None:
        out["descan"] = d.grad.numpy()
    if props is not None:
        out["propagators"] = props.grad.numpy()
    return out


def orthogonalize_probe_mgs(probe: torch.Tensor) -> torch.Tensor:
    """The reference's probe orthogonalisation (probe_models.py:484-525) in torch: modified Gram-Schmidt,
    each mode rescaled to its input norm, sorted by intensity, differentiable by autograd."""
    norms = torch.sqrt(torch.sum(probe.real.square() + probe.imag.square(), dim=(-2, -1), keepdim=True))
    out = []
    for i in range(probe.shape[0]):
        p = probe[i]
        for q in out:
            p = p - torch.sum(q.conj() * p) * q
        n = torch.sqrt(torch.sum(p.real.square() + p.imag.square())).clamp_min(1e-12)
        out.append(p / n)
    out = torch.stack(out) * norms.view(-1, 1, 1)
    order = torch.argsort(torch.sum(torch.abs(out).square(), dim=(-2, -1)), descending=True)
    return torch.complex(out.real[order], out.imag[order])
