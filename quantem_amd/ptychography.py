"""Drop-in forward/loss/gradient for quantem's multislice mixed-probe ptychography.

The reference runs, per mini-batch (quantem ``ptychography.py:213-245``)::

    patch_indices, _, frac, descan = dset.forward(batch_indices, obj_padding_px)  # dataset_models.py:1482
    shifted_probes = probe_model.forward(frac)                                     # probe_models.py:718
    obj_patches = obj_model.forward(patch_indices)                                 # object_models.py:678
    _, overlap = forward_operator(obj_patches, shifted_probes, descan)             # ptychography_base.py:892
    pred_intensities = detector_model.forward(overlap)                             # detector_models.py:22
    loss, targets = error_estimate(pred_intensities, batch_indices, loss_type)     # ptychography_base.py:908
    loss.backward()                                                                # ptychography.py:328

:func:`multislice_loss` replaces everything from ``probe_model.forward`` to ``loss.backward()``
with one ``torch.autograd.Function`` whose forward and backward are the HIP kernels of
``libqem_ptycho.so``.  It consumes the tensors the reference's models expose (``obj_model.obj``
after hard constraints, ``probe_model.probe`` after orthogonalisation, the dataset's scan
positions in pixels, the propagators, the gathered targets, the detector mask and the descan
shifts), so the caller's loop, constraints, optimizers and gradient rescaling stay unchanged.
"""

from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _native as nat

__all__ = [
    "multislice_loss",
    "forward_intensities",
    "MultisliceLoss",
    "fft2",
    "ifft2",
    "workspace_bytes",
    "fresnel_propagators",
]


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _need(t: torch.Tensor, name: str, dtype: torch.dtype, shape: tuple, device: torch.device) -> torch.Tensor:
    """Validates like quantem's validate_tensor (core/utils/validators.py:478-495)."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise TypeError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if t.device != device:
        raise ValueError(f"{name} must be on {device}, got {t.device}")
    return t.contiguous()


def _desc(S, M, B, N, H, W, num_gpts, loss_type, mean_intensity, flags, segment_slices=0) -> nat.QemDesc:
    """Descriptor for patterns of shape N = n or (Nr, Nc)."""
    if loss_type not in nat.LOSS_TYPES:
        raise ValueError(f"Unknown loss type {loss_type}, should be one of {sorted(nat.LOSS_TYPES)}")
    nr, nc = (N, N) if isinstance(N, int) else (int(N[0]), int(N[1]))
    return nat.QemDesc(
        num_slices=S,
        num_modes=M,
        batch=B,
        roi=nr,
        roi_cols=nc,
        obj_rows=H,
        obj_cols=W,
        num_gpts=int(num_gpts),
        loss_type=nat.LOSS_TYPES[loss_type],
        mean_intensity=float(mean_intensity),
        flags=flags,
        segment_slices=int(segment_slices),
    )


def workspace_bytes(desc: nat.QemDesc) -> int:
    lib = nat.load()
    out = ctypes.c_size_t(0)
    nat.check(lib.qem_ptycho_workspace_bytes(ctypes.byref(desc), ctypes.byref(out)), "workspace query")
    return int(out.value)


_ONES_MASKS: dict = {}


def _ones_mask(shape: tuple, device: torch.device) -> torch.Tensor:
    """The default detector mask (all ones, ``dataset_models.py:177-187``), one read-only tensor per
    (shape, device), so that a batch without a mask costs no fill kernel."""
    key = (tuple(shape), device.type, device.index)
    m = _ONES_MASKS.get(key)
    if m is None:
        m = torch.ones(tuple(shape), dtype=torch.float32, device=device)
        _ONES_MASKS[key] = m
    return m


class MultisliceLoss(torch.autograd.Function):
    """Loss of one batch; backward runs the hand-written adjoint.

    Inputs follow the layouts of the reference tensors (see ``include/qem_ptycho.h``).  Gradients
    are returned for ``obj``, ``probe``, ``positions``, ``descan`` and ``propagators`` (the last three
    only when they require grad: propagators do when the caller builds them from a learned probe tilt,
    ``ptychography_base.py:898-899``); targets and mask get ``None``.  The backward pass consumes the workspace of its forward: differentiating the
    same graph a second time (``retain_graph=True``) raises ``RuntimeError`` instead of returning
    wrong gradients, and the backward itself is not differentiable.
    """

    @staticmethod
    def forward(ctx, obj, probe, positions, propagators, targets, mask, descan, num_gpts, mean_intensity, loss_type,
                return_intensities, grad_enabled, segment_slices, full_propagators=False):
        device = obj.device
        if device.type != "cuda":
            raise ValueError(f"obj must be on a GPU device, got {device}")
        if obj.dim() != 3:
            raise ValueError(f"obj must have shape (S, H, W), got {tuple(obj.shape)}")
        if probe.dim() != 3:
            raise ValueError(f"probe must have shape (M, Nr, Nc), got {tuple(probe.shape)}")
        S, H, W = obj.shape
        M, NR, NC = probe.shape
        if not nat.roi_supported(NR, NC):
            raise ValueError(f"roi_shape ({NR}, {NC}) not supported: square {nat.SUPPORTED_ROI}, or "
                             f"2..{nat.GENERIC_ROI_MAX} per axis with prime factors 2, 3, 5 and 7 only")
        roi = (NR, NC)
        if positions.dim() != 2:
            raise ValueError(f"positions must have shape (B, 2), got {tuple(positions.shape)}")
        B = positions.shape[0]
        if B < 1:
            raise ValueError("batch must hold at least one scan position")
        obj_c = _need(obj, "obj", torch.complex64, (S, H, W), device)
        probe_c = _need(probe, "probe", torch.complex64, (M, NR, NC), device)
        pos_c = _need(positions, "positions", torch.float32, (B, 2), device)
        if S > 1:
            prop_c = _need(propagators, "propagators", torch.complex64, (S - 1, NR, NC), device)
        else:
            prop_c = None
        tgt_c = _need(targets, "targets", torch.float32, (B, NR, NC), device)
        if mask is None:
            mask_c = _ones_mask(roi, device)
        else:
            mask_c = _need(mask, "mask", torch.float32, roi, device)
        desc_c = None if descan is None else _need(descan, "descan", torch.float32, (B, 2), device)

        # grad mode is always off inside Function.forward: the caller's mode arrives as grad_enabled
        grad_prop = bool(grad_enabled) and prop_c is not None and ctx.needs_input_grad[3]
        need_grad = bool(grad_enabled) and (any(ctx.needs_input_grad[i] for i in (0, 1, 2, 6)) or grad_prop)
        flags = nat.QEM_FLAG_FULL_PROPAGATORS if full_propagators else 0
        if need_grad:
            flags |= nat.QEM_FLAG_STORE_WAVES
            if ctx.needs_input_grad[2]:
                flags |= nat.QEM_FLAG_GRAD_POSITIONS
            if desc_c is not None and ctx.needs_input_grad[6]:
                flags |= nat.QEM_FLAG_GRAD_DESCAN
            if grad_prop:
                flags |= nat.QEM_FLAG_GRAD_PROPAGATORS
        d = _desc(S, M, B, roi, H, W, num_gpts, loss_type, mean_intensity, flags, segment_slices)
        ws = torch.empty(workspace_bytes(d), dtype=torch.uint8, device=device)
        loss = torch.empty((), dtype=torch.float32, device=device)
        intens = torch.empty((B, NR, NC), dtype=torch.float32, device=device) if return_intensities else None
        lib = nat.load()
        rc = lib.qem_ptycho_forward_loss(
            ctypes.byref(d), _ptr(obj_c), _ptr(probe_c), _ptr(pos_c), _ptr(prop_c), _ptr(tgt_c), _ptr(mask_c),
            _ptr(desc_c), _ptr(loss), _ptr(intens), _ptr(ws), ws.numel(), _stream(device))
        nat.check(rc, "qem_ptycho_forward_loss")
        if need_grad:
            ctx.save_for_backward(obj_c, probe_c, pos_c, prop_c, desc_c, ws)
            ctx.desc = d
        ctx.consumed = False
        if intens is None:
            intens = torch.empty((0,), dtype=torch.float32, device=device)
        ctx.mark_non_differentiable(intens)
        return loss, intens

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_loss, grad_intens):
        # The adjoint overwrites the workspace the forward filled (gradient patches over the stored
        # waves, probe-gradient spectra over the far field), so a graph kept with retain_graph=True
        # can be differentiated once only.
        if ctx.consumed:
            raise RuntimeError("multislice_loss: the forward's workspace was consumed by an earlier backward "
                               "pass; run the forward again before differentiating a second time")
        ctx.consumed = True
        obj_c, probe_c, pos_c, prop_c, desc_c, ws = ctx.saved_tensors
        d = ctx.desc
        device = obj_c.device
        gl = grad_loss.detach().to(device=device, dtype=torch.float32).contiguous()
        g_obj = torch.empty_like(obj_c)
        g_probe = torch.empty_like(probe_c)
        g_pos = torch.empty_like(pos_c) if (d.flags & nat.QEM_FLAG_GRAD_POSITIONS) else None
        g_desc = torch.empty_like(desc_c) if (d.flags & nat.QEM_FLAG_GRAD_DESCAN) else None
        lib = nat.load()
        rc = lib.qem_ptycho_adjoint(
            ctypes.byref(d), _ptr(obj_c), _ptr(probe_c), _ptr(pos_c), _ptr(prop_c), _ptr(desc_c), _ptr(gl),
            _ptr(g_obj), _ptr(g_probe), _ptr(g_pos), _ptr(g_desc), _ptr(ws), ws.numel(), _stream(device))
        nat.check(rc, "qem_ptycho_adjoint")
        g_prop = None
        if d.flags & nat.QEM_FLAG_GRAD_PROPAGATORS:
            g_prop = torch.empty_like(prop_c)
            rc = lib.qem_ptycho_propagator_gradient(ctypes.byref(d), _ptr(g_prop), _ptr(ws), ws.numel(),
                                                    _stream(device))
            nat.check(rc, "qem_ptycho_propagator_gradient")
        return (g_obj if ctx.needs_input_grad[0] else None,
                g_probe if ctx.needs_input_grad[1] else None,
                g_pos, g_prop, None, None, g_desc, None, None, None, None, None, None, None)


def multislice_loss(
    obj: torch.Tensor,
    probe: torch.Tensor,
    positions: torch.Tensor,
    propagators: Optional[torch.Tensor],
    targets: torch.Tensor,
    mask: Optional[torch.Tensor] = None,
    descan: Optional[torch.Tensor] = None,
    *,
    num_gpts: int,
    mean_intensity: float,
    loss_type: str = "l2_amplitude",
    return_intensities: bool = False,
    segment_slices: int = 0,
    full_propagators: bool = False,
):
    """Batch loss of quantem's forward model (``ptychography.py:220-229``), differentiable.

    Args:
        obj: (S, H, W) complex64 transmission after hard constraints (``obj_model.obj``; a real
            potential V must be passed as ``torch.exp(1j * V)``, ``object_models.py:253-254``).
        probe: (M, Nr, Nc) complex64 probe modes (``probe_model.probe``), corner-centred; (Nr, Nc) is the
            dataset's roi_shape: square 32, 64, 128 or 256 (fixed-size kernels) or any 2..512 per axis
            with prime factors 2, 3, 5, 7 (generic-size kernels).
        positions: (B, 2) float32 scan positions in pixels (``dset.scan_positions_px[batch]``).
        propagators: (S-1, Nr, Nc) complex64 (``ptycho.propagators``), or None for one slice.
        targets: (B, Nr, Nc) float32 fftshifted targets (``dset.targets[batch]``).
        mask: (Nr, Nc) float32 detector mask (``dset.detector_mask``), default ones.
        descan: (B, 2) float32 descan shifts or None (``dataset_models.py:1495-1498``).
        num_gpts: total number of scan positions of the dataset.
        mean_intensity: ``dset.mean_diffraction_intensity``.
        loss_type: one of l2_amplitude, l1_amplitude, l2_intensity, l1_intensity, poisson.
        return_intensities: also return the (B, Nr, Nc) fftshifted predicted intensities.
        segment_slices: 0 keeps every slice's incident wave for the backward pass; k in [2, S) keeps
            only every k-th one and recomputes the others in the backward pass, one k-slice segment
            at a time (less memory, one extra propagation per slice).
        full_propagators: 256 x 256 patterns (and the 128 x 128 adjoint without descan or
            propagator gradient) apply a propagator slice that factors into a row and a column term
            (the reference's Fresnel propagators, tilted or not) as that product, after checking on
            the GPU that it reproduces every entry within 1e-6 of the largest; True applies every
            entry as given (``QEM_FLAG_FULL_PROPAGATORS``).

    Returns:
        loss (0-d float32 tensor), and intensities if requested.
    """
    if loss_type not in nat.LOSS_TYPES:
        raise ValueError(f"Unknown loss type {loss_type}, should be one of {sorted(nat.LOSS_TYPES)}")
    if int(segment_slices) < 0:
        raise ValueError(f"segment_slices must be >= 0, got {segment_slices}")
    loss, intens = MultisliceLoss.apply(obj, probe, positions, propagators, targets, mask, descan, num_gpts,
                                        mean_intensity, loss_type, return_intensities, torch.is_grad_enabled(),
                                        int(segment_slices), bool(full_propagators))
    return (loss, intens) if return_intensities else loss


def forward_intensities(
    obj: torch.Tensor,
    probe: torch.Tensor,
    positions: torch.Tensor,
    propagators: Optional[torch.Tensor],
    descan: Optional[torch.Tensor] = None,
) -> torch.Tensor:
    """Predicted fftshifted intensities (B, Nr, Nc) of the forward model (no gradient)."""
    B = positions.shape[0]
    tgt = torch.zeros((B,) + tuple(probe.shape[-2:]), dtype=torch.float32, device=obj.device)
    with torch.no_grad():
        _, intens = multislice_loss(obj, probe, positions, propagators, tgt, None, descan, num_gpts=B,
                                    mean_intensity=1.0, return_intensities=True)
    return intens


def fresnel_propagators(roi, sampling: float, energy: float, thicknesses, tilt_mrad: torch.Tensor,
                        device=None) -> torch.Tensor:
    """(S-1, Nr, Nc) complex64 Fresnel propagators with a probe tilt, differentiable in ``tilt_mrad``
    (``roi`` = N or (Nr, Nc)).

    The reference's ``ProbeBase._compute_propagator_arrays`` (``probe_models.py:376-427``), op for op:
    exp(-i pi lambda dz k^2) times, for a tilt component that is not exactly zero,
    exp(-2 pi i dz tan(theta / 1e3) k) along that axis; a zero component contributes no factor, so (as
    in the reference) it receives no gradient.  ``tilt_mrad`` is a (2,) float32 tensor (row, col) in mrad.
    """
    from .synth import electron_wavelength_angstrom

    dev = tilt_mrad.device if device is None else torch.device(device)
    nr, nc = (roi, roi) if isinstance(roi, int) else (int(roi[0]), int(roi[1]))
    kr = torch.fft.fftfreq(nr, sampling, device=dev)
    kc = torch.fft.fftfreq(nc, sampling, device=dev)
    k2 = (kr[:, None] ** 2 + kc[None] ** 2).to(torch.complex64)
    wavelength = electron_wavelength_angstrom(energy)
    dz = torch.tensor(list(thicknesses), device=dev, dtype=k2.dtype)
    props = torch.exp((-1.0j * torch.pi * wavelength * dz[:, None, None]) * k2)
    theta_r, theta_c = tilt_mrad[0], tilt_mrad[1]
    if theta_r != 0:
        kr_term = 1.0j * (-2 * torch.pi * dz[:, None, None] * torch.tan(theta_r / 1e3))
        props = props * torch.exp(kr_term * kr[None, :, None])
    if theta_c != 0:
        kc_term = 1.0j * (-2 * torch.pi * dz[:, None, None] * torch.tan(theta_c / 1e3))
        props = props * torch.exp(kc_term * kc[None, None, :])
    return props


def _fft2(x: torch.Tensor, inverse: bool) -> torch.Tensor:
    if x.dtype != torch.complex64:
        raise TypeError(f"fft2 needs complex64, got {x.dtype}")
    if x.dim() < 2 or x.shape[-1] != x.shape[-2] or x.shape[-1] not in nat.FFT_ROI:
        raise ValueError(f"fft2 needs (..., N, N) with N in {nat.FFT_ROI}, got {tuple(x.shape)}")
    if x.device.type != "cuda":
        raise ValueError(f"fft2 input must be on a GPU device, got {x.device}")
    xc = x.contiguous()
    out = torch.empty_like(xc)
    n = x.shape[-1]
    count = xc.numel() // (n * n)
    rc = nat.load().qem_fft2_c64(n, count, 1 if inverse else 0, _ptr(xc), _ptr(out), _stream(x.device))
    nat.check(rc, "qem_fft2_c64")
    return out


def fft2(x: torch.Tensor) -> torch.Tensor:
    """torch.fft.fft2 (default norm) over the last two dims, N in {32, 64, 128}."""
    return _fft2(x, False)


def ifft2(x: torch.Tensor) -> torch.Tensor:
    """torch.fft.ifft2 (default norm, 1/N^2) over the last two dims, N in {32, 64, 128}."""
    return _fft2(x, True)


def analytical_gradients(
    obj: torch.Tensor,
    probe: torch.Tensor,
    positions: torch.Tensor,
    propagators: Optional[torch.Tensor],
    targets: torch.Tensor,
    mask: Optional[torch.Tensor] = None,
    descan: Optional[torch.Tensor] = None,
    *,
    num_gpts: int,
    mean_intensity: float,
    loss_type: str = "l2_amplitude",
    obj_type: str = "complex",
):
    """The reference's autograd=False update for one batch (ptychography.py:339-372).

    Returns ``(loss, obj_grad, probe_grad)`` with the values the reference assigns to
    ``_obj.grad`` (``object_models.py:686-725``: minus the slice-replayed update divided by the summed
    per-mode maximum probe overlap) and ``_probe.grad`` (``probe_models.py:774-783``: minus the batch sum
    of the probe-plane gradient over the maximum patch energy and its ``ortho_norm``).  ``obj`` is the
    transmission; for ``obj_type="potential"`` the update is ``real(-i conj(O) ...)`` as in the reference.
    """
    device = obj.device
    S, H, W = obj.shape
    M, N, _ = probe.shape
    B = positions.shape[0]
    obj_c = _need(obj.detach(), "obj", torch.complex64, (S, H, W), device)
    probe_c = _need(probe.detach(), "probe", torch.complex64, (M, N, N), device)
    pos_c = _need(positions.detach(), "positions", torch.float32, (B, 2), device)
    prop_c = _need(propagators, "propagators", torch.complex64, (S - 1, N, N), device) if S > 1 else None
    tgt_c = _need(targets, "targets", torch.float32, (B, N, N), device)
    mask_c = _ones_mask((N, N), device) if mask is None else \
        _need(mask, "mask", torch.float32, (N, N), device)
    desc_c = None if descan is None else _need(descan.detach(), "descan", torch.float32, (B, 2), device)
    d = _desc(S, M, B, N, H, W, num_gpts, loss_type, mean_intensity, nat.QEM_FLAG_STORE_WAVES)
    ws = torch.empty(workspace_bytes(d), dtype=torch.uint8, device=device)
    loss = torch.empty((), dtype=torch.float32, device=device)
    lib = nat.load()
    st = _stream(device)
    nat.check(lib.qem_ptycho_forward_loss(ctypes.byref(d), _ptr(obj_c), _ptr(probe_c), _ptr(pos_c), _ptr(prop_c),
                                          _ptr(tgt_c), _ptr(mask_c), _ptr(desc_c), _ptr(loss), None, _ptr(ws),
                                          ws.numel(), st), "qem_ptycho_forward_loss")
    obj_upd = torch.empty_like(obj_c)
    probe_upd = torch.empty_like(probe_c)
    overlap = torch.empty((S, M, H, W), dtype=torch.float32, device=device)
    energy = torch.empty((S, B), dtype=torch.float32, device=device)
    nat.check(lib.qem_ptycho_projection_update(ctypes.byref(d), _ptr(obj_c), _ptr(probe_c), _ptr(pos_c),
                                               _ptr(prop_c), _ptr(tgt_c), _ptr(obj_upd), _ptr(probe_upd),
                                               _ptr(overlap), _ptr(energy), _ptr(ws), ws.numel(), st),
              "qem_ptycho_projection_update")
    probe_norm = overlap.amax(dim=(-2, -1)).sum(dim=1)  # (S,) sum over modes of the max overlap
    if obj_type == "potential":
        upd = torch.real(-1j * torch.conj(obj_c) * obj_upd)
    else:
        upd = obj_upd
    obj_grad = -upd / probe_norm.view(S, 1, 1)
    ortho_norm = 2.0 * N if M == 1 else 1.0 / (2.0 * N)
    probe_grad = -probe_upd / energy.max() / ortho_norm
    return loss, obj_grad, probe_grad
