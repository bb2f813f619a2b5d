"""Reconstruction loop around the fused kernels, mirroring quantem's ``Ptychography.reconstruct``.

The loop body is the reference's (``ptychography.py:208-245``): a fresh random permutation of the scan
positions each epoch (``SimpleBatcher``, ``ptycho_utils.py:16-117``), hard constraints on the object and
probe on the autograd graph (``object_models.py:289-339``, ``probe_models.py:452-525``), the batch loss
and its gradient (here one call of :func:`quantem_amd.ptychography.multislice_loss`), the object
gradient rescale by ``upsample_factor**2 / 2`` (``ptychography.py:330-332``) and an Adam step with the
reference's default learning rates (object 5e-3, probe 1e-3; ``object_models.py:44-48``,
``probe_models.py:52-54``); the rescale and the Adam step run as one kernel per parameter
(:class:`quantem_amd.optim.Adam`).  The object starts flat (``ObjectPixelated.from_uniform``,
``object_models.py:549-571, 653-657``) unless an initial object is given (the warm start).
"""

from __future__ import annotations

import math
import time
from math import ceil
from typing import Callable, Optional

import numpy as np
import torch

from .constraints import object_constraints
from .optim import make_optimizer, make_scheduler, step_scheduler
from .ptychography import analytical_gradients, multislice_loss


def _real_view(x: torch.Tensor) -> torch.Tensor:
    return torch.view_as_real(x) if x.is_complex() else x


def validation_split(indices: np.ndarray, val_ratio: float, val_mode: str, rng: np.random.Generator):
    """(train, validation) index arrays of the batcher's fixed split (``ptycho_utils.py:34-76``).

    ``val_ratio`` outside [0, 1) means no validation set; ``round(num * val_ratio)`` indices are held
    out.  "random" takes them from one permutation drawn from ``rng``.  "grid" takes every k-th index,
    k = round(1 / val_ratio), as validation for ratios up to 1/2; above 1/2 every k-th index,
    k = round(1 / (1 - val_ratio)), is the training set and the rest validation.  Either grid selection
    keeps at most ``round(num * val_ratio)`` of the strided indices.
    """
    empty = np.asarray([], dtype=int)
    if not 0.0 <= val_ratio < 1.0:
        val_ratio = 0.0
    n_val = int(round(len(indices) * val_ratio))
    if n_val == 0:
        return indices, empty
    if val_mode == "random":
        val = rng.permutation(indices)[:n_val]
        return np.setdiff1d(indices, val), val
    if val_ratio <= 0.5:
        val = indices[::max(1, int(round(1.0 / val_ratio)))][:n_val]
        return np.setdiff1d(indices, val), val
    train = indices[::max(1, int(round(1.0 / (1.0 - val_ratio))))][:n_val]
    return train, np.setdiff1d(indices, train)


class SimpleBatcher:
    """Random mini-batches of scan indices, a new permutation every epoch (ptycho_utils.py:16-117).

    An optional validation set, fixed for the batcher's lifetime, is either given explicitly
    (``train_indices`` with ``val_indices``) or split off by :func:`validation_split`; iteration
    covers the training indices only and :meth:`iter_val` the validation ones, in order.
    """

    def __init__(self, num: int, batch_size: Optional[int], shuffle: bool = True, rng=None,
                 val_ratio: float = 0.0, val_mode: str = "grid", train_indices=None, val_indices=None):
        self.indices = np.arange(num)
        self.batch_size = batch_size if batch_size is not None else num
        self.shuffle = shuffle
        if rng is None or isinstance(rng, (int, float, np.integer)):
            rng = np.random.default_rng(None if rng is None else int(rng))
        if not isinstance(rng, np.random.Generator):
            raise TypeError(f"rng should be a np.random.Generator or a seed, got {type(rng)}")
        self.rng = rng
        if train_indices is not None or val_indices is not None:
            if train_indices is None or val_indices is None:
                raise ValueError("Both train_indices and val_indices must be provided together.")
            self.train_indices = np.asarray(train_indices, dtype=int)
            self.val_indices = np.asarray(val_indices, dtype=int)
        else:
            self.train_indices, self.val_indices = validation_split(self.indices, val_ratio, val_mode, self.rng)

    def __iter__(self):
        order = self.rng.permutation(self.train_indices) if self.shuffle else self.train_indices
        for i in range(0, len(order), self.batch_size):
            yield order[i:i + self.batch_size]

    def __len__(self) -> int:
        return int(ceil(len(self.train_indices) / self.batch_size))

    def iter_val(self):
        v = self.val_indices
        return (v[i:i + self.batch_size] for i in range(0, len(v), self.batch_size))

    @property
    def has_validation(self) -> bool:
        return len(self.val_indices) > 0

    def val_len(self) -> int:
        return int(ceil(len(self.val_indices) / self.batch_size))


# the dataset's hard / soft constraint switches and their defaults (dataset_models.py:557-562)
DATASET_CONSTRAINTS = {
    "descan_tv_weight": 0.0,
    "descan_shifts_constant": False,
    "center_scan_positions": False,
    "clip_scan_positions": True,
}


# the object's constraint switches and their defaults (ObjectConstraints.DEFAULT_CONSTRAINTS,
# object_models.py:274-287): the hard part shapes the object every batch, the soft weights add losses
OBJECT_HARD_CONSTRAINTS = {
    "positivity": True,
    "fix_potential_baseline": False,
    "fix_potential_baseline_factor": 1.0,
    "identical_slices": False,
    "apply_fov_mask": False,
    "gaussian_sigma": None,  # pixels
    "butterworth_order": 4,
    "q_lowpass": None,  # 1/A
    "q_highpass": None,  # 1/A
}
OBJECT_SOFT_CONSTRAINTS = {"tv_weight_z": 0.0, "tv_weight_xy": 0.0, "surface_zero_weight": 0.0}
OBJECT_CONSTRAINTS = {**OBJECT_HARD_CONSTRAINTS, **OBJECT_SOFT_CONSTRAINTS}
# the probe's soft-constraint weight (probe_models.py:431-436)
PROBE_SOFT_CONSTRAINTS = {"tv_weight": 0.0}


def _tv(array: torch.Tensor, w_z: float, w_xy: float) -> torch.Tensor:
    """sum over axes with a positive weight of w * mean |diff| (axis 0 of a 3-D array takes w_z, the
    others w_xy), divided by the number of such axes (object_models.py:402-415)."""
    loss = torch.zeros((), dtype=torch.float32, device=array.device)
    n = 0
    for dim in range(array.ndim):
        w = w_z if (dim == 0 and array.ndim == 3) else w_xy
        if w > 0:
            n += 1
            loss = loss + w * torch.mean(torch.abs(array.diff(dim=dim)))
    return loss / n if n else loss


def object_soft_loss(obj: torch.Tensor, obj_type: str = "complex", tv_weight_z: float = 0.0,
                     tv_weight_xy: float = 0.0, surface_zero_weight: float = 0.0) -> torch.Tensor:
    """ObjectConstraints.apply_soft_constraints (object_models.py:341-451) of the hard-constrained object
    ((S, H, W) complex transmission, or the real potential for obj_type "potential"): total variation of
    the phase (and, for "complex", the amplitude), without the z term for one slice, plus the
    surface-zero term (first and last slice pulled to zero phase / unit amplitude, 3 slices or more)."""
    loss = torch.zeros((), dtype=torch.float32, device=obj.device)
    w_z = tv_weight_z if obj.shape[0] > 1 else 0.0
    if w_z > 0 or tv_weight_xy > 0:
        if obj.is_complex():
            loss = loss + _tv(obj.angle(), w_z, tv_weight_xy)
            if obj_type == "complex":
                loss = loss + _tv(obj.abs(), w_z, tv_weight_xy)
        else:
            loss = loss + _tv(obj, w_z, tv_weight_xy)
    if surface_zero_weight != 0 and obj.shape[0] >= 3:
        if obj.is_complex():
            ph = obj.angle().abs()
            if obj_type == "complex":
                amp = obj.abs()
                loss = loss + surface_zero_weight * (torch.mean(1.0 - amp[0]) + torch.mean(1.0 - amp[-1]))
            loss = loss + surface_zero_weight * (torch.mean(torch.abs(ph[0] - ph[0].mean())) +
                                                 torch.mean(torch.abs(ph[-1] - ph[-1].mean())))
        else:
            loss = loss + surface_zero_weight * (torch.mean(torch.abs(obj[0])) + torch.mean(torch.abs(obj[-1])))
    return loss


def probe_soft_loss(probe: torch.Tensor, tv_weight: float = 0.0) -> torch.Tensor:
    """ProbeConstraints.apply_soft_constraints (probe_models.py:441-467): tv_weight * (mean |diff| along
    columns + along rows) of the complex probe modes."""
    tv = torch.zeros((), dtype=torch.float32, device=probe.device)
    if tv_weight == 0:
        return tv
    for dim in (-1, -2):
        tv = tv + torch.mean(torch.abs(torch.diff(probe, dim=dim)))
    return tv_weight * tv


def gaussian_blur_2d(t: torch.Tensor, sigma: float) -> torch.Tensor:
    """Separable Gaussian blur of every (H, W) slice of an (S, H, W) tensor, real and imaginary parts
    alike: a normalised kernel of 2 ceil(3 sigma) + 1 taps, zero padding (object_models.py:437-479)."""
    k = int(2 * math.ceil(3 * sigma) + 1)
    if k % 2 == 0:
        k += 1
    ax = torch.arange(-k // 2 + 1.0, k // 2 + 1.0, device=t.device)
    g = torch.exp(-0.5 * (ax / sigma) ** 2)
    g = g / g.sum()
    kh, kv = g.view(1, 1, -1, 1), g.view(1, 1, 1, -1)

    def blur(x):
        x = torch.nn.functional.conv2d(x.unsqueeze(1), kh, padding=(k // 2, 0))
        return torch.nn.functional.conv2d(x, kv, padding=(0, k // 2)).squeeze(1)

    return torch.complex(blur(t.real), blur(t.imag)) if t.is_complex() else blur(t)


def butterworth_filter(t: torch.Tensor, sampling, q_lowpass=None, q_highpass=None, order: int = 4,
                       real_out: bool = False) -> torch.Tensor:
    """Butterworth low / high pass of every slice about its mean (object_models.py:481-521): envelope
    1 / (1 + (q/q_low)^(2 order)) times 1 - 1 / (1 + (q/q_high)^(2 order)) on |q| (1/A, fftfreq of the
    slice shape at ``sampling`` A/px); the real part is kept for a potential."""
    qx = torch.fft.fftfreq(t.shape[-2], sampling[0], device=t.device)
    qy = torch.fft.fftfreq(t.shape[-1], sampling[1], device=t.device)
    qya, qxa = torch.meshgrid(qy, qx, indexing="xy")
    qra = torch.sqrt(qxa ** 2 + qya ** 2)
    env = torch.ones_like(qra)
    if q_highpass:
        env = env * (1 - 1 / (1 + (qra / q_highpass) ** (2 * order)))
    if q_lowpass:
        env = env * (1 / (1 + (qra / q_lowpass) ** (2 * order)))
    mean = t.mean(dim=(-2, -1), keepdim=True)
    out = torch.fft.ifft2(torch.fft.fft2(t - mean) * env) + mean
    return out.real if real_out else out


def object_hard_constraints(obj: torch.Tensor, obj_type: str = "complex", constraints: Optional[dict] = None,
                            mask: Optional[torch.Tensor] = None, sampling=None) -> torch.Tensor:
    """ObjectConstraints.apply_hard_constraints (object_models.py:289-339) in torch, returning the object in
    its own type (a potential stays real):

    * "complex" / "pure_phase": amplitude clamp(|O|, 0, 1) (or 1), phase minus its mean; with
      ``apply_fov_mask`` and a mask, amp * mask * exp(i phase * mask);
    * "potential": minus a baseline (the minimum, or the mean below half the mask's maximum, times
      ``fix_potential_baseline_factor``, without gradient) when ``fix_potential_baseline``, then
      clamped at zero when ``positivity``;
    * then the FOV mask, a Gaussian blur (``gaussian_sigma`` px), a Butterworth filter (``q_lowpass`` /
      ``q_highpass``, needs ``sampling``) and, for ``identical_slices``, every slice set to the mean
      slice without gradient, in that order.

    The default constraint set of the complex types is also a HIP kernel pair
    (:func:`quantem_amd.constraints.object_constraints`), which the loop uses on the GPU.
    """
    c = dict(OBJECT_HARD_CONSTRAINTS)
    if constraints:
        c.update({k: v for k, v in constraints.items() if k in OBJECT_HARD_CONSTRAINTS})
    use_mask = mask is not None and c["apply_fov_mask"]
    if obj_type in ("complex", "pure_phase"):
        amp = torch.clamp(torch.abs(obj), 0.0, 1.0) if obj_type == "complex" else 1.0
        ang = obj.angle()
        phase = ang - ang.mean()
        out = amp * mask * torch.exp(1.0j * phase * mask) if use_mask else amp * torch.exp(1.0j * phase)
    elif obj_type == "potential":
        offset = 0.0
        if c["fix_potential_baseline"]:
            if mask is not None:
                background = mask < 0.5 * mask.max()
                offset = obj[..., background].mean() if background.any() else obj.min()
            else:
                offset = obj.min()
            offset = offset.detach() * c["fix_potential_baseline_factor"]
        out = torch.clamp(obj - offset, min=0.0) if c["positivity"] else obj - offset
    else:
        raise ValueError(f"unknown obj_type {obj_type}")
    return _object_filters(out, obj_type, c, mask, sampling)


def _object_filters(out: torch.Tensor, obj_type: str, c: dict, mask, sampling) -> torch.Tensor:
    """The steps of the hard constraints after the amplitude / phase or baseline step."""
    if c["apply_fov_mask"] and mask is not None:
        out = out * mask
    if c.get("gaussian_sigma") is not None:
        out = gaussian_blur_2d(out, c["gaussian_sigma"])
    if c["q_lowpass"] or c["q_highpass"]:
        if sampling is None:
            raise ValueError("q_lowpass / q_highpass need the object sampling (A/px)")
        out = butterworth_filter(out, sampling, c["q_lowpass"], c["q_highpass"], c["butterworth_order"],
                                 real_out=obj_type == "potential")
    if out.shape[0] > 1 and c["identical_slices"]:
        with torch.no_grad():
            out[:] = torch.mean(out, dim=0, keepdim=True)
    return out


def orthogonalize_probe(probe: torch.Tensor) -> torch.Tensor:
    """Gram-Schmidt keeping each mode's norm, modes sorted by power (probe_models.py:484-525), on the
    probe-orthogonalisation kernels (:func:`quantem_amd.constraints.orthogonalize_probe`)."""
    from .constraints import orthogonalize_probe as ortho

    return ortho(probe)


def center_probe(probe: torch.Tensor) -> torch.Tensor:
    """ProbeConstraints' centre-of-mass constraint (probe_models.py:469-481): each corner-centred mode is
    Fourier-shifted (ptycho_utils.py:128-176, fftfreq ramps) by minus the offset of the centre of mass of
    its fftshifted intensity from (N/2, N/2)."""
    nr, nc = probe.shape[-2:]
    inten = torch.fft.fftshift(probe.abs().square(), dim=(-2, -1))
    rows = torch.arange(nr, device=probe.device, dtype=inten.dtype)
    cols = torch.arange(nc, device=probe.device, dtype=inten.dtype)
    tot = inten.sum(dim=(-2, -1))
    d_r = (inten * rows[:, None]).sum(dim=(-2, -1)) / tot - nr // 2
    d_c = (inten * cols[None, :]).sum(dim=(-2, -1)) / tot - nc // 2
    kr = torch.fft.fftfreq(nr, d=1.0, device=probe.device).to(inten.dtype)
    kc = torch.fft.fftfreq(nc, d=1.0, device=probe.device).to(inten.dtype)
    ramp = (torch.exp(2.0j * math.pi * kr[None, :, None] * d_r[:, None, None]) *
            torch.exp(2.0j * math.pi * kc[None, None, :] * d_c[:, None, None]))
    return torch.fft.ifft2(torch.fft.fft2(probe) * ramp.to(probe.dtype))


def projected_phase(obj: torch.Tensor) -> torch.Tensor:
    """Sum over slices of the transmission phase."""
    return torch.sum(torch.angle(obj), dim=0)


def phase_nmse(est: np.ndarray, true: np.ndarray, crop: Optional[tuple] = None) -> float:
    """||(est - true) - mean(est - true)||^2 / ||true - mean(true)||^2 over the crop (constant phase removed)."""
    if crop is not None:
        (r0, r1), (c0, c1) = crop
        est = est[r0:r1, c0:c1]
        true = true[r0:r1, c0:c1]
    d = est - true
    d = d - d.mean()
    t = true - true.mean()
    return float(np.sum(d * d) / max(np.sum(t * t), 1e-300))


class Reconstruction:
    """Gradient-descent multislice reconstruction on one GPU (the reference's autograd=True mode)."""

    def __init__(self, positions, targets, propagators, probe, obj_shape, mean_intensity: float,
                 mask=None, obj_type: str = "complex", orthogonalize: bool = True, upsample_factor: float = 1.0,
                 obj_init: Optional[torch.Tensor] = None, descan=None, object_constraints: Optional[dict] = None,
                 fov_mask=None, sampling=None, center_probe: bool = False, probe_tilt=None,
                 propagator_params: Optional[dict] = None, device="cuda"):
        dev = torch.device(device)
        self.device = dev
        # probe tilt (mrad, row / col) and what the propagators are built from (keys sampling, energy,
        # thicknesses): with both the propagators follow the tilt, which the loop can learn
        self.probe_tilt = torch.nn.Parameter(torch.tensor([0.0, 0.0] if probe_tilt is None else list(probe_tilt),
                                                          dtype=torch.float32, device=dev), requires_grad=False)
        self.propagator_params = None if propagator_params is None else dict(propagator_params)
        self.positions = torch.as_tensor(positions, dtype=torch.float32, device=dev).clone().contiguous()
        self.targets = torch.as_tensor(targets, dtype=torch.float32, device=dev).contiguous()
        S = int(obj_shape[0])
        if propagators is None and self.propagator_params is not None and S > 1:
            with torch.no_grad():
                propagators = self.tilted_propagators()
        self.propagators = (torch.as_tensor(propagators, dtype=torch.complex64, device=dev).contiguous()
                            if S > 1 else None)
        self.mask = None if mask is None else torch.as_tensor(mask, dtype=torch.float32, device=dev)
        self.mean_intensity = float(mean_intensity)
        self.num_gpts = int(self.positions.shape[0])
        # descan shifts in pixels (P, 2); the forward model applies them only while they are learned
        self.descan = (torch.zeros((self.num_gpts, 2), dtype=torch.float32, device=dev) if descan is None else
                       torch.as_tensor(descan, dtype=torch.float32, device=dev).clone().contiguous())
        if self.positions.shape != (self.num_gpts, 2) or self.descan.shape != (self.num_gpts, 2):
            raise ValueError(f"positions and descan must be (P, 2), got {tuple(self.positions.shape)} and "
                             f"{tuple(self.descan.shape)}")
        self.obj_shape = tuple(int(x) for x in obj_shape)
        self.obj_type = obj_type
        # the object's constraints (hard and soft, the reference's keys), its FOV mask (H, W) and sampling
        self.object_constraints = dict(OBJECT_CONSTRAINTS)
        self.set_object_constraints(object_constraints)
        self.fov_mask = None if fov_mask is None else torch.as_tensor(fov_mask, dtype=torch.float32, device=dev)
        self.sampling = None if sampling is None else tuple(float(x) for x in sampling)
        self.orthogonalize = orthogonalize
        self.center_probe = center_probe
        self.upsample_factor = float(upsample_factor)
        # a potential object is real and starts at zero; the others are complex and start flat at 1
        odt = torch.float32 if obj_type == "potential" else torch.complex64
        if obj_init is None:
            obj0 = (torch.zeros if obj_type == "potential" else torch.ones)(
                tuple(int(x) for x in obj_shape), dtype=odt, device=dev)
        else:
            obj0 = torch.as_tensor(obj_init, dtype=odt, device=dev).clone()
        # complex parameters live in real (..., 2) storage, the view the optimizer updates
        self._obj = torch.nn.Parameter(_real_view(obj0).clone())
        self._probe = torch.nn.Parameter(
            _real_view(torch.as_tensor(probe, dtype=torch.complex64, device=dev)).clone())
        self.history: list = []
        self.val_history: list = []
        self.lr_history: dict = {}
        self.times: list = []

    def tilted_propagators(self) -> torch.Tensor:
        """The propagators at the current probe tilt (``probe_models.py:376-427``), differentiable in it."""
        from .ptychography import fresnel_propagators

        pp = self.propagator_params
        if pp is None:
            raise ValueError("learning the probe tilt needs propagator_params (sampling, energy, thicknesses)")
        return fresnel_propagators(tuple(int(x) for x in self.targets.shape[-2:]), float(pp["sampling"]),
                                   float(pp["energy"]),
                                   pp["thicknesses"], self.probe_tilt)

    @property
    def obj(self) -> torch.Tensor:
        return self._obj if self.obj_type == "potential" else torch.view_as_complex(self._obj)

    @property
    def probe(self) -> torch.Tensor:
        return torch.view_as_complex(self._probe)

    def set_object_constraints(self, constraints: Optional[dict]) -> None:
        """Updates the object's constraint switches / weights (keys of :data:`OBJECT_CONSTRAINTS`)."""
        if constraints:
            unknown = set(constraints) - set(OBJECT_CONSTRAINTS)
            if unknown:
                raise KeyError(f"unknown object constraints {sorted(unknown)}; known: {sorted(OBJECT_CONSTRAINTS)}")
            self.object_constraints.update(constraints)

    @property
    def obj_hard(self) -> torch.Tensor:
        """The object after its hard constraints, in its own type (``obj_model.obj`` of the reference)."""
        c = self.object_constraints
        if (self.device.type == "cuda" and self.obj_type in ("complex", "pure_phase")
                and not (c["apply_fov_mask"] and self.fov_mask is not None)):
            # the amplitude / phase step as the HIP kernel pair, the optional filters after it in torch
            return _object_filters(object_constraints(self.obj, self.obj_type), self.obj_type, c, self.fov_mask,
                                   self.sampling)
        return object_hard_constraints(self.obj, self.obj_type, c, self.fov_mask, self.sampling)

    @property
    def obj_constrained(self) -> torch.Tensor:
        """The transmission the forward model takes: the constrained object, exp(i V) for a potential."""
        o = self.obj_hard
        return torch.exp(1.0j * o) if self.obj_type == "potential" else o

    @property
    def probe_constrained(self) -> torch.Tensor:
        """ProbeConstraints.apply_hard_constraints (probe_models.py:452-459): Gram-Schmidt (default on),
        then centring (default off)."""
        p = orthogonalize_probe(self.probe) if (self.orthogonalize and self.probe.shape[0] > 1) else self.probe
        return center_probe(p) if self.center_probe else p

    def apply_dataset_constraints(self, constraints: dict) -> None:
        """The dataset's hard constraints on the scan positions (``dataset_models.py:597-613``), in place:
        clip to [0, object extent - 1] per axis, then (optionally) move the mean position to the
        object centre."""
        ext = torch.tensor(self.obj_shape[1:], dtype=torch.float32, device=self.device)
        with torch.no_grad():
            pos = self.positions
            if constraints["clip_scan_positions"]:
                pos.copy_(torch.clamp(pos, min=torch.zeros_like(ext), max=ext - 1))
            if constraints["center_scan_positions"]:
                pos.copy_(pos - pos.mean(dim=0, keepdim=True) + ext / 2)

    def descan_tv_loss(self, weight: float) -> torch.Tensor:
        """weight * (mean |diff d_r| + mean |diff d_c|) / 2 over the scan order (``dataset_models.py:579-587``)."""
        d = self.descan
        return weight * (torch.mean(torch.abs(d[:, 0].diff())) + torch.mean(torch.abs(d[:, 1].diff()))) / 2

    def reconstruct(self, num_iters: int, batch_size: int, obj_lr: float = 5e-3, probe_lr: float = 1e-3,
                    learn_probe: bool = True, loss_type: str = "l2_amplitude", seed: int = 0,
                    autograd: bool = True, segment_slices: int = 0, dataset_lr: Optional[float] = None,
                    learn_positions: bool = True, learn_descan: bool = True,
                    dataset_constraints: Optional[dict] = None, val_ratio: float = 0.0, val_mode: str = "grid",
                    schedulers: Optional[dict] = None, object_constraints: Optional[dict] = None,
                    probe_constraints: Optional[dict] = None, optimizers: Optional[dict] = None,
                    learn_probe_tilt: bool = False,
                    callback: Optional[Callable[[int, float, "Reconstruction"], None]] = None) -> list:
        """Runs `num_iters` epochs; returns the per-epoch mean batch loss (consistency + soft constraints).

        ``autograd=False`` selects the reference's analytical update (Fourier projection and slice
        replay, ``ptychography.py:339-372``) instead of the loss gradient; as in the reference, its
        gradients are assigned to the raw parameters without the upsample rescale.
        ``segment_slices`` k >= 2 keeps every k-th incident wave and recomputes the rest in the
        backward pass (less memory; see :func:`quantem_amd.ptychography.multislice_loss`).

        ``dataset_lr`` set gives the dataset an Adam optimizer over the scan positions and / or the
        descan shifts (``learn_positions``, ``learn_descan``; ``dataset_models.py:98-108``, default
        rate 1e-3 in the reference).  Descan shifts enter the forward model only then
        (``dataset_models.py:1495-1498``); ``dataset_constraints`` overrides :data:`DATASET_CONSTRAINTS`.
        Position and descan gradients come from the loss gradient, so they need ``autograd=True``.
        ``val_ratio`` > 0 holds out scan positions (:func:`validation_split`): they take no optimizer
        step, and their mean batch loss after each epoch goes to ``self.val_history``.

        Each model ("object", "probe", "dataset") has its own optimizer, as in the reference
        (``ptychography_opt.py:13-234``): Adam by default, or what ``optimizers`` names for it
        (``{"object": {"type": "adamw", "weight_decay": 1e-2}}``; types "adam", "adamw", "sgd" with their
        torch keyword arguments, ``optimizer_mixin.py:84-126``, on the HIP optimizer steps of
        :mod:`quantem_amd.optim`; the rates above are the defaults of ``lr``); ``schedulers`` maps a model name to the reference's scheduler
        parameters (:func:`quantem_amd.optim.make_scheduler`, built with ``num_iter = num_iters``),
        stepped once per epoch with the epoch's mean loss (``ptychography.py:278-279``).  The learning
        rate of every model after each epoch goes to ``self.lr_history``.

        ``learn_probe_tilt`` optimises the probe tilt with the probe's optimizer (the reference adds
        ``probe_tilt`` to the probe model's parameters, ``probe_models.py:300-303``): every batch the
        propagators are rebuilt from it (``ptychography_base.py:898-899``) and the loss gradient reaches it
        through the kernels' propagator gradient.  A tilt component that is exactly zero gets no gradient,
        as in the reference.

        ``object_constraints`` updates the object's constraints (hard switches and soft weights,
        :data:`OBJECT_CONSTRAINTS`; they persist on the instance), ``probe_constraints`` the probe's soft
        weight (:data:`PROBE_SOFT_CONSTRAINTS`).  The soft losses join every batch loss
        (``ptychography.py:231-232``), so with ``autograd=False`` they are reported but do not act, as in
        the reference.
        """
        def merged(defaults, given, what):
            out = dict(defaults)
            if given:
                unknown = set(given) - set(out)
                if unknown:
                    raise KeyError(f"unknown {what} constraints {sorted(unknown)}; known: {sorted(out)}")
                out.update(given)
            return out

        cons = merged(DATASET_CONSTRAINTS, dataset_constraints, "dataset")
        self.set_object_constraints(object_constraints)
        ocons = self.object_constraints
        pcons = merged(PROBE_SOFT_CONSTRAINTS, probe_constraints, "probe")
        obj_soft = any(ocons[k] != 0 for k in OBJECT_SOFT_CONSTRAINTS)
        learn_dset = dataset_lr is not None and (learn_positions or learn_descan)
        if learn_dset and not autograd:
            raise ValueError("learning positions or descan needs autograd=True (the analytical update gives "
                             "no position or descan gradient)")
        use_descan = learn_dset and learn_descan
        # the object-gradient rescale of the autograd path (ptychography.py:330-332) is applied inside
        # the optimizer step; the analytical update is not rescaled, as in the reference
        scale = self.upsample_factor ** 2 / 2 if autograd else 1.0
        opt_cfg = dict(optimizers or {})
        unknown = set(opt_cfg) - {"object", "probe", "dataset"}
        if unknown:
            raise KeyError(f"optimizers for unknown models {sorted(unknown)}; known: object, probe, dataset")
        opts = {"object": make_optimizer([{"params": [self._obj], "grad_scale": scale}], opt_cfg.get("object"),
                                         obj_lr)}
        if learn_probe_tilt and (self.propagators is None or self.propagator_params is None):
            raise ValueError("learning the probe tilt needs more than one slice and propagator_params")
        if learn_probe_tilt and not autograd:
            raise ValueError("learning the probe tilt needs autograd=True")
        probe_params = ([self._probe] if learn_probe else []) + ([self.probe_tilt] if learn_probe_tilt else [])
        if probe_params:
            opts["probe"] = make_optimizer([{"params": probe_params}], opt_cfg.get("probe"), probe_lr)
        self._probe.requires_grad_(learn_probe)
        self.probe_tilt.requires_grad_(learn_probe_tilt)
        self.tilt_history = []
        dset_params = []
        if learn_dset and learn_descan:
            dset_params.append(self.descan)
        if learn_dset and learn_positions:
            dset_params.append(self.positions)
        for t in (self.positions, self.descan):
            t.requires_grad_(any(t is q for q in dset_params))
        if dset_params:
            opts["dataset"] = make_optimizer([{"params": dset_params}], opt_cfg.get("dataset"), float(dataset_lr))
        sched_cfg = dict(schedulers or {})
        unknown = set(sched_cfg) - {"object", "probe", "dataset"}
        if unknown:
            raise KeyError(f"schedulers for unknown models {sorted(unknown)}; known: object, probe, dataset")
        scheds = {k: make_scheduler(opts[k], v, num_iters) for k, v in sched_cfg.items() if k in opts}
        for k in opts:
            self.lr_history.setdefault(k, [])
        batcher = SimpleBatcher(self.num_gpts, batch_size, rng=seed, val_ratio=val_ratio, val_mode=val_mode)
        tv_weight = float(cons["descan_tv_weight"]) if use_descan else 0.0

        def soft_loss(oh, pc):
            """the batch's soft-constraint loss (object, probe, descan), or None when every weight is 0"""
            terms = []
            if obj_soft:
                terms.append(object_soft_loss(oh, self.obj_type, ocons["tv_weight_z"], ocons["tv_weight_xy"],
                                              ocons["surface_zero_weight"]))
            if pcons["tv_weight"] != 0:
                terms.append(probe_soft_loss(pc, pcons["tv_weight"]))
            if tv_weight > 0:
                terms.append(self.descan_tv_loss(tv_weight))
            return sum(terms[1:], terms[0]) if terms else None

        def batch_inputs(idx):
            it = torch.from_numpy(idx).to(self.device)
            self.apply_dataset_constraints(cons)  # every batch, as the dataset's forward does
            desc = None
            if use_descan:
                desc = torch.zeros_like(self.descan) if cons["descan_shifts_constant"] else self.descan
                desc = desc[it]
            return it, self.positions[it], desc

        t0 = time.perf_counter()
        t_cb = 0.0  # time spent in callbacks, excluded from self.times
        try:
            for epoch in range(num_iters):
                total = torch.zeros((), dtype=torch.float32, device=self.device)
                for idx in batcher:
                    for o in opts.values():
                        o.zero_grad(set_to_none=True)
                    it, pos, desc = batch_inputs(idx)
                    if autograd:
                        oh, pc = self.obj_hard, self.probe_constrained
                        oc = torch.exp(1.0j * oh) if self.obj_type == "potential" else oh
                        props = self.tilted_propagators() if learn_probe_tilt else self.propagators
                        loss = multislice_loss(oc, pc, pos, props, self.targets[it], self.mask, desc,
                                               num_gpts=self.num_gpts, mean_intensity=self.mean_intensity,
                                               loss_type=loss_type, segment_slices=segment_slices)
                        soft = soft_loss(oh, pc)
                        if soft is not None:
                            loss = loss + soft
                        loss.backward()
                    else:
                        with torch.no_grad():
                            oh, pc = self.obj_hard, self.probe_constrained
                            oc = torch.exp(1.0j * oh) if self.obj_type == "potential" else oh
                            loss, og, pg = analytical_gradients(
                                oc, pc, pos, self.propagators, self.targets[it], self.mask, None,
                                num_gpts=self.num_gpts, mean_intensity=self.mean_intensity, loss_type=loss_type,
                                obj_type=self.obj_type)
                            soft = soft_loss(oh, pc)
                            if soft is not None:
                                loss = loss + soft
                        self._obj.grad = _real_view(og).contiguous()
                        if learn_probe:
                            self._probe.grad = _real_view(pg).contiguous()
                    for o in opts.values():
                        o.step()
                    total = total + loss.detach()
                if batcher.has_validation:
                    vtot = torch.zeros((), dtype=torch.float32, device=self.device)
                    with torch.no_grad():
                        for idx in batcher.iter_val():
                            it, pos, desc = batch_inputs(idx)
                            vtot = vtot + multislice_loss(
                                self.obj_constrained, self.probe_constrained, pos, self.propagators,
                                self.targets[it], self.mask, desc, num_gpts=self.num_gpts,
                                mean_intensity=self.mean_intensity, loss_type=loss_type)
                    self.val_history.append(float(vtot.item()) / batcher.val_len())
                mean_loss = float(total.item()) / len(batcher)  # one host sync per epoch
                self.history.append(mean_loss)
                if learn_probe_tilt:
                    with torch.no_grad():
                        self.propagators = self.tilted_propagators().contiguous()
                    self.tilt_history.append(self.probe_tilt.detach().cpu().tolist())
                for k, sc in scheds.items():
                    step_scheduler(sc, mean_loss)
                for k, o in opts.items():
                    self.lr_history[k].append(o.param_groups[0]["lr"])
                self.times.append(time.perf_counter() - t0 - t_cb)
                if callback is not None:
                    tc = time.perf_counter()
                    callback(epoch, mean_loss, self)
                    t_cb += time.perf_counter() - tc
        finally:
            self.positions.requires_grad_(False)
            self.descan.requires_grad_(False)
            self.probe_tilt.requires_grad_(False)
        return self.history


def scan_window(spec) -> tuple:
    """Rows / columns of the object covered by the scan positions (the NMSE crop)."""
    from .synth import scan_origin

    o = scan_origin(spec)
    ext = (np.asarray(spec.scan) - 1) * spec.scan_step
    return tuple((int(np.floor(o[i])), int(np.ceil(o[i] + ext[i])) + 1) for i in range(2))


def simulate_amplitudes(spec, prob: dict, device, seed: int = 0) -> tuple:
    """Measured amplitudes on the GPU: the forward kernel's intensities (batches of spec.batch) with
    Poisson counts at the probe's dose; returns (amplitudes (P, Nr, Nc) float32 on device, mean intensity)."""
    from .ptychography import forward_intensities

    dev = torch.device(device)
    obj = torch.from_numpy(prob["obj"]).to(dev)
    probe = torch.from_numpy(prob["probe"]).to(dev)
    pos = torch.from_numpy(prob["positions"]).to(dev)
    props = torch.from_numpy(prob["propagators"]).to(dev) if spec.num_slices > 1 else None
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed + 1)
    P = pos.shape[0]
    from .synth import _shape2

    amps = torch.empty((P,) + _shape2(spec.roi), dtype=torch.float32, device=dev)
    total = torch.zeros((), dtype=torch.float64, device=dev)
    for i in range(0, P, spec.batch):
        counts = torch.poisson(forward_intensities(obj, probe, pos[i:i + spec.batch], props).clamp_min(0),
                               generator=gen)
        total += counts.sum(dtype=torch.float64)
        amps[i:i + spec.batch] = torch.sqrt(counts)
    return amps, float(total.item()) / P


def main(argv=None) -> int:
    """Flat-start reconstruction of a named synthetic workload; prints per-epoch loss / NMSE and a JSON
    summary.  ``python -m quantem_amd.reconstruct --workload large30-1024px --epochs 50 --segment 6``."""
    import argparse
    import json
    import sys

    import scipy.ndimage

    from . import synth

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--workload", default="single-slice", choices=sorted(synth.WORKLOADS))
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=None, help="default: the workload's batch")
    ap.add_argument("--segment", type=int, default=0, help="segment_slices of the backward pass")
    ap.add_argument("--learn-probe", action="store_true", help="also optimise the probe (default: fixed)")
    ap.add_argument("--dataset-lr", type=float, default=None,
                    help="learn the scan positions and descan shifts at this Adam rate (default: fixed)")
    ap.add_argument("--val-ratio", type=float, default=0.0, help="fraction of positions held out for validation")
    ap.add_argument("--schedule", default=None, choices=["cyclic", "plateau", "exp", "linear"],
                    help="learning-rate schedule of the object and probe optimizers (reference defaults)")
    ap.add_argument("--obj-type", default="complex", choices=["complex", "pure_phase", "potential"],
                    help="the reference's object types (object_models.py:289-339)")
    ap.add_argument("--tv-z", type=float, default=0.0, help="object soft constraint tv_weight_z")
    ap.add_argument("--tv-xy", type=float, default=0.0, help="object soft constraint tv_weight_xy")
    ap.add_argument("--obj-lr", type=float, default=5e-3, help="object learning rate (the reference's 5e-3)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None, help="write the summary JSON here")
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("the reconstruction runs on a GPU", file=sys.stderr)
        return 1
    spec = synth.WORKLOADS[args.workload]
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    prob = synth.make_problem(spec, seed=args.seed)
    amps, mean_I = simulate_amplitudes(spec, prob, dev, args.seed)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    print(f"[reconstruct] {spec.name}: {amps.shape[0]} patterns simulated in {t_data:.1f} s", flush=True)
    true_phase = prob["true_phase"].sum(0)
    crop = scan_window(spec)
    rec = Reconstruction(prob["positions"], amps, prob["propagators"], prob["probe"],
                         (spec.num_slices,) + tuple(spec.obj_shape), mean_I, obj_type=args.obj_type,
                         upsample_factor=synth.spec_upsample_factor(spec),
                         object_constraints={"tv_weight_z": args.tv_z, "tv_weight_xy": args.tv_xy},
                         device=dev)
    nmse_hist, diag_hist = [], []
    (r0, r1), (c0, c1) = crop

    def report(epoch, loss, r):
        oc = r.obj_constrained.detach()
        est = projected_phase(oc).cpu().numpy()
        nmse_hist.append(phase_nmse(est, true_phase, crop))
        # the same error after a Gaussian low-pass (sigma 1.5 px, the atoms' width) of both phases:
        # separates the drift of weakly constrained high spatial frequencies from that of the atoms
        lp = lambda x: scipy.ndimage.gaussian_filter(x, 1.5, mode="wrap")  # noqa: E731
        lowpass_nmse = phase_nmse(lp(est), lp(true_phase), crop)
        # where the per-slice phases sit: values near +-pi wrap in the projected phase (a sum of angles)
        ang = torch.angle(oc[:, r0:r1, c0:c1])
        diag = {"lowpass_phase_nmse": lowpass_nmse, "max_abs_slice_phase": float(ang.abs().max()), "frac_abs_phase_gt_2_5": float((ang.abs() > 2.5).float().mean()),
                "mean_amplitude": float(oc[:, r0:r1, c0:c1].abs().mean()),
                "slice_phase_std": [round(float(x), 4) for x in ang.std(dim=(1, 2)).cpu()]}
        diag_hist.append(diag)
        print(f"[reconstruct] epoch {epoch + 1}: loss {loss:.5g}, phase NMSE {nmse_hist[-1]:.4f} "
              f"(low-pass {lowpass_nmse:.4f}), max |slice phase| {diag['max_abs_slice_phase']:.3f}, |phase| > 2.5 rad {diag['frac_abs_phase_gt_2_5']:.2e}, "
              f"mean amplitude {diag['mean_amplitude']:.4f}, {r.times[-1]:.1f} s", flush=True)

    torch.cuda.reset_peak_memory_stats(dev)
    rec.reconstruct(args.epochs, args.batch or spec.batch, obj_lr=args.obj_lr, learn_probe=args.learn_probe,
                    seed=args.seed,
                    segment_slices=args.segment, dataset_lr=args.dataset_lr, val_ratio=args.val_ratio,
                    schedulers=({"object": {"type": args.schedule}, "probe": {"type": args.schedule}}
                                if args.schedule else None),
                    callback=report)
    P = amps.shape[0]
    out = {"workload": spec.name, "patterns": P, "epochs": args.epochs, "batch": args.batch or spec.batch,
           "segment_slices": args.segment, "learn_probe": args.learn_probe, "schedule": args.schedule,
           "reconstruction_s": round(rec.times[-1], 3),
           "patterns_per_s": round(P * args.epochs / rec.times[-1], 1),
           "final_loss": rec.history[-1], "final_phase_nmse": nmse_hist[-1],
           "final_val_loss": rec.val_history[-1] if rec.val_history else None,
           "phase_nmse_per_epoch": [round(x, 5) for x in nmse_hist], "diagnostics_per_epoch": diag_hist,
           "obj_type": args.obj_type, "tv_weight_z": args.tv_z, "tv_weight_xy": args.tv_xy, "obj_lr": args.obj_lr,
           "true_slice_phase_std": [round(float(x), 4) for x in prob["true_phase"][:, r0:r1, c0:c1].std(axis=(1, 2))],
           "peak_hbm_bytes": int(torch.cuda.max_memory_allocated(dev)), "simulation_s": round(t_data, 2)}
    print(json.dumps(out), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
