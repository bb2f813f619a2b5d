"""FP64 CPU oracle of quantem's multislice mixed-probe ptychography forward model, loss and gradient.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import this module, and only as the checker.  The product path (quantem_amd) never calls it.

Each function restates the cited reference lines (quantem, src/quantem/diffractive_imaging/)
in float64/complex128 numpy.  How each one is pinned (independent checks in tests/test_oracle.py):

  patch_rows_cols        brute force vs the reference test's numpy recipe, wrap-around cases
  gather_patches         np.take on tiny objects
  translation_ramp       closed form; integer shift == np.roll; zero shift == identity
  shift_probes           integer fractional shift == np.roll; unitarity
  multislice             S=1 == the reference test's single-slice recipe; O == 1 => one Fresnel
                         propagation by the summed thickness; energy conservation (pure phase)
  far_field_intensities  Parseval; mode-sum linearity; fftshift placement
  batch_loss             hand-evaluated tiny cases; zero at targets == sqrt(I)
  adjoint                complex128 central finite differences; independent torch-autograd
                         restatement (oracle/torch_ref.py); dot-product tests per linear stage
                         (the propagator gradient: the same finite differences and autograd)
"""

from __future__ import annotations

from typing import Optional

import numpy as np

LOSS_TYPES = ("l2_amplitude", "l1_amplitude", "l2_intensity", "l1_intensity", "poisson")


def fftfreq_offsets(n: int) -> np.ndarray:
    """Integer pixel offsets fftfreq(n, 1/n) (dataset_models.py:519-520)."""
    return np.fft.fftfreq(n, d=1.0 / n).round().astype(np.int64)


def origin_and_fraction(positions: np.ndarray):
    """r0 = round-half-even(pos), frac = pos - r0 (dataset_models.py:164-167, 1488-1489, 516-517)."""
    pos = np.asarray(positions, dtype=np.float64)
    r0 = np.round(pos)  # numpy rounds half to even, like torch.round
    return r0.astype(np.int64), pos - r0


def _shape2(n) -> tuple:
    """(rows, cols) of a pattern size given as one int (square) or a pair."""
    return (int(n), int(n)) if np.ndim(n) == 0 else (int(n[0]), int(n[1]))


def patch_rows_cols(positions: np.ndarray, n, obj_shape) -> tuple:
    """Row/col object indices of every patch pixel, periodic (dataset_models.py:514-539); n = N or
    (Nr, Nc) (roi_shape)."""
    nr, nc = _shape2(n)
    r0, _ = origin_and_fraction(positions)
    rows = (r0[:, 0, None] + fftfreq_offsets(nr)[None, :]) % obj_shape[-2]
    cols = (r0[:, 1, None] + fftfreq_offsets(nc)[None, :]) % obj_shape[-1]
    return rows, cols  # (B, Nr), (B, Nc)


def gather_patches(obj: np.ndarray, rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
    """(S, B, n, n) patches obj[s, rows[b, i], cols[b, j]] (object_models.py:252-265)."""
    return obj[:, rows[:, :, None], cols[:, None, :]]


def ramp_freqs(n) -> tuple:
    """Per-axis ramp frequencies fftfreq(N, d=1) rounded to float32, as the reference builds them
    (ptycho_utils.py:167-168); exact for powers of two, 1-ulp-of-float32 off 1/N otherwise."""
    nr, nc = _shape2(n)
    return (np.fft.fftfreq(nr, d=1.0).astype(np.float32).astype(np.float64),
            np.fft.fftfreq(nc, d=1.0).astype(np.float32).astype(np.float64))


def translation_ramp(shifts: np.ndarray, n) -> np.ndarray:
    """exp(-2 pi i kr dr) exp(-2 pi i kc dc), k = fftfreq(N, d=1) per axis (ptycho_utils.py:156-176)."""
    s = np.asarray(shifts, dtype=np.float64)
    kr, kc = ramp_freqs(n)
    rr = np.exp(-2j * np.pi * kr[None, :, None] * s[:, 0, None, None])
    rc = np.exp(-2j * np.pi * kc[None, None, :] * s[:, 1, None, None])
    return rr * rc  # (B, Nr, Nc)


def shift_probes(probe: np.ndarray, frac: np.ndarray) -> np.ndarray:
    """(M, B, n, n) = ifft2(fft2(P_m) * ramp_b) (ptycho_utils.py:128-139; probe_models.py:718-720)."""
    ramp = translation_ramp(frac, probe.shape[-2:])
    return np.fft.ifft2(np.fft.fft2(probe)[:, None] * ramp[None])


def propagate(x: np.ndarray, prop: np.ndarray) -> np.ndarray:
    """ifft2(fft2(x) * prop) (ptychography_base.py:970-989)."""
    return np.fft.ifft2(np.fft.fft2(x) * prop)


def multislice(patches: np.ndarray, psi0: np.ndarray, props: np.ndarray) -> tuple:
    """Incident waves psi_s (S, M, B, n, n) and exit wave x_{S-1} (ptychography_base.py:934-947)."""
    S = patches.shape[0]
    psis = [psi0]
    x = patches[0][None] * psi0
    for s in range(1, S):
        psi = propagate(x, props[s - 1])
        psis.append(psi)
        x = patches[s][None] * psi
    return np.stack(psis), x


def descan_ramp(descan: np.ndarray, n) -> np.ndarray:
    """Real-space ramp applied to the exit wave (ptychography_base.py:903-905)."""
    return translation_ramp(descan, n)


def far_field_intensities(exit_wave: np.ndarray) -> np.ndarray:
    """fftshift(sum_m |fft2_ortho(x_m)|^2) (detector_models.py:22-30)."""
    X = np.fft.fft2(exit_wave, norm="ortho")
    return np.fft.fftshift(np.sum(np.abs(X) ** 2, axis=0), axes=(-2, -1))


def batch_loss(I: np.ndarray, targets: np.ndarray, mask: Optional[np.ndarray], num_gpts: int,
               mean_intensity: float, loss_type: str = "l2_amplitude") -> float:
    """PtychographyBase.error_estimate (ptychography_base.py:908-932)."""
    if loss_type not in LOSS_TYPES:
        raise ValueError(f"Unknown loss type {loss_type}")
    m = np.ones(I.shape[-2:]) if mask is None else np.asarray(mask, dtype=np.float64)
    pred = np.sqrt(I + 1e-9) if "amplitude" in loss_type else I
    B = I.shape[0]
    if loss_type == "poisson":
        err = np.sum(pred - targets * np.log(pred + 1e-6))
    else:
        d = pred * m - targets * m
        err = np.sum(np.abs(d)) if "l1" in loss_type else np.sum(np.abs(d) ** 2)
        err = err / (B / num_gpts)
    return float(err / mean_intensity)


def dloss_dI(I: np.ndarray, targets: np.ndarray, mask: Optional[np.ndarray], num_gpts: int,
             mean_intensity: float, loss_type: str = "l2_amplitude") -> np.ndarray:
    """Derivative of batch_loss with respect to the fftshifted intensities (torch's sign(0) = 0)."""
    B = I.shape[0]
    m = np.ones(I.shape[-2:]) if mask is None else np.asarray(mask, dtype=np.float64)
    amp = "amplitude" in loss_type
    pred = np.sqrt(I + 1e-9) if amp else I
    if loss_type == "poisson":
        dpred = 1.0 - targets / (pred + 1e-6)
        scale = 1.0 / mean_intensity
    else:
        d = pred * m - targets * m
        dpred = (2.0 * d * m) if "l2" in loss_type else (np.sign(d) * m)
        scale = num_gpts / B / mean_intensity
    dI = dpred * 0.5 / pred if amp else dpred
    return scale * dI


def forward(obj, probe, positions, propagators, descan=None) -> dict:
    """Full forward: returns patches, probe shifts, waves and fftshifted intensities."""
    obj = np.asarray(obj, dtype=np.complex128)
    probe = np.asarray(probe, dtype=np.complex128)
    n = probe.shape[-2:]  # (Nr, Nc)
    rows, cols = patch_rows_cols(positions, n, obj.shape)
    patches = gather_patches(obj, rows, cols)
    _, frac = origin_and_fraction(positions)
    psi0 = shift_probes(probe, frac)
    props = np.asarray(propagators, dtype=np.complex128) if propagators is not None else None
    psis, x = multislice(patches, psi0, props)
    if descan is not None:
        x = x * descan_ramp(descan, n)[None]
    I = far_field_intensities(x)
    return {"rows": rows, "cols": cols, "patches": patches, "frac": frac, "psis": psis, "exit": x, "I": I}


def adjoint(obj, probe, positions, propagators, targets, mask, num_gpts, mean_intensity,
            loss_type="l2_amplitude", descan=None, dloss: float = 1.0) -> dict:
    """Gradients in torch's convention (dL/dRe + i dL/dIm) of the batch loss.

    Restates the chain that torch autograd builds over ptychography.py:220-229 (no rescaling;
    the caller applies upsample_factor**2/2, ptychography.py:330-332).
    """
    obj = np.asarray(obj, dtype=np.complex128)
    probe = np.asarray(probe, dtype=np.complex128)
    fw = forward(obj, probe, positions, propagators, descan)
    n = probe.shape[-2:]  # (Nr, Nc)
    npix = n[0] * n[1]
    S = obj.shape[0]
    props = np.asarray(propagators, dtype=np.complex128) if S > 1 else None
    gI = dloss * dloss_dI(fw["I"], targets, mask, num_gpts, mean_intensity, loss_type)
    gI = np.fft.ifftshift(gI, axes=(-2, -1))  # adjoint of fftshift
    X = np.fft.fft2(fw["exit"], norm="ortho")
    gX = 2.0 * gI[None] * X
    gx = np.fft.ifft2(gX, norm="ortho")  # adjoint of the unitary ortho FFT
    out = {}
    if descan is not None:
        ramp = descan_ramp(descan, n)
        fr, fc = ramp_freqs(n)
        w = np.conj(gx) * fw["exit"]  # exit already includes the ramp
        out["descan"] = np.stack([2 * np.pi * np.sum(fr[None, None, :, None] * w.imag, axis=(0, 2, 3)),
                                  2 * np.pi * np.sum(fc[None, None, None, :] * w.imag, axis=(0, 2, 3))], axis=-1)
        gx = gx * np.conj(ramp)[None]
    g_obj = np.zeros_like(obj)
    g_prop = np.zeros((max(S - 1, 0),) + tuple(n), dtype=np.complex128)
    rows, cols = fw["rows"], fw["cols"]
    for s in range(S - 1, -1, -1):
        psi = fw["psis"][s]
        gpatch = np.sum(np.conj(psi) * gx, axis=0)  # (B, n, n), summed over modes
        for b in range(gpatch.shape[0]):
            np.add.at(g_obj[s], (rows[b][:, None], cols[b][None, :]), gpatch[b])
        gpsi = np.conj(fw["patches"][s])[None] * gx
        if s > 0:
            # psi_s = ifft2(fft2(x_{s-1}) * P): the adjoint of ifft2(fft2(.) * P) is ifft2(fft2(.) * conj(P));
            # d psi_s = ifft2(X dP) gives dL/dP = conj(X) fft2(g_psi) / n^2, X = fft2(x_{s-1}) (summed
            # over modes and patterns)
            G = np.fft.fft2(gpsi)
            X = np.fft.fft2(fw["patches"][s - 1][None] * fw["psis"][s - 1])
            g_prop[s - 1] = np.sum(np.conj(X) * G, axis=(0, 1)) / npix
            gx = np.fft.ifft2(G * np.conj(props[s - 1]))
    # probe: psi0 = ifft2(fft2(P) * R) => dP = sum_b ifft2(fft2(gpsi0) * conj(R))
    ramp = translation_ramp(fw["frac"], n)
    G = np.fft.fft2(gpsi)  # (M, B, n, n)
    out["probe"] = np.fft.ifft2(np.sum(G * np.conj(ramp)[None], axis=1))
    # positions: d psi0 / d frac_r = ifft2(fft2(P) R (-2 pi i kr)); <g, .> = (1/n^2) sum conj(G) Pk R (-2 pi i k)
    Pk = np.fft.fft2(probe)
    kr, kc = ramp_freqs(n)
    Z = np.conj(G) * Pk[:, None] * ramp[None]
    out["positions"] = np.stack([
        np.real(np.sum(Z * (-2j * np.pi * kr[None, None, :, None]), axis=(0, 2, 3))),
        np.real(np.sum(Z * (-2j * np.pi * kc[None, None, None, :]), axis=(0, 2, 3))),
    ], axis=-1) / npix
    out["obj"] = g_obj
    out["propagators"] = g_prop
    out["loss"] = dloss * 0 + batch_loss(fw["I"], targets, mask, num_gpts, mean_intensity, loss_type)
    out["I"] = fw["I"]
    return out


def analytical_update(obj, probe, positions, propagators, targets, descan=None, obj_type="complex"):
    """The reference's autograd=False gradients (ptychography.py:339-372, object_models.py:686-725,
    probe_models.py:774-783), float64: returns (obj_grad, probe_grad) as assigned to .grad."""
    obj = np.asarray(obj, dtype=np.complex128)
    probe = np.asarray(probe, dtype=np.complex128)
    fw = forward(obj, probe, positions, propagators, descan)
    n = probe.shape[-1]
    M = probe.shape[0]
    S = obj.shape[0]
    amps = np.fft.fftshift(np.asarray(targets, dtype=np.float64), axes=(-2, -1))  # corner centred
    X = np.fft.fft2(fw["exit"], norm="ortho")
    if M == 1:
        Xp = amps[None] * np.exp(1j * np.angle(X))
    else:
        a = np.sqrt(np.sum(np.abs(X + 1e-9) ** 2, axis=0))
        a[a == 0] = np.inf
        Xp = (amps / a)[None] * X
    g = np.fft.ifft2(Xp, norm="ortho") - fw["exit"]
    rows, cols = fw["rows"], fw["cols"]
    obj_grad = np.zeros_like(obj)
    props = None if propagators is None else np.asarray(propagators, dtype=np.complex128)
    for s in range(S - 1, -1, -1):
        psi = fw["psis"][s]
        norm = 0.0
        upd = np.zeros(obj.shape[1:], dtype=np.complex128)
        for m in range(M):
            ov = np.zeros(obj.shape[1:])
            for b in range(psi.shape[1]):
                np.add.at(ov, (rows[b][:, None], cols[b][None, :]), np.abs(psi[m, b]) ** 2)
                term = np.conj(psi[m, b]) * g[m, b]
                if obj_type == "potential":
                    term = np.real(-1j * np.conj(fw["patches"][s][b]) * term)
                np.add.at(upd, (rows[b][:, None], cols[b][None, :]), term)
            norm += ov.max()
        obj_grad[s] = upd / norm
        g = g * np.conj(fw["patches"][s])[None]
        if s > 0:
            g = np.fft.ifft2(np.fft.fft2(g) * np.conj(props[s - 1]))
    energy = np.sum(np.abs(fw["patches"]) ** 2, axis=(-2, -1)).max()
    ortho_norm = 2 * n if M == 1 else 1 / (2 * n)
    probe_grad = np.sum(g, axis=1) / energy / ortho_norm
    return -obj_grad, -probe_grad
