"""CPU pins of the probe synthesis (complex_probe.py:31-125, 164-215, 434-507; probe_models.py:484-525,
860-873) and of the object-gradient rescale factor (dataset_models.py:451-456, 877-879)."""

import numpy as np
import pytest

from quantem_amd import synth


@pytest.mark.parametrize("kv,lam", [(100e3, 0.037014), (200e3, 0.025079), (300e3, 0.019687)])
def test_electron_wavelength_textbook_values(kv, lam):
    # relativistic electron wavelengths as tabulated (e.g. Williams & Carter, Table 1.1)
    assert abs(synth.electron_wavelength_angstrom(kv) - lam) < 2e-6


@pytest.mark.parametrize("n,sampling,semiangle", [(128, 0.1, 20.0), (64, 0.2, 25.0), (256, 0.1, 30.0)])
def test_hard_aperture_radius_is_alpha_over_lambda(n, sampling, semiangle):
    """Without soft edges the aperture passes exactly the frequencies |k| <= alpha / lambda."""
    lam = synth.electron_wavelength_angstrom(300e3)
    f = synth.fourier_probe(n, sampling, 300e3, semiangle, 0.0, soft_edges=False)
    k = np.fft.fftfreq(n, sampling)
    kr = np.hypot(k[:, None], k[None, :])
    kcut = semiangle * 1e-3 / lam
    inside = kr <= kcut * (1 - 1e-9)
    outside = kr > kcut * (1 + 1e-9)
    assert np.all(np.abs(f[inside]) > 0)
    assert np.all(f[outside] == 0)
    # a flat disc: every passed frequency carries the same amplitude
    a = np.abs(f[inside])
    assert a.max() - a.min() < 1e-12 * a.max()


@pytest.mark.parametrize("n,sampling", [(128, 0.1), (96, 0.15)])
def test_soft_aperture_edge(n, sampling):
    """The soft edge is one angular pixel wide and centred on alpha / lambda: amplitude = that of the
    disc inside alpha - d_alpha / 2, zero beyond alpha + d_alpha / 2, crossing one half at the cut-off."""
    lam = synth.electron_wavelength_angstrom(300e3)
    f = synth.fourier_probe(n, sampling, 300e3, 20.0, 0.0, soft_edges=True)
    a = np.abs(f) / np.abs(f[0, 0])
    k = np.fft.fftfreq(n, sampling)
    dk = k[1]
    kr = np.hypot(k[:, None], k[None, :])
    kcut = 20e-3 / lam
    assert np.allclose(a[kr <= kcut - dk / 2 - 1e-9], 1.0, atol=1e-12)
    assert np.all(a[kr >= kcut + dk / 2 + 1e-9] == 0.0)
    # the partial pixels lie within half a pixel of the cut-off, above one half inside it, below outside
    ramp = (a > 0) & (a < 1)
    assert ramp.sum() > 0
    assert np.all(np.abs(kr[ramp] - kcut) < dk / 2 + 1e-9)
    assert np.all((a[ramp] > 0.5) == (kr[ramp] < kcut))


def test_probe_normalisation_unit_l2():
    f = synth.fourier_probe(128, 0.1, 300e3, 20.0, 100.0)
    assert abs(np.sum(np.abs(f) ** 2) - 1.0) < 1e-12
    p = synth.real_space_probe(128, 0.1, 300e3, 20.0, 100.0)
    assert abs(np.sum(np.abs(p) ** 2) - 1.0) < 1e-12
    # real space = ifft2 of the Fourier probe up to the renormalisation (complex_probe.py:502-505)
    q = np.fft.ifft2(f)
    assert np.allclose(p, q / np.linalg.norm(q), atol=1e-14)


def test_defocus_is_negative_c10_equals_fresnel_propagation():
    """C10 = -defocus (probe_models.py:178-180): the defocused Fourier probe is the focused one times
    the conjugate Fresnel propagator over the defocus distance (chi = pi lambda k^2 C10)."""
    n, samp, dz = 128, 0.1, 100.0
    f0 = synth.fourier_probe(n, samp, 300e3, 20.0, 0.0)
    fd = synth.fourier_probe(n, samp, 300e3, 20.0, dz)
    prop = synth.fresnel_propagators(n, samp, 300e3, [dz])[0].astype(np.complex128)
    assert np.abs(fd - f0 * np.conj(prop)).max() < 1e-6 * np.abs(f0).max()


def test_gram_schmidt_orthogonal_norms_kept_sorted():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(4, 16, 16)) + 1j * rng.normal(size=(4, 16, 16))
    x *= np.array([1.0, 3.0, 0.5, 2.0])[:, None, None]
    g = synth.gram_schmidt(x)
    gram = np.einsum("aij,bij->ab", np.conj(g), g)
    off = gram - np.diag(np.diag(gram))
    assert np.abs(off).max() < 1e-10 * np.abs(np.diag(gram)).max()
    pw = np.diag(gram).real
    assert np.all(np.diff(pw) <= 0)  # sorted by intensity, strongest first
    assert np.allclose(np.sort(pw)[::-1], np.sort(np.sum(np.abs(x) ** 2, axis=(1, 2)))[::-1])


def test_mixed_probe_weights_and_total_intensity():
    base = synth.real_space_probe(64, 0.2, 300e3, 20.0, 50.0)
    w = [0.7, 0.15, 0.1, 0.05]
    p = synth.mixed_probe(base, w, 1e5)
    # per-mode power in the far field (ortho FFT) = weight * total (probe_models.py:860-873)
    pw = np.sum(np.abs(np.fft.fft2(p, norm="ortho")) ** 2, axis=(1, 2))
    assert np.allclose(pw, np.asarray(w) * 1e5, rtol=1e-10)
    gram = np.einsum("aij,bij->ab", np.conj(p), p)
    assert np.abs(gram - np.diag(np.diag(gram))).max() < 1e-8 * 1e5


@pytest.mark.parametrize("name,value", [("headline", 252 / 64), ("depth20-256px", 188 / 48),
                                        ("large30-1024px", 508 / 128), ("single-slice", 124 / 32)])
def test_upsample_factor_of_the_workloads(name, value):
    """upsample_factor = floor(fov / sampling) (rounded up to even) / gpts; at the 64 x 64 scan with a
    4 px step: fov = 63 * 0.4 A = 25.2 A -> 252 px -> 252 / 64 = 3.9375."""
    assert synth.spec_upsample_factor(synth.WORKLOADS[name]) == pytest.approx(value, abs=1e-12)


@pytest.mark.parametrize("tilt", [(0.0, 0.0), (3.0, -2.0), (0.0, 5.0)])
def test_torch_propagator_builder_matches_numpy(tilt):
    """The differentiable propagator builder of the loop (torch, complex64) against the numpy one that
    simulates the data (probe_models.py:376-427)."""
    import torch

    from quantem_amd import ptychography as qp

    a = qp.fresnel_propagators(64, 0.1, 300e3, [2.0, 1.5], torch.tensor(tilt)).numpy()
    b = synth.fresnel_propagators(64, 0.1, 300e3, [2.0, 1.5], tilt)
    assert np.abs(a - b).max() < 1e-6


def test_zero_tilt_component_gets_no_gradient():
    """As in the reference, a tilt component that is exactly zero adds no factor and so no gradient."""
    import torch

    from quantem_amd import ptychography as qp

    th = torch.tensor([0.0, 2.0], requires_grad=True)
    p = qp.fresnel_propagators(32, 0.1, 300e3, [5.0], th)
    (p.real.sum() + p.imag.sum()).backward()
    assert th.grad[0] == 0 and th.grad[1] != 0
