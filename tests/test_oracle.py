"""Pins the FP64 oracle with checks that do not restate its formulas (CPU only).

Each check is chosen so that a plausible slip (dropped term, wrong sign or index, transposed
operand, wrong normalisation) in the oracle fails it:
  * the reference test's own numpy simulation recipe (tests/diffractive_imaging/test_ptychography.py:46-82)
  * closed forms: integer shifts == np.roll, zero shift == identity, propagator composition
  * invariants: Parseval, unitarity, energy conservation for pure-phase objects
  * complex128 central finite differences of the loss
  * an independent torch-autograd restatement of the reference op order (oracle/torch_ref.py)
  * the committed golden fixture of the reference test recipe (tests/golden/)
"""

import os

import numpy as np
import pytest

from oracle import ptycho_oracle as po
from oracle import torch_ref
from quantem_amd import synth
from tests._cases import attach_targets, make_case, rel

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _recipe_patch_indices(positions_px, roi_shape, obj_shape):
    # the reference test's return_patch_indices (test_ptychography.py:46-57), re-typed from its description
    x0 = np.round(positions_px[:, 0]).astype("int")
    y0 = np.round(positions_px[:, 1]).astype("int")
    x_ind = np.fft.fftfreq(roi_shape[0], d=1 / roi_shape[0]).astype("int")
    y_ind = np.fft.fftfreq(roi_shape[1], d=1 / roi_shape[1]).astype("int")
    row = (x0[:, None, None] + x_ind[None, :, None]) % obj_shape[0]
    col = (y0[:, None, None] + y_ind[None, None, :]) % obj_shape[1]
    return row, col


def test_patch_indices_match_reference_recipe():
    rng = np.random.default_rng(0)
    pos = np.concatenate([rng.uniform(0, 47, (20, 2)), [[0, 0], [47, 47], [2.5, 3.5], [1.5, 0.5]]])
    rows, cols = po.patch_rows_cols(pos, 16, (48, 40))
    r2, c2 = _recipe_patch_indices(pos, (16, 16), (48, 40))
    assert np.array_equal(rows, r2[:, :, 0])
    assert np.array_equal(cols, c2[:, 0, :])


def test_round_half_even():
    r0, frac = po.origin_and_fraction(np.array([[2.5, 3.5], [-0.5, 0.49]]))
    assert r0.tolist() == [[2, 4], [0, 0]]
    assert np.allclose(frac, [[0.5, -0.5], [-0.5, 0.49]])


def test_gather_matches_take():
    rng = np.random.default_rng(1)
    obj = rng.normal(size=(2, 9, 7)) + 1j * rng.normal(size=(2, 9, 7))
    rows = rng.integers(0, 9, (3, 4))
    cols = rng.integers(0, 7, (3, 4))
    p = po.gather_patches(obj, rows, cols)
    for s in range(2):
        for b in range(3):
            flat = np.take(obj[s].ravel(), rows[b][:, None] * 7 + cols[b][None, :])
            assert np.array_equal(p[s, b], flat)


def test_shift_integer_is_roll_and_zero_is_identity():
    rng = np.random.default_rng(2)
    probe = rng.normal(size=(2, 16, 16)) + 1j * rng.normal(size=(2, 16, 16))
    out = po.shift_probes(probe, np.array([[0.0, 0.0], [1.0, -2.0]]))
    assert np.allclose(out[:, 0], probe, atol=1e-12)
    assert np.allclose(out[:, 1], np.roll(probe, (1, -2), axis=(-2, -1)), atol=1e-12)
    frac = rng.uniform(-0.5, 0.5, (5, 2))
    sh = po.shift_probes(probe, frac)
    assert np.allclose(np.sum(np.abs(sh) ** 2, axis=(-2, -1)), np.sum(np.abs(probe) ** 2, axis=(-2, -1))[:, None])


def test_propagator_composition_and_unit_modulus():
    p1 = synth.fresnel_propagators(32, 0.1, 300e3, [1.0, 2.0, 3.0]).astype(np.complex128)
    assert np.allclose(np.abs(p1), 1.0, atol=1e-6)
    assert np.allclose(p1[0] * p1[1], p1[2], atol=1e-5)
    p0 = synth.fresnel_propagators(32, 0.1, 300e3, [0.0])
    assert np.allclose(p0, 1.0)


def test_multislice_single_slice_matches_reference_recipe():
    """S = 1, M = 1, integer positions: the reference test's simulate_intensities (fft2, unnormalised)."""
    rng = np.random.default_rng(42)
    N = 16
    obj = np.exp(1j * (rng.random((N, N)) - 0.5))
    probe = synth.real_space_probe(N, 0.5, 300e3, 20.0, 50.0)
    pos = np.stack(np.meshgrid(np.arange(4.0), np.arange(4.0), indexing="ij"), -1).reshape(-1, 2)
    row, col = _recipe_patch_indices(pos, (N, N), (N, N))
    recipe = np.fft.fftshift(np.abs(np.fft.fft2(obj[row, col] * probe)) ** 2, axes=(-2, -1))
    got = po.forward(obj[None], probe[None], pos, None)["I"]
    # the detector uses the ortho FFT: intensities are 1/N^2 of the unnormalised recipe
    assert rel(got * N * N, recipe) < 1e-12


def test_free_space_equals_single_propagation():
    """O == 1: S slices propagate the probe by the summed thickness."""
    N, S = 32, 4
    probe = synth.real_space_probe(N, 0.1, 300e3, 20.0, 10.0)[None]
    props = synth.fresnel_propagators(N, 0.1, 300e3, [2.0] * (S - 1)).astype(np.complex128)
    total = synth.fresnel_propagators(N, 0.1, 300e3, [2.0 * (S - 1)]).astype(np.complex128)
    obj = np.ones((S, 40, 40), dtype=np.complex128)
    fw = po.forward(obj, probe, np.array([[5.0, 7.0]]), props)
    direct = np.fft.ifft2(np.fft.fft2(probe[0]) * total[0])
    assert rel(fw["exit"][0, 0], direct) < 1e-5


def test_parseval_and_energy_conservation():
    case = make_case(N=32, S=4, M=3, B=5, seed=3, weak=True)
    fw = po.forward(case["obj"], case["probe"], case["positions"], case["propagators"])
    energy_probe = np.sum(np.abs(case["probe"].astype(np.complex128)) ** 2)
    assert np.allclose(fw["I"].sum(axis=(-2, -1)), energy_probe, rtol=1e-5)


def test_fftshift_places_dc_at_center():
    N = 16
    x = np.ones((1, 1, N, N), dtype=np.complex128)
    I = po.far_field_intensities(x)
    assert I[0, N // 2, N // 2] == pytest.approx(N * N)
    assert np.sum(I) == pytest.approx(N * N)


def test_loss_values_by_hand():
    I = np.array([[[4.0, 9.0], [16.0, 1.0]]])
    t = np.array([[[2.0, 2.0], [4.0, 0.0]]])
    # l2 amplitude: sqrt(I + 1e-9) - t = [0, 1, 0, 1] -> 2 ; * P/B / mean
    assert po.batch_loss(I, t, None, 10, 2.0, "l2_amplitude") == pytest.approx(2 * 10 / 2.0, rel=1e-8)
    assert po.batch_loss(I, t, None, 10, 2.0, "l1_intensity") == pytest.approx((2 + 7 + 12 + 1) * 10 / 2.0)
    pois = np.sum(I - t * np.log(I + 1e-6)) / 2.0
    assert po.batch_loss(I, t, None, 10, 2.0, "poisson") == pytest.approx(pois)
    assert po.batch_loss(I, np.sqrt(I + 1e-9), None, 10, 2.0, "l2_amplitude") == pytest.approx(0.0, abs=1e-20)


@pytest.mark.parametrize("loss_type", po.LOSS_TYPES)
def test_adjoint_matches_torch_autograd(loss_type):
    case = make_case(N=16, S=3, M=2, B=4, seed=5, descan=True, loss_type=loss_type)
    I = po.forward(case["obj"], case["probe"], case["positions"], case["propagators"], case["descan"])["I"]
    case = attach_targets(case, I)
    a = po.adjoint(case["obj"], case["probe"], case["positions"], case["propagators"], case["targets"],
                   case["mask"], case["num_gpts"], case["mean_intensity"], loss_type, descan=case["descan"])
    b = torch_ref.loss_and_grads(case["obj"], case["probe"], case["positions"], case["propagators"],
                                 case["targets"], case["mask"], case["num_gpts"], case["mean_intensity"],
                                 loss_type, descan=case["descan"])
    assert abs(a["loss"] - b["loss"]) <= 1e-12 * abs(b["loss"])
    for k in ("obj", "probe", "positions", "descan", "propagators"):
        assert rel(a[k], b[k]) < 1e-10, k


def test_adjoint_matches_finite_differences():
    case = make_case(N=16, S=2, M=2, B=3, seed=9, descan=True, half_integer=False)  # round() is flat away from .5
    I = po.forward(case["obj"], case["probe"], case["positions"], case["propagators"], case["descan"])["I"]
    case = attach_targets(case, I)
    obj = case["obj"].astype(np.complex128)
    probe = case["probe"].astype(np.complex128)
    pos = case["positions"].astype(np.float64)
    desc = case["descan"].astype(np.float64)
    props = case["propagators"].astype(np.complex128)

    def L(o=obj, p=probe, x=pos, d=desc):
        I = po.forward(o, p, x, props, d)["I"]
        return po.batch_loss(I, case["targets"], case["mask"], case["num_gpts"], case["mean_intensity"])

    g = po.adjoint(obj, probe, pos, props, case["targets"], case["mask"], case["num_gpts"],
                   case["mean_intensity"], descan=desc)
    rng = np.random.default_rng(0)
    h = 1e-4
    rows, cols = po.patch_rows_cols(pos, 16, obj.shape[1:])
    for _ in range(8):  # object pixels that a patch covers
        s, b = rng.integers(0, 2), rng.integers(0, 3)
        i, j = rows[b, rng.integers(0, 16)], cols[b, rng.integers(0, 16)]
        for unit in (1.0, 1j):
            e = np.zeros_like(obj)
            e[s, i, j] = h * unit
            fd = (L(o=obj + e) - L(o=obj - e)) / (2 * h)
            an = g["obj"][s, i, j].real if unit == 1.0 else g["obj"][s, i, j].imag
            assert fd == pytest.approx(an, rel=1e-5, abs=1e-9)
    for _ in range(6):
        m, i, j = rng.integers(0, 2), rng.integers(0, 16), rng.integers(0, 16)
        for unit in (1.0, 1j):
            e = np.zeros_like(probe)
            e[m, i, j] = h * unit
            fd = (L(p=probe + e) - L(p=probe - e)) / (2 * h)
            an = g["probe"][m, i, j].real if unit == 1.0 else g["probe"][m, i, j].imag
            assert fd == pytest.approx(an, rel=1e-5, abs=1e-9)
    def Lp(pp):
        I = po.forward(obj, probe, pos, pp, desc)["I"]
        return po.batch_loss(I, case["targets"], case["mask"], case["num_gpts"], case["mean_intensity"])

    for _ in range(6):  # propagator entries (the gradient that learning a probe tilt chains through)
        i, j = rng.integers(0, 16), rng.integers(0, 16)
        for unit in (1.0, 1j):
            e = np.zeros_like(props)
            e[0, i, j] = h * unit
            fd = (Lp(props + e) - Lp(props - e)) / (2 * h)
            an = g["propagators"][0, i, j].real if unit == 1.0 else g["propagators"][0, i, j].imag
            assert fd == pytest.approx(an, rel=1e-5, abs=1e-9)
    for b in range(3):
        for k in range(2):
            e = np.zeros_like(pos)
            e[b, k] = 1e-5
            fd = (L(x=pos + e) - L(x=pos - e)) / 2e-5
            assert fd == pytest.approx(g["positions"][b, k], rel=1e-4, abs=1e-8)
            e = np.zeros_like(desc)
            e[b, k] = 1e-5
            fd = (L(d=desc + e) - L(d=desc - e)) / 2e-5
            assert fd == pytest.approx(g["descan"][b, k], rel=1e-4, abs=1e-8)


def test_adjoint_dot_product_of_propagation():
    """<A x, y> == <x, A^H y> for the propagation stage used by the adjoint."""
    rng = np.random.default_rng(4)
    P = synth.fresnel_propagators(16, 0.1, 300e3, [3.0]).astype(np.complex128)[0]
    x = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
    y = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
    lhs = np.vdot(y, po.propagate(x, P))
    rhs = np.vdot(np.fft.ifft2(np.fft.fft2(y) * np.conj(P)), x)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_golden_reference_recipe_fixture():
    """Committed vectors of the reference test recipe (N = 64 white-noise object, C10 = 50 A)."""
    path = os.path.join(GOLDEN, "reference_recipe_n64.npz")
    z = np.load(path)
    got = po.forward(z["obj"][None], z["probe"][None], z["positions"], None)["I"]
    assert rel(got * 64 * 64, z["intensities_unnormalised"]) < 1e-6
    g = po.adjoint(z["obj"][None], z["probe"][None], z["positions"], None, z["targets"], None,
                   int(z["num_gpts"]), float(z["mean_intensity"]))
    assert abs(g["loss"] - float(z["loss"])) < 1e-9 * abs(float(z["loss"]))
    assert rel(g["obj"], z["grad_obj"]) < 1e-6
    assert rel(g["probe"], z["grad_probe"]) < 1e-6


@pytest.mark.parametrize("M", [1, 3])
def test_analytical_update_relates_to_autograd(M):
    """The projection residual is -1/(2 scale) times the l2-amplitude gradient at the exit wave
    (scale = P / (B mean_I), using a = sqrt(I + 1e-9) ~ |X|), and both replay the slices the same
    way, so the reference's normalised analytical updates map back onto torch-autograd gradients.
    Integer positions make the probe shift's adjoint the identity."""
    case = make_case(N=16, S=3, M=M, B=4, seed=11, half_integer=False)
    case["positions"] = np.round(case["positions"]).astype(np.float32)
    I = po.forward(case["obj"], case["probe"], case["positions"], case["propagators"])["I"]
    case = attach_targets(case, I)
    og, pg = po.analytical_update(case["obj"], case["probe"], case["positions"], case["propagators"],
                                  case["targets"])
    ag = torch_ref.loss_and_grads(case["obj"], case["probe"], case["positions"], case["propagators"],
                                  case["targets"], None, case["num_gpts"], case["mean_intensity"],
                                  grad_positions=False)
    scale = case["num_gpts"] / 4 / case["mean_intensity"]
    fw = po.forward(case["obj"], case["probe"], case["positions"], case["propagators"])
    rows, cols = fw["rows"], fw["cols"]
    for s in range(3):
        norm = 0.0
        for m in range(M):
            ov = np.zeros(case["obj"].shape[1:])
            for b in range(4):
                np.add.at(ov, (rows[b][:, None], cols[b][None, :]), np.abs(fw["psis"][s, m, b]) ** 2)
            norm += ov.max()
        assert rel(og[s] * (2 * scale * norm), ag["obj"][s]) < 1e-4
    energy = np.sum(np.abs(fw["patches"]) ** 2, axis=(-2, -1)).max()
    ortho = 2 * 16 if M == 1 else 1 / (2 * 16)
    assert rel(pg * (2 * scale * energy * ortho), ag["probe"]) < 1e-4


def test_analytical_potential_branch_is_the_phase_chain_rule():
    """For a potential V with O = exp(iV) the reference's update term is real(-i conj(O) x) of the complex
    term (object_models.py:701-715); the normalisation is shared, so pixel by pixel the potential update
    equals real(-i conj(O) * complex update), and the probe update does not change."""
    case = make_case(N=16, S=2, M=2, B=4, seed=12)
    v = np.random.default_rng(12).uniform(0.0, 0.6, case["obj"].shape)
    o = np.exp(1j * v).astype(np.complex64)
    I = po.forward(o, case["probe"], case["positions"], case["propagators"])["I"]
    case = attach_targets(case, I * 1.3)
    og_c, pg_c = po.analytical_update(o, case["probe"], case["positions"], case["propagators"], case["targets"])
    og_v, pg_v = po.analytical_update(o, case["probe"], case["positions"], case["propagators"], case["targets"],
                                      obj_type="potential")
    o128 = o.astype(np.complex128)
    assert np.abs(og_v.imag).max() == 0.0
    assert rel(og_v.real, np.real(-1j * np.conj(o128) * og_c)) < 1e-12
    assert rel(pg_v, pg_c) < 1e-14


def test_rectangular_patterns_adjoint_matches_autograd_and_finite_differences():
    """Non-square, non-power-of-two patterns (roi_shape (12, 20)): the oracle's gradients equal torch
    autograd of the reference-order restatement (complex128) and complex128 finite differences, and its
    patch indices follow fftfreq offsets per axis."""
    case = make_case(N=(12, 20), S=3, M=2, B=3, seed=12, descan=True, half_integer=False)
    I = po.forward(case["obj"], case["probe"], case["positions"], case["propagators"], case["descan"])["I"]
    assert I.shape == (3, 12, 20)
    case = attach_targets(case, I)
    a = po.adjoint(case["obj"], case["probe"], case["positions"], case["propagators"], case["targets"],
                   case["mask"], case["num_gpts"], case["mean_intensity"], descan=case["descan"])
    b = torch_ref.loss_and_grads(case["obj"], case["probe"], case["positions"], case["propagators"],
                                 case["targets"], case["mask"], case["num_gpts"], case["mean_intensity"],
                                 descan=case["descan"])
    assert abs(a["loss"] - b["loss"]) <= 1e-12 * abs(b["loss"])
    for k in ("obj", "probe", "positions", "descan", "propagators"):
        assert rel(a[k], b[k]) < 1e-10, k
    rows, cols = po.patch_rows_cols(case["positions"], (12, 20), case["obj"].shape)
    r0 = np.round(case["positions"]).astype(int)
    assert np.array_equal(rows[0], (r0[0, 0] + np.r_[0:6, -6:0]) % case["obj"].shape[1])
    assert np.array_equal(cols[0], (r0[0, 1] + np.r_[0:10, -10:0]) % case["obj"].shape[2])
    obj = case["obj"].astype(np.complex128)
    props = case["propagators"].astype(np.complex128)

    def L(o):
        I = po.forward(o, case["probe"], case["positions"], props, case["descan"])["I"]
        return po.batch_loss(I, case["targets"], case["mask"], case["num_gpts"], case["mean_intensity"])

    rng = np.random.default_rng(1)
    h = 1e-4
    for _ in range(6):
        s, bb = rng.integers(0, 3), rng.integers(0, 3)
        i, j = rows[bb, rng.integers(0, 12)], cols[bb, rng.integers(0, 20)]
        e = np.zeros_like(obj)
        e[s, i, j] = h
        fd = (L(obj + e) - L(obj - e)) / (2 * h)
        assert fd == pytest.approx(a["obj"][s, i, j].real, rel=1e-5, abs=1e-9)
