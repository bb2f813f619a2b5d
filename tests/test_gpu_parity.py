"""GPU parity of the HIP kernels (through the C ABI) against the FP64 oracle.

Tolerances (relative L2 error against the float64 oracle evaluated on the same complex64 inputs):
  FFT                          2e-6
  intensities, loss            1e-5
  object / probe / position /  1e-4
  descan gradients
"""

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from quantem_amd import ptychography as qp
from quantem_amd import synth
from tests._cases import attach_targets, make_case, rel

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
GRAD_TOL = 1e-4


def _t(x, dev):
    return None if x is None else torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _run(case, dev, grads=True, pos_grad=True, desc_grad=True, segment_slices=0, full_propagators=False):
    obj = _t(case["obj"], dev).requires_grad_(grads)
    probe = _t(case["probe"], dev).requires_grad_(grads)
    pos = _t(case["positions"], dev).requires_grad_(grads and pos_grad)
    props = _t(case["propagators"], dev) if case["obj"].shape[0] > 1 else None
    d = _t(case["descan"], dev)
    if d is not None:
        d.requires_grad_(grads and desc_grad)
    loss, intens = qp.multislice_loss(obj, probe, pos, props, _t(case["targets"], dev), _t(case["mask"], dev), d,
                                      num_gpts=case["num_gpts"], mean_intensity=case["mean_intensity"],
                                      loss_type=case["loss_type"], return_intensities=True,
                                      segment_slices=segment_slices, full_propagators=full_propagators)
    out = {"loss": loss.item(), "I": intens.cpu().numpy()}
    if grads:
        loss.backward()
        out["obj"] = obj.grad.cpu().numpy()
        out["probe"] = probe.grad.cpu().numpy()
        if pos_grad:
            out["positions"] = pos.grad.cpu().numpy()
        if d is not None and desc_grad:
            out["descan"] = d.grad.cpu().numpy()
    return out


def _oracle(case):
    return po.adjoint(case["obj"], case["probe"], case["positions"],
                      case["propagators"] if case["obj"].shape[0] > 1 else None, case["targets"], case["mask"],
                      case["num_gpts"], case["mean_intensity"], case["loss_type"], descan=case["descan"])


def _case_with_targets(**kw):
    case = make_case(**kw)
    I = po.forward(case["obj"], case["probe"], case["positions"], case["propagators"], case["descan"])["I"]
    return attach_targets(case, I)


@pytest.mark.parametrize("n", [32, 64, 128])
def test_fft2_matches_numpy(cuda, n):
    print(f"fft2 n={n}", flush=True)
    rng = np.random.default_rng(n)
    x = (rng.normal(size=(5, n, n)) + 1j * rng.normal(size=(5, n, n))).astype(np.complex64)
    xt = torch.from_numpy(x).to(cuda)
    f = qp.fft2(xt).cpu().numpy()
    assert rel(f, np.fft.fft2(x.astype(np.complex128))) < 2e-6
    g = qp.ifft2(xt).cpu().numpy()
    assert rel(g, np.fft.ifft2(x.astype(np.complex128))) < 2e-6


@pytest.mark.parametrize(
    "n,s,m,b,descan",
    [(32, 1, 1, 5, False), (32, 3, 2, 6, True), (64, 1, 1, 7, False), (64, 4, 3, 5, False),
     (128, 1, 1, 3, False), (128, 3, 2, 4, True), (128, 8, 4, 3, False), (32, 2, 6, 4, False),
     (256, 1, 1, 2, False), (256, 3, 2, 3, True)],
)
def test_forward_and_gradients_match_oracle(cuda, n, s, m, b, descan):
    print(f"parity n={n} s={s} m={m} b={b} descan={descan}", flush=True)
    case = _case_with_targets(N=n, S=s, M=m, B=b, seed=100 + n + s, descan=descan)
    got = _run(case, cuda)
    ref = _oracle(case)
    assert rel(got["I"], ref["I"]) < FWD_TOL
    assert abs(got["loss"] - ref["loss"]) / abs(ref["loss"]) < FWD_TOL
    for k in ("obj", "probe", "positions") + (("descan",) if descan else ()):
        assert rel(got[k], ref[k]) < GRAD_TOL, k


@pytest.mark.parametrize("loss_type", ["l2_amplitude", "l1_amplitude", "l2_intensity", "l1_intensity", "poisson"])
def test_loss_types(cuda, loss_type):
    print(f"loss {loss_type}", flush=True)
    case = _case_with_targets(N=32, S=2, M=2, B=6, seed=7, loss_type=loss_type)
    got = _run(case, cuda)
    ref = _oracle(case)
    assert abs(got["loss"] - ref["loss"]) / abs(ref["loss"]) < FWD_TOL
    for k in ("obj", "probe", "positions"):
        assert rel(got[k], ref[k]) < GRAD_TOL, k


@pytest.mark.parametrize("n", [64, 128, 256])
def test_no_grad_forward_matches(cuda, n):
    case = _case_with_targets(N=n, S=3, M=2, B=3, seed=3)
    with torch.no_grad():
        got = _run(case, cuda, grads=False)
    ref = _oracle(case)
    assert rel(got["I"], ref["I"]) < FWD_TOL


@pytest.mark.parametrize(
    "n,s,m,b,k,descan",
    [(32, 5, 2, 4, 2, True), (64, 7, 2, 3, 3, False), (128, 8, 4, 3, 3, True), (128, 3, 1, 2, 2, False),
     (256, 5, 2, 2, 2, True), (256, 7, 1, 2, 3, False)],
)
def test_segment_recompute_matches_stored_and_oracle(cuda, n, s, m, b, k, descan):
    """Checkpointed forward + per-segment recomputation in the backward pass gives the stored-wave
    results (the recomputed waves run the same arithmetic) and meets the oracle bar."""
    print(f"recompute n={n} s={s} m={m} b={b} k={k} descan={descan}", flush=True)
    case = _case_with_targets(N=n, S=s, M=m, B=b, seed=500 + n + s, descan=descan)
    full = _run(case, cuda)
    seg = _run(case, cuda, segment_slices=k)
    ref = _oracle(case)
    assert seg["loss"] == full["loss"]
    assert np.array_equal(seg["I"], full["I"])
    for key in ("obj", "probe", "positions") + (("descan",) if descan else ()):
        assert rel(seg[key], full[key]) < 1e-6, key
        assert rel(seg[key], ref[key]) < GRAD_TOL, key


def test_separable_propagators_256(cuda):
    """256 x 256 patterns apply a Fresnel propagator slice (a row factor times a column factor) through
    its factors by default: the intensities stay within 1e-6 of every entry applied as given
    (QEM_FLAG_FULL_PROPAGATORS) and differ from them in the last bits (the factor path ran), the
    gradients within 2e-5 (they carry the residual |psi| - target, here about a tenth of |psi|, which
    scales the float32 differences up), and both paths meet the oracle bar.  At 0.1 A sampling the largest propagator phase is 6 rad and the float32 entries factor
    to 4e-7; the other 256 x 256 cases here sample at 2/256 A, where phases near 1000 rad leave 5e-5 of
    float32 rounding that no product reproduces, so they run the entry-by-entry path.  A slice
    multiplied by a random phase does not factor: one such slice sends the whole launch down the
    entry-by-entry path (bitwise the results with QEM_FLAG_FULL_PROPAGATORS), where a factor applied to
    the perturbed slice would miss the oracle by the size of the perturbation (about 1e-1)."""
    print("separable propagators n=256 s=3 m=2 b=3", flush=True)
    case = _case_with_targets(N=256, S=3, M=2, B=3, seed=41, descan=True, sampling=0.1)
    fac = _run(case, cuda)
    full = _run(case, cuda, full_propagators=True)
    ref = _oracle(case)
    keys = ("obj", "probe", "positions", "descan")
    print("  factors vs entries: " + ", ".join(f"{k} {rel(fac[k], full[k]):.2e}" for k in ("I",) + keys), flush=True)
    for name, got in (("factors", fac), ("entries", full)):
        print(f"  {name} vs oracle: " + ", ".join(f"{k} {rel(got[k], ref[k]):.2e}" for k in ("I",) + keys), flush=True)
    assert not np.array_equal(fac["I"], full["I"])
    assert rel(fac["I"], full["I"]) < 1e-6
    for k in keys:
        assert rel(fac[k], full[k]) < 2e-5, k
    for got in (fac, full):
        assert rel(got["I"], ref["I"]) < FWD_TOL
        for k in keys:
            assert rel(got[k], ref[k]) < GRAD_TOL, k

    print("one slice not separable", flush=True)
    mixed = make_case(N=256, S=3, M=2, B=3, seed=42, sampling=0.1)
    rng = np.random.default_rng(43)
    mixed["propagators"][1] *= np.exp(0.3j * rng.uniform(-1.0, 1.0, (256, 256))).astype(np.complex64)
    I = po.forward(mixed["obj"], mixed["probe"], mixed["positions"], mixed["propagators"])["I"]
    mixed = attach_targets(mixed, I)
    got = _run(mixed, cuda)
    full = _run(mixed, cuda, full_propagators=True)
    ref = _oracle(mixed)
    errs = {k: rel(got[k], ref[k]) for k in ("I", "obj", "probe", "positions")}
    print("  " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert errs["I"] < FWD_TOL
    assert np.array_equal(got["I"], full["I"])
    for k in ("obj", "probe", "positions"):
        assert errs[k] < GRAD_TOL, k
        assert np.array_equal(got[k], full[k]), k


def test_separable_propagators_128(cuda):
    """The 128 x 128 adjoint (no descan, no propagator gradient) applies a separable propagator slice
    through its row and column factors by default.  The forward applies every entry in both runs, so
    the intensities are bitwise those of QEM_FLAG_FULL_PROPAGATORS; the gradients differ in the last
    bits (the factor path ran) and stay within 2e-5 of the entry-by-entry ones, and both paths meet the
    oracle bar.  One slice multiplied by a random phase sends the adjoint down the entry-by-entry path:
    gradients bitwise those of QEM_FLAG_FULL_PROPAGATORS."""
    print("separable propagators n=128 s=4 m=2 b=3", flush=True)
    case = _case_with_targets(N=128, S=4, M=2, B=3, seed=51, sampling=0.1)
    fac = _run(case, cuda, desc_grad=False)
    full = _run(case, cuda, desc_grad=False, full_propagators=True)
    ref = _oracle(case)
    keys = ("obj", "probe", "positions")
    print("  factors vs entries: " + ", ".join(f"{k} {rel(fac[k], full[k]):.2e}" for k in keys), flush=True)
    for name, got in (("factors", fac), ("entries", full)):
        print(f"  {name} vs oracle: " + ", ".join(f"{k} {rel(got[k], ref[k]):.2e}" for k in ("I",) + keys), flush=True)
    assert np.array_equal(fac["I"], full["I"])
    assert not np.array_equal(fac["obj"], full["obj"])
    for k in keys:
        assert rel(fac[k], full[k]) < 2e-5, k
    for got in (fac, full):
        assert rel(got["I"], ref["I"]) < FWD_TOL
        for k in keys:
            assert rel(got[k], ref[k]) < GRAD_TOL, k

    print("one slice not separable", flush=True)
    mixed = make_case(N=128, S=4, M=2, B=3, seed=52, sampling=0.1)
    rng = np.random.default_rng(53)
    mixed["propagators"][2] *= np.exp(0.3j * rng.uniform(-1.0, 1.0, (128, 128))).astype(np.complex64)
    I = po.forward(mixed["obj"], mixed["probe"], mixed["positions"], mixed["propagators"])["I"]
    mixed = attach_targets(mixed, I)
    got = _run(mixed, cuda, desc_grad=False)
    full = _run(mixed, cuda, desc_grad=False, full_propagators=True)
    ref = _oracle(mixed)
    errs = {k: rel(got[k], ref[k]) for k in ("I",) + keys}
    print("  " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert errs["I"] < FWD_TOL
    for k in keys:
        assert errs[k] < GRAD_TOL, k
        assert np.array_equal(got[k], full[k]), k


@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
def test_nonfinite_propagator_entry_takes_the_entry_path_256(cuda, bad):
    """A separable 256 x 256 propagator with one non-finite entry away from its first row and column
    (where the factors would not see it) is not applied through factors: the intensities equal, NaNs
    included, those with every entry applied as given."""
    print(f"non-finite propagator entry {bad}", flush=True)
    case = _case_with_targets(N=256, S=3, M=1, B=2, seed=44, sampling=0.1)
    case["propagators"][1, 5, 7] = bad
    with torch.no_grad():
        got = _run(case, cuda, grads=False)
        full = _run(case, cuda, grads=False, full_propagators=True)
    assert not np.all(np.isfinite(got["I"]))
    assert np.array_equal(got["I"], full["I"], equal_nan=True)


def test_depth_20_slices_256px_against_oracle_and_finite_differences(cuda):
    """The deep-adjoint configuration (20 slices of 512x512, 1 A, 256x256 patterns, 2 modes) on two
    scan positions, with every incident wave stored and with 5-slice recomputation segments:
    gradients against the FP64 oracle, against the reference-order torch autograd gradient, and
    against complex128 central finite differences of the oracle loss on 64 random perturbations
    (32 object pixels, 32 probe entries; h = 1e-4)."""
    print("depth-20 256px", flush=True)
    rng = np.random.default_rng(20)
    case = make_case(N=256, S=20, M=2, B=2, H=512, W=512, seed=20, dz=1.0, sampling=0.1, half_integer=False)
    # pure-phase Gaussian-atom slices as in that configuration (sigma 1.5 px, peaks U(0.1, 0.5) rad)
    phases = np.stack([synth.gaussian_atom_phase((512, 512), 530, 1.5, (0.1, 0.5), rng) for _ in range(20)])
    case["obj"] = np.exp(1j * phases).astype(np.complex64)
    case["positions"] = (np.array([[250.3, 240.6], [262.7, 251.2]])).astype(np.float32)
    fw = po.forward(case["obj"], case["probe"], case["positions"], case["propagators"])
    case = attach_targets(case, fw["I"])
    got = _run(case, cuda, desc_grad=False)
    seg = _run(case, cuda, desc_grad=False, segment_slices=5)
    ref = _oracle(case)
    assert rel(got["I"], ref["I"]) < FWD_TOL
    for k in ("obj", "probe", "positions"):
        print(f"  {k}: recompute vs stored {rel(seg[k], got[k]):.2e}", flush=True)
        assert rel(seg[k], got[k]) < 1e-6, k
    # At this depth float32 itself departs from float64 by more than 1e-4 (the reference's own op order
    # in complex64 gives ~2e-4 for the object gradient), so gradients are held to
    # max(1e-4, 2 x the reference-order complex64 error on the same inputs).
    from oracle import torch_ref

    f32 = torch_ref.loss_and_grads(case["obj"], case["probe"], case["positions"], case["propagators"], case["targets"],
                                   case["mask"], case["num_gpts"], case["mean_intensity"], dtype=torch.complex64)
    for k in ("obj", "probe", "positions"):
        tol = max(GRAD_TOL, 2.0 * rel(f32[k], ref[k]))
        print(f"  {k}: kernel {rel(got[k], ref[k]):.2e}, complex64 reference {rel(f32[k], ref[k]):.2e}", flush=True)
        assert rel(got[k], ref[k]) < tol, k
    obj = case["obj"].astype(np.complex128)
    probe = case["probe"].astype(np.complex128)
    props = case["propagators"].astype(np.complex128)

    def L(o=obj, p=probe):
        I = po.forward(o, p, case["positions"], props)["I"]
        return po.batch_loss(I, case["targets"], case["mask"], case["num_gpts"], case["mean_intensity"])

    # The 32 + 32 directional derivatives are compared as vectors: their relative L2 error stays at the
    # level of float32's own departure, while a sign or scale slip on any subset of them, small-gradient
    # entries included, moves it by the size of that subset's share.  The oracle's own gradient is held
    # to the finite differences far more tightly (float64 throughout).
    h = 1e-4
    rows, cols = po.patch_rows_cols(case["positions"], 256, obj.shape[1:])
    fd_o, k_o, r_o, fd_p, k_p, r_p = [], [], [], [], [], []
    for t in range(32):
        s = rng.integers(0, 20)
        i, j = rows[0, rng.integers(96, 160)], cols[0, rng.integers(96, 160)]
        e = np.zeros_like(obj)
        e[s, i, j] = h
        fd_o.append((L(o=obj + e) - L(o=obj - e)) / (2 * h))
        k_o.append(got["obj"][s, i, j].real)
        r_o.append(ref["obj"][s, i, j].real)
        mi, a, c = rng.integers(0, 2), rng.integers(0, 256), rng.integers(0, 256)
        e = np.zeros_like(probe)
        e[mi, a, c] = 1j * h
        fd_p.append((L(p=probe + e) - L(p=probe - e)) / (2 * h))
        k_p.append(got["probe"][mi, a, c].imag)
        r_p.append(ref["probe"][mi, a, c].imag)
    print(f"  finite differences: object kernel {rel(k_o, fd_o):.2e} oracle {rel(r_o, fd_o):.2e}; "
          f"probe kernel {rel(k_p, fd_p):.2e} oracle {rel(r_p, fd_p):.2e}", flush=True)
    assert rel(r_o, fd_o) < 1e-6 and rel(r_p, fd_p) < 1e-6
    assert rel(k_o, fd_o) < 5e-4 and rel(k_p, fd_p) < 5e-4


def test_deterministic(cuda):
    case = _case_with_targets(N=64, S=3, M=2, B=16, seed=11)
    a = _run(case, cuda)
    b = _run(case, cuda)
    for k in ("I", "obj", "probe", "positions"):
        assert np.array_equal(a[k], b[k]), k
    assert a["loss"] == b["loss"]


def test_object_equal_to_patch_size_wraps(cuda):
    """Object as large as the patch (the reference test's 64x64 object with 64x64 patches)."""
    case = _case_with_targets(N=64, S=1, M=2, B=5, H=64, W=64, seed=5)
    got = _run(case, cuda)
    ref = _oracle(case)
    assert rel(got["I"], ref["I"]) < FWD_TOL
    assert rel(got["obj"], ref["obj"]) < GRAD_TOL


def test_argument_errors(cuda):
    case = _case_with_targets(N=32, S=2, M=1, B=2, seed=1)
    dev = cuda
    args = dict(num_gpts=8, mean_intensity=1.0)
    obj, probe, pos = _t(case["obj"], dev), _t(case["probe"], dev), _t(case["positions"], dev)
    props, tg = _t(case["propagators"], dev), _t(case["targets"], dev)
    with pytest.raises(TypeError):
        qp.multislice_loss(obj.to(torch.complex128), probe, pos, props, tg, **args)
    with pytest.raises(ValueError):
        qp.multislice_loss(obj, probe, pos, props[:0], tg, **args)
    with pytest.raises(ValueError):
        qp.multislice_loss(obj, probe, pos, props, tg, loss_type="l3", **args)
    with pytest.raises(ValueError):
        qp.multislice_loss(obj, probe, pos, props, tg, num_gpts=8, mean_intensity=0.0)


@pytest.mark.parametrize("n,s,m,b,samp", [(32, 1, 1, 5, None), (64, 3, 2, 4, None), (128, 4, 3, 3, None),
                                          (256, 3, 2, 2, None), (256, 3, 2, 2, 0.1), (256, 1, 1, 2, None)])
def test_analytical_update_matches_oracle(cuda, n, s, m, b, samp):
    """The reference's autograd=False update (Fourier projection + slice replay); at 256 x 256 with
    0.1 A sampling the propagators run as row x column factors, at 2/256 A entry by entry."""
    print(f"analytical n={n} s={s} m={m} sampling={samp}", flush=True)
    case = _case_with_targets(N=n, S=s, M=m, B=b, seed=300 + n, descan=False, sampling=samp)
    dev = cuda
    props = _t(case["propagators"], dev) if s > 1 else None
    loss, og, pg = qp.analytical_gradients(_t(case["obj"], dev), _t(case["probe"], dev), _t(case["positions"], dev),
                                           props, _t(case["targets"], dev), _t(case["mask"], dev),
                                           num_gpts=case["num_gpts"], mean_intensity=case["mean_intensity"])
    ref_o, ref_p = po.analytical_update(case["obj"], case["probe"], case["positions"],
                                        case["propagators"] if s > 1 else None, case["targets"])
    assert rel(og.cpu().numpy(), ref_o) < GRAD_TOL
    assert rel(pg.cpu().numpy(), ref_p) < GRAD_TOL


@pytest.mark.parametrize("n", [64, 256])
def test_analytical_update_potential_matches_oracle(cuda, n):
    """The analytical update of a potential object (O = exp(iV)): real(-i conj(O) update)."""
    print(f"analytical potential n={n}", flush=True)
    case = _case_with_targets(N=n, S=2, M=2, B=3, seed=400 + n, descan=False)
    v = np.random.default_rng(n).uniform(0.0, 0.5, case["obj"].shape)
    o = np.exp(1j * v).astype(np.complex64)
    dev = cuda
    loss, og, pg = qp.analytical_gradients(_t(o, dev), _t(case["probe"], dev), _t(case["positions"], dev),
                                           _t(case["propagators"], dev), _t(case["targets"], dev),
                                           _t(case["mask"], dev), num_gpts=case["num_gpts"],
                                           mean_intensity=case["mean_intensity"], obj_type="potential")
    ref_o, ref_p = po.analytical_update(o, case["probe"], case["positions"], case["propagators"], case["targets"],
                                        obj_type="potential")
    assert og.dtype == torch.float32
    assert rel(og.cpu().numpy(), ref_o.real) < GRAD_TOL
    assert rel(pg.cpu().numpy(), ref_p) < GRAD_TOL


@pytest.mark.parametrize("n,s,m,b,h,w", [(32, 2, 2, 70, 56, 48), (64, 2, 2, 130, 96, 80), (32, 2, 2, 5, 40, 6400)])
def test_object_reduce_chunks_and_wide_rows(cuda, n, s, m, b, h, w):
    """Batches larger than the row reduce's N-pattern chunks, and rows wider than its LDS row
    (6400 columns: the 32 x 32 tile reduce)."""
    print(f"reduce n={n} b={b} obj=({h},{w})", flush=True)
    case = _case_with_targets(N=n, S=s, M=m, B=b, H=h, W=w, seed=700 + b)
    got = _run(case, cuda)
    ref = _oracle(case)
    assert rel(got["I"], ref["I"]) < FWD_TOL
    for k in ("obj", "probe", "positions"):
        assert rel(got[k], ref[k]) < GRAD_TOL, k


def test_headline_size_batch_is_the_sum_of_its_parts(cuda):
    """The headline shape (8 x 384^2, 128^2, 4 modes, batch 256): per-pattern intensities are
    bitwise those of four 64-pattern sub-batches, and loss and gradients are the sub-batch sums
    (the loss carries P / B, hence the factor B_sub / B).  The oracle pins two patterns' intensities."""
    print("headline-size batch decomposition", flush=True)
    case = make_case(N=128, S=8, M=4, B=256, H=384, W=384, seed=8)
    dev = cuda
    props = _t(case["propagators"], dev)
    obj_t, probe_t = _t(case["obj"], dev), _t(case["probe"], dev)
    pos_t = _t(case["positions"], dev)
    tg = qp.forward_intensities(obj_t * 0.98, probe_t, pos_t, props).clamp_min(0).sqrt()
    args = dict(num_gpts=4096, mean_intensity=float((tg ** 2).sum(dim=(-2, -1)).mean()))

    def run(sl):
        o = obj_t.clone().requires_grad_(True)
        p = probe_t.clone().requires_grad_(True)
        q = pos_t[sl].clone().requires_grad_(True)
        loss, I = qp.multislice_loss(o, p, q, props, tg[sl], None, None, return_intensities=True, **args)
        loss.backward()
        return loss.detach(), I, o.grad, p.grad, q.grad

    full = run(slice(0, 256))
    parts = [run(slice(k, k + 64)) for k in range(0, 256, 64)]
    f = 64 / 256
    assert torch.equal(full[1], torch.cat([pt[1] for pt in parts]))
    assert abs(float(full[0]) - f * sum(float(pt[0]) for pt in parts)) < 1e-5 * abs(float(full[0]))
    for i in (2, 3):
        ref = f * sum(pt[i] for pt in parts)
        assert rel(full[i].cpu().numpy(), ref.cpu().numpy()) < 1e-5, i
    assert rel(full[4].cpu().numpy(), (f * torch.cat([pt[4] for pt in parts])).cpu().numpy()) < 1e-5
    sub = po.forward(case["obj"], case["probe"], case["positions"][:2], case["propagators"])["I"]
    assert rel(full[1][:2].cpu().numpy(), sub) < FWD_TOL


@pytest.mark.parametrize("n,s,m,b,descan,samp", [(32, 3, 2, 4, True, None), (64, 4, 2, 3, False, None),
                                                 (128, 3, 2, 3, False, None), (128, 8, 4, 2, False, None),
                                                 (256, 3, 2, 2, True, None), (256, 3, 2, 2, True, 0.1)])
def test_propagator_gradient_matches_oracle(cuda, n, s, m, b, descan, samp):
    """dL/dpropagators (the gradient a learned probe tilt chains through) against the FP64 oracle; the
    other gradients are unchanged by the extra output.  At 256 x 256 with 0.1 A sampling the propagators
    factor, so the gradient comes from the kernels that apply them as row x column factors."""
    print(f"propagator gradient n={n} s={s} m={m} b={b} descan={descan} sampling={samp}", flush=True)
    case = _case_with_targets(N=n, S=s, M=m, B=b, seed=900 + n + s, descan=descan, sampling=samp)
    dev = cuda
    obj = _t(case["obj"], dev).requires_grad_(True)
    probe = _t(case["probe"], dev).requires_grad_(True)
    props = _t(case["propagators"], dev).requires_grad_(True)
    loss = qp.multislice_loss(obj, probe, _t(case["positions"], dev), props, _t(case["targets"], dev),
                              _t(case["mask"], dev), _t(case["descan"], dev), num_gpts=case["num_gpts"],
                              mean_intensity=case["mean_intensity"])
    loss.backward()
    ref = _oracle(case)
    e = rel(props.grad.cpu().numpy(), ref["propagators"])
    print(f"  propagators {e:.2e}", flush=True)
    assert e < GRAD_TOL
    assert rel(obj.grad.cpu().numpy(), ref["obj"]) < GRAD_TOL
    assert rel(probe.grad.cpu().numpy(), ref["probe"]) < GRAD_TOL
    if s > 2:  # with recomputation segments the propagator gradient is refused on the host
        with pytest.raises(ValueError):
            qp.multislice_loss(obj, probe, _t(case["positions"], dev), props, _t(case["targets"], dev),
                               _t(case["mask"], dev), None, num_gpts=case["num_gpts"],
                               mean_intensity=case["mean_intensity"], segment_slices=2)


def _tilt_props64(n, sampling, dz, tilt):
    """float64 Fresnel propagators with a tilt (the formula of probe_models.py:404-425)."""
    lam = synth.electron_wavelength_angstrom(300e3)
    k = np.fft.fftfreq(n, sampling)
    k2 = k[:, None] ** 2 + k[None, :] ** 2
    out = []
    for d in dz:
        pr = np.exp(-1j * np.pi * lam * d * k2)
        pr = pr * np.exp(-2j * np.pi * d * np.tan(tilt[0] / 1e3) * k)[:, None]
        pr = pr * np.exp(-2j * np.pi * d * np.tan(tilt[1] / 1e3) * k)[None, :]
        out.append(pr)
    return np.stack(out)


def test_probe_tilt_gradient_matches_finite_differences(cuda):
    """The probe-tilt gradient through the kernels' propagator gradient and the torch propagator builder
    against complex128 central finite differences of the oracle loss in the tilt (h = 1e-3 mrad)."""
    n, S, samp = 64, 4, 0.1
    dz = [8.0] * (S - 1)
    tilt = np.array([6.0, -4.0])
    case = make_case(N=n, S=S, M=2, B=4, seed=31, sampling=samp)
    case["propagators"] = _tilt_props64(n, samp, dz, tilt).astype(np.complex64)
    I = po.forward(case["obj"], case["probe"], case["positions"], case["propagators"])["I"]
    case = attach_targets(case, I)
    dev = cuda
    th = torch.tensor(tilt, dtype=torch.float32, device=dev, requires_grad=True)
    props = qp.fresnel_propagators(n, samp, 300e3, dz, th)
    loss = qp.multislice_loss(_t(case["obj"], dev), _t(case["probe"], dev), _t(case["positions"], dev), props,
                              _t(case["targets"], dev), _t(case["mask"], dev), num_gpts=case["num_gpts"],
                              mean_intensity=case["mean_intensity"])
    loss.backward()
    got = th.grad.cpu().numpy()

    def L(t):
        I = po.forward(case["obj"], case["probe"], case["positions"], _tilt_props64(n, samp, dz, t))["I"]
        return po.batch_loss(I, case["targets"], case["mask"], case["num_gpts"], case["mean_intensity"])

    h = 1e-3
    fd = np.array([(L(tilt + h * e) - L(tilt - h * e)) / (2 * h) for e in np.eye(2)])
    print(f"tilt gradient: kernels {got}, finite differences {fd}", flush=True)
    assert np.allclose(got, fd, rtol=2e-3, atol=1e-4 * np.abs(fd).max())
