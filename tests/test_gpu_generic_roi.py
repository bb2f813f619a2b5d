"""GPU parity of the generic-size kernels (Nr x Nc patterns: non-square, non-power-of-two) against the FP64
oracle, through the C ABI.

The reference takes any roi_shape (dataset_models.py:514-539 per-axis patch offsets, ptycho_utils.py:164-170
per-axis ramps, probe_models.py:404-407 per-axis propagator frequencies); the library runs square 32, 64,
128 and 256 on the fixed-size kernels and every other shape with prime factors 2, 3, 5, 7 (2..512 per
axis) on the mixed-radix LDS kernels tested here.  Tolerances as for the fixed sizes: intensities and
loss 1e-5, gradients 1e-4 relative L2 against float64 on the same complex64 inputs.
"""

import numpy as np
import pytest
import torch

from oracle import ptycho_oracle as po
from quantem_amd import ptychography as qp
from tests._cases import attach_targets, make_case, rel

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
GRAD_TOL = 1e-4


def _t(x, dev):
    return None if x is None else torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _case(**kw):
    case = make_case(**kw)
    I = po.forward(case["obj"], case["probe"], case["positions"], case["propagators"], case["descan"])["I"]
    return attach_targets(case, I)


def _run(case, dev, segment_slices=0, grad_props=False):
    obj = _t(case["obj"], dev).requires_grad_(True)
    probe = _t(case["probe"], dev).requires_grad_(True)
    pos = _t(case["positions"], dev).requires_grad_(True)
    props = _t(case["propagators"], dev) if case["obj"].shape[0] > 1 else None
    if props is not None and grad_props:
        props.requires_grad_(True)
    d = _t(case["descan"], dev)
    if d is not None:
        d.requires_grad_(True)
    loss, intens = qp.multislice_loss(obj, probe, pos, props, _t(case["targets"], dev), _t(case["mask"], dev), d,
                                      num_gpts=case["num_gpts"], mean_intensity=case["mean_intensity"],
                                      loss_type=case["loss_type"], return_intensities=True,
                                      segment_slices=segment_slices)
    loss.backward()
    out = {"loss": loss.item(), "I": intens.cpu().numpy(), "obj": obj.grad.cpu().numpy(),
           "probe": probe.grad.cpu().numpy(), "positions": pos.grad.cpu().numpy()}
    if d is not None:
        out["descan"] = d.grad.cpu().numpy()
    if props is not None and grad_props:
        out["propagators"] = props.grad.cpu().numpy()
    return out


def _oracle(case):
    return po.adjoint(case["obj"], case["probe"], case["positions"],
                      case["propagators"] if case["obj"].shape[0] > 1 else None, case["targets"], case["mask"],
                      case["num_gpts"], case["mean_intensity"], case["loss_type"], descan=case["descan"])


@pytest.mark.parametrize(
    "shape,s,m,b,descan",
    [((96, 96), 3, 2, 4, True), ((48, 80), 2, 1, 5, False), ((100, 120), 3, 3, 3, True), ((16, 16), 4, 2, 6, False),
     ((160, 192), 2, 2, 2, False), ((30, 42), 3, 2, 5, True), ((45, 75), 2, 1, 3, False), ((512, 20), 2, 1, 2, True),
     ((8, 8), 2, 1, 4, False), ((24, 300), 2, 2, 3, True), ((36, 50), 1, 2, 4, True)],
)
def test_generic_shape_matches_oracle(cuda, shape, s, m, b, descan):
    print(f"generic shape={shape} s={s} m={m} b={b} descan={descan}", flush=True)
    case = _case(N=shape, S=s, M=m, B=b, seed=300 + shape[0] + shape[1] + s, descan=descan)
    got = _run(case, cuda)
    ref = _oracle(case)
    assert got["I"].shape == (b,) + shape
    errs = {"I": rel(got["I"], ref["I"]), "loss": abs(got["loss"] - ref["loss"]) / abs(ref["loss"])}
    for k in ("obj", "probe", "positions") + (("descan",) if descan else ()):
        errs[k] = rel(got[k], ref[k])
    print("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert errs["I"] < FWD_TOL and errs["loss"] < FWD_TOL
    for k in ("obj", "probe", "positions") + (("descan",) if descan else ()):
        assert errs[k] < GRAD_TOL, k


@pytest.mark.parametrize("loss_type", ["l1_amplitude", "l2_intensity", "l1_intensity", "poisson"])
def test_generic_shape_loss_types(cuda, loss_type):
    print(f"generic loss {loss_type}", flush=True)
    case = _case(N=(36, 60), S=2, M=2, B=4, seed=41, loss_type=loss_type)
    got = _run(case, cuda)
    ref = _oracle(case)
    assert abs(got["loss"] - ref["loss"]) / abs(ref["loss"]) < FWD_TOL
    for k in ("obj", "probe", "positions"):
        assert rel(got[k], ref[k]) < GRAD_TOL, k


@pytest.mark.parametrize("shape,s,k", [((96, 80), 5, 2), ((60, 72), 7, 3)])
def test_generic_shape_segments_match_oracle(cuda, shape, s, k):
    """Checkpointed forward + per-segment recomputation on the generic kernels."""
    print(f"generic segments shape={shape} s={s} k={k}", flush=True)
    case = _case(N=shape, S=s, M=2, B=3, seed=77 + s, descan=True)
    got = _run(case, cuda, segment_slices=k)
    ref = _oracle(case)
    assert abs(got["loss"] - ref["loss"]) / abs(ref["loss"]) < FWD_TOL
    for key in ("obj", "probe", "positions", "descan"):
        assert rel(got[key], ref[key]) < GRAD_TOL, key


def test_generic_shape_propagator_gradient(cuda):
    print("generic propagator gradient (72, 90)", flush=True)
    case = _case(N=(72, 90), S=3, M=2, B=3, seed=5, descan=True)
    got = _run(case, cuda, grad_props=True)
    ref = _oracle(case)
    e = rel(got["propagators"], ref["propagators"])
    print(f"  propagators {e:.2e}", flush=True)
    assert e < GRAD_TOL
    assert rel(got["obj"], ref["obj"]) < GRAD_TOL


def test_generic_shape_batch_is_the_sum_of_its_parts(cuda):
    """A batch of 40 patterns of 120 x 96: the loss and gradients of the whole batch equal the sums over
    two halves (with the batch-size normalisation undone), and two patterns match the oracle."""
    print("generic batch decomposition (120, 96) B=40", flush=True)
    case = _case(N=(120, 96), S=3, M=2, B=40, seed=9, H=400, W=380)
    full = _run(case, cuda)
    halves = []
    for sl in (slice(0, 20), slice(20, 40)):
        c = dict(case)
        c["positions"] = case["positions"][sl]
        c["targets"] = case["targets"][sl]
        halves.append(_run(c, cuda))
    # the l2 loss is sum / (B / P) / mean: halves carry B = 20, the full batch B = 40
    assert abs(full["loss"] - 0.5 * (halves[0]["loss"] + halves[1]["loss"])) < 1e-5 * abs(full["loss"])
    g = 0.5 * (halves[0]["obj"] + halves[1]["obj"])
    assert rel(full["obj"], g) < 1e-5
    assert rel(full["probe"], 0.5 * (halves[0]["probe"] + halves[1]["probe"])) < 1e-5
    small = dict(case)
    small["positions"] = case["positions"][:2]
    small["targets"] = case["targets"][:2]
    I = po.forward(small["obj"], small["probe"], small["positions"], small["propagators"], None)["I"]
    assert rel(full["I"][:2], I) < FWD_TOL


def test_generic_shape_wide_object(cuda):
    """An object row wider than the row-owner reduce's LDS row (6200 px): the tile reduce takes over."""
    print("generic wide object (16, 24) W=6200", flush=True)
    case = _case(N=(16, 24), S=2, M=2, B=4, seed=8, H=40, W=6200)
    got = _run(case, cuda)
    ref = _oracle(case)
    assert abs(got["loss"] - ref["loss"]) / abs(ref["loss"]) < FWD_TOL
    for k in ("obj", "probe", "positions"):
        assert rel(got[k], ref[k]) < GRAD_TOL, k


def test_projection_update_refuses_generic_shape(cuda):
    """The analytical update runs on the fixed sizes only; a 96 x 96 request is refused by the library."""
    case = _case(N=(96, 96), S=2, M=1, B=2, seed=1)
    dev = cuda
    with pytest.raises(ValueError, match="projection update"):
        qp.analytical_gradients(_t(case["obj"], dev), _t(case["probe"], dev), _t(case["positions"], dev),
                             _t(case["propagators"], dev), _t(case["targets"], dev), num_gpts=case["num_gpts"],
                             mean_intensity=case["mean_intensity"])
