"""Pins of the warm-start network's CPU restatement (oracle/unet_oracle.py).

The network has no reference counterpart; the oracle is pinned by an
independent correlation routine, adjoint identities and central finite
differences of its own float64 loss, so that a transposed weight, a wrong tap
sign or a dropped skip-connection term fails here."""
import os
import sys

import numpy as np
import pytest
import scipy.ndimage

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import unet_oracle as uo  # noqa: E402


def test_bf16_rounding_bits():
    x = np.array([1.0, 1 + 2**-8, 1 + 3 * 2**-8, 1 + 2**-7 + 2**-9, -2.5, 65536.0 + 256.0],
                 np.float32)
    # ties go to the even mantissa; 65792 = 2^16 (1 + 2^-8) is a tie -> 65536
    want = np.array([1.0, 1.0, 1 + 2**-6, 1 + 2**-7, -2.5, 65536.0], np.float32)
    np.testing.assert_array_equal(uo.bf16_round(x), want)
    assert (uo.bf16_round(x).view(np.uint32) & 0xFFFF).max() == 0


def test_conv3_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((6, 5, 4, 3))
    w = rng.standard_normal((2, 27, 3))
    b = rng.standard_normal(2)
    got = uo.conv3(x, w, b)
    for co in range(2):
        ref = sum(scipy.ndimage.correlate(x[..., ci], w[co, :, ci].reshape(3, 3, 3),
                                          mode="constant", cval=0.0) for ci in range(3)) + b[co]
        np.testing.assert_allclose(got[..., co], ref, rtol=1e-12, atol=1e-12)


def test_adjoint_identities():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 6, 2, 3))
    w = rng.standard_normal((5, 27, 3))
    g = rng.standard_normal((4, 6, 2, 5))
    lhs = np.vdot(uo.conv3(x, w), g)
    assert np.isclose(lhs, np.vdot(x, uo.conv3_dgrad(g, w)), rtol=1e-12)
    assert np.isclose(lhs, np.vdot(w, uo.conv3_wgrad(x, g)), rtol=1e-12)
    y = rng.standard_normal((8, 4, 6, 2))
    c = rng.standard_normal((4, 2, 3, 2))
    assert np.isclose(np.vdot(uo.up(c), y), np.vdot(c, uo.sum_children(y)), rtol=1e-12)
    assert np.isclose(np.vdot(uo.up(c), y), 8 * np.vdot(c, uo.pool(y)), rtol=1e-12)


def _random_params(rng, head_scale=0.3):
    p = []
    for _, ci, co in uo.SPEC:
        p.append(rng.standard_normal(co * 27 * ci) * np.sqrt(2.0 / (27 * ci)))
        p.append(rng.uniform(0.0, 0.1, co))
    p.append(rng.standard_normal(uo.HEAD_C) * head_scale)
    p.append([0.05])
    return np.concatenate(p)


def test_param_count():
    assert uo.param_count() == _random_params(np.random.default_rng(0)).size
    assert uo.layer_slices()[-1][1].stop == uo.param_count()


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    params = _random_params(rng)
    x = rng.uniform(0.0, 2e-3, (8, 8, 8))
    tgt = x + rng.normal(0.0, 2e-4, x.shape)
    net = uo.UNet(params, scale=1000.0, bf16=False)
    loss, grad = net.loss_grad(x, tgt)
    assert loss > 0
    worst = 0.0
    for name, sl in uo.layer_slices():
        idx = rng.choice(np.arange(sl.start, sl.stop), size=min(3, sl.stop - sl.start),
                         replace=False)
        for i in idx:
            eps = 1e-6 * max(1.0, abs(params[i]))
            pp, pm = params.copy(), params.copy()
            pp[i] += eps
            pm[i] -= eps
            fd = (uo.UNet(pp, 1000.0, bf16=False).loss_grad(x, tgt)[0] -
                  uo.UNet(pm, 1000.0, bf16=False).loss_grad(x, tgt)[0]) / (2 * eps)
            err = abs(fd - grad[i]) / max(abs(fd), 1e-6 * np.abs(grad).max())
            worst = max(worst, err)
            assert err < 1e-4, (name, i, fd, grad[i])
    assert worst < 1e-4


def test_bf16_mode_tracks_float64():
    rng = np.random.default_rng(3)
    params = _random_params(rng)
    x = rng.uniform(0.0, 2e-3, (8, 8, 16))
    tgt = x * 0.9
    l64, g64 = uo.UNet(params, 1000.0, bf16=False).loss_grad(x, tgt)
    l16, g16 = uo.UNet(params, 1000.0, bf16=True).loss_grad(x, tgt)
    assert abs(l16 - l64) / l64 < 0.05
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 0.1


@pytest.mark.parametrize("mistake", ["flip_dgrad", "drop_skip"])
def test_planted_mistakes_fail_the_pin(mistake, monkeypatch):
    """A plausible mistake in the oracle's backward breaks the finite-difference pin."""
    rng = np.random.default_rng(4)
    params = _random_params(rng)
    x = rng.uniform(0.0, 2e-3, (8, 8, 8))
    tgt = x * 0.8
    if mistake == "flip_dgrad":
        def bad(g, w):
            g = g.astype(np.float64)
            out = np.zeros(g.shape[:3] + (w.shape[2],))
            for t, a in enumerate(uo.TAPS):
                out += uo.shift(g, a) @ w[:, t, :]  # offset sign not negated
            return out
        monkeypatch.setattr(uo, "conv3_dgrad", bad)
    else:
        orig = uo.up

        def bad_up(c):  # skip-gradient path loses the pooled contribution
            return orig(c) * (0.0 if c.shape[-1] == 16 and c.shape[0] == 4 else 1.0)
        monkeypatch.setattr(uo, "up", bad_up)
    _, grad = uo.UNet(params, 1000.0, bf16=False).loss_grad(x, tgt)
    monkeypatch.undo()
    i = uo.layer_slices()[0][1].start + 5  # a stem weight: every backward path reaches it
    eps = 1e-6
    pp, pm = params.copy(), params.copy()
    pp[i] += eps
    pm[i] -= eps
    fd = (uo.UNet(pp, 1000.0, bf16=False).loss_grad(x, tgt)[0] -
          uo.UNet(pm, 1000.0, bf16=False).loss_grad(x, tgt)[0]) / (2 * eps)
    assert abs(fd - grad[i]) > 1e-3 * abs(fd)
