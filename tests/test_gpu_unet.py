"""The warm-start network's HIP kernels (BF16 MFMA forward, backward, Adam)
against the CPU restatement in oracle/unet_oracle.py, and its use as an ICD
starting volume.

Tolerances: the oracle reproduces the GPU's bfloat16 rounding points and sums
in float64, so GPU and oracle differ only by FP32 summation order (and the
rare bf16 rounding flip it causes): relative L2 error <= 1e-2 for the network
output's residual part and <= 3e-2 for every parameter block's gradient,
1e-3 for the loss."""
import numpy as np
import pytest

import unet_oracle as uo
from conftest import SMALL, make_scan
from timbir_python_amd import xt_hip

pytestmark = pytest.mark.gpu

SHAPE = (16, 16, 8)
SCALE = 1000.0


def _params(seed, head_scale=0.3):
    rng = np.random.default_rng(seed)
    p = []
    for _, ci, co in uo.SPEC:
        p.append(rng.standard_normal(co * 27 * ci) * np.sqrt(2.0 / (27 * ci)))
        p.append(rng.uniform(0.0, 0.1, co))
    p.append(rng.standard_normal(uo.HEAD_C) * head_scale)
    p.append([0.05])
    return np.concatenate(p).astype(np.float32)


def _volume(seed, shape=SHAPE):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 2e-3, shape).astype(np.float32)
    return x, (x * 0.85 + 1e-4).astype(np.float32)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_oracle():
    print(f"[gpu] unet forward {SHAPE}", flush=True)
    p = _params(1)
    x, _ = _volume(2)
    with xt_hip.UNet(SHAPE, SCALE, seed=0) as net:
        assert net.n == uo.param_count()
        net.set_params(p)
        np.testing.assert_array_equal(net.get_params(), p)
        got = net.forward(x)
    want = uo.UNet(p, SCALE, bf16=True).forward(x)
    # compare the network's own contribution h = (out - x) * scale
    e = rel((got.astype(np.float64) - x) * SCALE, (want - x) * SCALE)
    print(f"  rel L2 of the residual: {e:.2e}")
    assert e <= 1e-2


def test_gradients_match_oracle():
    print(f"[gpu] unet loss/grad {SHAPE}", flush=True)
    p = _params(3)
    x, t = _volume(4)
    with xt_hip.UNet(SHAPE, SCALE, seed=0, train=True) as net:
        net.set_params(p)
        loss, g = net.loss_grad(x, t, want_grad=True)
    wl, wg = uo.UNet(p, SCALE, bf16=True).loss_grad(x, t)
    print(f"  loss gpu {loss:.6e} oracle {wl:.6e}")
    assert abs(loss - wl) / wl <= 1e-3
    worst = ("", 0.0)
    for name, sl in uo.layer_slices():
        e = rel(g[sl], wg[sl])
        if e > worst[1]:
            worst = (name, e)
        assert e <= 3e-2, (name, e)
    print(f"  worst gradient block {worst[0]}: rel L2 {worst[1]:.2e}")


def test_training_reduces_loss():
    print("[gpu] unet training 40 steps", flush=True)
    rng = np.random.default_rng(7)
    shape = (16, 16, 16)
    x = rng.uniform(0.0, 2e-3, shape).astype(np.float32)
    tgt = (0.5 * x + 0.5 * np.roll(x, 1, 0)).astype(np.float32)  # a local linear map
    with xt_hip.UNet(shape, SCALE, seed=3, train=True) as net:
        losses = [net.train_step(x, tgt, 1e-3) for _ in range(40)]
        f_ms, b_ms = net.last_ms()
    print(f"  loss {losses[0]:.4e} -> {losses[-1]:.4e}; fwd {f_ms:.2f} ms bwd {b_ms:.2f} ms")
    assert losses[-1] < 0.5 * losses[0]


def test_bad_shapes_are_errors():
    with pytest.raises(xt_hip.XtError):
        xt_hip.UNet((12, 16, 8))
    with xt_hip.UNet(SHAPE) as net:
        with pytest.raises(xt_hip.XtError):
            net.set_params(np.zeros(3, np.float32))


def test_icd_unet_start_is_network_on_fbp(scan_root):
    """xt_icd_unet_start leaves the network's output on the FBP volume in the
    ICD state (object layout [N_y][N_x][N_z], computed in place) and a
    resynchronised e."""
    cfg = SMALL
    path, _ = make_scan(cfg, scan_root, "unet")
    shape = (cfg.N_xy, cfg.N_xy, cfg.N_z)
    p = _params(8, head_scale=0.05)
    with xt_hip.Engine(path, cfg.cli_args()) as eng:
        with xt_hip.Icd(eng) as icd:
            icd.fbp()
            icd.download()
        fbp = eng.volume()[0].transpose(1, 2, 0).copy()  # [z][y][x] -> [y][x][z]
        with xt_hip.UNet(shape, SCALE) as net, xt_hip.Icd(eng) as icd:
            net.set_params(p)
            want = net.forward(fbp)
            xt_hip.icd_unet_start(icd, net)
            c_start = icd.cost()
            icd.download()
            got = eng.volume()[0].transpose(1, 2, 0)
            icd.resync()
            assert abs(icd.cost() - c_start) <= 1e-9 * abs(c_start)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-7 * np.abs(want).max())


def test_trained_network_gains_over_fbp_start(scan_root):
    """The committed weights (trained on 256^2x16, 720-view pairs) on a held-out
    128^2x16 scan with the same view density: the network's start is closer to
    the converged image than the FBP start, and stays closer after 3 ICD
    iterations."""
    import os
    import dataclasses
    from timbir_python_amd import synth
    weights = os.path.join(os.path.dirname(xt_hip.__file__), "weights",
                           "unet_warmstart_256x256x16.npz")
    cfg = dataclasses.replace(synth.C1, name="unet_gain", N_xy=128, N_z=16, N_p=360, I0=2e4,
                              zinger_frac=0.005, num_threads=16, seed=424242)
    from timbir_python_amd import warmstart
    path = os.path.join(str(scan_root), "unet_gain")
    synth.generate(cfg, path, warmstart.gpu_projector(cfg, path))
    shape = (cfg.N_xy, cfg.N_xy, cfg.N_z)
    xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 1)
    print("[gpu] unet gain: reference ICD from FBP, 80 iterations", flush=True)
    with xt_hip.Engine(path, cfg.cli_args()) as eng:
        with xt_hip.Icd(eng) as icd:
            icd.fbp()
            icd.iterate(80)
            icd.download()
        ref = eng.volume()[0]
        mean_att = float(ref[ref > 2.5e-4].mean())
        curves = {}
        with xt_hip.UNet.load(weights, shape) as net:
            for start in ("fbp", "unet"):
                with xt_hip.Icd(eng) as icd:
                    if start == "fbp":
                        icd.fbp()
                    else:
                        xt_hip.icd_unet_start(icd, net)
                    c = []
                    for it in range(4):
                        if it:
                            icd.iterate(1)
                        icd.download()
                        c.append(rel(eng.volume()[0], ref) * np.linalg.norm(ref) /
                                 np.sqrt(ref.size) / mean_att)
                    curves[start] = c
    print("  RMSE to the converged image after 0..3 iterations: "
          + "; ".join(f"{k} " + " ".join(f"{v:.2%}" for v in c) for k, c in curves.items()))
    assert curves["unet"][0] < 0.5 * curves["fbp"][0]
    assert curves["unet"][3] < curves["fbp"][3]
