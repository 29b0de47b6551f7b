"""Training data and training of the warm-start network.

Pairs come from the GPU ICD itself: for seeded 256x256x16 ellipsoid scans
with 720 views over [0,180), I0 drawn from [5e3, 5e4] and 0-1% zingers, the
input is the filtered backprojection (the ICD state's FBP start) and the target
is the volume after `icd_iters` ICD iterations from that start.  The network
(xt_hip.UNet, the library's BF16 MFMA kernels) is trained with Adam on the
L2 loss, over the eight symmetries of the square (y, x) plane.

    warmstart.main(["--pairs", "256", "--out", "weights.npz"])

Volumes are [N_y][N_x][N_z] (the ICD's device object layout), attenuation per
voxel; the network sees them times `scale`.
"""
from __future__ import annotations

import argparse
import dataclasses
import os
import shutil
import sys
import tempfile
import time

import numpy as np

from . import synth, xt_hip

TRAIN = synth.ScanConfig(name="warmstart_256x256x16", N_xy=256, N_z=16, N_p=720, I0=1e4,
                         zinger_frac=0.005, num_threads=16, iterations=30)
SCALE = 1000.0


def pair_config(i: int, seed0: int) -> synth.ScanConfig:
    rng = np.random.default_rng(seed0 + i)
    return dataclasses.replace(TRAIN, name=f"ws{seed0 + i}", seed=seed0 + i,
                               I0=float(rng.uniform(5e3, 5e4)),
                               zinger_frac=float(rng.uniform(0.0, 0.01)))


def gpu_projector(cfg: synth.ScanConfig, path: str):
    """Ax of a phantom with the library's forward projector (the engine's A)."""
    def project(ph):
        eng = xt_hip.Engine(path, cfg.cli_args(), read_sinogram=False)
        try:
            eng.set_volume(ph.astype(np.float32))
            return xt_hip.forward_project(eng, xt_hip.detector_response(eng))
        finally:
            eng.close()
    return project


def to_net_layout(vol_zyx: np.ndarray) -> np.ndarray:
    """[N_z][N_y][N_x] -> [N_y][N_x][N_z]."""
    return np.ascontiguousarray(np.transpose(vol_zyx, (1, 2, 0)), np.float32)


def generate_pairs(n: int, icd_iters: int = 30, group: int = 8, seed0: int = 5000, log=print):
    """Yields (fbp, target) volumes in the network layout, `group` scans at a
    time advanced together by one xt_icd_batch."""
    root = tempfile.mkdtemp(prefix="xt_ws_")
    xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 1)
    try:
        for g0 in range(0, n, group):
            cfgs = [pair_config(i, seed0) for i in range(g0, min(n, g0 + group))]
            engs, icds, fbps = [], [], []
            try:
                for c in cfgs:
                    p = os.path.join(root, c.name)
                    synth.generate(c, p, gpu_projector(c, p))
                    e = xt_hip.Engine(p, c.cli_args())
                    engs.append(e)
                    icd = xt_hip.Icd(e)
                    icds.append(icd)
                    icd.fbp()
                    icd.download()
                    fbps.append(to_net_layout(e.volume()[0]))
                with xt_hip.IcdBatch(icds) as b:
                    b.iterate(icd_iters)
                for c, e, icd, f in zip(cfgs, engs, icds, fbps):
                    icd.download()
                    yield f, to_net_layout(e.volume()[0])
            finally:
                for icd in icds:
                    icd.close()
                for e in engs:
                    e.close()
                for c in cfgs:
                    shutil.rmtree(os.path.join(root, c.name), ignore_errors=True)
            log(f"pairs {min(n, g0 + group)}/{n}")
    finally:
        shutil.rmtree(root, ignore_errors=True)


def symmetry(vol: np.ndarray, k: int) -> np.ndarray:
    """The k-th (0..7) symmetry of the square in the (y, x) plane."""
    v = np.rot90(vol, k % 4, axes=(0, 1))
    if k >= 4:
        v = v[:, ::-1]
    return np.ascontiguousarray(v)


def train(net: xt_hip.UNet, pairs, epochs: int, lr: float, seed: int = 0, log=print):
    """Adam over `epochs` passes of the pairs, one random symmetry per step,
    learning rate lr with a cosine decay to lr/20.  Returns the mean loss per
    epoch."""
    rng = np.random.default_rng(seed)
    steps = epochs * len(pairs)
    hist, s = [], 0
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        tot = 0.0
        for i in order:
            k = int(rng.integers(8))
            f, t = pairs[i]
            cur = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * s / max(1, steps))))
            tot += net.train_step(symmetry(f, k), symmetry(t, k), cur)
            s += 1
        hist.append(tot / len(pairs))
        log(f"epoch {ep + 1}/{epochs}: mean loss {hist[-1]:.4e}")
    return hist


def rmse(a, b, mean_att):
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - b) ** 2)) / mean_att)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--pairs", type=int, default=256)
    ap.add_argument("--holdout", type=int, default=8)
    ap.add_argument("--icd-iters", type=int, default=30)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    ap.add_argument("--log", default="", help="optional .npz for the loss curve and held-out "
                    "errors (kept apart from the weights)")
    a = ap.parse_args(argv)

    def log(*m):
        print("[warmstart]", *m, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    pairs = list(generate_pairs(a.pairs, a.icd_iters, log=log))
    hold = list(generate_pairs(a.holdout, a.icd_iters, seed0=900000, log=log))
    log(f"{len(pairs)} training and {len(hold)} held-out pairs in {time.perf_counter() - t0:.0f} s")
    shape = pairs[0][0].shape
    with xt_hip.UNet(shape, SCALE, seed=a.seed, train=True) as net:
        t0 = time.perf_counter()
        hist = train(net, pairs, a.epochs, a.lr, seed=a.seed, log=log)
        t_train = time.perf_counter() - t0
        f_ms, b_ms = net.last_ms()
        ev = []
        for f, t in hold:
            m = float(t[t > 2.5e-4].mean())  # mean attenuation inside the objects
            ev.append((rmse(f, t, m), rmse(net.forward(f), t, m)))
        net.save(a.out)
    ev = np.array(ev)
    log(f"trained {a.epochs} epochs x {len(pairs)} pairs in {t_train:.0f} s "
        f"(device fwd {f_ms:.2f} ms + bwd {b_ms:.2f} ms per step)")
    log(f"held-out RMSE to the ICD target, relative to mean attenuation: FBP "
        f"{ev[:, 0].mean():.3%}, network {ev[:, 1].mean():.3%}")
    if a.log:
        np.savez(a.log, loss=np.array(hist), holdout=ev, pairs=a.pairs, icd_iters=a.icd_iters,
                 epochs=a.epochs, lr=a.lr, train_seconds=t_train)
    return 0


if __name__ == "__main__":
    sys.exit(main())
