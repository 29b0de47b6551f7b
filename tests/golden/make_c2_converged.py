"""Writes tests/golden/c2_converged_sample.npz: the MAP image of the production
slab (512x512x32, 1024 views; synth.C2) reached by running the GPU ICD to
convergence from the filtered-backprojection start (600 iterations, FP64 resync
every iteration; it moves by 0.04% RMSE over its last 50 iterations and ends at a
lower MAP cost than any zero-start run reaches in 400 iterations), sampled at the
same 262144 seeded voxels as c2_cpu_icd_sample.npz.  The benchmark's time to
converge is measured against it.

    python tests/golden/make_c2_converged.py [--iters 600] [--zero-start] [--from-volume FILE]

Needs the GPU (or, with --from-volume, a float32 [N_z][N_y][N_x] volume that such
a run wrote).
"""
import argparse
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

N_SAMPLE, SEED = 262144, 20261202

ap = argparse.ArgumentParser()
ap.add_argument("--iters", type=int, default=600)
ap.add_argument("--zero-start", action="store_true")
ap.add_argument("--from-volume", default=None)
a = ap.parse_args()
bench.load_pkg()
from timbir_python_amd import synth, xt_hip  # noqa: E402

cfg = synth.C2
if a.from_volume:
    vol = np.fromfile(a.from_volume, np.float32)
else:
    work = tempfile.mkdtemp(prefix="xt_conv_")
    bench.make_workload(cfg, work, xt_hip, synth)
    xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 1)
    with xt_hip.Engine(work, cfg.cli_args()) as eng, xt_hip.Icd(eng) as icd:
        if not a.zero_start:
            icd.fbp()
        icd.iterate(a.iters)
        cost = icd.cost()
        icd.download()
        vol = eng.volume()[0].reshape(-1)
    print(f"MAP cost after {a.iters} iterations: {cost:.10g}")
nvox = cfg.N_xy * cfg.N_xy * cfg.N_z
assert vol.size == nvox
idx = np.sort(np.random.default_rng(SEED).choice(nvox, N_SAMPLE, replace=False)).astype(np.int32)
ph = synth.phantom(cfg)[0]
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_converged_sample.npz"),
                    idx=idx, values=vol[idx], mean_att=np.float64(ph[ph > 0].mean()),
                    iterations=np.int32(a.iters), fbp_start=np.int32(0 if a.zero_start else 1),
                    layout=np.array("[N_z][N_y][N_x] float32, frame 0"))
