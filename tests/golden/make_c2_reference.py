"""Writes tests/golden/c2_cpu_icd_sample.npz: the CPU ICD (the oracle's C
restatement of the reference engine, float build, OpenMP) run for 30 iterations
on the production slab (512x512x32, 1024 views, 768x32 bins; synth.C2), sampled
at 262144 seeded voxel positions.  The benchmark measures the GPU's time to
come within 1% RMSE (relative to the phantom's mean attenuation) of it.

    python tests/golden/make_c2_reference.py RUN_DIR [--threads 8] [--iters 30]

RUN_DIR receives the scan and the per-iteration object snapshots; a folder that
already holds object_n0_time_0_iter_<iters>.bin is sampled without running
again.  The run takes hours on 8 cores.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import conftest  # noqa: E402,F401  (registers the package)
import py_oracle  # noqa: E402
from timbir_python_amd import synth  # noqa: E402

N_SAMPLE = 262144
SEED = 20261202

ap = argparse.ArgumentParser()
ap.add_argument("run_dir")
ap.add_argument("--threads", type=int, default=8)
ap.add_argument("--iters", type=int, default=30)
a = ap.parse_args()
cfg = synth.C2
d = a.run_dir
args = cfg.cli_args(iterations=a.iters)
args[args.index("--num_threads") + 1] = str(a.threads)
final = os.path.join(d, f"object_n0_time_0_iter_{a.iters}.bin")
if not os.path.exists(final):
    os.makedirs(d, exist_ok=True)
    if not os.path.exists(os.path.join(d, "weight_n0.bin")):
        o64a = py_oracle.Oracle(True, float_amatrix=True)
        synth.generate(cfg, d, o64a.projector(d, args))
    o = py_oracle.Oracle(False)
    n, arr = py_oracle.argv_array(args)
    costs = np.full(a.iters + 1, np.nan)
    it = C.c_int32()
    t0 = time.time()
    rc = o.lib.oracle_engine_run(d.encode(), n, arr, 1 | 4, costs.ctypes.data, C.byref(it))
    assert rc == 0, rc
    print("oracle ICD", it.value, "iterations in", time.time() - t0, "s")

nvox = cfg.N_xy * cfg.N_xy * cfg.N_z
idx = np.sort(np.random.default_rng(SEED).choice(nvox, N_SAMPLE, replace=False)).astype(np.int32)
ph = synth.phantom(cfg)[0]
mean_att = float(ph[ph > 0].mean())
vol = lambda k: np.fromfile(os.path.join(d, f"object_n0_time_0_iter_{k}.bin"), np.float32)
ref = vol(a.iters)
assert ref.size == nvox
# the CPU run's own distance to its last iteration, as RMSE / mean attenuation
own = {}
for k in range(1, a.iters):
    p = os.path.join(d, f"object_n0_time_0_iter_{k}.bin")
    if os.path.exists(p):
        own[k] = float(np.sqrt(np.mean((vol(k) - ref) ** 2)) / mean_att)
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_cpu_icd_sample.npz"),
                    idx=idx, values=ref[idx], mean_att=np.float64(mean_att),
                    iterations=np.int32(a.iters), threads=np.int32(a.threads),
                    layout=np.array("[N_z][N_y][N_x] float32, frame 0"),
                    own_rmse=np.array(json.dumps(own)))
print("mean attenuation", mean_att, "own RMSE to last:", {k: round(v, 5) for k, v in own.items()})
