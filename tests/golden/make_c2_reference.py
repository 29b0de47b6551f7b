"""Writes tests/golden/c2_cpu_icd_sample.npz: the CPU ICD (the oracle's C
restatement of the reference engine, float build, OpenMP) run for --iters
iterations (the committed file: 20, 32 threads) on the production slab
(512x512x32, 1024 views, 768x32 bins; synth.C2), sampled at 262144 seeded voxel
positions, with the MAP cost of every iteration (the float build's own
evaluator), the FP64 MAP cost of the final object and ring offsets (the f64a
build: the float system matrix in double arithmetic, e = y - Ax - d recomputed)
and the distance of every iteration's object to the converged image of
c2_converged_sample.npz.  The benchmark compares the GPU's object after the same
number of iterations, and its FP64 MAP cost, with these.

    python tests/golden/make_c2_reference.py RUN_DIR [--threads 32] [--iters 30]

--threads sets the engine's num_threads, from which the reference derives its
z-block count (2 * floor(num_threads / N_time), capped at N_z,
Source_Code_4D_Fast/XT_Init.c:396-400): 32 threads give 32 blocks of one slice,
the block count the GPU engine runs the slab with.  RUN_DIR receives the scan
and the per-iteration object snapshots; a folder that already holds
object_n0_time_0_iter_<iters>.bin is sampled without running again.  The run
takes hours on 8 cores.
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
ap.add_argument("--threads", type=int, default=32)
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
    print("oracle ICD", it.value, "iterations in", time.time() - t0, "s", flush=True)
    np.save(os.path.join(d, "costs_f32.npy"), costs)

nvox = cfg.N_xy * cfg.N_xy * cfg.N_z
idx = np.sort(np.random.default_rng(SEED).choice(nvox, N_SAMPLE, replace=False)).astype(np.int32)
ph = synth.phantom(cfg)[0]
mean_att = float(ph[ph > 0].mean())
vol = lambda k: np.fromfile(os.path.join(d, f"object_n0_time_0_iter_{k}.bin"), np.float32)
ref = vol(a.iters)
assert ref.size == nvox
# the CPU run's own distance to its last iteration, as RMSE / mean attenuation, and
# (where the converged-image fixture samples the same voxels) its distance to the
# converged image after every iteration
conv_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_converged_sample.npz")
conv = np.load(conv_path) if os.path.exists(conv_path) else None
if conv is not None and not np.array_equal(conv["idx"], idx):
    conv = None
own, to_conv = {}, {}
for k in range(1, a.iters + 1):
    p = os.path.join(d, f"object_n0_time_0_iter_{k}.bin")
    if os.path.exists(p):
        v = vol(k)
        if k < a.iters:
            own[k] = float(np.sqrt(np.mean((v - ref) ** 2)) / mean_att)
        if conv is not None:
            to_conv[k] = float(np.sqrt(np.mean((v[idx] - conv["values"]) ** 2)) / mean_att)


def map_cost64(vol, offset):
    """FP64 MAP cost of (object, ring offsets) on this scan, f64a oracle."""
    o64 = py_oracle.Oracle(True, float_amatrix=True)
    d64 = d.rstrip("/") + "_f64"
    os.makedirs(d64, exist_ok=True)
    for f in ("projection_n0.bin", "weight_n0.bin", "proj_offset_n0.bin"):
        np.fromfile(os.path.join(d, f), np.float32).astype(np.float64).tofile(os.path.join(d64, f))
    with open(os.path.join(d, "view_info.txt")) as src, open(os.path.join(d64, "view_info.txt"), "w") as dst:
        dst.write(src.read())
    S, O, T = o64.setup(d64, args)
    try:
        o64.read_sinogram(d64, S, T)
        o64.set_object(O, vol.reshape(1, cfg.N_z, cfg.N_xy, cfg.N_xy).astype(np.float64))
        o64.offset_view(S)[:] = offset.reshape(cfg.N_r, cfg.N_t)
        H = o64.detector_response(S, O, T)
        E = np.zeros((S.N_p, S.N_r, S.N_t), np.float64)
        et, ht = py_oracle.xt_abi.PtrTree(E), py_oracle.xt_abi.PtrTree(H)
        o64.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
        return float(o64.lib.oracle_compute_cost(C.byref(S), C.byref(O), C.byref(T), et.ptr))
    finally:
        o64.free(S, O, T)


offset = np.fromfile(os.path.join(d, "proj_offset_n0.bin"), np.float32)
cost64 = map_cost64(ref, offset)
cpath = os.path.join(d, "costs_f32.npy")
costs = np.load(cpath) if os.path.exists(cpath) else np.full(a.iters + 1, np.nan)
print("FP64 MAP cost of the final object", repr(cost64), flush=True)
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_cpu_icd_sample.npz"),
                    idx=idx, values=ref[idx], mean_att=np.float64(mean_att),
                    iterations=np.int32(a.iters), threads=np.int32(a.threads),
                    num_z_blocks=np.int32(min(cfg.N_z, max(2, 2 * (a.threads // cfg.N_time)))),
                    layout=np.array("[N_z][N_y][N_x] float32, frame 0"),
                    own_rmse=np.array(json.dumps(own)), cost64=np.float64(cost64),
                    rmse_to_converged=np.array(json.dumps(to_conv)),
                    costs_f32=costs)
print("mean attenuation", mean_att, "own RMSE to last:", {k: round(v, 5) for k, v in own.items()})
