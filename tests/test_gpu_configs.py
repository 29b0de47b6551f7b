"""The benchmark workloads on the GPU at their real sizes.

Plain 3D (128x128x8, 256 views, 192x8 bins, synth.C1): against the CPU oracle's
own runs on the same seeded scan (tests/golden/c1_cpu_icd.npz, made by
tests/golden/make_c1_reference.py) -- after 30 iterations from the reference's
zero start, and converged from a common warm start.

The production slab (512^2x32, 1024 views), the 4D interlaced scan (256^2x16 x 16
frames, 2048 views), the 24 uneven slabs and the learned warm start are too large
for the oracle; there the checks are the properties the mathematics fixes: the
first iteration updates every voxel exactly once (no zero-skipping before the
second iteration, F/XT_VoxUpdate.c:208-232), the MAP cost falls, the FP32 error
sinogram after an iteration agrees with its FP64 recompute, a slab batched with
others gives what it gives alone, and the network start lands nearer the
converged image than the filtered backprojection it refines.
"""
import dataclasses
import os
import shutil
import sys

import numpy as np
import pytest

from conftest import ROOT, make_scan
from timbir_python_amd import synth, xt_hip

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from c1_warm_start import warm_start_args, write_warm_start  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden")


def _rmse(a, b, mean_att):
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - b) ** 2)) / mean_att)


def _cost_of(path, cfg, vol, offset):
    """FP64 MAP cost of (object, offsets) on the GPU: e = y - Ax - d recomputed in
    FP64 at create, then the cost (xt_icd_cost; within 1e-5 of the oracle's FP64
    evaluator, tests/test_gpu_parity.py::test_cost)."""
    saved = xt_hip.get_option(xt_hip.OPT_AUTOTUNE)
    xt_hip.set_option(xt_hip.OPT_AUTOTUNE, 0)
    try:
        with xt_hip.Engine(path, cfg.cli_args()) as eng:
            eng.set_volume(vol)
            eng.offset()[:] = offset
            with xt_hip.Icd(eng) as icd:
                return icd.cost()
    finally:
        xt_hip.set_option(xt_hip.OPT_AUTOTUNE, saved)


def _read(path, cfg):
    vol = np.stack([np.fromfile(os.path.join(path, f"object_n0_time_{i}.bin"), np.float32)
                    .reshape(cfg.N_z, cfg.N_xy, cfg.N_xy) for i in range(cfg.N_time)])
    off = np.fromfile(os.path.join(path, "proj_offset_n0.bin"), np.float32).reshape(cfg.N_r, cfg.N_t)
    return vol, off


@pytest.fixture(scope="module")
def c1_ref():
    return np.load(os.path.join(GOLD, "c1_cpu_icd.npz"))


@pytest.mark.parametrize("nhicd_lines", [0, 1], ids=["nhicd-per-block", "nhicd-per-parity"])
def test_c1_zero_start_30_iterations(nhicd_lines, scan_root, c1_ref):
    """30 iterations from zero (the reference's initICD 0) on the GPU and in the
    oracle (both with 8 one-slice z-blocks), with the reference's NHICD rule and
    with the per-parity one: the GPU's image is as near the converged image as
    the oracle's (its own distance to it is the yardstick), its MAP cost is not
    above the oracle's, and the two images differ by less than the oracle still
    has to travel from its 30th iteration to its 300th (8.1%): the two runs take
    different random line orders along the same descent."""
    ref = c1_ref
    cfg = dataclasses.replace(synth.C1, iterations=int(ref["early_iterations"]))
    path, info = make_scan(cfg, scan_root, "c1zero")
    print(f"[gpu] C1 {cfg.N_xy}^2x{cfg.N_z}, {cfg.N_p} views: {cfg.iterations} iterations "
          f"from zero, NHICD lines {nhicd_lines}", flush=True)
    xt_hip.set_option(xt_hip.OPT_NHICD_LINES, nhicd_lines)
    try:
        xt_hip.engine_run(path, cfg.cli_args())
    finally:
        xt_hip.set_option(xt_hip.OPT_NHICD_LINES, 1)
    vol, off = _read(path, cfg)
    v = vol[0].reshape(-1)[ref["idx"]]
    m = float(ref["mean_att"])
    r_cpu = _rmse(v, ref["early_values"], m)
    d_gpu = _rmse(v, ref["fbp_values"], m)
    d_cpu = _rmse(ref["early_values"], ref["fbp_values"], m)
    c_gpu = _cost_of(path, cfg, vol, off)
    c_cpu = float(ref["early_cost64"])
    print(f"  RMSE(gpu - oracle) after {cfg.iterations}: {r_cpu:.3%}; distance to the converged "
          f"image: gpu {d_gpu:.3%}, oracle {d_cpu:.3%}; FP64 cost gpu {c_gpu:.9g} oracle "
          f"{c_cpu:.9g} (rel {(c_gpu - c_cpu) / c_cpu:+.2e})", flush=True)
    own30 = float(ref["own_rmse_to_final"][cfg.iterations - 1])
    assert np.all(np.isfinite(vol))
    assert r_cpu <= own30
    assert d_gpu <= 1.1 * d_cpu
    assert c_gpu <= c_cpu * (1 + 1e-4)


def test_c1_converged_from_common_warm_start(scan_root, c1_ref):
    """From the same starting files (FBP with its ring-offset guess, initICD 1,
    updateProjOffset 3) the GPU and the oracle converge to the same image: within
    1% RMSE of the mean attenuation, MAP costs within 1e-4."""
    ref = c1_ref
    n = int(ref["fbp_iterations"])
    cfg = dataclasses.replace(synth.C1, iterations=n)
    path, info = make_scan(cfg, scan_root, "c1fbp")
    write_warm_start(path, cfg)
    print(f"[gpu] C1 from the warm start, {n} iterations", flush=True)
    xt_hip.engine_run(path, warm_start_args(cfg, n))
    vol, off = _read(path, cfg)
    v = vol[0].reshape(-1)[ref["idx"]]
    m = float(ref["mean_att"])
    rmse = _rmse(v, ref["fbp_values"], m)
    c_gpu = _cost_of(path, cfg, vol, off)
    c_cpu = float(ref["fbp_cost64"])
    print(f"  RMSE(gpu - oracle) {rmse:.3%} (oracle's own change over its last third "
          f"{float(ref['fbp_own_rmse']):.3%}); FP64 cost gpu {c_gpu:.9g} oracle {c_cpu:.9g} "
          f"(rel {(c_gpu - c_cpu) / c_cpu:+.2e})", flush=True)
    assert rmse <= 0.01
    assert abs(c_gpu - c_cpu) <= 1e-4 * c_cpu


# -- full sizes: properties ---------------------------------------------------
def _gpu_scan(cfg, root, tag):
    path = os.path.join(str(root), f"full_{cfg.name}_{tag}")
    if not os.path.exists(os.path.join(path, "weight_n0.bin")):
        synth.generate(cfg, path, xt_hip.projector(path, cfg.cli_args()))
    return path


@pytest.fixture(scope="module")
def c2_scan(tmp_path_factory):
    root = tmp_path_factory.mktemp("c2")
    print(f"[gpu] generating {synth.C2.name} with the GPU projector", flush=True)
    p = _gpu_scan(synth.C2, root, "c2")
    yield p
    shutil.rmtree(root, ignore_errors=True)


def _iteration_properties(cfg, path):
    """Iteration 1 updates every voxel once; the cost falls over two iterations;
    the FP32 e agrees with its FP64 recompute after each."""
    saved = xt_hip.get_option(xt_hip.OPT_RESYNC_EVERY)
    xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 0)  # e accumulates in FP32 over the iteration
    try:
        with xt_hip.Engine(path, cfg.cli_args()) as eng, xt_hip.Icd(eng) as icd:
            c = [icd.cost()]
            nvox = cfg.N_xy * cfg.N_xy * cfg.N_z * cfg.N_time
            for it in (1, 2):
                s0 = icd.stats()
                icd.iterate(1)
                s1 = icd.stats()
                c_fp32 = icd.cost()
                icd.resync()
                c.append(icd.cost())
                upd = s1.voxel_updates - s0.voxel_updates
                print(f"  iteration {it}: {upd:.0f} voxel updates ({nvox} voxels), cost "
                      f"{c[-1]:.9g}, FP32-e cost {c_fp32:.9g} (rel {(c_fp32 - c[-1]) / c[-1]:+.1e})",
                      flush=True)
                if it == 1:
                    assert upd == nvox
                else:
                    assert upd > nvox  # homogeneous share plus the NHICD lines
                assert abs(c_fp32 - c[-1]) <= 1e-4 * abs(c[-1])
            assert c[2] < c[1] < c[0]
    finally:
        xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, saved)


def test_c2_slab_properties(c2_scan):
    print(f"[gpu] {synth.C2.name}: two iterations", flush=True)
    _iteration_properties(synth.C2, c2_scan)


def test_c3_interlaced_4d_properties(tmp_path):
    cfg = synth.C3
    print(f"[gpu] generating {cfg.name} with the GPU projector", flush=True)
    path = _gpu_scan(cfg, tmp_path, "c3")
    _iteration_properties(cfg, path)


def test_c4_batch_properties(tmp_path):
    """The 24 uneven slabs advanced by one graph per iteration: each slab's first
    iteration updates every voxel once and its cost falls.  Then three of them
    (the smallest) batched and alone with one line in flight, where a run is
    deterministic: identical objects."""
    cfgs = synth.batch_configs()
    print(f"[gpu] generating {len(cfgs)} slabs with the GPU projector", flush=True)
    paths = [_gpu_scan(c, tmp_path, "c4") for c in cfgs]
    engs = [xt_hip.Engine(p, c.cli_args()) for c, p in zip(cfgs, paths)]
    icds = [xt_hip.Icd(e) for e in engs]
    try:
        c0 = [i.cost() for i in icds]
        with xt_hip.IcdBatch(icds) as b:
            print("[gpu] batch iteration 1", flush=True)
            b.iterate(1)
            u1 = [i.stats().voxel_updates for i in icds]
            b.iterate(1)
        c2 = [i.cost() for i in icds]
    finally:
        for i in icds:
            i.close()
        for e in engs:
            e.close()
    for c, u, a, b2 in zip(cfgs, u1, c0, c2):
        assert u == c.N_xy * c.N_xy * c.N_z, c.name
        assert b2 < a, c.name
    sizes = [c.N_xy * c.N_xy * c.N_z * c.N_p for c in cfgs]
    sample = sorted(range(len(cfgs)), key=lambda k: sizes[k])[:3]
    xt_hip.set_option(xt_hip.OPT_LINES, 1)
    try:
        alone = []
        for k in sample:
            with xt_hip.Engine(paths[k], cfgs[k].cli_args()) as eng, xt_hip.Icd(eng) as icd:
                icd.iterate(2)
                icd.download()
                alone.append(eng.volume().copy())
        engs = [xt_hip.Engine(paths[k], cfgs[k].cli_args()) for k in sample]
        icds = [xt_hip.Icd(e) for e in engs]
        try:
            with xt_hip.IcdBatch(icds) as b:
                b.iterate(2)
            batched = []
            for i, e in zip(icds, engs):
                i.download()
                batched.append(e.volume().copy())
        finally:
            for i in icds:
                i.close()
            for e in engs:
                e.close()
    finally:
        xt_hip.set_option(xt_hip.OPT_LINES, 0)
    for k, a, b2 in zip(sample, alone, batched):
        c = cfgs[k]
        print(f"  {c.name} {c.N_xy}^2x{c.N_z}/{c.N_p} views: max |batched - alone| "
              f"{np.abs(a - b2).max():.3e}", flush=True)
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b2)


def test_c5_unet_start_on_slab(c2_scan):
    """The learned start on the production slab: network on the FBP volume, then
    the FP64 recompute of e.  Its MAP cost is below the FBP start's, which is below
    the zero start's, and it is nearer the converged image than the FBP volume."""
    cfg = synth.C2
    weights = os.path.join(ROOT, "timbir-python_amd", "weights", "unet_warmstart_256x256x16.npz")
    conv = np.load(os.path.join(GOLD, "c2_converged_sample.npz"))
    idx, want, m = conv["idx"], conv["values"], float(conv["mean_att"])
    res = {}
    with xt_hip.UNet.load(weights, (cfg.N_xy, cfg.N_xy, cfg.N_z)) as net:
        for start in ("zero", "fbp", "unet"):
            with xt_hip.Engine(c2_scan, cfg.cli_args()) as eng, xt_hip.Icd(eng) as icd:
                if start == "fbp":
                    icd.fbp()
                elif start == "unet":
                    xt_hip.icd_unet_start(icd, net)
                c = icd.cost()
                icd.resync()
                c_again = icd.cost()
                icd.download()
                v = eng.volume()[0].reshape(-1)[idx]
                res[start] = (c, _rmse(v, want, m))
                assert abs(c_again - c) <= 1e-9 * abs(c)  # the start already ends with the recompute
    print("  start: (MAP cost, RMSE to the converged image) " +
          ", ".join(f"{k} ({c:.9g}, {r:.2%})" for k, (c, r) in res.items()), flush=True)
    assert res["unet"][0] < res["fbp"][0] < res["zero"][0]
    assert res["unet"][1] < 0.5 * res["fbp"][1]
