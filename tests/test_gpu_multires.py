"""Multi-resolution initialisation (initICD 2/3, F/XT_ICD_update.c:316-512) on the
GPU: the upsampling bit-exact against the C restatement in oracle/xt_oracle.c
(pinned by closed forms in tests/test_oracle_pins.py), ICD_BackProject's
initICD 2 path through the run folder, and the stage loop as a warm start."""
import dataclasses
import os

import numpy as np
import pytest

import py_oracle
from conftest import SMALL, make_scan
from timbir_python_amd import multires, xt_hip

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("init_icd", [2, 3])
@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 5, 7), (8, 64, 48)], ids=str)
def test_upsample_bitexact(init_icd, shape):
    print(f"[gpu] upsample initICD {init_icd} {shape}", flush=True)
    rng = np.random.default_rng(sum(shape) + init_icd)
    c = rng.uniform(0.0, 2e-3, shape).astype(np.float32)
    got = xt_hip.upsample_object(init_icd, c)
    want = py_oracle.Oracle(False).upsample_object(init_icd, c)
    assert got.shape == want.shape
    np.testing.assert_array_equal(got, want)


def test_upsample_rejects_small_sizes():
    with pytest.raises(xt_hip.XtError):
        xt_hip.upsample_object(2, np.zeros((2, 1, 4), np.float32))


def test_icd_backproject_initicd2(scan_root):
    """Coarse stage (voxels twice the size) through the engine, then the fine
    stage with --initICD 2 and no iterations: its object is the coarse object
    upsampled exactly as the reference's initObject does."""
    cfg = SMALL
    path, _ = make_scan(cfg, scan_root, "multires")
    coarse = multires.stage_configs(cfg, 2)[0]
    print("[gpu] initICD 2 through the engine", flush=True)
    xt_hip.engine_run(path, coarse.cli_args(iterations=3))
    c = np.fromfile(os.path.join(path, "object_n0_time_0.bin"), np.float32).reshape(
        cfg.N_z, cfg.N_xy // 2, cfg.N_xy // 2)
    assert np.abs(c).max() > 0
    xt_hip.engine_run(path, cfg.cli_args(iterations=0, init_icd=2))
    f = np.fromfile(os.path.join(path, "object_n0_time_0.bin"), np.float32).reshape(
        cfg.N_z, cfg.N_xy, cfg.N_xy)
    np.testing.assert_array_equal(f, py_oracle.Oracle(False).upsample_object(2, c))


def test_multires_start_beats_zero_start(scan_root):
    """Two coarse stages then the final one: after 3 final-resolution iterations
    the volume is closer to the converged image than 3 iterations from zero."""
    cfg = dataclasses.replace(SMALL, N_xy=64, seed=12)
    path, _ = make_scan(cfg, scan_root, "mrstart")
    xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 1)
    with xt_hip.Engine(path, cfg.cli_args()) as eng:
        with xt_hip.Icd(eng) as icd:
            icd.fbp()
            icd.iterate(60)
            icd.download()
        ref = eng.volume()[0].copy()
    vol, off = multires.coarse_stages(path, cfg, [10, 5])
    err = {}
    for start in ("zero", "multires"):
        with xt_hip.Engine(path, cfg.cli_args()) as eng, xt_hip.Icd(eng) as icd:
            if start == "multires":
                multires.start_from_coarse(icd, vol, off)
            icd.iterate(3)
            icd.download()
            err[start] = float(np.sqrt(np.mean((eng.volume()[0] - ref) ** 2)))
    print(f"  RMSE to the converged image after 3 iterations: {err}")
    assert err["multires"] < 0.5 * err["zero"]


def test_coarse_stages_from_open_sinogram(scan_root):
    """coarse_stages copying the projection and weights of an open Engine on the
    host, or y and w of an open state on the device (xt_icd_create_from; the
    final stage's, as the benchmark does), gives the stages the same inputs as
    reading the run folder's files: with one line in flight (deterministic) the
    coarse object and offsets are bit-identical."""
    cfg = dataclasses.replace(SMALL, N_xy=64, seed=12)
    path, _ = make_scan(cfg, scan_root, "mrcopy")
    xt_hip.set_option(xt_hip.OPT_LINES, 1)
    try:
        print("[gpu] coarse stages from files, from an open engine and from an open state",
              flush=True)
        a = multires.coarse_stages(path, cfg, [3, 2], fbp_first=True)
        with xt_hip.Engine(path, cfg.cli_args()) as eng:
            b = multires.coarse_stages(path, cfg, [3, 2], fbp_first=True, sino_from=eng)
            with xt_hip.Icd(eng) as icd:
                c = multires.coarse_stages(path, cfg, [3, 2], fbp_first=True, icd_from=icd)
    finally:
        xt_hip.set_option(xt_hip.OPT_LINES, 0)
    assert np.abs(a[0]).max() > 0
    for r in (b, c):
        np.testing.assert_array_equal(a[0], r[0])
        np.testing.assert_array_equal(a[1], r[1])


def test_create_from_refuses_another_sinogram(scan_root):
    """xt_icd_create_from takes y and w only from a state of the same sinogram
    size and layout."""
    cfg = dataclasses.replace(SMALL, N_xy=64, seed=12)
    other = dataclasses.replace(cfg, N_p=48, seed=13)
    p1, _ = make_scan(cfg, scan_root, "cf1")
    p2, _ = make_scan(other, scan_root, "cf2")
    print("[gpu] xt_icd_create_from with a source of another size", flush=True)
    with xt_hip.Engine(p1, cfg.cli_args()) as e1, xt_hip.Icd(e1) as i1, \
            xt_hip.Engine(p2, other.cli_args()) as e2:
        with pytest.raises(xt_hip.XtError, match="sinogram source"):
            xt_hip.Icd(e2, sino_from=i1)


def test_two_stage_converged_parity(scan_root, oracle32, oracle64a):
    """The launcher's two-stage run (voxels twice as wide, 40 iterations, then
    --initICD 2 at full resolution, 120 iterations) on the GPU and in the C
    oracle: converged objects within 1% RMSE of the phantom's mean attenuation
    and MAP costs within 1e-4."""
    from test_gpu_icd import map_cost, read_object
    cfg = dataclasses.replace(SMALL, iterations=120)
    coarse = dataclasses.replace(multires.stage_configs(cfg, 2)[0], iterations=40)
    p_gpu, info = make_scan(cfg, scan_root, "mr-gpu")
    p_cpu, _ = make_scan(cfg, scan_root, "mr-cpu")
    print("[gpu] two-stage engine runs, GPU and oracle", flush=True)
    for stage, init in ((coarse, 0), (cfg, 2)):
        args = stage.cli_args(init_icd=init)
        xt_hip.engine_run(p_gpu, args)
        oracle32.run(p_cpu, args)
    xg, xc = read_object(p_gpu, cfg), read_object(p_cpu, cfg)
    ph = info["phantom"]
    rmse = np.sqrt(np.mean((xg - xc) ** 2)) / ph[ph > 0].mean()
    dg = np.fromfile(os.path.join(p_gpu, "proj_offset_n0.bin"), np.float32).reshape(cfg.N_r, cfg.N_t)
    dc = np.fromfile(os.path.join(p_cpu, "proj_offset_n0.bin"), np.float32).reshape(cfg.N_r, cfg.N_t)
    cg = map_cost(oracle64a, p_gpu, cfg, xg, dg)
    cc = map_cost(oracle64a, p_cpu, cfg, xc, dc)
    print(f"  RMSE(gpu-oracle)/mean = {rmse:.4%}, cost gpu {cg:.9g} oracle {cc:.9g}")
    assert rmse <= 0.01
    assert abs(cg - cc) <= 1e-4 * abs(cc)


def test_three_stage_coarse_z_converged_parity(scan_root, oracle32, oracle64a):
    """The launcher's policy with a coarse-z level (XT_Initialize.py:175-190):
    voxels 4x and 2x wider and two detector rows deep (delta_z 2, z_overlap_num
    2), then --initICD 3 at full resolution (object upsampled in z, y and x), on
    the GPU and in the C oracle: converged objects within 1% RMSE and MAP costs
    within 1e-4."""
    from test_gpu_icd import map_cost, read_object
    cfg = dataclasses.replace(SMALL, iterations=120)
    stages = multires.stage_configs(cfg, 3, z_stages=2)
    inits = multires.init_icd_of(3, z_stages=2)
    assert inits == [0, 2, 3] and [c.mult_z for c in stages] == [2, 2, 1]
    p_gpu, info = make_scan(cfg, scan_root, "mr3-gpu")
    p_cpu, _ = make_scan(cfg, scan_root, "mr3-cpu")
    print("[gpu] three-stage engine runs (delta_z 2, 2, 1), GPU and oracle", flush=True)
    for stage, init, n in zip(stages, inits, (30, 30, cfg.iterations)):
        args = stage.cli_args(iterations=n, init_icd=init)
        xt_hip.engine_run(p_gpu, args)
        oracle32.run(p_cpu, args)
    xg, xc = read_object(p_gpu, cfg), read_object(p_cpu, cfg)
    ph = info["phantom"]
    rmse = np.sqrt(np.mean((xg - xc) ** 2)) / ph[ph > 0].mean()
    dg = np.fromfile(os.path.join(p_gpu, "proj_offset_n0.bin"), np.float32).reshape(cfg.N_r, cfg.N_t)
    dc = np.fromfile(os.path.join(p_cpu, "proj_offset_n0.bin"), np.float32).reshape(cfg.N_r, cfg.N_t)
    cg = map_cost(oracle64a, p_gpu, cfg, xg, dg)
    cc = map_cost(oracle64a, p_cpu, cfg, xc, dc)
    print(f"  RMSE(gpu-oracle)/mean = {rmse:.4%}, cost gpu {cg:.9g} oracle {cc:.9g}")
    assert rmse <= 0.01
    assert abs(cg - cc) <= 1e-4 * abs(cc)


def test_coarse_z_stage_start_beats_zero_start(scan_root):
    """The stage loop in memory with a coarse-z level: after 3 final-resolution
    iterations from the upsampled (initICD 3) start the volume is closer to the
    converged image than 3 iterations from zero."""
    cfg = dataclasses.replace(SMALL, N_xy=64, seed=14)
    path, _ = make_scan(cfg, scan_root, "mrz")
    xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 1)
    with xt_hip.Engine(path, cfg.cli_args()) as eng:
        with xt_hip.Icd(eng) as icd:
            icd.fbp()
            icd.iterate(60)
            icd.download()
        ref = eng.volume()[0].copy()
    vol, off = multires.coarse_stages(path, cfg, [10, 5], z_stages=2)
    assert vol.shape == (cfg.N_z // 2, cfg.N_xy // 2, cfg.N_xy // 2)
    err = {}
    for start in ("zero", "multires"):
        with xt_hip.Engine(path, cfg.cli_args()) as eng, xt_hip.Icd(eng) as icd:
            if start == "multires":
                multires.start_from_coarse(icd, vol, off, init_icd=3)
            icd.iterate(3)
            icd.download()
            err[start] = float(np.sqrt(np.mean((eng.volume()[0] - ref) ** 2)))
    print(f"  RMSE to the converged image after 3 iterations: {err}")
    assert err["multires"] < 0.5 * err["zero"]
