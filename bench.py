"""Benchmark of the ICD hot path on one MI355X.

Workload: the production slab -- 512x512x32 voxels, 1024 views over [0,180),
768x32 detector bins, Poisson counts at I0=1e4 with 0.5% zingers and ring
offsets, qGGMRF prior (p=1.2, q=2, c=0.01, sigma_s=1e-4), FP32 footprint sums
and error sinogram, FP64 recompute of e = y - Ax - d every iteration.

A step is one ICD iteration: 10 sub-iterations of voxel-line updates over every
line, NHICD lines from the second iteration on, four ring-offset updates per
sub-iteration and the FP64 resync.  Inputs are resident in HBM before the timed
region.  value = performed voxel updates per second.

Also reported: the roofline of the voxel-update kernel (algorithmic bytes =
12 B per sinogram-footprint entry of every performed update, over the kernel's
summed duration from HIP events on its stream), time to converge (iterations
until the MAP cost is within 1e-4 of the cost after --converge-iters
iterations), and the CPU oracle timed on a bounded sample of the same
workload on this host (the reference's OpenMP structure, all cores).
"""
import argparse
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "timbir-python_amd")

METRIC = ("time-to-converge (≤1% RMSE vs CPU ICD) 512²×32 slab; "
          "voxel-updates/s; GB/s")
HBM_PEAK_GBS = 8000.0


def load_pkg():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "timbir_python_amd", os.path.join(PKG, "__init__.py"), submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["timbir_python_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_workload(cfg, path, xt_hip, synth):
    """Seeded scan in the reference's file formats; Ax by the GPU projector."""
    os.makedirs(path, exist_ok=True)
    ang, t = cfg.views()
    synth.write_view_info(path, ang, t)
    ph = synth.phantom(cfg)
    eng = xt_hip.Engine(path, cfg.cli_args(), read_sinogram=False)
    try:
        eng.set_volume(ph)
        ax = xt_hip.forward_project(eng, xt_hip.detector_response(eng))
    finally:
        eng.close()
    y, w, d = synth.measure(cfg, ax)
    y.astype(np.float32).tofile(os.path.join(path, "projection_n0.bin"))
    w.astype(np.float32).tofile(os.path.join(path, "weight_n0.bin"))
    np.zeros((cfg.N_r, cfg.N_t), np.float32).tofile(os.path.join(path, "proj_offset_n0.bin"))
    return ph


def gpu_timed(cfg, path, xt_hip, steps, warmup, args=None, volume=None):
    """steps timed ICD iterations after warmup ones on the stepwise API; volume:
    the starting object (the structs' own, zero, otherwise)."""
    xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 1)
    with xt_hip.Engine(path, args or cfg.cli_args()) as eng:
        if volume is not None:
            eng.set_volume(volume)
        with xt_hip.Icd(eng) as icd:
            return _timed_steps(icd, steps, warmup)


KERNEL_TIMING_STEPS = 3


def _timed_steps(icd, steps, warmup):
    """The timed region: steps iterations as the library runs them (from the
    second iteration on, one captured hipGraph per iteration), wall clock with the
    stream synchronised on both sides.  Then KERNEL_TIMING_STEPS more iterations
    with per-launch HIP events on the library's stream (launched eagerly: HIP gives
    no elapsed times for events inside a graph replay) for the voxel-update
    kernel's average launch duration."""
    icd.iterate(warmup)
    icd.synchronize()
    s0 = icd.stats()
    t0 = time.perf_counter()
    icd.iterate(steps)
    icd.synchronize()
    t1 = time.perf_counter()
    s1 = icd.stats()
    icd.timing(True)
    icd.iterate(KERNEL_TIMING_STEPS)
    icd.synchronize()
    icd.timing(False)
    s2 = icd.stats()
    dt = t1 - t0
    return dict(seconds=dt, updates=s1.voxel_updates - s0.voxel_updates,
                footprint=s1.footprint_entries - s0.footprint_entries,
                launches=s1.icd_launches - s0.icd_launches, iters=steps,
                visits=s1.line_visits - s0.line_visits, lines_in_flight=s1.lines_in_flight,
                kt_kernel_ms=s2.icd_kernel_ms - s1.icd_kernel_ms,
                kt_launches=s2.icd_launches - s1.icd_launches,
                kt_footprint=s2.footprint_entries - s1.footprint_entries)


CONVERGED = os.path.join(ROOT, "tests", "golden", "c2_converged_sample.npz")
CPU_ICD = os.path.join(ROOT, "tests", "golden", "c2_cpu_icd_sample.npz")


WEIGHTS = os.path.join(PKG, "weights", "unet_warmstart_256x256x16.npz")
# ICD iterations of the 128^2 and 256^2 stages before the 512^2 one: the reference
# launcher's policy (zero start, offsets from the second stage) and the FBP-first form
MULTIRES_ITERS = (10, 5)
MULTIRES_FBP_ITERS = (8, 2)


def gpu_converge(cfg, path, xt_hip, max_iter, warm, tol=0.01):
    """GPU ICD one iteration at a time from a zero start (warm "zero"), from the
    filtered backprojection ("fbp"), from the warm-start network applied to
    it ("unet"), from the reference launcher's coarser stages ("multires":
    128^2 and 256^2 reconstructions of the same sinogram, each stage upsampled
    into the next, their set-up and iterations timed) or from the same stages
    with the coarsest one started from its FBP ("multires_fbp"): wall time
    (start included; the RMSE evaluations between
    iterations are not timed) until the volume is within tol RMSE, relative to
    the phantom's mean attenuation, of the converged MAP image (sampled voxels of
    the reference file), and the FP64 MAP cost of the last iterate."""
    if not os.path.exists(CONVERGED):
        return None
    ref = np.load(CONVERGED)
    idx, want, mean_att = ref["idx"], ref["values"], float(ref["mean_att"])
    xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 1)
    times, rmse = [], []
    net = None
    if warm == "unet":
        if not os.path.exists(WEIGHTS):
            return None
        net = xt_hip.UNet.load(WEIGHTS, (cfg.N_xy, cfg.N_xy, cfg.N_z))
    with xt_hip.Engine(path, cfg.cli_args()) as eng, xt_hip.Icd(eng) as icd:
        icd.synchronize()
        t0 = time.perf_counter()
        if warm == "fbp":
            icd.fbp()
        elif warm == "unet":
            xt_hip.icd_unet_start(icd, net)
        elif warm in ("multires", "multires_fbp"):
            from timbir_python_amd import multires
            fbp_first = warm == "multires_fbp"
            vol, off = multires.coarse_stages(
                path, cfg, MULTIRES_FBP_ITERS if fbp_first else MULTIRES_ITERS,
                fbp_first=fbp_first, icd_from=icd)
            multires.start_from_coarse(icd, vol, off)
        icd.synchronize()
        t_init = time.perf_counter() - t0
        if net is not None:
            net.close()
        for _ in range(max_iter):
            icd.synchronize()
            t0 = time.perf_counter()
            icd.iterate(1)
            icd.synchronize()
            times.append(time.perf_counter() - t0)
            icd.download()
            v = eng.volume()[0].reshape(-1)
            rmse.append(float(np.sqrt(np.mean((v[idx] - want) ** 2)) / mean_att))
            if len(times) % 10 == 0:
                log(f"{warm} start, iteration {len(times)}: "
                    f"RMSE to the converged image {rmse[-1]:.3%}")
            if rmse[-1] <= tol:
                break
        final_cost = icd.cost()  # FP64 MAP cost of the endpoint (e recomputed every iteration)
    hit = len(rmse) if rmse and rmse[-1] <= tol else None
    return dict(iters=hit, seconds=(t_init + float(np.sum(times))) if hit else None,
                init_seconds=t_init, rmse=rmse, times=times,
                ref_iters=int(ref["iterations"]), final_cost64=final_cost)


def cpu_comparison(cfg, path, xt_hip):
    """Like for like with the CPU ICD fixture (the oracle's float build from a zero
    start, its z-block count fixed by its thread count): the GPU from the same
    zero start for the same number of iterations, with the reference's NHICD rule
    (per z-block, XT_OPT_NHICD_LINES 0) and with the default per-parity rule;
    RMSE to the CPU's image after every iteration and the FP64 MAP cost at the
    end, against the CPU's FP64 cost of its own image; where the converged-image
    sample holds the same voxels, both sides' distance to it after every
    iteration."""
    if not os.path.exists(CPU_ICD):
        return None
    cpu = np.load(CPU_ICD)
    n = int(cpu["iterations"])
    idx, want, m = cpu["idx"], cpu["values"], float(cpu["mean_att"])
    conv = np.load(CONVERGED) if os.path.exists(CONVERGED) else None
    if conv is not None and not np.array_equal(conv["idx"], idx):
        conv = None
    out = {"iterations": n, "cpu_threads": int(cpu["threads"]),
           "cpu_num_z_blocks": int(cpu["num_z_blocks"]) if "num_z_blocks" in cpu else None,
           "cpu_cost64": float(cpu["cost64"]) if "cost64" in cpu else None, "gpu": {}}
    if "rmse_to_converged" in cpu:
        tc = json.loads(str(cpu["rmse_to_converged"]))
        out["cpu_rmse_to_converged_by_iteration"] = [round(tc[str(k)], 5) for k in range(1, n + 1)
                                                     if str(k) in tc]
    xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 1)
    for rule in (0, 1):
        xt_hip.set_option(xt_hip.OPT_NHICD_LINES, rule)
        try:
            with xt_hip.Engine(path, cfg.cli_args()) as eng, xt_hip.Icd(eng) as icd:
                r, rc = [], []
                for _ in range(n):
                    icd.iterate(1)
                    icd.download()
                    v = eng.volume()[0].reshape(-1)
                    r.append(float(np.sqrt(np.mean((v[idx] - want) ** 2)) / m))
                    if conv is not None:
                        rc.append(float(np.sqrt(np.mean((v[idx] - conv["values"]) ** 2)) / m))
                c = icd.cost()
                nb = eng.T.num_z_blocks
        finally:
            xt_hip.set_option(xt_hip.OPT_NHICD_LINES, 1)
        key = "nhicd_per_block" if rule == 0 else "nhicd_per_parity"
        out["gpu"][key] = {"rmse_vs_cpu_by_iteration": [round(x, 5) for x in r],
                           "rmse_to_converged_by_iteration": [round(x, 5) for x in rc],
                           "cost64": c, "num_z_blocks": nb,
                           "cost_rel_to_cpu": ((c - out["cpu_cost64"]) / out["cpu_cost64"]
                                               if out["cpu_cost64"] else None)}
        log(f"like-for-like ({key}): RMSE to the CPU ICD after {n} iterations {r[-1]:.3%}, "
            f"FP64 cost {c:.9g} (CPU {out['cpu_cost64']})")
    return out


def bench_batch(a, synth, xt_hip):
    """The batched workload: every slab's scan, its own ICD state (autotuned at
    create), all advanced by one hipGraph per iteration; value = voxel updates per
    second summed over the slabs."""
    cfgs = synth.batch_configs()
    root = a.workdir or tempfile.mkdtemp(prefix="xt_batch_")
    engs, icds = [], []
    try:
        t0 = time.perf_counter()
        for c in cfgs:
            p = os.path.join(root, c.name)
            make_workload(c, p, xt_hip, synth)
        log(f"{len(cfgs)} slabs ready in {time.perf_counter() - t0:.1f}s: " +
            ", ".join(f"{c.N_xy}^2x{c.N_z}/{c.N_p}v" for c in cfgs))
        xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 1)
        for c in cfgs:
            e = xt_hip.Engine(os.path.join(root, c.name), c.cli_args())
            engs.append(e)
            icds.append(xt_hip.Icd(e))
        with xt_hip.IcdBatch(icds) as b:
            b.iterate(a.warmup)
            for i in icds:
                i.synchronize()
            s0 = [i.stats() for i in icds]
            t0 = time.perf_counter()
            b.iterate(a.steps)
            for i in icds:
                i.synchronize()
            dt = time.perf_counter() - t0
            s1 = [i.stats() for i in icds]
    finally:
        for i in icds:
            i.close()
        for e in engs:
            e.close()
        if a.workdir is None:
            shutil.rmtree(root, ignore_errors=True)
    upd = sum(y.voxel_updates - x.voxel_updates for x, y in zip(s0, s1))
    foot = sum(y.footprint_entries - x.footprint_entries for x, y in zip(s0, s1))
    log(f"batch: {dt / a.steps * 1e3:.1f} ms per iteration of all {len(cfgs)} slabs, "
        f"{upd / dt:.3e} updates/s")
    print(json.dumps({
        "metric": "batched uneven slabs: voxel-updates/s (all slabs)", "value": upd / dt,
        "unit": "voxel-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded ellipsoid phantoms, Poisson I0=1e4, 0.3% zingers",
        "config": {"workload": "batch24_uneven",
                   "slabs": [[c.N_xy, c.N_xy, c.N_z, c.N_p] for c in cfgs],
                   "step": "one ICD iteration of every slab, one hipGraph launch",
                   "lines_in_flight": [x.lines_in_flight for x in s1]},
        "algorithmic_gbs": 12.0 * foot / dt / 1e9}), flush=True)


BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def unet_flops(shape):
    """Forward flops of the warm-start network on one volume: 2*27*C_in*C_out per
    voxel of each 3x3x3 conv at its level, plus the 1x1x1 head."""
    spec = [(0, 1, 16), (0, 16, 16), (1, 16, 32), (1, 32, 32), (2, 32, 64), (2, 64, 64),
            (3, 64, 128), (3, 128, 128), (2, 192, 64), (2, 64, 64), (1, 96, 32), (1, 32, 32),
            (0, 48, 16), (0, 16, 16)]
    v0 = shape[0] * shape[1] * shape[2]
    return sum(2.0 * 27 * ci * co * v0 / 8 ** lvl for lvl, ci, co in spec) + 2.0 * 16 * v0


def bench_unet(a, synth, xt_hip):
    """The warm-start network's kernels: inference on the 512^2x32 slab with the
    committed weights (device time per forward from HIP events), and training
    steps at the 256^2x16 training shape (forward + backward device time)."""
    rng = np.random.default_rng(7)
    inf_shape, tr_shape = (512, 512, 32), (256, 256, 16)
    x = rng.uniform(0.0, 2e-3, inf_shape).astype(np.float32)
    with xt_hip.UNet.load(WEIGHTS, inf_shape) as net:
        for _ in range(a.warmup):
            net.forward(x)
        fwd = []
        for _ in range(a.steps):
            net.forward(x)
            fwd.append(net.last_ms()[0])
    xt = rng.uniform(0.0, 2e-3, tr_shape).astype(np.float32)
    tt = (0.9 * xt).astype(np.float32)
    with xt_hip.UNet(tr_shape, 1000.0, seed=1, train=True) as net:
        for _ in range(a.warmup):
            net.train_step(xt, tt, 1e-4)
        tr = []
        t0 = time.perf_counter()
        for _ in range(a.steps):
            net.train_step(xt, tt, 1e-4)
            tr.append(net.last_ms())
        wall = (time.perf_counter() - t0) / a.steps
    f_ms = float(np.mean(fwd))
    inf_tf = unet_flops(inf_shape) / (f_ms * 1e-3) / 1e12
    tf_ms, tb_ms = float(np.mean([t[0] for t in tr])), float(np.mean([t[1] for t in tr]))
    train_tf = 3 * unet_flops(tr_shape) / ((tf_ms + tb_ms) * 1e-3) / 1e12
    log(f"unet: inference {f_ms:.2f} ms ({inf_tf:.0f} TFLOP/s); training step "
        f"{tf_ms:.2f} + {tb_ms:.2f} ms ({train_tf:.0f} TFLOP/s)")
    print(json.dumps({
        "metric": "warm-start U-Net inference on the 512^2x32 slab (TFLOP/s)", "value": inf_tf,
        "unit": "TFLOP/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": f_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "data": "synthetic volumes; committed weights for inference",
        "config": {"workload": "unet3d_4lvl_16-128", "inference": list(inf_shape),
                   "training": list(tr_shape)},
        "roofline": {"bound": "mfma", "achieved": inf_tf, "peak": BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": inf_tf / BF16_PEAK_TFLOPS, "traffic": None,
                     "achieved_basis": "2*27*C_in*C_out flops per voxel and conv / device time "
                                       "of the whole forward (HIP events)"},
        "training_step": {"forward_ms": tf_ms, "backward_ms": tb_ms, "wall_ms": wall * 1e3,
                          "tflops": train_tf}}), flush=True)


def cpu_sample(cfg, path, target_s, phase_contrast=False):
    """Oracle (C restatement, -O3, OpenMP) on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C
    import py_oracle
    from timbir_python_amd import xt_abi
    ncpu = len(os.sched_getaffinity(0))
    threads = min(ncpu, 16)
    args = cfg.cli_args()
    args[args.index("--num_threads") + 1] = str(threads)
    o = py_oracle.Oracle(False)
    o.set_phase_contrast(phase_contrast)
    S, O, T = o.setup(path, args)
    try:
        o.read_sinogram(path, S, T)
        y = o.sino_view(S.Projection, S)
        E = np.ascontiguousarray(y - o.offset_view(S)[None])  # e = y - A*0 - d
        et = xt_abi.PtrTree(E)
        sec, tasks = C.c_double(), C.c_int32()
        max_lines = O.N_y * O.N_x

        def run(n):
            u = o.lib.oracle_sample_updates(C.byref(S), C.byref(O), C.byref(T), et.ptr, n,
                                            threads, C.byref(sec), C.byref(tasks))
            return u, sec.value

        lines = 2
        upd, s_ = run(lines)
        while s_ < 0.5 and lines < max_lines:  # calibrate
            lines = min(max_lines, lines * 4)
            upd, s_ = run(lines)
        lines = max(1, min(max_lines, int(lines * target_s / max(s_, 1e-6))))
        upd, s_ = run(lines)  # the measured sample
        return dict(rate=upd / s_, updates=int(upd), seconds=s_, lines=lines,
                    threads=threads, tasks=tasks.value, nb=T.num_z_blocks)
    finally:
        o.free(S, O, T)
        o.set_phase_contrast(False)


def _ratio(conv, slow, fast):
    """slow start's time over the fast one's; when the slow start did not reach
    the criterion within its iteration cap, the time of its capped run gives a
    lower bound, returned as {"at_least": r}."""
    a, b = conv.get(slow), conv.get(fast)
    if not a or not b or not b["seconds"]:
        return None
    if a["seconds"]:
        return a["seconds"] / b["seconds"]
    capped = a["init_seconds"] + float(np.sum(a["times"]))
    return {"at_least": capped / b["seconds"], "capped_iterations": len(a["times"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--converge-iters", type=int, default=200,
                    help="iteration cap of each time-to-converge run")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--starts", default="multires_fbp,unet,fbp,multires,zero",
                    help="comma-separated time-to-converge starts to run (the default: all; "
                         "the speedups need fbp, the like-for-like CPU comparison runs with "
                         "the default set only)")
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--workload", choices=("slab", "4d", "batch", "unet", "pc", "edge"),
                    default="slab",
                    help="slab: the 512^2x32 production slab (the metric's workload); 4d: the "
                         "256^2x16 x 16-frame interlaced reconstruction; batch: 24 uneven slabs "
                         "advanced together; unet: the warm-start network's BF16 MFMA kernels "
                         "(inference on the slab, a training step at 256^2x16); pc: the "
                         "slab as a phase-contrast scan (Laplacian footprints); edge: the slab "
                         "with edge-only updates from the phantom (initICD 1) "
                         "(timing only for all but slab)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (calibrates the "
                         "FETCH_SIZE/WRITE_SIZE counters for the roofline traffic figure)")
    a = ap.parse_args()
    if a.copy_probe:
        import torch
        n = a.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4}), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py: this engine runs one slab on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU z-slab sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    load_pkg()
    from timbir_python_amd import synth, xt_hip
    if a.workload == "batch":
        return bench_batch(a, synth, xt_hip)
    if a.workload == "unet":
        return bench_unet(a, synth, xt_hip)
    cfg = synth.C3 if a.workload == "4d" else synth.C2
    if a.workload != "slab":
        a.no_converge = True  # the converged-image reference exists for the slab only
    # phase contrast: the scan's projections and the reconstruction use Laplacian footprints
    xt_hip.set_option(xt_hip.OPT_PHASE_CONTRAST, 1 if a.workload == "pc" else 0)
    args = cfg.cli_args()
    if a.workload == "edge":
        args = cfg.cli_args(init_icd=1) + ["--only_Edge_Updates"]
    work = a.workdir or tempfile.mkdtemp(prefix="xt_bench_")
    try:
        log(f"generating {cfg.name}: {cfg.N_xy}^2 x {cfg.N_z}, {cfg.N_p} views, "
            f"{cfg.N_r}x{cfg.N_t} bins in {work}")
        t0 = time.perf_counter()
        ph = make_workload(cfg, work, xt_hip, synth)
        if a.workload == "edge":
            # the starting object the edge mask is taken from: the phantom
            for f in range(cfg.N_time):
                ph[f].astype(np.float32).tofile(os.path.join(work, f"object_n0_time_{f}.bin"))
            np.array([1.0], np.float32).tofile(os.path.join(work, "variance_estimate_n0.bin"))
        log(f"workload ready in {time.perf_counter() - t0:.1f}s; timing {a.steps} iterations "
            f"after {a.warmup} warm-up")
        g = gpu_timed(cfg, work, xt_hip, a.steps, a.warmup, args,
                      volume=ph if a.workload == "edge" else None)
        log(f"gpu: {g['seconds'] / a.steps * 1e3:.1f} ms/iteration, "
            f"{g['updates'] / g['seconds']:.3e} updates/s")
        conv = None
        like = None
        if not a.no_converge:
            conv = {}
            starts = [w for w in a.starts.split(",") if w]
            for warm in starts:
                if warm not in ("multires_fbp", "unet", "fbp", "multires", "zero"):
                    raise SystemExit(f"bench.py: unknown start {warm!r}")
                c = gpu_converge(cfg, work, xt_hip, a.converge_iters, warm)
                if c is None:
                    log(f"{warm} start not measured (no converged reference or weights)")
                    continue
                conv[warm] = c
                log(f"{warm} start: within 1% RMSE of the converged image "
                    f"after {c['iters']} iterations, {c['seconds']} s")
            partial = len(starts) != 5
            conv = conv if "fbp" in conv else None
            like = (cpu_comparison(cfg, work, xt_hip) if a.workload == "slab" and not partial
                    else None)
        # the ceiling of the line kernel's e update: its atomic write pattern alone
        atomic_rate = None
        if a.workload == "slab":
            atomic_rate = xt_hip.probe_footprint_atomics()
            log(f"footprint atomic probe: {atomic_rate:.3e} FP32 atomic lane-ops/s")
        cpu = None
        if not a.no_cpu:
            cpu = cpu_sample(cfg, work, a.cpu_seconds, phase_contrast=a.workload == "pc")
            log(f"cpu oracle: {cpu['rate']:.3e} updates/s on {cpu['threads']} threads "
                f"({cpu['updates']} updates in {cpu['seconds']:.1f}s)")
    finally:
        if a.workdir is None:
            shutil.rmtree(work, ignore_errors=True)
    # the dominant kernel's rate: algorithmic bytes of its launches over their
    # summed duration (HIP events), from the kernel-timing iterations
    achieved = (12.0 * g["kt_footprint"] / (g["kt_kernel_ms"] * 1e-3) / 1e9
                if g["kt_kernel_ms"] > 0 else None)
    # the same bytes of the timed iterations over their whole wall time
    achieved_wall = 12.0 * g["footprint"] / g["seconds"] / 1e9
    # HBM bytes per launch of the voxel-update kernel measured with rocprofv3
    # FETCH_SIZE / WRITE_SIZE on this workload, scaled by a 1 GiB copy
    traffic = None
    tf = os.path.join(ROOT, "profiles", "k_icd_lines_slab512x512x32_traffic.json")
    if os.path.exists(tf) and a.workload == "slab":
        traffic = json.load(open(tf))["traffic_bytes_per_launch"]
    out = {
        "metric": METRIC,
        "value": g["updates"] / g["seconds"],
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": g["seconds"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 12-ellipsoid phantom, Poisson I0=1e4, 0.5% zingers at 5x max, "
                "ring offsets N(0,0.01)",
        "config": {"workload": cfg.name + {"pc": "_phase_contrast",
                                           "edge": "_edge_only"}.get(a.workload, ""),
                   "volume": [cfg.N_xy, cfg.N_xy, cfg.N_z],
                   "views": cfg.N_p, "bins": [cfg.N_r, cfg.N_t], "step": "one ICD iteration",
                   "resync": "fp64 every iteration", "prior": "qGGMRF p=1.2 q=2 c=0.01",
                   "offset_update": cfg.update_offset},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                     "copy-calibrated; profiles/)",
                     "achieved_basis": "12 B per footprint entry of every performed voxel update "
                                       "/ summed launch duration (HIP events on the library "
                                       f"stream, {KERNEL_TIMING_STEPS} iterations after the "
                                       "timed ones)",
                     "kernel": "k_icd_lines_pc" if a.workload == "pc" else "k_icd_lines",
                     "kernel_ms_per_launch": g["kt_kernel_ms"] / max(1, g["kt_launches"]),
                     "bytes_per_launch": 12.0 * g["kt_footprint"] / max(1, g["kt_launches"]),
                     "launches_per_step": g["launches"] / a.steps,
                     "bytes_per_step": 12.0 * g["footprint"] / a.steps,
                     "achieved_wall": achieved_wall,
                     "achieved_wall_basis": "the same bytes of the timed iterations / their wall "
                                            "time (every kernel of the step included)",
                     "atomic_ceiling": {
                         "kernel_lane_ops_per_s": (g["kt_footprint"] / (g["kt_kernel_ms"] * 1e-3)
                                                   if g["kt_kernel_ms"] > 0 else None),
                         "probe_lane_ops_per_s": atomic_rate,
                         "frac": (g["kt_footprint"] / (g["kt_kernel_ms"] * 1e-3) / atomic_rate
                                  if g["kt_kernel_ms"] > 0 and atomic_rate else None),
                         "basis": "one FP32 atomic per footprint entry of a performed update, "
                                  "against xt_probe_footprint_atomics (the same write pattern "
                                  "alone, same run): the line kernel's binding ceiling"}},
        "voxel_updates_per_step": g["updates"] / a.steps,
        "line_visits_per_step": g["visits"] / a.steps,
        "lines_in_flight_per_frame": g["lines_in_flight"],
        "time_to_converge": ({
            "criterion": "RMSE <= 1% of the phantom's mean attenuation vs the converged MAP "
                         f"image ({conv['fbp']['ref_iters']} GPU ICD iterations from the FBP "
                         "start, the lowest MAP cost reached; 262144-voxel sample)",
            "starts": "zero: plain ICD; fbp: filtered backprojection with ring-offset guess; "
                      "unet: the warm-start 3D U-Net (BF16 MFMA, weights in "
                      "timbir-python_amd/weights/) on the FBP volume, inference timed; "
                      f"multires: the reference launcher's stages ({MULTIRES_ITERS[0]} ICD "
                      f"iterations at 128^2 from zero without offsets, {MULTIRES_ITERS[1]} at "
                      "256^2, initICD-2 upsampling between them; coarse states with the default "
                      "lines in flight and no FP64 recompute of e), stage set-up and iterations "
                      f"timed; multires_fbp: the same stages ({MULTIRES_FBP_ITERS[0]} and "
                      f"{MULTIRES_FBP_ITERS[1]} iterations) with the 128^2 one started from its "
                      "FBP and ring-offset guess",
            "speedup_multires_fbp_vs_fbp": _ratio(conv, "fbp", "multires_fbp"),
            "speedup_multires_fbp_vs_zero": _ratio(conv, "zero", "multires_fbp"),
            "speedup_unet_vs_fbp": _ratio(conv, "fbp", "unet"),
            "speedup_unet_vs_zero": _ratio(conv, "zero", "unet"),
            **{k: {"seconds": c["seconds"], "iterations": c["iters"],
                   "init_seconds": round(c["init_seconds"], 4),
                   "rmse_by_iteration": [round(r, 5) for r in c["rmse"]],
                   "final_cost64": c["final_cost64"]}
               for k, c in conv.items()}} if conv else None),
        "like_for_like_vs_cpu_icd": like,
        "cpu_baseline": ({"value": cpu["rate"], "unit": "voxel-updates/s", "cores": cpu["threads"],
                          "kind": "port",
                          "sample": f"one homogeneous half-phase of the oracle's ICD schedule: "
                                    f"{cpu['tasks']} OpenMP tasks (z-blocks of {cfg.N_z // cpu['nb']} "
                                    f"slices), {cpu['lines']} lines each, {cpu['updates']} voxel "
                                    f"updates in {cpu['seconds']:.1f}s"}
                         if cpu else None),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
