"""End-to-end GPU ICD against the CPU oracle.

Both engines run the reference command line on identical run folders.  ICD is
Gauss-Seidel in a randomized order on both sides, so trajectories differ; the
bar is at convergence: the reconstructions agree to within 1% RMSE relative to
the phantom's mean attenuation, and their MAP costs -- both evaluated by the
same FP64 oracle evaluator from (object, offset, variance) -- agree to 1e-4
relative.
"""
import ctypes as C
import dataclasses
import os

import numpy as np
import pytest

from conftest import SMALL, SMALL4D, TINY, make_scan, oracle_path
from timbir_python_amd import xt_abi, xt_hip

pytestmark = pytest.mark.gpu


def read_object(path, cfg):
    return np.stack([np.fromfile(os.path.join(path, f"object_n0_time_{i}.bin"), np.float32)
                     .reshape(cfg.N_z, cfg.N_xy, cfg.N_xy) for i in range(cfg.N_time)])


def map_cost(oracle64, path, cfg, vol, offset, var_est=1.0):
    """FP64 MAP cost of a result (oracle f64a build: the float system matrix in
    double arithmetic): e = y - Ax - d recomputed, then computeCost."""
    path = oracle_path(path, oracle64)
    S, O, T = oracle64.setup(path, cfg.cli_args())
    try:
        oracle64.read_sinogram(path, S, T)
        oracle64.set_object(O, vol.astype(np.float64))
        oracle64.offset_view(S)[:] = offset
        H = oracle64.detector_response(S, O, T)
        E = np.zeros((S.N_p, S.N_r, S.N_t), np.float64)
        et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(H)
        oracle64.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
        return float(oracle64.lib.oracle_compute_cost(C.byref(S), C.byref(O), C.byref(T), et.ptr))
    finally:
        oracle64.free(S, O, T)


def run_both(cfg, scan_root, oracle32):
    p_gpu, info = make_scan(cfg, scan_root, "gpu")
    p_cpu, _ = make_scan(cfg, scan_root, "cpu")
    args = cfg.cli_args()
    print(f"[gpu] engine_run {cfg.name} iters {cfg.iterations}", flush=True)
    xt_hip.engine_run(p_gpu, args)
    oracle32.run(p_cpu, args)
    return p_gpu, p_cpu, info


S150 = dataclasses.replace(SMALL, iterations=150)
S4D30 = dataclasses.replace(SMALL4D, iterations=30)
GROUPS_OFF = pytest.mark.skip(reason="grouped line updates (XT_OPT_LINE_GROUP 2, 4) are switched "
                              "off: their first GPU run took a memory fault whose cause is not "
                              "found")


@pytest.mark.parametrize("cfg,nhicd_lines,line_threads,group",
                         [(S150, 1, 0, 1), (S150, 0, 0, 1), (S150, 1, 64, 1), (S4D30, 1, 0, 1),
                          (S4D30, 1, 64, 1),
                          pytest.param(S150, 1, 0, 2, marks=GROUPS_OFF),
                          pytest.param(S150, 1, 0, 4, marks=GROUPS_OFF),
                          pytest.param(S150, 0, 0, 4, marks=GROUPS_OFF),
                          pytest.param(S4D30, 1, 0, 4, marks=GROUPS_OFF),
                          pytest.param(S4D30, 1, 64, 4, marks=GROUPS_OFF)],
                         ids=["small3d", "small3d-nhicd-per-block", "small3d-one-wave-lines",
                              "small4d", "small4d-one-wave-lines", "small3d-groups2",
                              "small3d-groups4", "small3d-nhicd-per-block-groups4",
                              "small4d-groups4", "small4d-one-wave-groups4"])
def test_converged_parity(cfg, nhicd_lines, line_threads, group, scan_root, oracle32, oracle64a):
    xt_hip.set_option(xt_hip.OPT_NHICD_LINES, nhicd_lines)
    xt_hip.set_option(xt_hip.OPT_LINE_THREADS, line_threads)
    xt_hip.set_option(xt_hip.OPT_LINE_GROUP, group)
    try:
        p_gpu, p_cpu, info = run_both(cfg, scan_root, oracle32)
    finally:
        xt_hip.set_option(xt_hip.OPT_NHICD_LINES, 1)
        xt_hip.set_option(xt_hip.OPT_LINE_THREADS, 0)
        xt_hip.set_option(xt_hip.OPT_LINE_GROUP, 1)
    xg, xc = read_object(p_gpu, cfg), read_object(p_cpu, cfg)
    ph = info["phantom"]
    mean_att = ph[ph > 0].mean()
    rmse = np.sqrt(np.mean((xg - xc) ** 2)) / mean_att
    dg = np.fromfile(os.path.join(p_gpu, "proj_offset_n0.bin"), np.float32).reshape(cfg.N_r, cfg.N_t)
    dc = np.fromfile(os.path.join(p_cpu, "proj_offset_n0.bin"), np.float32).reshape(cfg.N_r, cfg.N_t)
    cg = map_cost(oracle64a, p_gpu, cfg, xg, dg)
    cc = map_cost(oracle64a, p_cpu, cfg, xc, dc)
    print(f"  RMSE(gpu-oracle)/mean = {rmse:.4%}, cost gpu {cg:.9g} oracle {cc:.9g}, "
          f"rel {abs(cg - cc) / abs(cc):.2e}")
    assert np.all(np.isfinite(xg))
    assert rmse <= 0.01
    assert abs(cg - cc) <= 1e-4 * abs(cc)


def test_stepwise_and_resync(scan_root):
    cfg = dataclasses.replace(TINY, iterations=5)
    path, info = make_scan(cfg, scan_root, "step")
    print(f"[gpu] stepwise {cfg.name}", flush=True)
    with xt_hip.Engine(path, cfg.cli_args()) as eng:
        with xt_hip.Icd(eng) as icd:
            icd.timing(True)
            c0 = icd.cost()
            icd.iterate(3)
            c3 = icd.cost()
            icd.resync()
            c3r = icd.cost()
            st = icd.stats()
            print(f"  cost {c0:.6g} -> {c3:.6g} (resync {c3r:.6g}), updates {st.voxel_updates}, "
                  f"icd ms {st.icd_kernel_ms:.3f}, launches {st.icd_launches}")
            assert c3 < c0
            assert abs(c3r - c3) <= 1e-4 * abs(c3)  # FP32 e drift vs FP64 recompute
            assert st.iterations == 3
            assert st.voxel_updates > 0 and st.footprint_entries > st.voxel_updates
            assert st.icd_kernel_ms > 0
            icd.download()
        vol = eng.volume()
        assert np.all(np.isfinite(vol)) and np.abs(vol).max() > 0


def test_graph_replay_matches_eager(scan_root):
    """Iterations replayed from one captured hipGraph against the same iterations
    enqueued eagerly, with eager cost evaluations and synchronisations between the
    replays: same launches, and every replay's cost agrees with its eager
    counterpart.  Two eager runs already differ (the order of the error-sinogram
    atomics and of concurrent reads of it varies from run to run, and ICD
    amplifies such differences in its first iterations), so the bar is that
    run-to-run spread: 1e-4 after the first iteration, 5e-3 after later ones."""
    cfg = dataclasses.replace(TINY, iterations=6)
    path, info = make_scan(cfg, scan_root, "graph")
    costs = {}
    for graph in (0, 1):
        xt_hip.set_option(xt_hip.OPT_GRAPH, graph)
        try:
            with xt_hip.Engine(path, cfg.cli_args()) as eng, xt_hip.Icd(eng) as icd:
                c = []
                for it in range(5):
                    icd.iterate(1)
                    c.append(icd.cost())
                    icd.synchronize()
                st = icd.stats()
                assert st.iterations == 5
                costs[graph] = (np.array(c), st.voxel_updates, st.icd_launches)
        finally:
            xt_hip.set_option(xt_hip.OPT_GRAPH, 1)
    ce, ue, le = costs[0]
    cg, ug, lg = costs[1]
    print(f"  eager {ce}\n  graph {cg}")
    assert abs(cg[0] - ce[0]) <= 1e-4 * abs(ce[0])
    assert np.all(np.abs(cg - ce) <= 5e-3 * np.abs(ce))
    assert np.all(np.diff(cg) < 0) and np.all(np.diff(ce) < 0)
    assert lg == le and ug > 0 and ue > 0


def test_icd_backproject_struct_boundary(scan_root, tmp_path):
    """The drop-in call itself: structs in, Projection freed, outputs written."""
    cfg = dataclasses.replace(TINY, iterations=3)
    path, info = make_scan(cfg, scan_root, "struct")
    eng = xt_hip.Engine(path, cfg.cli_args())
    try:
        cwd = os.getcwd()
        os.chdir(tmp_path)
        try:
            xt_hip.ICD_BackProject(eng.S, eng.O, eng.T)
        finally:
            os.chdir(cwd)
        assert not eng.S.Projection  # released by the callee
        assert os.path.exists(tmp_path / "variance_estimate_n0.bin")
        mag = np.fromfile(tmp_path / "mag_update_map_n0.bin", np.float32)
        assert mag.size == cfg.N_time * eng.T.num_z_blocks * cfg.N_xy * cfg.N_xy
        vol = eng.volume()
        assert np.abs(vol).max() > 0
        sel = eng.proj_select()
        assert 0.9 < sel.mean() <= 1.0
    finally:
        eng.close()


def test_positivity_option(scan_root):
    cfg = dataclasses.replace(TINY, iterations=4)
    path, info = make_scan(cfg, scan_root, "pos")
    xt_hip.set_option(xt_hip.OPT_POSITIVITY, 1)
    try:
        xt_hip.engine_run(path, cfg.cli_args())
    finally:
        xt_hip.set_option(xt_hip.OPT_POSITIVITY, 0)
    x = read_object(path, cfg)
    assert x.min() >= 0.0


def test_bad_input_is_an_error(scan_root):
    cfg = TINY
    path, info = make_scan(cfg, scan_root, "bad")
    os.remove(os.path.join(path, "weight_n0.bin"))
    with pytest.raises(xt_hip.XtError):
        xt_hip.engine_run(path, cfg.cli_args())


def test_packed_lanes_match_one_lane_per_slice(scan_root):
    """A pass over one z-block parity packs that parity's slices into half the
    lanes; with one line at a time (a deterministic sequential order) it must
    reproduce the one-lane-per-slice kernel to float rounding of the theta sums.
    NHICD is off: its order-statistic threshold turns rounding-level differences
    of the update maps into different line sets, and early iterations amplify
    those."""
    cfg = dataclasses.replace(TINY, iterations=3, no_nhicd=True)
    path, info = make_scan(cfg, scan_root, "pack")
    vols = {}
    xt_hip.set_option(xt_hip.OPT_LINES, 1)
    try:
        for pack in (0, 1):
            xt_hip.set_option(xt_hip.OPT_PACK, pack)
            with xt_hip.Engine(path, cfg.cli_args()) as eng:
                with xt_hip.Icd(eng) as icd:
                    icd.iterate(3)
                    icd.download()
                vols[pack] = eng.volume().copy()
    finally:
        xt_hip.set_option(xt_hip.OPT_PACK, 1)
        xt_hip.set_option(xt_hip.OPT_LINES, 0)
    err = np.abs(vols[1] - vols[0]).max() / np.abs(vols[0]).max()
    print(f"  packed vs one lane per slice: max rel diff {err:.2e}")
    assert np.abs(vols[0]).max() > 0
    assert err <= 1e-4


def test_autotune_restores_state(scan_root):
    """The launch-shape autotuner runs trial launches at create time and restores
    the reconstruction state: the FP64 cost right after create is the same with
    and without it, and the chosen number of lines in flight is a candidate."""
    cfg = dataclasses.replace(SMALL, iterations=2)
    path, info = make_scan(cfg, scan_root, "tune")
    res = {}
    try:
        for tune in (0, 1):
            xt_hip.set_option(xt_hip.OPT_AUTOTUNE, tune)
            with xt_hip.Engine(path, cfg.cli_args()) as eng, xt_hip.Icd(eng) as icd:
                c0 = icd.cost()
                w = icd.stats().lines_in_flight
                icd.iterate(2)
                res[tune] = (c0, w, icd.cost())
    finally:
        xt_hip.set_option(xt_hip.OPT_AUTOTUNE, 1)
    print(f"  without autotune {res[0]}, with {res[1]}")
    assert res[1][0] == res[0][0]
    assert res[0][1] == 64 and res[1][1] in (16, 32, 64)
    assert res[1][2] < res[1][0]


def test_footprint_atomic_probe():
    """The probe of the line kernel's write pattern runs and reports a rate between
    1 G and 10 T FP32 atomic lane operations per second."""
    print("[gpu] footprint atomic probe", flush=True)
    r = xt_hip.probe_footprint_atomics(2048)
    print(f"  {r:.3e} lane-ops/s")
    assert 1e9 < r < 1e13
    with pytest.raises(xt_hip.XtError):
        xt_hip.probe_footprint_atomics(0)

