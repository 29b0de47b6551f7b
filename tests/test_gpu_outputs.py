"""Files ICD_BackProject writes besides the final object: --WritePerIter (the
object and ring offsets after every iteration, write_ObjectProjOff2TiffBinPerIter,
Source_Code_4D_Fast/XT_ICD_update.c:1019-1020) and the MAP-cost file of a build
without NO_COST_CALCULATE (cost_n<rank>.bin: the cost before ICD, then one value
per iteration, Real_t each, :1001-1007,1024-1027)."""
import dataclasses
import os

import numpy as np
import pytest

from conftest import TINY, make_scan
from timbir_python_amd import xt_hip

pytestmark = pytest.mark.gpu


def _read(path, name, n=None):
    return np.fromfile(os.path.join(path, name), np.float32)


def test_write_per_iter_and_cost_file(scan_root):
    cfg = dataclasses.replace(TINY, iterations=4)
    p_iter, _ = make_scan(cfg, scan_root, "wpi")
    p_end, _ = make_scan(cfg, scan_root, "end")
    # one line in flight: a run is deterministic, so the two runs must agree bit for bit
    xt_hip.set_option(xt_hip.OPT_LINES, 1)
    xt_hip.set_option(xt_hip.OPT_COST, 1)
    try:
        print("[gpu] engine_run --WritePerIter and a plain run, 4 iterations", flush=True)
        xt_hip.engine_run(p_iter, cfg.cli_args() + ["--WritePerIter"])
        xt_hip.engine_run(p_end, cfg.cli_args())
    finally:
        xt_hip.set_option(xt_hip.OPT_LINES, 0)
        xt_hip.set_option(xt_hip.OPT_COST, 0)
    n_obj = cfg.N_z * cfg.N_xy * cfg.N_xy
    # with WritePerIter the engine skips its final write (F/XT_Engine.c:161-162), so
    # these files come from ICD_BackProject's per-iteration writes
    o_iter, o_end = _read(p_iter, "object_n0_time_0.bin"), _read(p_end, "object_n0_time_0.bin")
    d_iter, d_end = _read(p_iter, "proj_offset_n0.bin"), _read(p_end, "proj_offset_n0.bin")
    assert o_iter.size == n_obj and d_iter.size == cfg.N_r * cfg.N_t
    assert np.abs(o_iter).max() > 0
    assert np.array_equal(o_iter, o_end) and np.array_equal(d_iter, d_end)
    for p in (p_iter, p_end):
        cost = _read(p, "cost_n0.bin")
        print(f"  costs {cost}", flush=True)
        assert cost.size == cfg.iterations + 1
        assert np.all(np.isfinite(cost)) and cost[-1] < cost[0]
    assert np.array_equal(_read(p_iter, "cost_n0.bin"), _read(p_end, "cost_n0.bin"))


def test_sinobin3_run_equals_binned_files(tmp_path):
    """ICD_BackProject on a sinobin-3 folder (counts binned by the engine over pairs
    of detector rows) equals the run on a sinobin-1 folder holding the binned
    projection and weights with the binned detector (N_t halved over the same
    length: rows of pitch 2), one line in flight (deterministic).  Binning along the
    bins instead would also rescale the prior (mult_xy enters sigma and the filter
    weights, F/XT_Init.c:215,391), so it has no sinobin-1 twin."""
    import test_host
    d3, args3, _, _ = test_host._counts_run(tmp_path / "s3", 1, 2, n_p=24, nr_full=32, nt_full=8)
    with xt_hip.Engine(d3, args3) as eng:
        y = eng.sino("Projection").copy()
        w = eng.sino("Weight").copy()
        n_t = eng.S.N_t
        assert n_t == 4 and eng.S.delta_t == 2.0
    d1 = str(tmp_path / "s1")
    os.makedirs(d1)
    with open(os.path.join(d3, "view_info.txt")) as src, \
            open(os.path.join(d1, "view_info.txt"), "w") as dst:
        dst.write(src.read())
    y.tofile(os.path.join(d1, "projection_n0.bin"))
    w.tofile(os.path.join(d1, "weight_n0.bin"))
    args1 = list(args3)
    for k, v in (("--sinobin", "1"), ("--delta_z", "1"), ("--N_t", str(n_t)), ("--iter", "4")):
        args1[args1.index(k) + 1] = v
    args3 = list(args3)
    args3[args3.index("--iter") + 1] = "4"
    xt_hip.set_option(xt_hip.OPT_LINES, 1)
    try:
        print("[gpu] engine_run sinobin 3 and sinobin 1, 4 iterations", flush=True)
        xt_hip.engine_run(d3, args3)
        xt_hip.engine_run(d1, args1)
    finally:
        xt_hip.set_option(xt_hip.OPT_LINES, 0)
    o3, o1 = _read(d3, "object_n0_time_0.bin"), _read(d1, "object_n0_time_0.bin")
    print(f"  max |sinobin3 - sinobin1| {np.abs(o3 - o1).max():.3e}, max {np.abs(o1).max():.3e}")
    assert o3.size == o1.size == 32 * 32 * 4
    assert np.abs(o1).max() > 0
    np.testing.assert_array_equal(o3, o1)
