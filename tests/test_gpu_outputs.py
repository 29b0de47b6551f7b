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
