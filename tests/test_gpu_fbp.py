"""GPU filtered backprojection (xt_icd_fbp) against the float64 restatement in
oracle/fbp_oracle.py on the same run folder, and as an ICD warm start."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

import fbp_oracle
from conftest import SMALL, SMALL4D, TINY, make_scan
from timbir_python_amd import xt_hip

pytestmark = pytest.mark.gpu


def host_fbp(eng):
    S, O = eng.S, eng.O
    n = S.N_p
    arr = lambda ptr: np.ctypeslib.as_array((C.c_float * n).from_address(ptr)).astype(np.float64)
    angles, cosv, sinv = arr(S.ViewPtr), arr(S.cosine), arr(S.sine)
    top = C.cast(O.ProjIdxPtr, C.POINTER(C.c_void_p))
    num = np.ctypeslib.as_array((C.c_int32 * O.N_time).from_address(O.ProjNum))
    frames = [list(np.ctypeslib.as_array((C.c_int32 * int(num[f])).from_address(top[f])))
              for f in range(O.N_time)]
    geom = dict(N_x=O.N_x, N_y=O.N_y, x0=O.x0, y0=O.y0, delta_xy=O.delta_xy, R0=S.R0,
                delta_r=S.delta_r)
    y = eng.sino("Projection")
    d = fbp_oracle.offset_guess(y, bool(eng.T.EnforceZeroMeanOffset)) if eng.T.updateProjOffset > 0 \
        else eng.offset()
    return fbp_oracle.fbp(y, d, angles, cosv, sinv, frames, geom)


@pytest.mark.parametrize("cfg", [TINY, SMALL4D], ids=lambda c: c.name)
def test_fbp_matches_oracle(cfg, scan_root):
    path, info = make_scan(cfg, scan_root, "fbp")
    print(f"[gpu] fbp {cfg.name}", flush=True)
    with xt_hip.Engine(path, cfg.cli_args()) as eng:
        want = host_fbp(eng)
        with xt_hip.Icd(eng) as icd:
            icd.fbp()
            icd.download()
        got = eng.volume()
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"  max |gpu - oracle| / max |oracle| = {err:.2e}")
    assert err <= 1e-4
    # (its attenuation scale is pinned by the disk closed form, tests/test_oracle_fbp.py)


def test_fbp_warm_start_helps(scan_root):
    """From the FBP start, ICD's MAP cost after 3 iterations is below the cost
    after 3 iterations from zero."""
    cfg = dataclasses.replace(SMALL, iterations=3)
    path, info = make_scan(cfg, scan_root, "fbpws")
    costs = {}
    for warm in (0, 1):
        with xt_hip.Engine(path, cfg.cli_args()) as eng, xt_hip.Icd(eng) as icd:
            c0 = icd.cost()
            if warm:
                icd.fbp()
            c1 = icd.cost()
            icd.iterate(3)
            costs[warm] = (c0, c1, icd.cost())
    print(f"  zero start {costs[0]}, fbp start {costs[1]}")
    assert costs[1][1] < costs[1][0]          # FBP lowers the cost of the zero object
    assert costs[1][2] < costs[0][2]
