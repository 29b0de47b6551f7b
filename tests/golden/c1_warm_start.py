"""The common warm start of the plain 3D workload's converged comparison: the
filtered backprojection of oracle/fbp_oracle.py (float64 numpy) with its
ring-offset guess, handed to an engine the way the reference takes a starting
object -- object_n0_time_0.bin with --initICD 1 (Source_Code_4D_Fast/
XT_ICD_update.c:443-483), the offsets in proj_offset_n0.bin with
--updateProjOffset 3 (read, then estimated; XT_Init.c), variance 1.  Used by
make_c1_reference.py (the oracle's run) and tests/test_gpu_configs.py (the
GPU's), so both start from the same files."""
import os

import numpy as np

import fbp_oracle
from timbir_python_amd import xt_hip


def warm_start_args(cfg, iterations):
    args = cfg.cli_args(iterations=iterations, init_icd=1)
    args[args.index("--updateProjOffset") + 1] = "3"
    return args


def write_warm_start(path, cfg):
    """Start files for the scan in path (projection, weight, view_info present)."""
    import ctypes as C
    with xt_hip.Engine(path, cfg.cli_args()) as eng:
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
        y = eng.sino("Projection").copy()
    d = fbp_oracle.offset_guess(y, cfg.zero_mean_offset)
    x = fbp_oracle.fbp(y, d, angles, cosv, sinv, frames, geom)
    for f in range(cfg.N_time):
        x[f].astype(np.float32).tofile(os.path.join(path, f"object_n0_time_{f}.bin"))
    d.astype(np.float32).tofile(os.path.join(path, "proj_offset_n0.bin"))
    np.array([1.0], np.float32).tofile(os.path.join(path, "variance_estimate_n0.bin"))
    return x, d
