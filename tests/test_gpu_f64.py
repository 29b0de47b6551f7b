"""The double build on the GPU (libxt_hip_f64.so through xt_hip_f64: Real_t =
double, the reference's REAL_IS_DOUBLE build, Source_Code_4D_Fast/
XT_Constants.h:45-53).  Its structs and .bin files are double; it keeps the float
build's system matrix, so its comparator is the oracle's double-arithmetic build
on that matrix (liboracle_f64a.so), read from the same double run folders: the
detector table bit-exact, Ax within 1e-5, per-voxel updates within 1e-4, and an
engine run through ICD_BackProject within 1% RMSE and 1e-4 MAP cost."""
import ctypes as C
import dataclasses
import os

import numpy as np
import pytest

from conftest import SMALL, TINY, as_double, make_scan
from timbir_python_amd import xt_abi, xt_hip_f64

pytestmark = pytest.mark.gpu


def _pair(cfg, scan_root, oracle, tag="f64"):
    path, info = make_scan(cfg, scan_root, tag)
    dpath = as_double(path)
    args = cfg.cli_args()
    eng = xt_hip_f64.Engine(dpath, args)
    S, O, T = oracle.setup(dpath, args)
    oracle.read_sinogram(dpath, S, T)
    return dpath, info, eng, (S, O, T)


def test_f64_detector_table_and_projection(scan_root, oracle64a):
    print("[gpu] double build: detector table, Ax (tiny)", flush=True)
    path, info, eng, (S, O, T) = _pair(TINY, scan_root, oracle64a)
    try:
        Hg = xt_hip_f64.detector_response(eng)
        Ho = oracle64a.detector_response(S, O, T)
        assert Hg.dtype == np.float64
        assert np.array_equal(Hg, Ho)
        vol = info["phantom"].astype(np.float64)
        oracle64a.set_object(O, vol)
        eng.set_volume(vol)
        ax_ref = oracle64a.forward_project(S, O, Ho)
        axg = xt_hip_f64.forward_project(eng, Hg)
        rel = np.abs(axg - ax_ref).max() / np.abs(ax_ref).max()
        print(f"  max|Ax_gpu - Ax_ref| / max|Ax| = {rel:.3e}")
        assert rel <= 1e-5
    finally:
        eng.close()
        oracle64a.free(S, O, T)


@pytest.mark.parametrize("Iter", [1, 2])
def test_f64_update_lines_per_voxel(Iter, scan_root, oracle64a):
    print(f"[gpu] double build: per-voxel update_lines Iter={Iter} (tiny)", flush=True)
    path, info, eng, (S, O, T) = _pair(TINY, scan_root, oracle64a)
    try:
        vol = 0.7 * info["phantom"].astype(np.float64)
        vol[:, :, :4, :] = 0
        oracle64a.set_object(O, vol)
        eng.set_volume(vol)
        H = oracle64a.detector_response(S, O, T)
        E = np.zeros((S.N_p, S.N_r, S.N_t), np.float64)
        et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(H)
        oracle64a.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
        Eg = E.copy()
        rng = np.random.default_rng(Iter)
        n = 40
        ys = rng.integers(0, O.N_y, n).astype(np.int32)
        xs = rng.integers(0, O.N_x, n).astype(np.int32)
        ys[:4] = 1
        mag_o = np.zeros((O.N_y, O.N_x), np.float64)
        mt = xt_abi.PtrTree(mag_o)
        zc = C.c_int64(0)
        tot_o = oracle64a.lib.oracle_update_lines(C.byref(S), C.byref(O), C.byref(T), et.ptr,
                                                  ht.ptr, 0, 0, O.N_z - 1, n, xs.ctypes.data,
                                                  ys.ctypes.data, Iter, C.byref(zc), mt.ptr, 0)
        mag_g, tot_g = xt_hip_f64.update_lines(eng, Eg, 0, xs, ys, Iter)
        xo = oracle64a.get_object(O)
        xg = eng.volume()
        scale = np.abs(xo).max()
        print(f"  max |dx| {np.abs(xg - xo).max() / scale:.3e}  e {np.abs(Eg - E).max():.3e}")
        assert xg.dtype == np.float64 and Eg.dtype == np.float64
        assert np.abs(xg - xo).max() <= 1e-4 * scale
        assert np.abs(Eg - E).max() <= 1e-4 * np.abs(E).max()
        assert np.abs(mag_g - mag_o).max() <= 1e-4 * max(1e-12, np.abs(mag_o).max())
        assert abs(tot_g - tot_o) <= 1e-4 * max(1e-12, abs(tot_o))
    finally:
        eng.close()
        oracle64a.free(S, O, T)


def _cost(oracle, path, cfg, vol, off):
    """FP64 MAP cost of (vol, off) on the double run folder `path`."""
    S, O, T = oracle.setup(path, cfg.cli_args())
    try:
        oracle.read_sinogram(path, S, T)
        oracle.set_object(O, vol)
        oracle.offset_view(S)[:] = off
        H = oracle.detector_response(S, O, T)
        E = np.zeros((S.N_p, S.N_r, S.N_t), np.float64)
        et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(H)
        oracle.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
        return float(oracle.lib.oracle_compute_cost(C.byref(S), C.byref(O), C.byref(T), et.ptr))
    finally:
        oracle.free(S, O, T)


def test_f64_engine_run_converged(scan_root, oracle64a):
    """The reference command line on double run folders: xt_engine_run (its
    ICD_BackProject) on the GPU and the oracle's engine, 150 iterations on the
    64^2 x 8 slab: objects (float64 files) within 1% RMSE of the phantom's mean
    attenuation, FP64 MAP costs within 1e-4."""
    cfg = dataclasses.replace(SMALL, iterations=150)
    pg, info = make_scan(cfg, scan_root, "f64gpu")
    pc, _ = make_scan(cfg, scan_root, "f64cpu")
    pg, pc = as_double(pg), as_double(pc)
    args = cfg.cli_args()
    print(f"[gpu] double build: engine_run {cfg.name} {cfg.iterations} iterations", flush=True)
    xt_hip_f64.engine_run(pg, args)
    oracle64a.run(pc, args)

    def read(p):
        vol = np.stack([np.fromfile(os.path.join(p, f"object_n0_time_{i}.bin"), np.float64)
                        .reshape(cfg.N_z, cfg.N_xy, cfg.N_xy) for i in range(cfg.N_time)])
        off = np.fromfile(os.path.join(p, "proj_offset_n0.bin"), np.float64)
        return vol, off.reshape(cfg.N_r, cfg.N_t)
    (xg, dg), (xc, dc) = read(pg), read(pc)
    ph = info["phantom"]
    rmse = np.sqrt(np.mean((xg - xc) ** 2)) / ph[ph > 0].mean()
    cg, cc = _cost(oracle64a, pg, cfg, xg, dg), _cost(oracle64a, pc, cfg, xc, dc)
    print(f"  RMSE(gpu-oracle)/mean = {rmse:.4%}, cost gpu {cg:.9g} oracle {cc:.9g}, "
          f"rel {abs(cg - cc) / abs(cc):.2e}")
    assert np.all(np.isfinite(xg)) and np.abs(xg).max() > 0
    assert rmse <= 0.01
    assert abs(cg - cc) <= 1e-4 * abs(cc)
