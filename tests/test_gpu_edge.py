"""Edge-only updates (--only_Edge_Updates: F/XT_ImageProc.c:83-170,
F/XT_ICD_update.c:996-997, the Mask of F/XT_VoxUpdate.c:138) on the GPU against
the C oracle (its edge mask pinned against scipy.ndimage in
tests/test_oracle_pins.py).

Bars:
  bit-exact        the edge mask (integer output of float filters summed in the
                   reference's order)
  1e-4 relative    per-voxel updates of masked lines; voxels outside the mask
                   bit-identical to the input
  1% RMSE, 1e-4    converged edge-only run from a given object (RMSE over the
                   phantom's mean attenuation, FP64 MAP cost); voxels outside the
                   mask unchanged
"""
import ctypes as C
import dataclasses
import os

import numpy as np
import pytest
from scipy import ndimage

from conftest import SMALL4D, TINY, make_scan
from timbir_python_amd import xt_abi, xt_hip

pytestmark = pytest.mark.gpu

RADIUS = "16"


def _args(cfg, **kw):
    a = cfg.cli_args(**kw)
    a[a.index("--radius_obj") + 1] = RADIUS
    return a


def _object(info, seed):
    """The phantom plus, in every slice, a smooth random field scaled so that
    its gradient straddles the 0.0002 threshold (exercises the comparison)."""
    ph = info["phantom"].astype(np.float64)
    rng = np.random.default_rng(seed)
    f = ndimage.gaussian_filter(rng.normal(size=ph.shape), (0, 0, 2.0, 2.0))
    return (ph + 2e-4 * f / np.abs(f).max()).astype(np.float32)


@pytest.mark.parametrize("cfg", [TINY, SMALL4D], ids=lambda c: c.name)
def test_edge_detect_bitexact(cfg, scan_root, oracle32):
    print(f"[gpu] edge mask {cfg.name}", flush=True)
    path, info = make_scan(cfg, scan_root, "edge")
    vol = _object(info, 3)
    with xt_hip.Engine(path, _args(cfg)) as eng:
        eng.set_volume(vol)
        got = xt_hip.edge_detect(eng)
    S, O, T = oracle32.setup(path, _args(cfg))
    try:
        oracle32.set_object(O, vol)
        want = oracle32.edge_detect(O, T)
    finally:
        oracle32.free(S, O, T)
    print(f"  masked fraction {want.mean():.3f}")
    assert 0.05 < want.mean() < 0.95
    np.testing.assert_array_equal(got, want)


def test_update_lines_masked_per_voxel(scan_root, oracle32):
    cfg = TINY
    print("[gpu] update_lines with the edge mask", flush=True)
    path, info = make_scan(cfg, scan_root, "edgeupd")
    args = _args(cfg)
    eng = xt_hip.Engine(path, args)
    S, O, T = oracle32.setup(path, args)
    oracle32.read_sinogram(path, S, T)
    try:
        vol = (0.7 * info["phantom"]).astype(np.float32)
        oracle32.set_object(O, vol)
        eng.set_volume(vol)
        mask = oracle32.edge_detect(O, T)
        H = oracle32.detector_response(S, O, T)
        E = np.zeros((S.N_p, S.N_r, S.N_t), np.float32)
        et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(H)
        oracle32.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
        Eg = E.copy()
        rng = np.random.default_rng(21)
        n = 60
        ys = rng.integers(0, O.N_y, n).astype(np.int32)
        xs = rng.integers(0, O.N_x, n).astype(np.int32)
        mag_o = np.zeros((O.N_y, O.N_x), np.float32)
        mt = xt_abi.PtrTree(mag_o)
        mk = xt_abi.PtrTree(np.ascontiguousarray(mask[0]))
        zc = C.c_int64(0)
        tot_o = oracle32.lib.oracle_update_lines_mask(
            C.byref(S), C.byref(O), C.byref(T), et.ptr, ht.ptr, 0, 0, O.N_z - 1, n,
            xs.ctypes.data, ys.ctypes.data, 1, C.byref(zc), mt.ptr, mk.ptr, 0)
        mag_g, tot_g = xt_hip.update_lines(eng, Eg, 0, xs, ys, 1, mask=mask)
        xo = oracle32.get_object(O)
        xg = eng.volume()
        scale = np.abs(xo).max()
        off = mask.transpose(0, 3, 1, 2) == 0  # [T][Z][Y][X]
        print(f"  max |dx| {np.abs(xg - xo).max() / scale:.3e}; unmasked voxels changed: "
              f"{int((xg != vol)[off].sum())}")
        assert np.array_equal(xg[off], vol[off])
        assert np.array_equal(xo[off], vol[off])
        assert (xg != vol).sum() > 0
        assert np.abs(xg - xo).max() <= 1e-4 * scale
        assert np.abs(Eg - E).max() <= 1e-4 * np.abs(E).max()
        assert np.abs(mag_g - mag_o).max() <= 1e-4 * max(1e-12, np.abs(mag_o).max())
        assert abs(tot_g - tot_o) <= 1e-4 * max(1e-12, abs(tot_o))
    finally:
        eng.close()
        oracle32.free(S, O, T)


def _start_files(path, cfg, vol):
    for f in range(cfg.N_time):
        vol[f].astype(np.float32).tofile(os.path.join(path, f"object_n0_time_{f}.bin"))
    np.array([1.0], np.float32).tofile(os.path.join(path, "variance_estimate_n0.bin"))


@pytest.mark.parametrize("cfg", [
    dataclasses.replace(TINY, iterations=100, update_offset=0),
    dataclasses.replace(TINY, name="tiny_nb4", iterations=100, update_offset=0, num_threads=8)],
    ids=["tiny", "tiny-4-blocks"])
def test_edge_only_converged_parity(cfg, scan_root, oracle32, oracle64a):
    """ICD_BackProject with --only_Edge_Updates from the phantom (initICD 1),
    ring offsets held (updateProjOffset 0), 100 iterations on the GPU and in the
    float oracle (the oracle moves 0.13% of the mean attenuation between 80 and
    160 iterations on this scan), with two z-blocks and with four (num_threads 8:
    each z-block parity's NHICD share is counted from the lines with a masked voxel
    in some block of that parity); about half the voxels are masked.  (A four-frame
    edge-only run from the phantom is still 7% from its own 200-iteration image
    after 100 iterations in the oracle, too slow for a converged comparison.)"""
    from test_gpu_icd import map_cost, read_object
    p_gpu, info = make_scan(cfg, scan_root, "edgegpu")
    p_cpu, _ = make_scan(cfg, scan_root, "edgecpu")
    vol = info["phantom"].astype(np.float32)
    _start_files(p_gpu, cfg, vol)
    _start_files(p_cpu, cfg, vol)
    args = _args(cfg, init_icd=1) + ["--only_Edge_Updates"]
    print("[gpu] edge-only engine_run, GPU and oracle", flush=True)
    xt_hip.engine_run(p_gpu, args)
    oracle32.run(p_cpu, args)
    xg, xc = read_object(p_gpu, cfg), read_object(p_cpu, cfg)
    S, O, T = oracle32.setup(p_cpu, _args(cfg))
    try:
        oracle32.set_object(O, vol)
        off = oracle32.edge_detect(O, T).transpose(0, 3, 1, 2) == 0
    finally:
        oracle32.free(S, O, T)
    assert 0.2 < off.mean() < 0.8
    assert np.array_equal(xg[off], vol[off]) and np.array_equal(xc[off], vol[off])
    ph = info["phantom"]
    rmse = np.sqrt(np.mean((xg - xc) ** 2)) / ph[ph > 0].mean()
    d0 = np.zeros((cfg.N_r, cfg.N_t), np.float32)
    cg = map_cost(oracle64a, p_gpu, cfg, xg, d0)
    cc = map_cost(oracle64a, p_cpu, cfg, xc, d0)
    moved = np.sqrt(np.mean((xg - vol) ** 2)) / ph[ph > 0].mean()
    print(f"  RMSE(gpu-oracle)/mean = {rmse:.4%} (moved {moved:.2%} from the start), "
          f"cost gpu {cg:.9g} oracle {cc:.9g}")
    assert moved > 0.01
    assert rmse <= 0.01
    assert abs(cg - cc) <= 1e-4 * abs(cc)
