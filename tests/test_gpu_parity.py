"""GPU parity tests: every hot-path seam of libxt_hip.so against the CPU oracle
on the same seeded inputs (and against the reference's own buildable sources
in oracle/_ref where they cover the function).

Tolerances (written per test):
  bit-exact        detector table, A-matrix columns (bins and values), NHICD line sets
  1e-5 relative    forward projection Ax (max |GPU - Ax_ref| / max |Ax|), Ax_ref = the
                   float build's system matrix summed in FP64 (oracle f64a build)
  1e-5 relative    FP64 MAP cost, ring offsets, variance
  1e-4 relative    per-voxel updates (FP32 theta sums in a different order)
"""
import ctypes as C
import dataclasses
import os

import numpy as np
import pytest

from conftest import SMALL, SMALL4D, TINY, make_scan, oracle_path
from timbir_python_amd import xt_abi, xt_hip
import py_oracle

pytestmark = pytest.mark.gpu

# voxels 3 and 6 detector bins wide (the coarse stages of a multi-resolution run):
# A columns of up to 8 and 14 bins, the line kernel's wide-column variant
WIDE3 = dataclasses.replace(TINY, name="wide3", N_xy=16, mult_xy=3.0)
WIDE6 = dataclasses.replace(TINY, name="wide6", N_xy=8, mult_xy=6.0)
# the production slab's z shape (32 slices in 32 one-slice z-blocks: 32 z lanes,
# 16 packed per parity) and the deepest slab a GPU takes (64 slices, 64 lanes)
NZ32 = dataclasses.replace(TINY, name="nz32", N_xy=32, N_z=32, N_p=48, num_threads=32, seed=21)
NZ64 = dataclasses.replace(TINY, name="nz64", N_xy=24, N_z=64, N_p=48, num_threads=64, seed=23)
# slices of two detector rows (delta_z = 2, z_overlap_num 2: a coarse-z multi-resolution stage)
MZ2 = dataclasses.replace(TINY, name="mz2", N_z=2, mult_z=2, seed=25)


def _print_case(name, cfg):
    print(f"[gpu] {name}: {cfg.name} {cfg.N_xy}x{cfg.N_xy}x{cfg.N_z} x{cfg.N_time} views {cfg.N_p}",
          flush=True)


def _pair(cfg, scan_root, oracle):
    path, info = make_scan(cfg, scan_root)
    args = cfg.cli_args()
    eng = xt_hip.Engine(path, args)
    opath = oracle_path(path, oracle)
    S, O, T = oracle.setup(opath, args)
    oracle.read_sinogram(opath, S, T)
    return path, info, eng, (S, O, T)


@pytest.mark.parametrize("cfg", [TINY, SMALL4D], ids=lambda c: c.name)
def test_detector_response_bitexact(cfg, scan_root, oracle32):
    _print_case("detector_response", cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle32)
    try:
        Hg = xt_hip.detector_response(eng)
        Ho = oracle32.detector_response(S, O, T)
        assert np.array_equal(Hg.view(np.uint32), Ho.view(np.uint32))
        if py_oracle.RefLib.available():
            ref = py_oracle.RefLib(False)
            Hr = np.zeros_like(Ho)
            ht = xt_abi.PtrTree(Hr)
            Ht = np.zeros(65, np.float32)
            ref.lib.DetectorResponseProfile(ht.ptr, C.c_void_p(Ht.ctypes.data), C.byref(S),
                                            C.byref(O), C.byref(T))
            assert np.array_equal(Hg.view(np.uint32), Hr.view(np.uint32))
    finally:
        eng.close()
        oracle32.free(S, O, T)


@pytest.mark.parametrize("cfg", [TINY, WIDE3, WIDE6], ids=lambda c: c.name)
def test_amatrix_columns_bitexact(cfg, scan_root, oracle32):
    _print_case("amatrix_columns", cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle32)
    try:
        H = oracle32.detector_response(S, O, T)
        rng = np.random.default_rng(0)
        n = 4000
        rows = rng.integers(0, O.N_y, n)
        cols = rng.integers(0, O.N_x, n)
        views = rng.integers(0, S.N_p, n)
        # include the image corners and every view
        rows[:S.N_p] = 0
        cols[:S.N_p] = O.N_x - 1
        views[:S.N_p] = np.arange(S.N_p)
        cnt, idx, val = xt_hip.amatrix_columns(eng, H, rows, cols, views, max_entries=16)
        AM = oracle32.AMatrixCol
        vals = np.zeros(16, np.float32)
        inds = np.zeros(16, np.int32)
        a = AM()
        a.values = vals.ctypes.data
        a.index = inds.ctypes.data
        ht = xt_abi.PtrTree(H)
        for i in range(n):
            oracle32.lib.oracle_amatrix_column(C.byref(S), C.byref(O), ht.ptr, C.byref(a),
                                               int(rows[i]), int(cols[i]), int(views[i]))
            assert cnt[i] == a.count
            k = a.count
            assert np.array_equal(idx[i, :k], inds[:k])
            assert np.array_equal(val[i, :k].view(np.uint32), vals[:k].view(np.uint32))
    finally:
        eng.close()
        oracle32.free(S, O, T)


@pytest.mark.parametrize("atable", [1, 0], ids=["atable", "onthefly"])
@pytest.mark.parametrize("cfg", [TINY, SMALL4D, WIDE3, WIDE6, NZ32, NZ64, MZ2],
                         ids=lambda c: c.name)
def test_forward_project(cfg, atable, scan_root, oracle32, oracle64a):
    """Ax within 1e-5 relative of the float reference system matrix (FP64 sums),
    with the columns read from the A-column table (the ICD state's resync) and
    built on the fly.  The reference's double build uses a different detector
    table (its float and double profiles differ by ~3e-3), so the float build is
    the target."""
    _print_case(f"forward_project atable={atable}", cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle32)
    xt_hip.set_option(xt_hip.OPT_ATABLE, atable)
    try:
        vol = info["phantom"].astype(np.float32)
        oracle32.set_object(O, vol)
        eng.set_volume(vol)
        H = oracle32.detector_response(S, O, T)
        ax32 = oracle32.forward_project(S, O, H).astype(np.float64)
        S2, O2, T2 = oracle64a.setup(oracle_path(path, oracle64a), cfg.cli_args())
        try:
            oracle64a.set_object(O2, vol.astype(np.float64))
            ax_ref = oracle64a.forward_project(S2, O2, oracle64a.detector_response(S2, O2, T2))
        finally:
            oracle64a.free(S2, O2, T2)
        axg = xt_hip.forward_project(eng, xt_hip.detector_response(eng))
        rel = np.abs(axg - ax_ref).max() / np.abs(ax_ref).max()
        rel32 = np.abs(ax32 - ax_ref).max() / np.abs(ax_ref).max()
        print(f"  max|Ax_gpu - Ax_ref| / max|Ax| = {rel:.3e} (float-summed oracle: {rel32:.3e})")
        assert rel <= 1e-5
    finally:
        xt_hip.set_option(xt_hip.OPT_ATABLE, 1)
        eng.close()
        oracle32.free(S, O, T)


def test_init_error_sino(scan_root, oracle32):
    cfg = TINY
    _print_case("init_error_sino", cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle32)
    try:
        vol = (0.5 * info["phantom"]).astype(np.float32)
        oracle32.set_object(O, vol)
        eng.set_volume(vol)
        H = oracle32.detector_response(S, O, T)
        E = np.zeros((S.N_p, S.N_r, S.N_t), np.float32)
        et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(H)
        oracle32.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
        sel_o = oracle32.sino_view(S.ProjSelect, S, bool).copy()
        Eg = xt_hip.init_error_sino(eng, H)
        sel_g = eng.proj_select().astype(bool)
        scale = np.abs(eng.sino("Projection")).max()
        assert np.abs(Eg - E).max() <= 1e-5 * scale
        assert np.mean(sel_g != sel_o) <= 1e-4
    finally:
        eng.close()
        oracle32.free(S, O, T)


def _error_sino_f64(oracle64, S, O, T, vol):
    oracle64.set_object(O, vol)
    H = oracle64.detector_response(S, O, T)
    E = np.zeros((S.N_p, S.N_r, S.N_t), np.float64)
    et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(H)
    oracle64.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
    return E, et


def test_cost(scan_root, oracle64a):
    cfg = SMALL4D
    _print_case("compute_cost", cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle64a)
    try:
        vol = (0.8 * info["phantom"]).astype(np.float32)
        E, et = _error_sino_f64(oracle64a, S, O, T, vol.astype(np.float64))
        c64 = oracle64a.lib.oracle_compute_cost(C.byref(S), C.byref(O), C.byref(T), et.ptr)
        eng.set_volume(vol)
        cg = xt_hip.compute_cost(eng, E.astype(np.float32))
        print(f"  cost gpu {cg:.10g} oracle_f64 {c64:.10g}")
        assert abs(cg - c64) <= 1e-5 * abs(c64)
    finally:
        eng.close()
        oracle64a.free(S, O, T)


def test_offset_update(scan_root, oracle64a):
    cfg = TINY
    _print_case("offset_update", cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle64a)
    try:
        vol = (0.9 * info["phantom"]).astype(np.float32)
        E, et = _error_sino_f64(oracle64a, S, O, T, vol.astype(np.float64))
        T.EnforceZeroMeanOffset = 1
        eng.T.EnforceZeroMeanOffset = 1
        Eg = E.astype(np.float32)
        oracle64a.lib.oracle_offset_update(C.byref(S), C.byref(T), et.ptr)
        xt_hip.offset_update(eng, Eg)
        d_o = oracle64a.offset_view(S)
        d_g = eng.offset()
        assert np.abs(d_g - d_o).max() <= 1e-5 * max(1e-3, np.abs(d_o).max())
        assert np.abs(Eg - E).max() <= 1e-5 * np.abs(E).max()
    finally:
        eng.close()
        oracle64a.free(S, O, T)


def test_variance_update(scan_root, oracle64a):
    cfg = TINY
    _print_case("variance_update", cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle64a)
    try:
        vol = info["phantom"].astype(np.float32)
        E, et = _error_sino_f64(oracle64a, S, O, T, vol.astype(np.float64))
        Eg = E.astype(np.float32)
        oracle64a.lib.oracle_variance_update(C.byref(S), C.byref(T), et.ptr)
        xt_hip.variance_update(eng, Eg)
        print(f"  var gpu {eng.T.var_est:.8g} oracle {T.var_est:.8g}")
        assert abs(eng.T.var_est - T.var_est) <= 1e-5 * T.var_est
        wo = oracle64a.sino_view(T.Weight, S)
        assert np.abs(eng.sino("Weight") - wo).max() <= 1e-5 * wo.max()
    finally:
        eng.close()
        oracle64a.free(S, O, T)


# 96x96 maps: 9216 lines, three chunks of the line-list compaction (the last partial)
VSC96 = dataclasses.replace(TINY, name="vsc96", N_xy=96, N_p=48, seed=17)
# the production slab's maps: 512x512 lines, 64 histogram workgroups and 64
# compaction chunks per map
VSC512 = dataclasses.replace(TINY, name="vsc512", N_xy=512, N_p=16, seed=19)


@pytest.mark.parametrize("cfg,zero_rows", [(SMALL4D, 3), (VSC96, 3), (VSC96, 90), (VSC512, 3),
                                           (VSC512, 480)],
                         ids=["small4d", "vsc96", "vsc96-threshold-at-zero", "vsc512",
                              "vsc512-threshold-at-zero"])
def test_vsc_select_bitexact(cfg, zero_rows, scan_root, oracle32):
    """Line sets of the GPU selection against the oracle's (pinned to the
    reference's compiled XT_NHICD.c in test_oracle_pins.py) and, where the
    reference build is present, against XT_NHICD.c itself.  With 90 of 96 rows
    zero the map's top tenth reaches into the zero-valued lines, so the
    threshold falls on a tie at zero."""
    _print_case("vsc_select", cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle32)
    try:
        nb = T.num_z_blocks
        rng = np.random.default_rng(7)
        mag = rng.exponential(1e-4, (O.N_time, nb, O.N_y, O.N_x)).astype(np.float32)
        mag[..., :zero_rows, :] = 0  # ties at zero
        usn = np.full((O.N_time, nb), O.N_y * O.N_x, np.int32)
        got = xt_hip.vsc_select(eng, mag, usn)
        usn_o = np.ctypeslib.as_array((C.c_int32 * (O.N_time * nb)).from_address(
            C.cast(T.UpdateSelectNum, C.POINTER(C.c_void_p))[0]))
        usn_o[:] = usn.ravel()
        mt = xt_abi.PtrTree(mag.copy())
        oracle32.lib.oracle_srandom2(761521)
        oracle32.lib.oracle_vsc_select(C.byref(O), C.byref(T), mt.ptr)
        nxy = O.N_y * O.N_x
        num = py_oracle.deref_data(T.NHICDSelectNum, 1, C.c_int32, O.N_time * nb).reshape(O.N_time, nb)
        xs = py_oracle.deref_data(T.x_NHICD_select, 2, C.c_int32, O.N_time * nb * nxy).reshape(O.N_time, nb, nxy)
        ys = py_oracle.deref_data(T.y_NHICD_select, 2, C.c_int32, O.N_time * nb * nxy).reshape(O.N_time, nb, nxy)
        sets = [(num.copy(), xs.copy(), ys.copy())]
        if py_oracle.RefLib.available():
            ref = py_oracle.RefLib(False)
            ref.lib.srandom2(761521)
            mt_ref = xt_abi.PtrTree(mag.copy())  # kept alive across the call
            ref.lib.VSC_based_Voxel_Line_Select(C.byref(O), C.byref(T), mt_ref.ptr)
            sets.append((num.copy(), xs.copy(), ys.copy()))
        for n_o, x_o, y_o in sets:
            for i in range(O.N_time):
                for b in range(nb):
                    k = n_o[i, b]
                    gy, gx = got[i][b]
                    assert len(gy) == k
                    assert np.array_equal(gy, y_o[i, b, :k]) and np.array_equal(gx, x_o[i, b, :k])
        if zero_rows > O.N_y // 2:
            # the threshold is 0 (a tie): only lines strictly above it are taken
            # (F/XT_NHICD.c:189), fewer than the requested share
            assert 0 < num.max() < (O.N_y * O.N_x) // O.NHICD_Iterations
    finally:
        eng.close()
        oracle32.free(S, O, T)


@pytest.mark.parametrize("cfg,atable,nt", [(TINY, 1, 0), (TINY, 0, 0), (WIDE3, 0, 0), (WIDE6, 0, 0),
                                            (TINY, 1, 64), (WIDE3, 0, 64), (NZ32, 1, 0),
                                            (NZ64, 1, 0), (MZ2, 1, 0), (MZ2, 0, 64)],
                         ids=["atable", "onthefly", "wide3", "wide6", "atable-wave", "wide3-wave",
                              "nz32", "nz64", "mz2", "mz2-onthefly-wave"])
@pytest.mark.parametrize("Iter", [1, 2])
def test_update_lines_per_voxel(Iter, cfg, atable, nt, scan_root, oracle32):
    _print_case(f"update_lines Iter={Iter} atable={atable} threads={nt or 256}", cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle32)
    xt_hip.set_option(xt_hip.OPT_ATABLE, atable)
    xt_hip.set_option(xt_hip.OPT_LINE_THREADS, nt)
    try:
        vol = (0.7 * info["phantom"]).astype(np.float32)
        vol[:, :, :4, :] = 0  # zero rows: zero-skipping paths
        oracle32.set_object(O, vol)
        eng.set_volume(vol)
        H = oracle32.detector_response(S, O, T)
        E = np.zeros((S.N_p, S.N_r, S.N_t), np.float32)
        et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(H)
        oracle32.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
        Eg = E.copy()
        rng = np.random.default_rng(Iter)
        n = 40
        ys = rng.integers(0, O.N_y, n).astype(np.int32)
        xs = rng.integers(0, O.N_x, n).astype(np.int32)
        ys[:4] = 1  # lines in the zero rows
        mag_o = np.zeros((O.N_y, O.N_x), np.float32)
        mt = xt_abi.PtrTree(mag_o)
        zc = C.c_int64(0)
        tot_o = oracle32.lib.oracle_update_lines(C.byref(S), C.byref(O), C.byref(T), et.ptr, ht.ptr,
                                                 0, 0, O.N_z - 1, n, xs.ctypes.data, ys.ctypes.data,
                                                 Iter, C.byref(zc), mt.ptr, 0)
        mag_g, tot_g = xt_hip.update_lines(eng, Eg, 0, xs, ys, Iter)
        xo = oracle32.get_object(O)
        xg = eng.volume()
        scale = np.abs(xo).max()
        print(f"  max |dx| {np.abs(xg - xo).max() / scale:.3e}  e {np.abs(Eg - E).max():.3e}")
        assert np.abs(xg - xo).max() <= 1e-4 * scale
        assert np.abs(Eg - E).max() <= 1e-4 * np.abs(E).max()
        assert np.abs(mag_g - mag_o).max() <= 1e-4 * max(1e-12, np.abs(mag_o).max())
        assert abs(tot_g - tot_o) <= 1e-4 * max(1e-12, abs(tot_o))
    finally:
        xt_hip.set_option(xt_hip.OPT_ATABLE, 1)
        xt_hip.set_option(xt_hip.OPT_LINE_THREADS, 0)
        eng.close()
        oracle32.free(S, O, T)


@pytest.mark.parametrize("cfg,atable,nt", [(NZ32, 1, 256), (NZ32, 0, 256), (NZ32, 1, 64),
                                            (NZ64, 1, 256), (SMALL, 1, 256)],
                         ids=["nz32-atable", "nz32-onthefly", "nz32-wave", "nz64-atable",
                              "small-atable"])
@pytest.mark.parametrize("Iter", [1, 2])
@pytest.mark.parametrize("parity", [0, 1])
def test_update_lines_parity_per_voxel(parity, Iter, cfg, atable, nt, scan_root, oracle32):
    """The pass the ICD iterations run -- one z-block parity, the sinogram split by
    parity, the parity's slices packed into half the lanes (k_icd_lines<32, TAB,
    PACK=true, 6, 256> for the production slab's 32 one-slice blocks) -- against
    the oracle's sequential reference loop over the same lines, block by block of
    that parity (F/XT_ICD_update.c:799-881, F/XT_VoxUpdate.c:85-254)."""
    _print_case(f"update_lines_parity parity={parity} Iter={Iter} atable={atable} threads={nt}",
                cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle32)
    xt_hip.set_option(xt_hip.OPT_ATABLE, atable)
    xt_hip.set_option(xt_hip.OPT_LINE_THREADS, nt)
    try:
        nb = T.num_z_blocks
        bs = O.N_z // nb
        assert nb % 2 == 0 and O.N_z % nb == 0
        vol = (0.7 * info["phantom"]).astype(np.float32)
        vol[:, :, :4, :] = 0  # zero rows: zero-skipping paths
        oracle32.set_object(O, vol)
        eng.set_volume(vol)
        H = oracle32.detector_response(S, O, T)
        E = np.zeros((S.N_p, S.N_r, S.N_t), np.float32)
        et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(H)
        oracle32.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
        Eg = E.copy()
        rng = np.random.default_rng(Iter + 10 * parity)
        n = 40
        ys = rng.integers(0, O.N_y, n).astype(np.int32)
        xs = rng.integers(0, O.N_x, n).astype(np.int32)
        ys[:4] = 1  # lines in the zero rows
        mag_o = np.zeros((O.N_y, O.N_x), np.float32)
        tot_o = 0.0
        for b in range(parity, nb, 2):
            mb = np.zeros((O.N_y, O.N_x), np.float32)
            mt = xt_abi.PtrTree(mb)
            zc = C.c_int64(0)
            tot_o += oracle32.lib.oracle_update_lines(
                C.byref(S), C.byref(O), C.byref(T), et.ptr, ht.ptr, 0, b * bs, (b + 1) * bs - 1,
                n, xs.ctypes.data, ys.ctypes.data, Iter, C.byref(zc), mt.ptr, 0)
            mag_o += mb
        mag_g, tot_g = xt_hip.update_lines_parity(eng, Eg, xs, ys, Iter, parity)
        xo = oracle32.get_object(O)
        xg = eng.volume()
        scale = np.abs(xo).max()
        moved = np.abs(xo - vol).max() / scale
        print(f"  max |dx| {np.abs(xg - xo).max() / scale:.3e} (moved {moved:.3e})  "
              f"e {np.abs(Eg - E).max():.3e}")
        assert moved > 1e-3
        # slices of the other parity are not touched
        other = np.array([((z // bs) & 1) != parity for z in range(O.N_z)])
        assert np.array_equal(xg[:, other], vol[:, other])
        assert np.abs(xg - xo).max() <= 1e-4 * scale
        assert np.abs(Eg - E).max() <= 1e-4 * np.abs(E).max()
        assert np.abs(mag_g - mag_o).max() <= 1e-4 * max(1e-12, np.abs(mag_o).max())
        assert abs(tot_g - tot_o) <= 1e-4 * max(1e-12, abs(tot_o))
    finally:
        xt_hip.set_option(xt_hip.OPT_ATABLE, 1)
        xt_hip.set_option(xt_hip.OPT_LINE_THREADS, 0)
        eng.close()
        oracle32.free(S, O, T)


# the production slab's image size, 8 slices in 8 one-slice z-blocks, and its full
# z shape (32 slices in 32 one-slice blocks: the benchmark's kernel instantiation)
PROD512 = dataclasses.replace(TINY, name="prod512", N_xy=512, N_z=8, N_p=256, num_threads=4,
                              seed=27)
PROD512Z32 = dataclasses.replace(PROD512, name="prod512z32", N_z=32, num_threads=32)
_HAMMING5 = np.array([[0.0064, 0.0432, 0.0800, 0.0432, 0.0064],
                      [0.0432, 0.2916, 0.5400, 0.2916, 0.0432],
                      [0.0800, 0.5400, 1.0000, 0.5400, 0.0800],
                      [0.0432, 0.2916, 0.5400, 0.2916, 0.0432],
                      [0.0064, 0.0432, 0.0800, 0.0432, 0.0064]], np.float32)


@pytest.mark.parametrize("cfg", [PROD512, PROD512Z32], ids=["z8", "z32"])
@pytest.mark.parametrize("parity", [0, 1])
def test_update_lines_parity_raster_runs_at_512(parity, cfg, scan_root, oracle32):
    """An NHICD-shaped pass at the production image size: the lines of the top
    4096 values of the parity's Hamming-filtered |x| map in raster order -- runs of
    neighbouring lines whose columns share most of their bins, each line reading
    the previous one's changes -- one after another, the GPU's packed parity pass
    against the oracle's loop over the parity's blocks (F/XT_ICD_update.c:844-881,
    F/XT_VoxUpdate.c:85-254).  Sequential dependence amplifies the float
    summation-order differences along the runs, so the bar is on the RMS
    difference relative to the pass's own change, with a looser maximum."""
    from scipy.ndimage import correlate
    _print_case(f"update_lines_parity raster runs parity={parity}", cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle32)
    try:
        nb = T.num_z_blocks
        assert nb == cfg.N_z
        # near the data (residuals inside the generalized-Huber threshold): data-driven steps
        vol = (0.97 * info["phantom"]).astype(np.float32)
        oracle32.set_object(O, vol)
        eng.set_volume(vol)
        H = oracle32.detector_response(S, O, T)
        E = np.zeros((S.N_p, S.N_r, S.N_t), np.float32)
        et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(H)
        oracle32.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
        Eg = E.copy()
        filt = correlate(np.abs(vol[0, parity::2]).sum(0), _HAMMING5, mode="constant")
        th = np.sort(filt.ravel())[-4097]
        ys, xs = (a.astype(np.int32) for a in np.nonzero(filt > th))
        n = len(ys)
        assert 2048 < n <= 4096
        # most consecutive lines are horizontal neighbours
        assert np.mean((ys[1:] == ys[:-1]) & (xs[1:] == xs[:-1] + 1)) > 0.8
        for b in range(parity, nb, 2):
            mb = np.zeros((O.N_y, O.N_x), np.float32)
            mt = xt_abi.PtrTree(mb)
            zc = C.c_int64(0)
            oracle32.lib.oracle_update_lines(C.byref(S), C.byref(O), C.byref(T), et.ptr, ht.ptr,
                                             0, b, b, n, xs.ctypes.data, ys.ctypes.data, 2,
                                             C.byref(zc), mt.ptr, 0)
        for c0 in range(0, n, 256):
            xt_hip.update_lines_parity(eng, Eg, xs[c0:c0 + 256], ys[c0:c0 + 256], 2, parity)
        xo = oracle32.get_object(O)[0]
        xg = eng.volume()[0]
        # the updated voxels: the listed lines in the parity's slices
        sel = (slice(parity, None, 2), ys, xs)
        d_o = xo[sel] - vol[0][sel]
        d_g = xg[sel] - xo[sel]
        change, rms = np.sqrt(np.mean(d_o ** 2)), np.sqrt(np.mean(d_g ** 2))
        print(f"  {n} lines: change rms {change:.3e} (max {np.abs(d_o).max():.3e}); GPU - oracle "
              f"rms {rms / change:.3e} of it, max {np.abs(d_g).max() / np.abs(d_o).max():.3e} of the "
              f"largest change; e {np.abs(Eg - E).max() / np.abs(E).max():.3e}")
        assert np.array_equal(xg[1 - parity::2], vol[0][1 - parity::2])
        assert np.abs(d_o).max() > 5e-3 * np.abs(vol).max()
        assert rms <= 2e-4 * change
        assert np.abs(d_g).max() <= 2e-3 * np.abs(d_o).max()
        assert np.abs(Eg - E).max() <= 1e-4 * np.abs(E).max()
    finally:
        eng.close()
        oracle32.free(S, O, T)


def _group_lines(O, rng):
    """A line list for the grouped kernel: 2 x 2 and 1 x 2 tiles of adjacent lines
    (columns that share bins), a line listed twice in a row (identical columns),
    lines in the zero rows, and random lines (columns mostly apart)."""
    pts = []
    for (y, x) in ((5, 6), (9, 3), (O.N_y - 2, O.N_x - 2)):
        pts += [(y, x), (y, x + 1), (y + 1, x), (y + 1, x + 1)]
    pts += [(12, 12), (12, 13), (7, 7), (7, 7), (1, 2), (1, 3)]
    pts += [(int(a), int(b)) for a, b in zip(rng.integers(0, O.N_y, 17), rng.integers(0, O.N_x, 17))]
    ys = np.array([p[0] for p in pts], np.int32)
    xs = np.array([p[1] for p in pts], np.int32)
    return xs, ys


GROUPS_OFF = ("grouped line updates (XT_OPT_LINE_GROUP 2, 4) are switched off: their first "
              "GPU run took a memory fault whose cause is not found")


@pytest.mark.skip(reason=GROUPS_OFF)
@pytest.mark.parametrize("group", [2, 4])
@pytest.mark.parametrize("cfg,parity,nt", [(TINY, -1, 256), (NZ32, 0, 256), (NZ32, 1, 256),
                                           (NZ32, 1, 64), (SMALL, 0, 256)],
                         ids=["tiny-all-blocks", "nz32-even", "nz32-odd", "nz32-odd-wave",
                              "small-even"])
@pytest.mark.parametrize("Iter", [1, 2])
def test_update_lines_groups_per_voxel(Iter, cfg, parity, nt, group, scan_root, oracle32):
    """Lines updated in groups of 2 or 4 consecutive list entries by one workgroup
    (k_icd_groups: e written once per group, each line reading e less the group's
    earlier changes) against the oracle's sequential loop over the same list
    (F/XT_VoxUpdate.c:85-254): every block (ascending slices, parity -1) or one
    z-block parity of the split sinogram (F/XT_ICD_update.c:799-881)."""
    _print_case(f"update_lines groups={group} parity={parity} Iter={Iter} threads={nt}", cfg)
    path, info, eng, (S, O, T) = _pair(cfg, scan_root, oracle32)
    xt_hip.set_option(xt_hip.OPT_LINE_GROUP, group)
    xt_hip.set_option(xt_hip.OPT_LINE_THREADS, nt)
    try:
        vol = (0.7 * info["phantom"]).astype(np.float32)
        vol[:, :, :4, :] = 0  # zero rows: zero-skipping paths
        oracle32.set_object(O, vol)
        eng.set_volume(vol)
        H = oracle32.detector_response(S, O, T)
        E = np.zeros((S.N_p, S.N_r, S.N_t), np.float32)
        et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(H)
        oracle32.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
        Eg = E.copy()
        xs, ys = _group_lines(O, np.random.default_rng(Iter + 7 * group))
        n = len(xs)
        mag_o = np.zeros((O.N_y, O.N_x), np.float32)
        tot_o = 0.0
        nb = T.num_z_blocks
        bs = O.N_z // nb
        ranges = [(0, O.N_z - 1)] if parity < 0 else \
            [(b * bs, (b + 1) * bs - 1) for b in range(parity, nb, 2)]
        for z0, z1 in ranges:
            mb = np.zeros((O.N_y, O.N_x), np.float32)
            mt = xt_abi.PtrTree(mb)
            zc = C.c_int64(0)
            tot_o += oracle32.lib.oracle_update_lines(
                C.byref(S), C.byref(O), C.byref(T), et.ptr, ht.ptr, 0, z0, z1, n, xs.ctypes.data,
                ys.ctypes.data, Iter, C.byref(zc), mt.ptr, 0)
            mag_o += mb
        if parity < 0:
            mag_g, tot_g = xt_hip.update_lines(eng, Eg, 0, xs, ys, Iter)
        else:
            mag_g, tot_g = xt_hip.update_lines_parity(eng, Eg, xs, ys, Iter, parity)
        xo = oracle32.get_object(O)
        xg = eng.volume()
        scale = np.abs(xo).max()
        moved = np.abs(xo - vol).max() / scale
        print(f"  max |dx| {np.abs(xg - xo).max() / scale:.3e} (moved {moved:.3e})  "
              f"e {np.abs(Eg - E).max():.3e}")
        assert moved > 1e-3
        assert np.abs(xg - xo).max() <= 1e-4 * scale
        assert np.abs(Eg - E).max() <= 1e-4 * np.abs(E).max()
        assert np.abs(mag_g - mag_o).max() <= 1e-4 * max(1e-12, np.abs(mag_o).max())
        assert abs(tot_g - tot_o) <= 1e-4 * max(1e-12, abs(tot_o))
    finally:
        xt_hip.set_option(xt_hip.OPT_LINE_GROUP, 1)
        xt_hip.set_option(xt_hip.OPT_LINE_THREADS, 0)
        eng.close()
        oracle32.free(S, O, T)
