"""CPU tests: the oracle pinned against the reference's golden vectors, against
the reference's own buildable sources (oracle/_ref) and against closed forms.

Each test names what pins the function; a plausible mistake (dropped term,
wrong sign, swapped index) fails at least one of them.
"""
import ctypes as C
import dataclasses
import json
import math
import os

import numpy as np
import pytest

from conftest import ROOT, SMALL4D, TINY, make_scan, oracle_path
from timbir_python_amd import synth, xt_abi
import py_oracle

GOLD = os.path.join(ROOT, "tests", "golden")
needs_ref = pytest.mark.skipif(not py_oracle.RefLib.available(),
                               reason="oracle/_ref not built (reference sources absent)")


# -- random numbers ----------------------------------------------------------
def test_rng_known_answer_table(oracle32):
    """Park-Miller states from seed 1: the table in randlib.c:69-91."""
    gold = json.load(open(os.path.join(GOLD, "randlib_seed1.json")))
    out = np.zeros(20)
    oracle32.lib.oracle_random2_seq(1, 20, out.ctypes.data)
    states = np.rint(out * 2147483647).astype(np.int64)
    assert states.tolist() == gold["states"]


@needs_ref
def test_rng_bitexact_vs_reference(oracle32):
    ref = py_oracle.RefLib(False)
    n = 20000
    out = np.zeros(n)
    oracle32.lib.oracle_random2_seq(761521, n, out.ctypes.data)
    ref.lib.srandom2(761521)
    r = np.array([ref.lib.random2() for _ in range(n)])
    assert np.array_equal(out, r)


# -- view order and view files ----------------------------------------------
def test_interlaced_views_golden_log(oracle32):
    """Bit-reversed interlaced order + clipping: the 399 deleted views and their
    times of the launcher log info.txt:1-175 (K=32, N_theta=1536, 25 cycles,
    any clip threshold in [2.8125, 3.75) degrees)."""
    gold = json.load(open(os.path.join(GOLD, "interlaced_clip.json")))
    K, Nt = gold["K"], gold["N_theta"]
    Np = Nt * 25
    views, times = np.zeros(Np), np.zeros(Np)
    dv, dt = np.zeros(Np), np.zeros(Np)
    nd = C.c_int32()
    kept = oracle32.lib.oracle_interlaced_views(K, Nt, Np, 0.03, 100.0, views.ctypes.data,
                                                times.ctypes.data, dv.ctypes.data, dt.ctypes.data,
                                                C.byref(nd))
    assert nd.value == 399 and kept == Np - 399
    assert np.array_equal(dv[:399], np.array(gold["deleted_views"]))
    assert np.array_equal(dt[:399], np.array(gold["deleted_times"]))
    # the hand check of one entry: 360.9375 deg = index 3080 = 96*32 + bitrev5(2)
    assert dv[0] == 3080 * 180.0 / Nt and dt[0] == 3080 / 32


def test_synth_launcher_views_golden_log():
    """The product generator's launcher view list (interlaced order, then
    clip_list_of_views) reproduces the 399 deleted views and times of info.txt:1-175
    (K=32, N_theta=1536, rotation 100 deg/s, 0.03 s between views: 3 degrees), and
    keeps everything else in order."""
    gold = json.load(open(os.path.join(GOLD, "interlaced_clip.json")))
    K, Nt = gold["K"], gold["N_theta"]
    kv, kt, dv, dt = synth.launcher_views(K, Nt, 25, 0.03, 100.0)
    assert len(dv) == 399 and len(kv) == Nt * 25 - 399
    assert np.array_equal(dv, np.array(gold["deleted_views"]))
    assert np.array_equal(dt, np.array(gold["deleted_times"]))
    ang, t = synth.interlaced_views(K, Nt, Nt * 25)
    assert np.array_equal(np.sort(np.concatenate([kv, dv])), np.sort(ang))
    assert kt[0] == 0.0 and np.array_equal(kt[1:], t[np.isin(ang, kv)][1:])
    # no clipping below the smallest step: every view kept
    kv0, _, dv0, _ = synth.launcher_views(K, Nt, 2, 0.0, 100.0)
    assert len(dv0) == 0 and len(kv0) == 2 * Nt


def test_synth_interlaced_views_match_oracle(oracle32):
    cfg = dataclasses.replace(synth.C3, N_p=1000)
    ang, t = cfg.views()
    Np = cfg.N_p
    views, times = np.zeros(Np), np.zeros(Np)
    dv, dt = np.zeros(Np), np.zeros(Np)
    nd = C.c_int32()
    oracle32.lib.oracle_interlaced_views(cfg.interlaced_K, cfg.interlaced_N_theta, Np, 0.0, 0.0,
                                         views.ctypes.data, times.ctypes.data, dv.ctypes.data,
                                         dt.ctypes.data, C.byref(nd))
    assert nd.value == 0
    assert np.array_equal(ang, views) and np.array_equal(t[1:], times[1:])


def test_parse_reference_view_info(oracle32):
    """The view_info.txt data file shipped with the engine: 2000 '<time> - <deg>,' lines."""
    path = os.path.join(GOLD, "view_info_reference.txt")
    views = np.zeros(2000, np.float32)
    times = np.zeros(2000, np.float32)
    n = oracle32.lib.oracle_parse_view_info(path.encode(), 2000, views.ctypes.data,
                                            times.ctypes.data)
    assert n == 2000
    lines = open(path).read().split("\n")
    t0, a0 = lines[1].split(" - ")
    assert times[1] == np.float32(float(t0))
    # (Real_t)atof(s) * M_PI / 180.0: float operand promoted to double
    assert views[1] == np.float32(float(np.float32(float(a0.rstrip(",")))) * math.pi / 180.0)
    assert np.all(np.diff(views) > 0)


def test_parse_view_info_rejects_bad_format(oracle32, tmp_path):
    p = tmp_path / "view_info.txt"
    p.write_text("0.0 - 1.0,\n1.0 ; 2.0,\n")
    v = np.zeros(4, np.float32)
    assert oracle32.lib.oracle_parse_view_info(str(p).encode(), 4, v.ctypes.data,
                                               v.ctypes.data) == -1


def test_bin_views_closed_form(oracle32):
    """Frames of width delta starting at Rtime0; times past the last frame stay in it."""
    t = np.sort(np.random.default_rng(0).uniform(0, 10, 200)).astype(np.float32)
    idx = np.zeros(4 * 200, np.int32)
    num = np.zeros(4, np.int32)
    oracle32.lib.oracle_bin_views(t.ctypes.data, 200, 4, C.c_float(0.0), C.c_float(2.0),
                                  idx.ctypes.data, num.ctypes.data)
    expect = np.minimum(np.floor(t / 2.0).astype(int), 3)
    for f in range(4):
        assert np.array_equal(idx[f * 200: f * 200 + num[f]], np.nonzero(expect == f)[0])


# -- detector table and system matrix --------------------------------------
def _setup(oracle, cfg, scan_root):
    path, info = make_scan(cfg, scan_root)
    path = oracle_path(path, oracle)
    S, O, T = oracle.setup(path, cfg.cli_args())
    oracle.read_sinogram(path, S, T)
    return path, info, (S, O, T)


@needs_ref
@pytest.mark.parametrize("double", [False, True])
def test_detector_response_bitexact_vs_reference(double, scan_root):
    o = py_oracle.Oracle(double)
    ref = py_oracle.RefLib(double)
    path, info, (S, O, T) = _setup(o, SMALL4D, scan_root)
    try:
        H = o.detector_response(S, O, T)
        Hr = np.zeros_like(H)
        ht = xt_abi.PtrTree(Hr)
        Ht = np.zeros(65, H.dtype)
        ref.lib.DetectorResponseProfile(ht.ptr, C.c_void_p(Ht.ctypes.data), C.byref(S), C.byref(O),
                                        C.byref(T))
        assert np.array_equal(H, Hr)
    finally:
        o.free(S, O, T)


def test_detector_response_closed_forms(oracle64, scan_root):
    """H_r[.][64] = 0; theta and theta+90 deg give the same row; the table is
    the trapezoid (unit-area beam) so H_r[0] equals the trapezoid's plateau
    value delta_xy/cos(theta) near theta = 0 up to the beam smoothing."""
    cfg = dataclasses.replace(TINY, N_p=64)
    path, info, (S, O, T) = _setup(oracle64, cfg, scan_root)
    try:
        H = oracle64.detector_response(S, O, T)
        assert np.all(H[:, 64] == 0)
        # views are i*180/64: view i+32 is i + 90 degrees
        assert np.allclose(H[:32], H[32:], rtol=1e-12, atol=1e-15)
        assert abs(H[0, 0] - O.delta_xy) < 1e-9  # beam inside the plateau at theta=0
        assert np.all(np.diff(H[:, :64], axis=1) <= 1e-12)  # non-increasing with distance
    finally:
        oracle64.free(S, O, T)


@needs_ref
def test_amatrix_bitexact_vs_reference(oracle32, scan_root):
    ref = py_oracle.RefLib(False)
    path, info, (S, O, T) = _setup(oracle32, TINY, scan_root)
    try:
        H = oracle32.detector_response(S, O, T)
        ht = xt_abi.PtrTree(H)
        cols = []
        for lib in (oracle32.lib.oracle_amatrix_column, ref.lib.calcAMatrixColumnforAngle):
            vals = np.zeros(8, np.float32)
            inds = np.zeros(8, np.int32)
            a = oracle32.AMatrixCol()
            a.values, a.index = vals.ctypes.data, inds.ctypes.data
            out = []
            for v in range(0, S.N_p, 5):
                for yy in range(O.N_y):
                    for xx in range(0, O.N_x, 3):
                        lib(C.byref(S), C.byref(O), ht.ptr, C.byref(a), yy, xx, v)
                        out.append((a.count, inds[:a.count].tolist(),
                                    vals[:a.count].view(np.uint32).tolist()))
            cols.append(out)
        assert cols[0] == cols[1]
    finally:
        oracle32.free(S, O, T)


def test_amatrix_area_invariant(oracle64, scan_root):
    """The footprint integrates to the voxel's area: the table integral over
    +-64*OffsetR equals delta_xy^2 for every view, and the column sum, averaged
    over voxels (sub-bin phases), equals delta_xy^2 / delta_r.  A single
    column's sum aliases with its phase (the bins sample a ~2-bin footprint)."""
    path, info, (S, O, T) = _setup(oracle64, TINY, scan_root)
    try:
        H = oracle64.detector_response(S, O, T)
        d = np.arange(65) * S.OffsetR
        area = O.delta_xy ** 2
        for v in range(S.N_p):
            assert 2 * np.trapezoid(H[v], d) == pytest.approx(area, rel=3e-3)
        ht = xt_abi.PtrTree(H)
        vals = np.zeros(8, np.float64)
        inds = np.zeros(8, np.int32)
        a = oracle64.AMatrixCol()
        a.values, a.index = vals.ctypes.data, inds.ctypes.data
        for v in range(0, S.N_p, 7):
            sums = []
            for yy in range(4, O.N_y - 4):
                for xx in range(4, O.N_x - 4):
                    oracle64.lib.oracle_amatrix_column(C.byref(S), C.byref(O), ht.ptr, C.byref(a),
                                                       yy, xx, v)
                    sums.append(vals[:a.count].sum())
            assert np.mean(sums) == pytest.approx(area / S.delta_r, rel=0.01)
    finally:
        oracle64.free(S, O, T)


@needs_ref
def test_forward_project_vs_reference_scatter(oracle32, scan_root):
    """Ax of the oracle = the reference's forward_project_voxel summed over every
    voxel in the same order (F/XT_ICD_update.c:552-569), bit for bit."""
    ref = py_oracle.RefLib(False)
    path, info, (S, O, T) = _setup(oracle32, TINY, scan_root)
    try:
        vol = info["phantom"].astype(np.float32)
        oracle32.set_object(O, vol)
        H = oracle32.detector_response(S, O, T)
        ax = oracle32.forward_project(S, O, H)
        ht = xt_abi.PtrTree(H)
        E = np.zeros_like(ax)
        et = xt_abi.PtrTree(E)
        vals = np.zeros(8, np.float32)
        inds = np.zeros(8, np.int32)
        a = oracle32.AMatrixCol()
        a.values, a.index = vals.ctypes.data, inds.ctypes.data
        lists, num = oracle32.proj_lists(O)
        for yy in range(O.N_y):
            for xx in range(O.N_x):
                for p in lists[0]:
                    ref.lib.calcAMatrixColumnforAngle(C.byref(S), C.byref(O), ht.ptr, C.byref(a),
                                                      yy, xx, int(p))
                    for z in range(O.N_z):
                        ref.lib.forward_project_voxel(C.byref(S), C.c_float(vol[0, z, yy, xx]),
                                                      et.ptr, C.byref(a), int(p), z)
        assert np.array_equal(ax, E)
    finally:
        oracle32.free(S, O, T)


# -- NHICD selection ---------------------------------------------------------
VSC96 = dataclasses.replace(TINY, name="vsc96", N_xy=96, N_p=48, seed=17)


@needs_ref
@pytest.mark.parametrize("cfg,zero_rows,full", [(SMALL4D, 2, False), (VSC96, 3, True),
                                                (VSC96, 90, True)],
                         ids=["small4d", "vsc96", "vsc96-threshold-at-zero"])
def test_vsc_select_bitexact_vs_reference(cfg, zero_rows, full, oracle32, scan_root):
    """Same line sets and the same random numbers consumed (F/XT_NHICD.c:164-216),
    on 32x32 maps of four frames and on 96x96 maps, one with 90 of 96 rows zero so
    that the order statistic is a tie at zero."""
    ref = py_oracle.RefLib(False)
    path, info, (S, O, T) = _setup(oracle32, cfg, scan_root)
    try:
        nb, nxy = T.num_z_blocks, O.N_y * O.N_x
        rng = np.random.default_rng(3)
        mag = rng.exponential(1.0, (O.N_time, nb, O.N_y, O.N_x)).astype(np.float32)
        mag[:, :, :zero_rows] = 0
        usn = py_oracle.deref_data(T.UpdateSelectNum, 1, C.c_int32, O.N_time * nb)
        usn[:] = nxy if full else rng.integers(nxy // 2, nxy + 1, usn.size)
        results = []
        for fn, seed in ((oracle32.lib.oracle_vsc_select, oracle32.lib.oracle_srandom2),
                         (ref.lib.VSC_based_Voxel_Line_Select, ref.lib.srandom2)):
            seed(4242)
            mt = xt_abi.PtrTree(mag.copy())
            fn(C.byref(O), C.byref(T), mt.ptr)
            num = py_oracle.deref_data(T.NHICDSelectNum, 1, C.c_int32, O.N_time * nb).copy()
            xs = py_oracle.deref_data(T.x_NHICD_select, 2, C.c_int32, O.N_time * nb * nxy).copy()
            ys = py_oracle.deref_data(T.y_NHICD_select, 2, C.c_int32, O.N_time * nb * nxy).copy()
            nxt = (oracle32.lib.oracle_random2() if fn is oracle32.lib.oracle_vsc_select
                   else ref.lib.random2())
            sel = [(num[k], xs.reshape(-1, nxy)[k, :num[k]].tolist(),
                    ys.reshape(-1, nxy)[k, :num[k]].tolist()) for k in range(O.N_time * nb)]
            results.append((sel, nxt))
        assert results[0] == results[1]
    finally:
        oracle32.free(S, O, T)


# -- voxel update: closed form of the quadratic case -------------------------
QUAD = dataclasses.replace(synth.C1, name="quad", N_xy=6, N_z=2, N_p=12, p=2.0, c_s=1.0,
                           sigma_s=2e-3, alpha=1.0, zinger_T=1e9, zinger_frac=0.0,
                           update_offset=0, zero_mean_offset=False, iterations=400,
                           num_threads=2, n_ellipsoids=3, seed=2)


def dense_map_solution(oracle64, path, cfg):
    """Minimiser of 1/2 (y-Ax)'W(y-Ax) + sum_pairs b (x_j-x_k)^2/(sigma^2 (c+1)),
    the MAP cost when p = q = 2, by a dense solve."""
    S, O, T = oracle64.setup(path, cfg.cli_args())
    try:
        oracle64.read_sinogram(path, S, T)
        H = oracle64.detector_response(S, O, T)
        ht = xt_abi.PtrTree(H)
        nz, ny, nx = O.N_z, O.N_y, O.N_x
        nvox = nz * ny * nx
        nsino = S.N_p * S.N_r * S.N_t
        A = np.zeros((nsino, nvox))
        vals = np.zeros(8)
        inds = np.zeros(8, np.int32)
        a = oracle64.AMatrixCol()
        a.values, a.index = vals.ctypes.data, inds.ctypes.data
        for yy in range(ny):
            for xx in range(nx):
                for p in range(S.N_p):
                    oracle64.lib.oracle_amatrix_column(C.byref(S), C.byref(O), ht.ptr, C.byref(a),
                                                       yy, xx, p)
                    for q in range(a.count):
                        for z in range(nz):
                            A[(p * S.N_r + inds[q]) * S.N_t + z, (z * ny + yy) * nx + xx] = vals[q]
        y = oracle64.sino_view(S.Projection, S).reshape(-1).copy()
        w = oracle64.sino_view(T.Weight, S).reshape(-1).copy()
        filt = np.array(T.Spatial_Filter).reshape(27)
        kappa = 2.0 / (T.Sigma_S_Q * (O.C_S + 1.0))
        L = np.zeros((nvox, nvox))
        for z in range(nz):
            for yy in range(ny):
                for xx in range(nx):
                    i = (z * ny + yy) * nx + xx
                    for dz in (-1, 0, 1):
                        for dy in (-1, 0, 1):
                            for dx in (-1, 0, 1):
                                z2, y2, x2 = z + dz, yy + dy, xx + dx
                                if (dz, dy, dx) == (0, 0, 0) or not (
                                        0 <= z2 < nz and 0 <= y2 < ny and 0 <= x2 < nx):
                                    continue
                                k = (z2 * ny + y2) * nx + x2
                                b = filt[((dz + 1) * 3 + dy + 1) * 3 + dx + 1] * kappa
                                L[i, i] += b
                                L[i, k] -= b
        M = A.T @ (w[:, None] * A) + L
        x = np.linalg.solve(M, A.T @ (w * y))
        return x.reshape(nz, ny, nx)
    finally:
        oracle64.free(S, O, T)


def test_icd_converges_to_dense_map_solution(oracle64, scan_root):
    """With p = q = 2, no zinger rejection and alpha = 1, ICD is Gauss-Seidel on
    the normal equations: the whole driver (theta sums, surrogate, prior weights,
    e updates, NHICD schedule) must land on the dense solution."""
    path, info = make_scan(QUAD, scan_root)
    path = oracle_path(path, oracle64)
    xs = dense_map_solution(oracle64, path, QUAD)
    oracle64.run(path, QUAD.cli_args())
    x = np.fromfile(os.path.join(path, "object_n0_time_0.bin"), np.float64).reshape(xs.shape)
    assert np.abs(x - xs).max() <= 1e-6 * np.abs(xs).max()


def test_surrogate_matches_potential_derivative(oracle64, scan_root):
    """b' = b * rho'(delta)/delta inside the update equals the finite-difference
    derivative of the qGGMRF potential rho (F/XT_ICD_update.c:79-82)."""
    path, info, (S, O, T) = _setup(oracle64, TINY, scan_root)
    try:
        p, c = O.MRF_P, O.C_S
        sq, sqp = T.Sigma_S_Q, T.Sigma_S_Q_P
        rho = lambda d: (abs(d) ** 2 / sq) / (c + abs(d) ** (2 - p) / sqp)
        flt = np.zeros(27, np.uint8)
        flt[0] = 1
        nb = np.zeros(27)
        tn = np.zeros(2)
        tf = np.zeros(2, np.uint8)
        b0 = T.Spatial_Filter[0][0][0]
        for delta in [3e-6, -2e-5, 1e-4, -7e-4, 2e-3]:
            V = 1e-3
            nb[0] = V - delta
            u = oracle64.lib.oracle_functional_substitution(
                V, 0.0, 1.0, C.byref(O), C.byref(T), nb.ctypes.data, flt.ctypes.data,
                tn.ctypes.data, tf.ctypes.data, 0)
            bprime = (u - V) / (nb[0] - u)  # alpha = T.alpha
            bprime = (T.alpha * (nb[0] - V) - (u - V)) and bprime
            h = abs(delta) * 1e-4
            drho = (rho(delta + h) - rho(delta - h)) / (2 * h)
            # u = V + alpha*(u* - V), u* = (b' x_n + V)/(b' + 1)
            ustar = V + (u - V) / T.alpha
            b_est = (ustar - V) / (nb[0] - ustar)
            assert b_est == pytest.approx(b0 * drho / delta, rel=1e-5)
    finally:
        oracle64.free(S, O, T)


# -- offset, variance, cost: closed forms ------------------------------------
def _rand_sino(rng, S):
    shape = (S.N_p, S.N_r, S.N_t)
    return rng.normal(0, 0.05, shape), rng.uniform(1e3, 2e4, shape)


def test_offset_update_closed_forms(oracle64, scan_root):
    """All entries selected: the update makes the weighted view-mean of e zero at
    every detector pixel; with the zero-mean constraint the offsets sum to 0."""
    path, info, (S, O, T) = _setup(oracle64, TINY, scan_root)
    try:
        rng = np.random.default_rng(1)
        E0, W = _rand_sino(rng, S)
        oracle64.sino_view(T.Weight, S)[:] = W
        T.ErrorSinoThresh = 1e9
        for zero_mean in (0, 1):
            T.EnforceZeroMeanOffset = zero_mean
            oracle64.offset_view(S)[:] = 0
            E = E0.copy()
            et = xt_abi.PtrTree(E)
            oracle64.lib.oracle_offset_update(C.byref(S), C.byref(T), et.ptr)
            d = oracle64.offset_view(S)
            if zero_mean:
                assert abs(d.sum()) <= 1e-9 * np.abs(d).sum()
            else:
                assert np.abs((E * W).sum(0) / W.sum(0)).max() <= 1e-12
                assert np.allclose(d, (E0 * W).sum(0) / W.sum(0), rtol=1e-12, atol=1e-15)
    finally:
        oracle64.free(S, O, T)


def test_variance_update_closed_form(oracle64, scan_root):
    path, info, (S, O, T) = _setup(oracle64, TINY, scan_root)
    try:
        rng = np.random.default_rng(2)
        E, W = _rand_sino(rng, S)
        oracle64.sino_view(T.Weight, S)[:] = W
        T.ErrorSinoThresh = 1e9
        T.var_est = 2.0
        et = xt_abi.PtrTree(E)
        oracle64.lib.oracle_variance_update(C.byref(S), C.byref(T), et.ptr)
        expect = np.mean(E * E * W * 2.0)
        assert T.var_est == pytest.approx(expect, rel=1e-12)
        assert np.allclose(oracle64.sino_view(T.Weight, S), W * 2.0 / expect, rtol=1e-12)
    finally:
        oracle64.free(S, O, T)


def test_cost_closed_form(oracle64, scan_root):
    """Zero object: prior 0; data term = sum of rho_huber over entries / 2 + log term."""
    path, info, (S, O, T) = _setup(oracle64, TINY, scan_root)
    try:
        rng = np.random.default_rng(3)
        E, W = _rand_sino(rng, S)
        oracle64.sino_view(T.Weight, S)[:] = W
        Tt, dl = T.ErrorSinoThresh, T.ErrorSinoDelta
        sel = np.abs(E * np.sqrt(W)) < Tt
        oracle64.sino_view(S.ProjSelect, S, bool)[:] = sel
        oracle64.set_object(O, np.zeros((O.N_time, O.N_z, O.N_y, O.N_x)))
        et = xt_abi.PtrTree(E)
        c = oracle64.lib.oracle_compute_cost(C.byref(S), C.byref(O), C.byref(T), et.ptr)
        u = E * np.sqrt(W)
        rho = np.where(sel, u * u, 2 * dl * Tt * np.abs(u) + Tt * Tt * (1 - 2 * dl))
        expect = rho.sum() / 2 + E.size * math.log(T.var_est) / 2
        assert c == pytest.approx(expect, rel=1e-12)
    finally:
        oracle64.free(S, O, T)


# ---- multi-resolution upsampling (initObject, F/XT_ICD_update.c:316-435) ----
def _ramp(shape, a, coef):
    idx = np.indices(shape).astype(np.float64)
    return a + sum(c * i for c, i in zip(coef, idx))


@pytest.mark.parametrize("init_icd", [2, 3])
def test_upsample_closed_forms(init_icd):
    """Constants stay exact; a ramp a + b*z + c*y + d*x with dyadic coefficients
    comes back as the ramp at the fine positions k -/+ 1/4 (interior) and clamped
    at the ends -- every weight and index of the three passes."""
    o = py_oracle.Oracle(False)
    nz, ny, nx = 3, 4, 5
    const = np.full((nz, ny, nx), 0.75, np.float32)
    np.testing.assert_array_equal(o.upsample_object(init_icd, const), 0.75)
    coef = (0.5, 0.25, 2.0)
    ramp = _ramp((nz, ny, nx), 1.0, coef).astype(np.float32)
    got = o.upsample_object(init_icd, ramp).astype(np.float64)

    def pos(n):  # fine index -> coarse coordinate, clamped at the ends
        p = np.arange(2 * n) / 2.0 - 0.25
        return np.clip(p, 0.0, n - 1.0)
    pz = pos(nz) if init_icd == 3 else np.arange(nz, dtype=np.float64)
    want = (1.0 + coef[0] * pz[:, None, None] + coef[1] * pos(ny)[None, :, None] +
            coef[2] * pos(nx)[None, None, :])
    np.testing.assert_array_equal(got, want)


def test_upsample_rounding_paths():
    """The interior odd rows of the y pass sum in float (the reference's integer
    3), every other sample in double: on random floats the two paths differ in
    the last bit somewhere, and the oracle follows each where the text says."""
    o = py_oracle.Oracle(False)
    rng = np.random.default_rng(5)
    c = rng.uniform(0.0, 2e-3, (2, 16, 16)).astype(np.float32)
    got = o.upsample_object(2, c)
    # x pass in double, as the reference
    bx = np.zeros((2, 16, 32), np.float32)
    c64 = c.astype(np.float64)
    bx[..., 0], bx[..., -1] = c[..., 0], c[..., -1]
    bx[..., 1] = ((3.0 * c64[..., 0] + c64[..., 1]) / 4.0).astype(np.float32)
    bx[..., -2] = ((c64[..., -2] + 3.0 * c64[..., -1]) / 4.0).astype(np.float32)
    bx[..., 2:-2:2] = ((c64[..., :-2] + 3.0 * c64[..., 1:-1]) / 4.0)[..., :14].astype(np.float32)
    bx[..., 3:-2:2] = ((3.0 * c64[..., 1:-1] + c64[..., 2:]) / 4.0)[..., :14].astype(np.float32)
    b64 = bx.astype(np.float64)
    odd_real = ((np.float32(3) * bx[:, 1:-1] + bx[:, 2:]) / np.float32(4)).astype(np.float32)
    odd_double = ((3.0 * b64[:, 1:-1] + b64[:, 2:]) / 4.0).astype(np.float32)
    np.testing.assert_array_equal(got[:, 3:-2:2], odd_real[:, :14])
    assert (odd_real != odd_double).any()  # the distinction is real on these inputs
    even = ((b64[:, :-2] + 3.0 * b64[:, 1:-1]) / 4.0).astype(np.float32)
    np.testing.assert_array_equal(got[:, 2:-2:2], even[:, :14])


def test_oracle_initicd2_starts_from_upsampled(scan_root):
    """The oracle's engine with --initICD 2 and no iterations writes the
    previous stage's half-size object upsampled by oracle_upsample_object."""
    cfg = dataclasses.replace(TINY, iterations=2)
    coarse = dataclasses.replace(TINY, name="tiny_coarse", N_xy=TINY.N_xy // 2,
                                 mult_xy=2 * TINY.mult_xy, update_offset=0)
    path, _ = make_scan(cfg, scan_root, "orc-init2")
    o = py_oracle.Oracle(False)
    o.run(path, coarse.cli_args(iterations=2))
    c = np.fromfile(os.path.join(path, "object_n0_time_0.bin"), np.float32).reshape(
        cfg.N_z, cfg.N_xy // 2, cfg.N_xy // 2)
    assert np.abs(c).max() > 0
    o.run(path, cfg.cli_args(iterations=0, init_icd=2))
    f = np.fromfile(os.path.join(path, "object_n0_time_0.bin"), np.float32).reshape(
        cfg.N_z, cfg.N_xy, cfg.N_xy)
    np.testing.assert_array_equal(f, o.upsample_object(2, c))


# ---- counts to sinogram (gen_projection_4m_HDF, F/XT_HDFIO.c:160-199) ----
def _counts_case(seed=0, n_p=5, t_in=6, r_in=12, n_w=4, n_d=3):
    rng = np.random.default_rng(seed)
    white = rng.integers(3800, 4200, (n_w, t_in, r_in)).astype(np.uint16)
    dark = rng.integers(90, 110, (n_d, t_in, r_in)).astype(np.uint16)
    mu = rng.uniform(0.0, 2.0, (n_p, t_in, r_in))
    counts = (dark.mean(0) + (white.mean(0) - dark.mean(0)) * np.exp(-mu)).round().astype(np.uint16)
    return counts, white, dark


def test_counts_to_sinogram_closed_forms():
    o = py_oracle.Oracle(False)
    # flat fields: counts equal to the white level, no dark -> zero projection
    white = np.full((3, 4, 8), 2000, np.uint16)
    dark = np.zeros((2, 4, 8), np.uint16)
    counts = np.full((2, 4, 8), 2000, np.uint16)
    p, w, b = o.counts_to_sinogram(counts, white, dark, N_t=2, N_r=4, bh=0.3)
    np.testing.assert_array_equal(p, 0.0)
    np.testing.assert_array_equal(w, 2000 * 4)  # 2 x 2 pixels per bin
    np.testing.assert_array_equal(b, 2000 * 4)
    # Beer-Lambert: block sums of (counts - dark) against (white - dark), BH quadratic
    counts, white, dark = _counts_case()
    N_t, N_r, bh = 3, 4, 0.25
    p, w, b = o.counts_to_sinogram(counts, white, dark, N_t=N_t, N_r=N_r, bh=bh)
    wm, dm = white.astype(np.float64).mean(0), dark.astype(np.float64).mean(0)
    rt, rr = counts.shape[1] // N_t, counts.shape[2] // N_r
    blk = lambda a: a.reshape(a.shape[:-2] + (N_t, rt, N_r, rr)).sum((-3, -1))
    b_ref = blk(np.abs(wm - dm)).T
    w_ref = np.transpose(blk(np.abs(counts - dm)), (0, 2, 1))
    t = np.log(b_ref[None] / w_ref)
    np.testing.assert_allclose(b, b_ref, rtol=1e-6)
    np.testing.assert_allclose(w, w_ref, rtol=1e-6)
    np.testing.assert_allclose(p, bh * t * t + t, rtol=2e-5, atol=2e-6)


def test_counts_to_sinogram_pixel_mapping():
    """One count changed at detector pixel (t, r) of view 1 moves only bin
    (r // ratio_r, t // ratio_t) of that view."""
    o = py_oracle.Oracle(False)
    counts, white, dark = _counts_case(1)
    p0, w0, _ = o.counts_to_sinogram(counts, white, dark, N_t=3, N_r=4)
    c2 = counts.copy()
    c2[1, 4, 7] += 50
    p1, w1, _ = o.counts_to_sinogram(c2, white, dark, N_t=3, N_r=4)
    changed = np.argwhere(w1 != w0)
    assert changed.tolist() == [[1, 7 // 3, 4 // 2]]
    assert np.isclose(w1[1, 2, 2] - w0[1, 2, 2], 50.0)


# ---- phase contrast (F/XT_AMatrix.c:113-187, F/XT_ForwardProject.c:27-44) ----
def _lap_filter(ax, dr):
    """The detector Laplacian (F/XT_Init.c:284) of every view, zero outside."""
    k = np.array([[-0.5, -1, -0.5], [-1, 6, -1], [-0.5, -1, -0.5]]) / (dr * dr)
    out = np.zeros_like(ax)
    n_r, n_t = ax.shape[1:]
    for a in range(3):
        for b in range(3):
            src = np.zeros_like(ax)
            r0, r1 = max(0, a - 1), min(n_r, n_r + a - 1)
            t0, t1 = max(0, b - 1), min(n_t, n_t + b - 1)
            src[:, r0 - (a - 1):r1 - (a - 1), t0 - (b - 1):t1 - (b - 1)] = ax[:, r0:r1, t0:t1]
            out += k[a, b] * src
    return out


def test_phase_contrast_forward_is_laplacian_of_absorption(scan_root):
    """Every voxel's footprint is the Laplacian of its absorption footprint
    trimmed to the detector, so Ax_pc = Laplacian(Ax) on the detector (a
    transposed kernel or a dropped edge row fails this), and an interior voxel's
    footprint sums to zero."""
    cfg = dataclasses.replace(TINY, name="pc_pin")
    path, info = make_scan(cfg, scan_root, "pc")
    o = py_oracle.Oracle(True)
    S, O, T = o.setup(oracle_path(path, o), cfg.cli_args())
    try:
        vol = info["phantom"].astype(np.float64)
        o.set_object(O, vol)
        H = o.detector_response(S, O, T)
        ax = o.forward_project(S, O, H)
        o.set_phase_contrast(True)
        try:
            axp = o.forward_project(S, O, H)
            one = np.zeros_like(vol)
            one[0, 2, 16, 15] = 1.0
            o.set_object(O, one)
            foot = o.forward_project(S, O, H)
        finally:
            o.set_phase_contrast(False)
        want = _lap_filter(ax, S.delta_r)
        np.testing.assert_allclose(axp, want, rtol=0, atol=1e-12 * np.abs(want).max())
        assert abs(foot.sum()) <= 1e-12 * np.abs(foot).sum()
        assert np.count_nonzero(foot[:, :, 1]) > 0 and np.count_nonzero(foot[:, :, 3]) > 0
    finally:
        o.free(S, O, T)


def _edge_mask_ndimage(img, length_x, delta_xy, radius, band):
    """The edge-only mask of one slice with scipy.ndimage in float64: pixels
    whose |gradient| is within `band` (relative) of the threshold are returned
    as undecided (lo = 0, hi = 1) and so are the dilated pixels they reach."""
    from scipy import ndimage
    k = np.arange(-2, 3)
    hb = np.exp(-(k[:, None] ** 2 + k[None, :] ** 2) / 6.0)
    hb /= hb.sum()
    hx = np.array([[3, 0, -3], [10, 0, -10], [3, 0, -3]], np.float64)
    hy = np.array([[3, 10, 3], [0, 0, 0], [-3, -10, -3]], np.float64)
    blur = ndimage.correlate(img, hb, mode="constant", cval=0.0)
    mag = np.hypot(ndimage.correlate(blur, hx, mode="constant"),
                   ndimage.correlate(blur, hy, mode="constant"))
    n_y, n_x = img.shape
    xc = -length_x / 2 + (np.arange(n_x) + 0.5) * delta_xy
    yc = -length_x / 2 + (np.arange(n_y) + 0.5) * delta_xy
    inside = (xc[None, :] ** 2 + yc[:, None] ** 2) < radius ** 2
    thr = 0.0002
    lo = inside & (mag > thr * (1 + band))
    hi = inside & (mag > thr * (1 - band))
    sq = np.ones((5, 5), bool)
    return (ndimage.binary_dilation(lo, structure=sq).astype(np.uint8),
            ndimage.binary_dilation(hi, structure=sq).astype(np.uint8), mag)


def test_edge_detect_matches_ndimage(scan_root):
    """oracle_edge_detect (F/XT_ImageProc.c:83-149) against scipy.ndimage on the
    same steps in float64: a normalised 5x5 Gaussian (variance 3), the 3x3
    gradient pair, |gradient| > 0.0002 inside radius_obj, a 5x5 dilation.  The
    slices are smooth random fields scaled so that about half their pixels lie
    above the threshold, so a missing normaliser, another variance or a 3x3
    dilation moves many pixels; each slice has its own field, so reading the
    wrong plane fails.  Pixels within 0.1% of the threshold are left open."""
    from scipy import ndimage
    cfg = dataclasses.replace(TINY, name="edge_pin")
    path, _ = make_scan(cfg, scan_root, "edge")
    o = py_oracle.Oracle(False)
    S, O, T = o.setup(path, cfg.cli_args())
    try:
        T.radius_obj = 12.0
        rng = np.random.default_rng(5)
        vol = np.zeros((O.N_time, O.N_z, O.N_y, O.N_x))
        for m in range(O.N_z):
            f = ndimage.gaussian_filter(rng.normal(size=(O.N_y, O.N_x)), 1.5 + 0.5 * m)
            _, _, mag = _edge_mask_ndimage(f, O.Length_X, O.delta_xy, 1e9, 0.0)
            vol[0, m] = f * (0.0002 / np.median(mag))
        vol = vol.astype(np.float32)
        o.set_object(O, vol)
        got = o.edge_detect(O, T)
        n_open = 0
        for m in range(O.N_z):
            lo, hi, _ = _edge_mask_ndimage(vol[0, m].astype(np.float64), O.Length_X, O.delta_xy,
                                           T.radius_obj, 1e-3)
            g = got[0, :, :, m]
            assert np.all(g >= lo) and np.all(g <= hi), f"slice {m}"
            n_open += int((lo != hi).sum())
            assert 0 < g.sum() < g.size
        assert n_open < 0.02 * got.size
        # outside the circle grown by the 2-pixel dilation nothing is set
        xc = -O.Length_X / 2 + (np.arange(O.N_x) + 0.5) * O.delta_xy
        r = np.hypot(xc[None, :], xc[:, None])
        assert not got[0][r > T.radius_obj + 3 * O.delta_xy].any()
    finally:
        o.free(S, O, T)


# ---- pins of the oracle's rejection branches, prior cost, temporal term ----
# The generalized-Huber data term rho(u) (u = e sqrt(w)) is pinned by
# test_cost_closed_form; the checks below derive what the voxel, offset and
# variance updates must compute from that pinned cost -- by finite differences
# and by the minimisation each update performs -- with entries on both sides of
# the rejection threshold, so a dropped term, a wrong branch or a wrong weight in
# a rejected entry fails them.
THETA = dataclasses.replace(TINY, name="theta", p=2.0, alpha=1.0, zinger_T=2.0,
                            update_offset=0, zero_mean_offset=False, seed=31)


def _error_sino(o, S, O, T, vol):
    """e = y - Ax - d of the object vol (float64), ProjSelect set from it."""
    o.set_object(O, vol)
    H = o.detector_response(S, O, T)
    E = np.zeros((S.N_p, S.N_r, S.N_t))
    et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(H)
    o.lib.oracle_init_error_sino(C.byref(S), C.byref(O), C.byref(T), ht.ptr, et.ptr)
    return E, et, H, ht


def _cost(o, S, O, T, vol):
    E, et, _, _ = _error_sino(o, S, O, T, vol)
    return o.lib.oracle_compute_cost(C.byref(S), C.byref(O), C.byref(T), et.ptr)


def _theta_object(info, O):
    rng = np.random.default_rng(8)
    vol = 0.8 * info["phantom"].astype(np.float64)
    return vol + rng.uniform(0.0, 2e-4, vol.shape)


def test_voxel_theta_from_cost_derivative(oracle64, scan_root):
    """theta1, theta2 of one voxel update (F/XT_VoxUpdate.c:20-58), read back from
    the update itself: with p = q = 2 and alpha = 1 the new value is
    u = (sum b' x_j + theta2 V - theta1) / (sum b' + theta2) with b' = filter *
    2 / (sigma^2 (c + 1)) known, so three prior strengths give theta1 and theta2
    (and the residual checks that form).  Then theta1 must be the derivative of
    the pinned data cost along that voxel (central differences), and the
    quadratic theta1 D + theta2 D^2 / 2 must majorize the data cost's change,
    touching at D = 0 -- with rejected entries in the voxel's footprint (T = 2)."""
    path, info = make_scan(THETA, scan_root)
    path = oracle_path(path, oracle64)
    zv, yv, xv = 1, 13, 17
    vol0 = None
    rows = []
    for sigma in (2e-3, 4e-3, 8e-3):
        cfg = dataclasses.replace(THETA, sigma_s=sigma)
        S, O, T = oracle64.setup(path, cfg.cli_args())
        try:
            oracle64.read_sinogram(path, S, T)
            if vol0 is None:
                vol0 = _theta_object(info, O)
            E, et, H, ht = _error_sino(oracle64, S, O, T, vol0)
            sel = oracle64.sino_view(S.ProjSelect, S, bool)
            assert 0.001 < 1.0 - sel.mean() < 0.5  # entries on both sides of the threshold
            kappa = 2.0 / (T.Sigma_S_Q * (O.C_S + 1.0))
            filt = np.array(T.Spatial_Filter).reshape(3, 3, 3)
            V = vol0[0, zv, yv, xv]
            B = Sx = 0.0
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        if (dz, dy, dx) != (0, 0, 0):
                            b = filt[dz + 1, dy + 1, dx + 1] * kappa
                            B += b
                            Sx += b * vol0[0, zv + dz, yv + dy, xv + dx]
            xs, ys = np.array([xv], np.int32), np.array([yv], np.int32)
            mag = np.zeros((O.N_y, O.N_x))
            mt = xt_abi.PtrTree(mag)
            zc = C.c_int64(0)
            oracle64.lib.oracle_update_lines(C.byref(S), C.byref(O), C.byref(T), et.ptr, ht.ptr,
                                             0, zv, zv, 1, xs.ctypes.data, ys.ctypes.data, 1,
                                             C.byref(zc), mt.ptr, 0)
            u = oracle64.get_object(O)[0, zv, yv, xv]
            rows.append((u - V, 1.0, Sx - u * B))
        finally:
            oracle64.free(S, O, T)
    M = np.array([r[:2] for r in rows])
    rhs = np.array([r[2] for r in rows])
    (theta2, theta1), res, _, _ = np.linalg.lstsq(M, rhs, rcond=None)
    assert np.abs(M @ [theta2, theta1] - rhs).max() <= 1e-9 * np.abs(rhs).max()
    # the data cost along the voxel, with the prior made negligible (sigma = 1e3)
    cfg = dataclasses.replace(THETA, sigma_s=1e3)
    S, O, T = oracle64.setup(path, cfg.cli_args())
    try:
        oracle64.read_sinogram(path, S, T)

        def D(delta):
            v = vol0.copy()
            v[0, zv, yv, xv] += delta
            return _cost(oracle64, S, O, T, v)
        c0 = D(0.0)
        h = 1e-8
        d1 = (D(h) - D(-h)) / (2 * h)
        assert d1 == pytest.approx(theta1, rel=1e-5)
        assert theta2 > 0
        for delta in (-3e-3, -1e-3, -2e-4, -3e-5, 3e-5, 2e-4, 1e-3, 3e-3):
            q = theta1 * delta + 0.5 * theta2 * delta * delta
            assert q >= D(delta) - c0 - 1e-9 * abs(c0), delta
        # touching: the gap closes as D^2 (a quadratic of this curvature, not larger)
        small = 1e-5
        gap = theta1 * small + 0.5 * theta2 * small * small - (D(small) - c0)
        assert 0.0 <= gap <= 0.5 * theta2 * small * small + 1e-9 * abs(c0)
    finally:
        oracle64.free(S, O, T)


def _per_entry_derivatives(o, S, O, T, E, et, h=1e-7):
    """dC/de for every sinogram entry of the pinned cost (central differences of
    oracle_compute_cost with one entry moved; ProjSelect held)."""
    flat = E.reshape(-1)
    out = np.zeros_like(flat)
    for k in range(flat.size):
        e0 = flat[k]
        flat[k] = e0 + h
        cp = o.lib.oracle_compute_cost(C.byref(S), C.byref(O), C.byref(T), et.ptr)
        flat[k] = e0 - h
        cm = o.lib.oracle_compute_cost(C.byref(S), C.byref(O), C.byref(T), et.ptr)
        flat[k] = e0
        out[k] = (cp - cm) / (2 * h)
    return out.reshape(E.shape)


OFFS = dataclasses.replace(TINY, name="offs", N_xy=8, N_z=2, N_p=16, zinger_T=2.0, seed=33)


@pytest.mark.parametrize("zero_mean", [0, 1])
def test_offset_update_minimises_majorizer(zero_mean, oracle64, scan_root):
    """update_Sinogram_Offset (F/XT_ICD_update.c:666-759) with rejected entries:
    the step of each detector pixel minimises the symmetric quadratic majorizer
    of the pinned data cost in that pixel's offset -- sum_k wt_k (e_k - s)^2 / 2
    with wt_k = (dC/de_k) / e_k (the weight whose quadratic has the cost's slope at
    e_k) -- and with the zero-mean constraint the steps are the constrained
    minimiser, solved here from its KKT system (so gamma is checked, not only
    sum d = 0)."""
    path, info, (S, O, T) = _setup(oracle64, OFFS, scan_root)
    try:
        rng = np.random.default_rng(4)
        shape = (S.N_p, S.N_r, S.N_t)
        W = rng.uniform(1e3, 2e4, shape)
        E0 = rng.normal(0.0, 0.012, shape)
        E0[rng.random(shape) < 0.05] *= 20.0  # beyond the threshold
        oracle64.sino_view(T.Weight, S)[:] = W
        sel = np.abs(E0 * np.sqrt(W)) < T.ErrorSinoThresh
        assert 0.02 < 1.0 - sel.mean() < 0.3
        oracle64.sino_view(S.ProjSelect, S, bool)[:] = sel
        oracle64.set_object(O, np.zeros((O.N_time, O.N_z, O.N_y, O.N_x)))
        d0 = rng.normal(0.0, 0.01, (S.N_r, S.N_t))
        E = E0.copy()
        et = xt_abi.PtrTree(E)
        g = _per_entry_derivatives(oracle64, S, O, T, E, et)
        wt = g / E0
        num = (wt * E0).sum(0)
        den = wt.sum(0)
        n = num.size
        if zero_mean:
            K = np.zeros((n + 1, n + 1))
            K[:n, :n] = np.diag(den.reshape(-1))
            K[:n, n] = 1.0
            K[n, :n] = 1.0
            sol = np.linalg.solve(K, np.concatenate([num.reshape(-1), [-d0.sum()]]))
            step = sol[:n].reshape(num.shape)
        else:
            step = num / den
        T.EnforceZeroMeanOffset = zero_mean
        oracle64.offset_view(S)[:] = d0
        oracle64.lib.oracle_offset_update(C.byref(S), C.byref(T), et.ptr)
        d = oracle64.offset_view(S)
        np.testing.assert_allclose(d - d0, step, rtol=2e-5, atol=1e-12)
        np.testing.assert_allclose(E, E0 - (d - d0)[None], rtol=1e-12, atol=1e-15)
        if zero_mean:
            assert abs(d.sum()) <= 1e-9 * np.abs(d).sum()
        # ProjSelect follows the new errors
        assert np.array_equal(oracle64.sino_view(S.ProjSelect, S, bool),
                              np.abs(E * np.sqrt(W)) < T.ErrorSinoThresh)
    finally:
        oracle64.free(S, O, T)


def test_variance_fixed_point_minimises_cost(oracle64, scan_root):
    """update_variance_parameter (F/XT_ICD_update.c:633-664) with rejected entries:
    repeated, it settles where the pinned MAP cost, as a function of the variance
    (weights w0 / var, the N/2 log var term), is stationary -- central differences
    of oracle_compute_cost there vanish and the cost is at a minimum."""
    path, info, (S, O, T) = _setup(oracle64, OFFS, scan_root)
    try:
        rng = np.random.default_rng(6)
        shape = (S.N_p, S.N_r, S.N_t)
        w0 = rng.uniform(1e3, 2e4, shape)  # weights for unit variance
        E = rng.normal(0.0, 0.015, shape)
        E[rng.random(shape) < 0.05] *= 15.0
        et = xt_abi.PtrTree(E)
        oracle64.set_object(O, np.zeros((O.N_time, O.N_z, O.N_y, O.N_x)))
        var = 1.0
        oracle64.sino_view(T.Weight, S)[:] = w0 / var
        oracle64.sino_view(S.ProjSelect, S, bool)[:] = np.abs(E * np.sqrt(w0 / var)) < T.ErrorSinoThresh
        T.var_est = var
        for _ in range(200):
            oracle64.lib.oracle_variance_update(C.byref(S), C.byref(T), et.ptr)
        var = T.var_est
        sel = oracle64.sino_view(S.ProjSelect, S, bool).copy()
        assert 0.01 < 1.0 - sel.mean() < 0.3
        # the last update left the weights as the reference does: w = w0 / var
        np.testing.assert_allclose(oracle64.sino_view(T.Weight, S), w0 / var, rtol=1e-9)

        def cost_at(v):
            oracle64.sino_view(T.Weight, S)[:] = w0 / v
            oracle64.sino_view(S.ProjSelect, S, bool)[:] = sel
            T.var_est = v
            return oracle64.lib.oracle_compute_cost(C.byref(S), C.byref(O), C.byref(T), et.ptr)
        h = 1e-6 * var
        slope = (cost_at(var + h) - cost_at(var - h)) / (2 * h)
        scale = E.size / var  # the size of either part of the slope
        assert abs(slope) <= 1e-6 * scale
        c = cost_at(var)
        assert c < cost_at(1.05 * var) and c < cost_at(0.95 * var)
    finally:
        oracle64.free(S, O, T)


def _qggmrf(delta, p, c, sigma_q):
    """The qGGMRF potential with q = 2 (Thibault et al., Med. Phys. 2007):
    |d/s|^2 / (c + |d/s|^(2-p)), s = sigma_q^(1/2)."""
    s = math.sqrt(sigma_q)
    a = np.abs(delta) / s
    return a ** 2 / (c + a ** (2.0 - p))


PRIOR3D = dataclasses.replace(TINY, name="prior3d", N_xy=5, N_z=3, N_p=12, seed=35)
PRIOR4D = dataclasses.replace(SMALL4D, name="prior4d", N_xy=5, N_z=3, N_time=2, N_p=32,
                              interlaced_K=2, interlaced_N_theta=16, seed=37)


@pytest.mark.parametrize("cfg", [PRIOR3D, PRIOR4D], ids=["3d", "4d-two-frames"])
def test_prior_cost_brute_force(cfg, oracle64, scan_root):
    """The prior half of computeCost (F/XT_ICD_update.c:205-300) on a random object
    with zero error sinogram (data term 0, log term 0 at unit variance): the sum
    over unordered pairs of 26-neighbours inside the object of filter * rho_s plus,
    between consecutive frames, Time_Filter * rho_t -- enumerated here from all 26
    offsets of every voxel (each pair twice, halved), not the reference's 13
    causal ones."""
    path, info, (S, O, T) = _setup(oracle64, cfg, scan_root)
    try:
        rng = np.random.default_rng(10)
        vol = rng.uniform(0.0, 3e-3, (O.N_time, O.N_z, O.N_y, O.N_x))
        oracle64.set_object(O, vol)
        E = np.zeros((S.N_p, S.N_r, S.N_t))
        et = xt_abi.PtrTree(E)
        oracle64.sino_view(S.ProjSelect, S, bool)[:] = True
        T.var_est = 1.0
        got = oracle64.lib.oracle_compute_cost(C.byref(S), C.byref(O), C.byref(T), et.ptr)
        filt = np.array(T.Spatial_Filter).reshape(3, 3, 3)
        nt, nz, ny, nx = vol.shape
        want = 0.0
        for t in range(nt):
            for z in range(nz):
                for y in range(ny):
                    for x in range(nx):
                        for dz in (-1, 0, 1):
                            for dy in (-1, 0, 1):
                                for dx in (-1, 0, 1):
                                    z2, y2, x2 = z + dz, y + dy, x + dx
                                    if (dz, dy, dx) == (0, 0, 0) or not (
                                            0 <= z2 < nz and 0 <= y2 < ny and 0 <= x2 < nx):
                                        continue
                                    want += 0.5 * filt[dz + 1, dy + 1, dx + 1] * _qggmrf(
                                        vol[t, z, y, x] - vol[t, z2, y2, x2], O.MRF_P, O.C_S,
                                        T.Sigma_S_Q)
                        if t + 1 < nt:
                            want += T.Time_Filter[0] * _qggmrf(
                                vol[t, z, y, x] - vol[t + 1, z, y, x], O.MRF_P, O.C_T,
                                T.Sigma_T_Q)
        assert got == pytest.approx(want, rel=1e-12)
        # a constant shift of the object leaves the prior unchanged
        oracle64.set_object(O, vol + 1e-3)
        assert oracle64.lib.oracle_compute_cost(C.byref(S), C.byref(O), C.byref(T), et.ptr) == \
            pytest.approx(want, rel=1e-9)
    finally:
        oracle64.free(S, O, T)


def test_init_filter_inverse_distance(oracle64, scan_root):
    """initFilter (F/XT_Init.c:210-256): the 26 spatial weights are inversely
    proportional to the neighbour's distance (ratios), the centre weight is 0, the
    temporal weight equals a face neighbour's (both at distance 1), and the 26
    spatial plus the two temporal weights sum to mult_xy^3 * delta_Rtime."""
    for cfg in (TINY, SMALL4D):
        path, info, (S, O, T) = _setup(oracle64, cfg, scan_root)
        try:
            filt = np.array(T.Spatial_Filter).reshape(3, 3, 3)
            off = np.indices((3, 3, 3)).reshape(3, -1).T - 1
            dist = np.sqrt((off ** 2).sum(1)).reshape(3, 3, 3)
            assert filt[1, 1, 1] == 0.0
            nz = dist > 0
            prod = filt[nz] * dist[nz]  # constant if weight ~ 1/distance
            np.testing.assert_allclose(prod, prod[0], rtol=1e-12)
            assert T.Time_Filter[0] == pytest.approx(filt[1, 1, 2], rel=1e-12)
            total = filt.sum() + 2 * T.Time_Filter[0]
            assert total == pytest.approx(O.mult_xy ** 3 * O.delta_Rtime, rel=1e-12)
        finally:
            oracle64.free(S, O, T)


QUAD4D = dataclasses.replace(QUAD, name="quad4d", N_time=2, N_p=24, interlaced_K=2,
                             interlaced_N_theta=12, sigma_t=3e-3, c_t=1.0, iterations=600,
                             seed=4)


def test_icd_converges_to_dense_map_solution_4d(oracle64, scan_root):
    """Two frames, p = q = 2, no rejection, alpha = 1: the MAP problem with the
    temporal term (F/XT_ICD_update.c:154-169, pairs x(t) - x(t+1) weighted by
    Time_Filter) is the linear system (A'WA + L_s + L_t) x = A'W y over both
    frames, each frame seeing only its own views; ICD must reach its dense
    solution."""
    path, info = make_scan(QUAD4D, scan_root)
    path = oracle_path(path, oracle64)
    S, O, T = oracle64.setup(path, QUAD4D.cli_args())
    try:
        assert O.N_time == 2 and T.Time_Filter[0] > 0
        oracle64.read_sinogram(path, S, T)
        H = oracle64.detector_response(S, O, T)
        ht = xt_abi.PtrTree(H)
        frames, num = oracle64.proj_lists(O)
        nt, nz, ny, nx = O.N_time, O.N_z, O.N_y, O.N_x
        nv = nz * ny * nx
        nsino = S.N_p * S.N_r * S.N_t
        A = np.zeros((nsino, nt * nv))
        vals = np.zeros(8)
        inds = np.zeros(8, np.int32)
        a = oracle64.AMatrixCol()
        a.values, a.index = vals.ctypes.data, inds.ctypes.data
        for t in range(nt):
            for p in frames[t]:
                for yy in range(ny):
                    for xx in range(nx):
                        oracle64.lib.oracle_amatrix_column(C.byref(S), C.byref(O), ht.ptr,
                                                           C.byref(a), yy, xx, int(p))
                        for q in range(a.count):
                            for z in range(nz):
                                A[(int(p) * S.N_r + inds[q]) * S.N_t + z,
                                  t * nv + (z * ny + yy) * nx + xx] = vals[q]
        y = oracle64.sino_view(S.Projection, S).reshape(-1).copy()
        w = oracle64.sino_view(T.Weight, S).reshape(-1).copy()
        filt = np.array(T.Spatial_Filter).reshape(27)
        ks = 2.0 / (T.Sigma_S_Q * (O.C_S + 1.0))
        kt = 2.0 / (T.Sigma_T_Q * (O.C_T + 1.0))
        L = np.zeros((nt * nv, nt * nv))
        for t in range(nt):
            for z in range(nz):
                for yy in range(ny):
                    for xx in range(nx):
                        i = t * nv + (z * ny + yy) * nx + xx
                        for dz in (-1, 0, 1):
                            for dy in (-1, 0, 1):
                                for dx in (-1, 0, 1):
                                    z2, y2, x2 = z + dz, yy + dy, xx + dx
                                    if (dz, dy, dx) == (0, 0, 0) or not (
                                            0 <= z2 < nz and 0 <= y2 < ny and 0 <= x2 < nx):
                                        continue
                                    k = t * nv + (z2 * ny + y2) * nx + x2
                                    b = filt[((dz + 1) * 3 + dy + 1) * 3 + dx + 1] * ks
                                    L[i, i] += b
                                    L[i, k] -= b
                        for t2 in (t - 1, t + 1):
                            if 0 <= t2 < nt:
                                k = t2 * nv + (z * ny + yy) * nx + xx
                                b = T.Time_Filter[0] * kt
                                L[i, i] += b
                                L[i, k] -= b
        xs = np.linalg.solve(A.T @ (w[:, None] * A) + L, A.T @ (w * y)).reshape(nt, nz, ny, nx)
    finally:
        oracle64.free(S, O, T)
    oracle64.run(path, QUAD4D.cli_args())
    x = np.stack([np.fromfile(os.path.join(path, f"object_n0_time_{t}.bin"), np.float64)
                  .reshape(nz, ny, nx) for t in range(nt)])
    # the temporal coupling matters on this problem: dropping it moves the solution
    assert np.abs(x - xs).max() <= 1e-6 * np.abs(xs).max()


# -- random line selection ----------------------------------------------------
SEL = dataclasses.replace(TINY, name="sel", N_xy=8, N_z=4, N_p=12, num_threads=2, seed=31)


def test_random_select_is_a_uniform_permutation(oracle32, scan_root):
    """randomly_select_x_y (F/XT_ICD_update.c:599-631) on a random voxel mask, per
    (frame, z-block): every line holding a masked voxel of the block is selected
    exactly once and no other line is; exactly N_y*N_x numbers of the pinned RNG
    are drawn per (frame, block); and over many seeds every selected line comes
    first, and lands in the first half of the order, as often as a uniform shuffle
    makes it (the raster-last line included, which a draw scaled by one less than
    the remaining count could never pick first)."""
    path, info, (S, O, T) = _setup(oracle32, SEL, scan_root)
    try:
        nb, nxy, nz = T.num_z_blocks, O.N_y * O.N_x, O.N_z
        bs = nz // nb
        rng = np.random.default_rng(8)
        mask = (rng.random((O.N_time, O.N_y, O.N_x, nz)) < 0.3).astype(np.uint8)
        mask[:, 2] = 0  # a row of lines with no masked voxel
        mask[:, -1, -1, :] = 1  # the raster-last line in every block
        mt = xt_abi.PtrTree(mask)
        sel = [mask[0, :, :, b * bs:(b + 1) * bs].any(axis=2).reshape(-1) for b in range(nb)]
        first = np.zeros((nb, nxy))
        firsthalf = np.zeros((nb, nxy))
        n_seeds = 4000
        for seed in range(1, n_seeds + 1):
            oracle32.lib.oracle_srandom2(seed)
            oracle32.lib.oracle_random_select(C.byref(O), C.byref(T), mt.ptr)
            num = py_oracle.deref_data(T.UpdateSelectNum, 1, C.c_int32, O.N_time * nb).copy()
            xs = py_oracle.deref_data(T.x_rand_select, 2, C.c_int32, O.N_time * nb * nxy)
            ys = py_oracle.deref_data(T.y_rand_select, 2, C.c_int32, O.N_time * nb * nxy)
            for k in range(O.N_time * nb):
                f, b = divmod(k, nb)
                got = (ys.reshape(-1, nxy)[k, :num[k]] * O.N_x + xs.reshape(-1, nxy)[k, :num[k]])
                if seed <= 50:
                    want = np.nonzero(mask[f, :, :, b * bs:(b + 1) * bs].any(axis=2).reshape(-1))[0]
                    assert len(set(got.tolist())) == num[k]
                    assert sorted(got.tolist()) == want.tolist()
                if f == 0:
                    first[b, got[0]] += 1
                    firsthalf[b, got[: num[k] // 2]] += 1
            if seed <= 50:
                nxt = oracle32.lib.oracle_random2()
                draws = np.zeros(O.N_time * nb * nxy + 1)
                oracle32.lib.oracle_random2_seq(seed, draws.size, draws.ctypes.data)
                assert nxt == draws[-1]  # N_y*N_x draws per (frame, block)
        for b in range(nb):
            n = int(sel[b].sum())
            for counts, p in ((first[b], 1.0 / n), (firsthalf[b], (n // 2) / n)):
                freq = counts[sel[b]] / n_seeds
                sd = math.sqrt(p * (1 - p) / n_seeds)
                assert np.abs(freq - p).max() < 4.5 * sd, (b, p, freq.min(), freq.max())
            assert first[b, ~sel[b]].sum() == 0
    finally:
        oracle32.free(S, O, T)
