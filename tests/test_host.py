"""CPU tests of the host side of libxt_hip.so (no GPU calls): it loads and
exports every include/xt_icd.h and include/xt_unet.h symbol, its engine set-up reproduces the
reference's argsParser + initStructures (checked against the oracle's own
restatement), its pointer trees have the reference layout, and bad input is an
error return rather than a process exit."""
import ctypes as C
import dataclasses
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, SMALL4D, TINY, as_double, make_scan
from timbir_python_amd import synth, xt_hip, xt_hip_f64


@pytest.mark.parametrize("mod,real", [(xt_hip, "float"), (xt_hip_f64, "double")],
                         ids=["libxt_hip", "libxt_hip_f64"])
def test_library_exports_every_declared_symbol(mod, real):
    declared = set()
    for h in ("xt_icd.h", "xt_unet.h"):
        header = open(os.path.join(ROOT, "include", h)).read()
        declared |= set(re.findall(r"\b(ICD_BackProject|xt_[a-z0-9_]+)\s*\(", header))
    assert declared == set(mod.EXPORTS)
    lib = mod.lib()
    for name in declared:
        assert hasattr(lib, name), name
    assert "gfx950" in mod.version() and mod.version().endswith("Real_t=" + real)
    assert C.sizeof(mod.Sinogram) == C.sizeof(xt_abi_types(real == "double")[0])


def xt_abi_types(double):
    from timbir_python_amd import xt_abi
    return xt_abi.abi_types(double)


def _scalar_fields(st):
    out = {}
    for name, ty in st._fields_:
        if ty is C.c_void_p:
            continue  # pointer fields: compared as arrays below
        v = getattr(st, name)
        if isinstance(v, (int, float)):
            out[name] = v
    return out


@pytest.mark.parametrize("double", [False, True], ids=["float", "double"])
@pytest.mark.parametrize("cfg", [TINY, SMALL4D,
                                 dataclasses.replace(TINY, name="tiny_mz2", N_z=2, mult_z=2)],
                         ids=lambda c: c.name)
def test_engine_setup_matches_oracle(cfg, double, scan_root, oracle32, oracle64):
    """Float build against the float oracle; the double build (Real_t = double,
    double .bin files) against the oracle's REAL_IS_DOUBLE build: every scalar,
    array and list equal."""
    path, info = make_scan(cfg, scan_root)
    mod, oracle = (xt_hip_f64, oracle64) if double else (xt_hip, oracle32)
    if double:
        path = as_double(path)
    args = cfg.cli_args()
    eng = mod.Engine(path, args)
    S, O, T = oracle.setup(path, args)
    oracle.read_sinogram(path, S, T)
    try:
        for a, b in ((eng.S, S), (eng.O, O), (eng.T, T)):
            fa, fb = _scalar_fields(a), _scalar_fields(b)
            for k in fa:
                if k in ("debug_file_ptr",):
                    continue
                assert fa[k] == fb[k], k
        assert np.array_equal(np.array(eng.T.Spatial_Filter), np.array(T.Spatial_Filter))
        assert eng.T.Time_Filter[0] == T.Time_Filter[0]
        n = S.N_p
        for fld in ("ViewPtr", "TimePtr", "cosine", "sine"):
            ga = np.ctypeslib.as_array((mod.REAL * n).from_address(getattr(eng.S, fld)))
            gb = oracle.vec_view(getattr(S, fld), n)
            assert np.array_equal(ga, gb), fld
        lists_o, num_o = oracle.proj_lists(O)
        num_g = np.ctypeslib.as_array((C.c_int32 * O.N_time).from_address(eng.O.ProjNum))
        assert np.array_equal(num_g, num_o)
        top = C.cast(eng.O.ProjIdxPtr, C.POINTER(C.c_void_p))
        for f in range(O.N_time):
            g = np.ctypeslib.as_array((C.c_int32 * int(num_o[f])).from_address(top[f]))
            assert np.array_equal(g, lists_o[f])
        assert eng.sino("Projection").dtype == mod.NP_REAL
        assert np.array_equal(eng.sino("Projection"), oracle.sino_view(S.Projection, S))
        assert np.array_equal(eng.sino("Weight"), oracle.sino_view(T.Weight, S))
        assert eng.proj_select().all()
        assert np.array_equal(eng.offset(), oracle.offset_view(S))
    finally:
        eng.close()
        oracle.free(S, O, T)


def test_engine_sigma_hu_roundtrip(scan_root):
    cfg = TINY
    path, info = make_scan(cfg, scan_root)
    with xt_hip.Engine(path, cfg.cli_args()) as eng:
        assert eng.O.Sigma_S == pytest.approx(cfg.sigma_s, rel=1e-6)
        assert eng.O.N_x == cfg.N_xy and eng.S.N_r == cfg.N_r
        # prior weights: 26 + 2 neighbours normalised to mult_xy^3 * delta_Rtime
        tot = np.sum(eng.T.Spatial_Filter) + 2 * eng.T.Time_Filter[0]
        assert tot == pytest.approx(cfg.mult_xy ** 3, rel=1e-6)


@pytest.mark.parametrize("bad", [["--no_such_option", "1"], ["--iter"]])
def test_engine_setup_rejects_bad_cli(bad, scan_root):
    cfg = TINY
    path, info = make_scan(cfg, scan_root)
    with pytest.raises(xt_hip.XtError):
        xt_hip.Engine(path, cfg.cli_args() + bad)


def test_engine_setup_rejects_view_count_mismatch(scan_root):
    cfg = TINY
    path, info = make_scan(cfg, scan_root)
    args = cfg.cli_args()
    args[args.index("--num_projections") + 1] = str(cfg.N_p + 1)
    with pytest.raises(xt_hip.XtError):
        xt_hip.Engine(path, args, read_sinogram=False)


def test_engine_setup_rejects_short_sinogram(scan_root):
    cfg = TINY
    path, info = make_scan(cfg, scan_root)
    w = np.fromfile(os.path.join(path, "weight_n0.bin"), np.float32)
    w[: w.size // 2].tofile(os.path.join(path, "weight_n0.bin"))
    with pytest.raises(xt_hip.XtError):
        xt_hip.Engine(path, cfg.cli_args())


def test_multialloc_reference_layout():
    lib = xt_hip.lib()
    dims = (C.c_int64 * 3)(2, 3, 4)
    top = lib.xt_multialloc(4, 3, dims)
    assert top
    lvl0 = C.cast(top, C.POINTER(C.c_void_p))
    rows = [C.cast(lvl0[i], C.POINTER(C.c_void_p)) for i in range(2)]
    data0 = rows[0][0]
    for i in range(2):
        for j in range(3):
            assert rows[i][j] == data0 + ((i * 3 + j) * 4) * 4  # one contiguous block
    lib.xt_multifree(top, 3)


def test_cli_binary_reports_errors(tmp_path):
    exe = os.path.join(ROOT, "timbir-python_amd", "xt_engine")
    assert os.path.exists(exe)
    r = subprocess.run([exe, "--no_such_option", "1"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0 and "xt_engine" in r.stderr


def test_ingest_hdf5_needs_h5py():
    from timbir_python_amd import ingest
    try:
        import h5py  # noqa: F401
    except ImportError:
        with pytest.raises(RuntimeError, match="h5py"):
            ingest.read_hdf5("a.h5", "b.h5", "c.h5", N_p=1, proj_start=0, slice_begin=0,
                             slice_num=1, x_start=0, x_width=8, N_r=8)


def test_output_writers(tmp_path):
    """The launcher's result formats: air -> 0 HU and water -> 1000 HU, the
    8-bit window clipped at both ends, TIFF slices that read back."""
    from PIL import Image
    from timbir_python_amd import output
    att, den = 100.0 ** 4, 0.001 ** 4
    air = 0.496372335005353 * att * 0.001205 * den
    water = 0.521225397034623 * att * 1.0 * den
    np.testing.assert_allclose(output.um_to_hu([air, water]), [0.0, 1000.0], atol=1e-9)
    np.testing.assert_array_equal(output.hu_to_uint8(np.array([-50.0, 0.0, 500.0, 1000.0, 2e3]),
                                                     1000.0, 0.0), [0, 0, 127, 255, 255])
    obj = np.zeros((2, 3, 4, 4), np.float32)
    obj[1, 2, 1, 1] = water
    for j in range(2):
        obj[j].tofile(tmp_path / f"object_n0_time_{j}.bin")
    names = output.write_tiff_slices(str(tmp_path), str(tmp_path / "tif"), 2, 3, 4, 1000.0, -1000.0)
    assert len(names) == 6
    img = np.asarray(Image.open(tmp_path / "tif" / "object_time_z_1_2.tif"))
    assert img.dtype == np.uint8 and img[1, 1] == 255 and img[0, 0] == 127


REF_F = "/root/reference/Source_Code_4D_Fast"


@pytest.mark.skipif(not os.path.isdir(REF_F), reason="reference sources absent")
def test_dropin_links_with_reference_host_objects():
    """The reference host sources that stay on the engine's link line when the
    library replaces XT_ICD_update.c, XT_VoxUpdate.c, XT_NHICD.c and
    XT_ForwardProject.c (XT_AMatrix.c and XT_Profile.c stay: XT_genSinogram.c:135,
    144,168 call them) link against libxt_hip.so with --no-undefined, and
    ICD_BackProject resolves from the library."""
    r = subprocess.run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "linkcheck",
                        f"REF={REF_F}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    so = os.path.join(ROOT, "oracle", "_ref", "libxt_host_link.so")
    dyn = subprocess.run(["readelf", "-d", so], capture_output=True, text=True).stdout
    assert "[libxt_hip.so]" in dyn
    und = subprocess.run(["nm", "-D", "-u", so], capture_output=True, text=True).stdout.split()
    assert "ICD_BackProject" in und
    # without XT_AMatrix.c and XT_Profile.c the link must fail on exactly their symbols
    r = subprocess.run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "linkcheck",
                        f"REF={REF_F}", "HOST_SRCS=XT_genSinogram.c XT_IOMisc.c XT_ImageProc.c "
                        "XT_Filter.c XT_HDFIO.c allocate.c randlib.c tiff.c"],
                       capture_output=True, text=True)
    assert r.returncode != 0
    missing = set(re.findall(r"undefined reference to `(\w+)'", r.stderr))
    assert missing == {"DetectorResponseProfile", "storeVoxelLineResponse",
                       "calcAMatrixColumnforAngle"}
    subprocess.run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "linkcheck",
                    f"REF={REF_F}"], capture_output=True, text=True, check=True)


@pytest.mark.skipif(not os.path.isdir(REF_F), reason="reference sources absent")
def test_dropin_links_double_host_with_double_library():
    """A host built with the reference's -DREAL_IS_DOUBLE (double structs) links
    against libxt_hip_f64.so the same way, with --no-undefined."""
    r = subprocess.run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "linkcheck_f64",
                        f"REF={REF_F}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    so = os.path.join(ROOT, "oracle", "_ref", "libxt_host_link_f64.so")
    dyn = subprocess.run(["readelf", "-d", so], capture_output=True, text=True).stdout
    assert "[libxt_hip_f64.so]" in dyn


@pytest.mark.parametrize("n_p,n_r,n_t", [(16384, 2048, 16), (8192, 2048, 32), (4096, 8192, 16)])
def test_oversize_sinogram_refused(n_p, n_r, n_t):
    """The voxel-update kernels address the sinogram through 32-bit byte offsets
    with 2^31 as the absent-bin value, so a sinogram of 2^29 entries or more is an
    argument error, raised before any device call (here: with no GPU)."""
    S, O, T = xt_hip.Sinogram(), xt_hip.ScannedObject(), xt_hip.TomoInputs()
    S.N_p, S.N_r, S.N_t, S.z_overlap_num, S.delta_r = n_p, n_r, n_t, 1, 1.0
    O.N_x = O.N_y = n_r * 2 // 3
    O.N_z, O.N_time, O.delta_xy = n_t, 1, 1.5
    assert n_p * n_r * n_t >= 1 << 29
    lib = xt_hip.lib()
    state = C.c_void_p()
    rc = lib.xt_icd_create(C.byref(S), C.byref(O), C.byref(T), C.byref(state))
    assert rc == 1  # XT_E_ARG
    assert b"at most 2^29 - 1 entries" in lib.xt_last_error()
    assert not state.value


def test_create_from_needs_a_source():
    """xt_icd_create_from without a source state is an argument error, raised
    before any device call (here: with no GPU)."""
    S, O, T = xt_hip.Sinogram(), xt_hip.ScannedObject(), xt_hip.TomoInputs()
    lib = xt_hip.lib()
    state = C.c_void_p()
    rc = lib.xt_icd_create_from(C.byref(S), C.byref(O), C.byref(T), None, C.byref(state))
    assert rc == 1  # XT_E_ARG
    assert b"null argument" in lib.xt_last_error()
    assert not state.value


def test_line_groups_switched_off():
    """Grouped line updates (XT_OPT_LINE_GROUP 2 and 4) are switched off in this
    build: the option refuses them and stays at 1."""
    lib = xt_hip.lib()
    for v in (2, 4, 3):
        assert lib.xt_set_option(xt_hip.OPT_LINE_GROUP, float(v)) == 1
        assert b"only 1" in lib.xt_last_error()
    assert lib.xt_get_option(xt_hip.OPT_LINE_GROUP) == 1
    assert lib.xt_set_option(xt_hip.OPT_LINE_GROUP, 1.0) == 0


def _counts_run(tmp_path, mxy, mz, n_p=8, nr_full=24, nt_full=8, seed=0):
    """A run folder with sinobin-3 inputs: bright_n0.bin [nr_full][nt_full] and
    counts weight_n0.bin [N_p][nr_full][nt_full] at full detector resolution."""
    rng = np.random.default_rng(seed)
    d = str(tmp_path)
    os.makedirs(d, exist_ok=True)
    ang = np.arange(n_p) * 180.0 / n_p
    synth.write_view_info(d, ang, np.arange(n_p, dtype=np.float64))
    bright = rng.uniform(9e3, 1.1e4, (nr_full, nt_full)).astype(np.float32)
    counts = (bright[None] * np.exp(-rng.uniform(0.0, 2.0, (n_p, nr_full, nt_full)))
              ).round().astype(np.float32)
    bright.tofile(os.path.join(d, "bright_n0.bin"))
    counts.tofile(os.path.join(d, "weight_n0.bin"))
    args = ["--p", "1.2", "--sigma_s", synth.sigma_to_hu(1e-4), "--c_s", "0.01",
            "--sigma_t", synth.sigma_to_hu(2e-4), "--c_t", "0.01", "--delta_xy", str(mxy),
            "--delta_z", str(mz), "--length_r", str(nr_full), "--length_t", str(nt_full),
            "--voxel_thresh", "0", "--cost_thresh", "0", "--iter", "1",
            "--rotation_center", str(nr_full / 2.0), "--alpha", "1.5", "--sinobin", "3",
            "--initICD", "0", "--Rtime0", "0", "--Rtime_delta", "1", "--Rtime_num", "1",
            "--num_projections", str(n_p), "--N_r", str(nr_full), "--N_t", str(nt_full),
            "--num_threads", "2", "--radius_obj", str(nr_full), "--updateProjOffset", "0",
            "--writeTiff", "0", "--zingerT", "4.0", "--zingerDel", "0.1", "--Est_of_Var", "1"]
    return d, args, bright, counts


def test_engine_sinobin3_binning_closed_form(tmp_path):
    """sinobin 3 (genSinogram_subsampleCounts, F/XT_genSinogram.c:303-370): the
    detector binned by delta_xy x delta_z (N_r, N_t, rotation centre divided,
    voxels of one binned pixel, slices of one row), weight = the summed counts,
    projection = log(summed bright / summed counts) with no beam hardening."""
    d, args, bright, counts = _counts_run(tmp_path, 2, 2)
    xt_hip.set_option(xt_hip.OPT_BH_QUAD_COEF, 0.0)
    with xt_hip.Engine(d, args) as eng:
        S, O = eng.S, eng.O
        assert (S.N_r, S.N_t, O.N_x, O.N_z, S.z_overlap_num) == (12, 4, 12, 4, 1)
        assert O.delta_xy == pytest.approx(S.delta_r) and S.delta_r == pytest.approx(2.0)
        w = eng.sino("Weight").astype(np.float64)
        y = eng.sino("Projection").astype(np.float64)
    cb = counts.astype(np.float64).reshape(8, 12, 2, 4, 2).sum(axis=(2, 4))
    bb = bright.astype(np.float64).reshape(12, 2, 4, 2).sum(axis=(1, 3))
    np.testing.assert_allclose(w, cb, rtol=1e-6)
    np.testing.assert_allclose(y, np.log(bb[None] / cb), rtol=1e-5, atol=1e-6)


@pytest.mark.skipif(not os.path.isdir(REF_F), reason="reference sources absent")
@pytest.mark.parametrize("double", [False, True], ids=["float", "double"])
@pytest.mark.parametrize("mxy,mz", [(1, 1), (2, 2), (3, 1)])
def test_engine_sinobin3_matches_reference(mxy, mz, double, tmp_path, monkeypatch):
    """The engine's counts reader against the reference's own compiled
    genSinogram_subsampleCounts (oracle/_ref, built with BH_QUAD_COEF 0.05, and
    with -DREAL_IS_DOUBLE for the double build) on the same structs and files:
    projection and weight bit for bit."""
    import py_oracle
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "gen", f"REF={REF_F}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    mod = xt_hip_f64 if double else xt_hip
    ref = C.CDLL(os.path.join(ROOT, "oracle", "_ref",
                              "libxtref_gen_f64.so" if double else "libxtref_gen_f32.so"))
    ref.genSinogram_subsampleCounts.argtypes = [C.c_void_p] * 3
    d, args, bright, counts = _counts_run(tmp_path, mxy, mz, nr_full=24, nt_full=6, seed=mxy)
    if double:
        for f in ("bright_n0.bin", "weight_n0.bin"):
            a = np.fromfile(os.path.join(d, f), np.float32)
            a.astype(np.float64).tofile(os.path.join(d, f))
    mod.set_option(mod.OPT_BH_QUAD_COEF, 0.05)
    try:
        with mod.Engine(d, args) as eng:
            y_ours = eng.sino("Projection").copy()
            w_ours = eng.sino("Weight").copy()
            eng.sino("Projection")[:] = 0
            eng.sino("Weight")[:] = 0
            eng.T.debug_file_ptr = py_oracle.devnull_file()
            monkeypatch.chdir(d)  # the reference reads bright_n0.bin, weight_n0.bin from the cwd
            assert ref.genSinogram_subsampleCounts(*[C.addressof(x) for x in (eng.S, eng.O, eng.T)]) == 0
            y_ref = eng.sino("Projection").copy()
            w_ref = eng.sino("Weight").copy()
    finally:
        mod.set_option(mod.OPT_BH_QUAD_COEF, 0.0)
    assert np.abs(y_ref).max() > 0
    np.testing.assert_array_equal(w_ours, w_ref)
    np.testing.assert_array_equal(y_ours, y_ref)


def test_engine_sinobin3_coarse_offsets_repeated(tmp_path):
    """With sinobin 3 a later multi-resolution stage reads the previous stage's
    ring offsets on its coarser detector (initICD 2: half the bins; 3: half the
    bins and rows) and repeats each over the pixels it covers (F/XT_Init.c:340-361)."""
    for init, (pxy, pz) in ((2, (2, 1)), (3, (2, 2))):
        d, args, _, _ = _counts_run(tmp_path / f"i{init}", 1, 1, nr_full=24, nt_full=8)
        args = list(args)
        args[args.index("--initICD") + 1] = str(init)
        args[args.index("--updateProjOffset") + 1] = "3"
        coarse = np.arange((24 // pxy) * (8 // pz), dtype=np.float32).reshape(24 // pxy, 8 // pz)
        coarse.tofile(os.path.join(d, "proj_offset_n0.bin"))
        np.array([1.0], np.float32).tofile(os.path.join(d, "variance_estimate_n0.bin"))
        with xt_hip.Engine(d, args, read_sinogram=False) as eng:
            got = eng.offset().copy()
        np.testing.assert_array_equal(got, np.repeat(np.repeat(coarse, pxy, 0), pz, 1))
