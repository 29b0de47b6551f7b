"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded only by tests/, the smoke check and the benchmark's CPU-baseline leg.
Wraps liboracle_f32.so / liboracle_f64.so (the C restatement in xt_oracle.c)
and, when it has been built, oracle/_ref/libxtref_*.so (the reference's own
self-contained sources).
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _load_pkg_module(name):
    path = os.path.join(ROOT, "timbir-python_amd", name + ".py")
    key = "timbir_python_amd." + name
    if key in sys.modules:
        return sys.modules[key]
    spec = importlib.util.spec_from_file_location(key, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[key] = mod
    spec.loader.exec_module(mod)
    return mod


xt_abi = _load_pkg_module("xt_abi")

_libc = C.CDLL(None)
_libc.fopen.restype = C.c_void_p
_libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
_libc.fclose.argtypes = [C.c_void_p]


def devnull_file():
    return _libc.fopen(b"/dev/null", b"w")


def argv_array(args):
    arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
    return len(args), arr


def deref_data(ptr, levels, ctype, n):
    """Follow first pointers `levels` times and view n elements of ctype as numpy."""
    addr = ptr if isinstance(ptr, int) else ptr
    for _ in range(levels):
        addr = C.cast(addr, C.POINTER(C.c_void_p))[0]
    buf = (ctype * n).from_address(addr)
    return np.ctypeslib.as_array(buf)


class Oracle:
    """One precision build of the oracle."""

    def __init__(self, real_is_double: bool = False, float_amatrix: bool = False):
        """real_is_double: the reference's -DREAL_IS_DOUBLE build; float_amatrix
        (with real_is_double): FP64 arithmetic on the float build's system matrix."""
        self.double = real_is_double
        if real_is_double and float_amatrix:
            name = "liboracle_f64a.so"
        else:
            name = "liboracle_f64.so" if real_is_double else "liboracle_f32.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
        self.lib = C.CDLL(path)
        self.R = C.c_double if real_is_double else C.c_float
        self.np_real = np.float64 if real_is_double else np.float32
        (self.Sinogram, self.ScannedObject, self.AMatrixCol,
         self.TomoInputs) = xt_abi.abi_types(real_is_double)
        L = self.lib
        P = C.c_void_p
        L.oracle_srandom2.argtypes = [C.c_uint32]
        L.oracle_random2.restype = C.c_double
        L.oracle_random2_seq.argtypes = [C.c_uint32, C.c_int32, P]
        L.oracle_engine_setup.argtypes = [C.c_char_p, C.c_int32, P, P, P, P]
        L.oracle_engine_setup.restype = C.c_int32
        L.oracle_engine_read_sinogram.argtypes = [C.c_char_p, P, P]
        L.oracle_engine_read_sinogram.restype = C.c_int32
        L.oracle_engine_free.argtypes = [P, P, P]
        L.oracle_engine_run.argtypes = [C.c_char_p, C.c_int32, P, C.c_int32, P, P]
        L.oracle_engine_run.restype = C.c_int32
        L.oracle_detector_response.argtypes = [P, P, P, P]
        L.oracle_amatrix_column.argtypes = [P, P, P, P, C.c_int32, C.c_int32, C.c_int32]
        L.oracle_forward_project.argtypes = [P, P, P, P]
        L.oracle_init_error_sino.argtypes = [P, P, P, P, P]
        L.oracle_compute_cost.argtypes = [P, P, P, P]
        L.oracle_compute_cost.restype = self.R
        L.oracle_offset_update.argtypes = [P, P, P]
        L.oracle_variance_update.argtypes = [P, P, P]
        L.oracle_functional_substitution.argtypes = [self.R, self.R, self.R, P, P, P, P, P, P,
                                                     C.c_int32]
        L.oracle_functional_substitution.restype = self.R
        L.oracle_update_lines.argtypes = [P, P, P, P, P, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32, P, P, C.c_int32, P, P, C.c_int32]
        L.oracle_update_lines.restype = self.R
        L.oracle_update_lines_mask.argtypes = [P, P, P, P, P, C.c_int32, C.c_int32, C.c_int32,
                                               C.c_int32, P, P, C.c_int32, P, P, P, C.c_int32]
        L.oracle_update_lines_mask.restype = self.R
        L.oracle_edge_detect.argtypes = [P, P, P]
        L.oracle_vsc_select.argtypes = [P, P, P]
        L.oracle_random_select.argtypes = [P, P, P]
        L.oracle_icd_backproject.argtypes = [P, P, P, C.c_char_p, C.c_int32, P]
        L.oracle_icd_backproject.restype = C.c_int32
        L.oracle_parse_view_info.argtypes = [C.c_char_p, C.c_int32, P, P]
        L.oracle_parse_view_info.restype = C.c_int32
        L.oracle_bin_views.argtypes = [P, C.c_int32, C.c_int32, self.R, self.R, P, P]
        L.oracle_upsample_object.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, P, P]
        L.oracle_counts_to_sinogram.argtypes = [C.c_int32] * 5 + [P, C.c_int32, P, C.c_int32, P,
                                                                  C.c_double, P, P, P]
        L.oracle_counts_to_sinogram.restype = None
        L.oracle_set_phase_contrast.argtypes = [C.c_int32]
        L.oracle_set_phase_contrast.restype = None
        L.oracle_upsample_object.restype = None
        L.oracle_interlaced_views.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                              C.c_double, P, P, P, P, P]
        L.oracle_interlaced_views.restype = C.c_int32
        L.oracle_sample_updates.argtypes = [P, P, P, P, C.c_int32, C.c_int32,
                                            C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.oracle_sample_updates.restype = C.c_int64

    # -- engine -----------------------------------------------------------
    def upsample_object(self, init_icd, coarse):
        """initObject's upsampling (initICD 2: y, x; 3: z, y, x) of a coarse
        [Nzc][Nyc][Nxc] volume."""
        c = np.ascontiguousarray(coarse, self.np_real)
        nz, ny, nx = c.shape
        out = np.zeros((2 * nz if init_icd == 3 else nz, 2 * ny, 2 * nx), self.np_real)
        self.lib.oracle_upsample_object(init_icd, nz, ny, nx, c.ctypes.data, out.ctypes.data)
        return out

    def edge_detect(self, O, T):
        """oracle_edge_detect: the edge-only mask [N_time][N_y][N_x][N_z] (uint8) of O."""
        mask = np.zeros((O.N_time, O.N_y, O.N_x, O.N_z), np.uint8)
        mt = xt_abi.PtrTree(mask)
        self.lib.oracle_edge_detect(C.byref(O), C.byref(T), mt.ptr)
        return mask

    def set_phase_contrast(self, on):
        """The -DPHASE_CONTRAST_TOMOGRAPHY build's Laplacian footprints on/off."""
        self.lib.oracle_set_phase_contrast(1 if on else 0)

    def counts_to_sinogram(self, counts, white, dark, N_t, N_r, bh=0.0):
        """gen_projection_4m_HDF's arithmetic on uint16 counts [N_p][T_in][R_in] and
        white/dark frames [n][T_in][R_in]: (projection, weight, bright)."""
        counts = np.ascontiguousarray(counts, np.uint16)
        white = np.ascontiguousarray(white, np.uint16)
        dark = np.ascontiguousarray(dark, np.uint16)
        n_p, t_in, r_in = counts.shape
        proj = np.zeros((n_p, N_r, N_t), self.np_real)
        wt = np.zeros_like(proj)
        bright = np.zeros((N_r, N_t), self.np_real)
        self.lib.oracle_counts_to_sinogram(n_p, t_in, r_in, N_t, N_r, counts.ctypes.data,
                                           white.shape[0], white.ctypes.data, dark.shape[0],
                                           dark.ctypes.data, float(bh), proj.ctypes.data,
                                           wt.ctypes.data, bright.ctypes.data)
        return proj, wt, bright

    def setup(self, path, args):
        S, O, T = self.Sinogram(), self.ScannedObject(), self.TomoInputs()
        n, arr = argv_array(args)
        rc = self.lib.oracle_engine_setup(path.encode(), n, arr, C.byref(S), C.byref(O),
                                          C.byref(T))
        if rc != 0:
            raise RuntimeError(f"oracle_engine_setup failed ({rc})")
        T.debug_file_ptr = devnull_file()
        return S, O, T

    def read_sinogram(self, path, S, T):
        if self.lib.oracle_engine_read_sinogram(path.encode(), C.byref(S), C.byref(T)) != 0:
            raise RuntimeError("oracle_engine_read_sinogram failed")

    def free(self, S, O, T):
        self.lib.oracle_engine_free(C.byref(S), C.byref(O), C.byref(T))
        if T.debug_file_ptr:
            _libc.fclose(T.debug_file_ptr)
            T.debug_file_ptr = None

    def run(self, path, args, cost=False, positivity=False, n_iter=None):
        """Whole engine on the files in `path`; returns (costs, iterations)."""
        n, arr = argv_array(args)
        niter = n_iter if n_iter is not None else int(args[args.index("--iter") + 1])
        costs = np.full(niter + 1, np.nan)
        iters = C.c_int32(0)
        flags = (1 if cost else 0) | (2 if positivity else 0)
        rc = self.lib.oracle_engine_run(path.encode(), n, arr, flags, costs.ctypes.data,
                                        C.byref(iters))
        if rc != 0:
            raise RuntimeError(f"oracle_engine_run failed ({rc})")
        return costs, iters.value

    # -- array access -----------------------------------------------------
    def object_view(self, O, frame):
        """numpy view of Object[frame] ([N_z+2][N_y][N_x], halo slices included)."""
        top = C.cast(O.Object, C.POINTER(C.c_void_p))[frame]
        n = (O.N_z + 2) * O.N_y * O.N_x
        return deref_data(top, 2, self.R, n).reshape(O.N_z + 2, O.N_y, O.N_x)

    def sino_view(self, ptr, S, dtype=None):
        ct = C.c_bool if dtype is bool else self.R
        n = S.N_p * S.N_r * S.N_t
        return deref_data(ptr, 2, ct, n).reshape(S.N_p, S.N_r, S.N_t)

    def offset_view(self, S):
        return deref_data(S.ProjOffset, 1, self.R, S.N_r * S.N_t).reshape(S.N_r, S.N_t)

    def proj_lists(self, O):
        top = C.cast(O.ProjIdxPtr, C.POINTER(C.c_void_p))
        num = np.ctypeslib.as_array((C.c_int32 * O.N_time).from_address(O.ProjNum)).copy()
        return [np.ctypeslib.as_array((C.c_int32 * int(num[i])).from_address(top[i])).copy()
                for i in range(O.N_time)], num

    def vec_view(self, ptr, n):
        return np.ctypeslib.as_array((self.R * n).from_address(ptr))

    # -- seams ------------------------------------------------------------
    def detector_response(self, S, O, T):
        H = np.zeros((S.N_p, 65), self.np_real)
        tree = xt_abi.PtrTree(H)
        self.lib.oracle_detector_response(tree.ptr, C.byref(S), C.byref(O), C.byref(T))
        return H

    def forward_project(self, S, O, H):
        AX = np.zeros((S.N_p, S.N_r, S.N_t), self.np_real)
        ht, at = xt_abi.PtrTree(np.ascontiguousarray(H, self.np_real)), xt_abi.PtrTree(AX)
        self.lib.oracle_forward_project(C.byref(S), C.byref(O), ht.ptr, at.ptr)
        return AX

    def set_object(self, O, vol):
        """vol: [N_time][N_z][N_y][N_x] -> Object slices 1..N_z (halo set to zero)."""
        for f in range(O.N_time):
            v = self.object_view(O, f)
            v[:] = 0
            v[1:O.N_z + 1] = vol[f]

    def get_object(self, O):
        return np.stack([self.object_view(O, f)[1:O.N_z + 1].copy() for f in range(O.N_time)])

    def projector(self, path, args):
        """Callable phantom -> Ax using this build's geometry for the scan in `path`."""
        def project(vol):
            S, O, T = self.setup(path, args)
            try:
                self.set_object(O, vol)
                H = self.detector_response(S, O, T)
                return self.forward_project(S, O, H).astype(np.float64)
            finally:
                self.free(S, O, T)
        return project


class RefLib:
    """The reference's own buildable subset (oracle/_ref), same struct ABI."""

    def __init__(self, real_is_double: bool = False):
        name = "libxtref_f64.so" if real_is_double else "libxtref_f32.so"
        self.path = os.path.join(HERE, "_ref", name)
        if not os.path.exists(self.path):
            raise FileNotFoundError(self.path)
        self.lib = C.CDLL(self.path)
        self.R = C.c_double if real_is_double else C.c_float
        self.np_real = np.float64 if real_is_double else np.float32
        P = C.c_void_p
        L = self.lib
        L.srandom2.argtypes = [C.c_uint32]
        L.random2.restype = C.c_double
        L.DetectorResponseProfile.argtypes = [P, P, P, P, P]
        L.calcAMatrixColumnforAngle.argtypes = [P, P, P, P, C.c_int32, C.c_int32, C.c_int32]
        L.forward_project_voxel.argtypes = [P, self.R, P, P, C.c_int32, C.c_int32]
        L.VSC_based_Voxel_Line_Select.argtypes = [P, P, P]
        L.VSC_based_Voxel_Line_Select.restype = C.c_int32

    @staticmethod
    def available(real_is_double=False):
        name = "libxtref_f64.so" if real_is_double else "libxtref_f32.so"
        return os.path.exists(os.path.join(HERE, "_ref", name))
