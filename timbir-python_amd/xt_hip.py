"""ctypes binding of libxt_hip.so (include/xt_icd.h).

Mirrors the reference's engine interface: ``ICD_BackProject(S, O, T)`` over the
Source_Code_4D_Fast/XT_Structures.h structs, the engine set-up around it, and
one seam per hot-path function.  There is no CPU fallback: loading fails with
an error when the library has not been built, and every call raises
``XtError`` with the library's message when it returns non-zero.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import xt_abi

HERE = os.path.dirname(os.path.abspath(__file__))
# Real_t of the library this module binds: float (libxt_hip.so, the reference's
# production build) unless xt_hip_f64 loads this module for the double build
REAL_IS_DOUBLE = bool(globals().get("_XT_REAL_IS_DOUBLE", False))
LIB_PATH = os.path.join(HERE, "libxt_hip_f64.so" if REAL_IS_DOUBLE else "libxt_hip.so")
REAL = C.c_double if REAL_IS_DOUBLE else C.c_float  # host arrays of the structs
NP_REAL = np.float64 if REAL_IS_DOUBLE else np.float32

(OPT_POSITIVITY, OPT_COST, OPT_RESYNC_EVERY, OPT_SEED, OPT_LINES, OPT_DEVICE, OPT_VERBOSE,
 OPT_OUT_DIR, OPT_ATABLE, OPT_GRAPH, OPT_PACK, OPT_NHICD_LINES, OPT_INIT_FBP,
 OPT_AUTOTUNE, OPT_BATCH_FUSE, OPT_LINE_THREADS, OPT_PHASE_CONTRAST, OPT_BH_QUAD_COEF,
 OPT_LINE_GROUP) = range(1, 20)

# every symbol include/xt_icd.h and include/xt_unet.h declare
EXPORTS = [
    "xt_last_error", "xt_version", "ICD_BackProject", "xt_set_option", "xt_get_option",
    "xt_icd_create", "xt_icd_create_from", "xt_icd_iterate", "xt_icd_set_timing",
    "xt_icd_get_stats", "xt_icd_cost",
    "xt_icd_resync", "xt_icd_fbp", "xt_icd_download", "xt_icd_synchronize", "xt_icd_destroy",
    "xt_icd_batch_create", "xt_icd_batch_iterate", "xt_icd_batch_destroy",
    "xt_engine_setup", "xt_engine_read_sinogram", "xt_engine_write_outputs", "xt_engine_free",
    "xt_engine_run", "xt_multialloc", "xt_multifree", "xt_detector_response",
    "xt_amatrix_columns", "xt_forward_project", "xt_init_error_sino", "xt_compute_cost",
    "xt_offset_update", "xt_variance_update", "xt_vsc_select", "xt_update_lines",
    "xt_update_lines_parity", "xt_edge_detect", "xt_probe_footprint_atomics",
    "xt_offset_update_noalloc",
    "xt_unet_create", "xt_unet_destroy", "xt_unet_param_count", "xt_unet_get_params",
    "xt_unet_set_params", "xt_unet_get_scale", "xt_unet_forward", "xt_unet_loss_grad",
    "xt_unet_adam_step", "xt_unet_train_step", "xt_unet_last_ms", "xt_icd_unet_start",
    "xt_unet_forward_host", "xt_unet_loss_grad_host", "xt_unet_train_step_host",
    "xt_upsample_object", "xt_icd_upload", "xt_counts_to_sinogram",
]


class XtError(RuntimeError):
    pass


class IcdStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("voxel_updates", C.c_double),
                ("footprint_entries", C.c_double), ("icd_kernel_ms", C.c_double),
                ("icd_launches", C.c_int64), ("last_cost", C.c_double),
                ("stop_percentage", C.c_double), ("stopped", C.c_int32),
                ("line_visits", C.c_double), ("lines_in_flight", C.c_int32),
                ("line_threads", C.c_int32)]


Sinogram, ScannedObject, AMatrixCol, TomoInputs = xt_abi.abi_types(REAL_IS_DOUBLE)

_lib = None


def lib():
    """The loaded library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise XtError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.xt_last_error.restype = C.c_char_p
        L.xt_version.restype = C.c_char_p
        L.ICD_BackProject.argtypes = [P, P, P]
        L.xt_set_option.argtypes = [C.c_int, C.c_double]
        L.xt_get_option.argtypes = [C.c_int]
        L.xt_get_option.restype = C.c_double
        L.xt_icd_create.argtypes = [P, P, P, C.POINTER(P)]
        L.xt_icd_create_from.argtypes = [P, P, P, P, C.POINTER(P)]
        L.xt_icd_iterate.argtypes = [P, C.c_int32, C.POINTER(C.c_int32)]
        L.xt_icd_set_timing.argtypes = [P, C.c_int32]
        L.xt_icd_get_stats.argtypes = [P, C.POINTER(IcdStats)]
        L.xt_icd_cost.argtypes = [P, C.POINTER(C.c_double)]
        L.xt_icd_resync.argtypes = [P]
        L.xt_icd_fbp.argtypes = [P]
        L.xt_icd_batch_create.argtypes = [C.c_int32, P, P]
        L.xt_icd_batch_iterate.argtypes = [P, C.c_int32]
        L.xt_icd_batch_destroy.argtypes = [P]
        L.xt_icd_batch_destroy.restype = None
        L.xt_icd_download.argtypes = [P, P, P, P]
        L.xt_icd_synchronize.argtypes = [P]
        L.xt_icd_destroy.argtypes = [P]
        L.xt_icd_destroy.restype = None
        L.xt_engine_setup.argtypes = [C.c_char_p, C.c_int, P, P, P, P]
        L.xt_engine_read_sinogram.argtypes = [C.c_char_p, P, P, P]
        L.xt_engine_write_outputs.argtypes = [C.c_char_p, P, P, P]
        L.xt_engine_free.argtypes = [P, P, P]
        L.xt_engine_free.restype = None
        L.xt_engine_run.argtypes = [C.c_char_p, C.c_int, P]
        L.xt_multialloc.argtypes = [C.c_size_t, C.c_int, P]
        L.xt_multialloc.restype = P
        L.xt_multifree.argtypes = [P, C.c_int]
        L.xt_multifree.restype = None
        L.xt_detector_response.argtypes = [P, P, P, P]
        L.xt_amatrix_columns.argtypes = [P, P, P, C.c_int32, P, P, P, C.c_int32, P, P, P]
        L.xt_forward_project.argtypes = [P, P, P, P]
        L.xt_init_error_sino.argtypes = [P, P, P, P, P]
        L.xt_compute_cost.argtypes = [P, P, P, P, C.POINTER(C.c_double)]
        L.xt_offset_update.argtypes = [P, P, P]
        L.xt_variance_update.argtypes = [P, P, P]
        L.xt_vsc_select.argtypes = [P, P, P]
        L.xt_update_lines.argtypes = [P, P, P, P, C.c_int32, C.c_int32, P, P, C.c_int32, P,
                                      C.POINTER(C.c_double), P]
        L.xt_offset_update_noalloc.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                               C.c_float, C.c_int32, P, P, P]
        L.xt_update_lines_parity.argtypes = [P, P, P, P, C.c_int32, P, P, C.c_int32, C.c_int32,
                                             P, C.POINTER(C.c_double)]
        L.xt_unet_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                     C.c_int32, C.POINTER(P)]
        L.xt_unet_destroy.argtypes = [P]
        L.xt_unet_destroy.restype = None
        L.xt_unet_param_count.argtypes = [P]
        L.xt_unet_param_count.restype = C.c_int64
        L.xt_unet_get_params.argtypes = [P, P, C.c_int64]
        L.xt_unet_set_params.argtypes = [P, P, C.c_int64]
        L.xt_unet_get_scale.argtypes = [P, C.POINTER(C.c_float)]
        L.xt_unet_forward.argtypes = [P, P, P]
        L.xt_unet_loss_grad.argtypes = [P, P, P, C.POINTER(C.c_double), P]
        L.xt_unet_adam_step.argtypes = [P, C.c_float, C.c_float, C.c_float, C.c_float]
        L.xt_unet_train_step.argtypes = [P, P, P, C.c_float, C.POINTER(C.c_double)]
        L.xt_unet_last_ms.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.xt_icd_unet_start.argtypes = [P, P]
        L.xt_upsample_object.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, P, P]
        L.xt_icd_upload.argtypes = [P, P, P, P]
        L.xt_counts_to_sinogram.argtypes = [C.c_int32] * 5 + [P, C.c_int32, P, C.c_int32, P,
                                                              C.c_double, P, P, P]
        L.xt_unet_forward_host.argtypes = [P, P, P]
        L.xt_unet_loss_grad_host.argtypes = [P, P, P, C.POINTER(C.c_double), P]
        L.xt_unet_train_step_host.argtypes = [P, P, P, C.c_float, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise XtError(f"libxt_hip error {rc}: {lib().xt_last_error().decode()}")


def version() -> str:
    return lib().xt_version().decode()


def set_option(opt: int, value: float) -> None:
    _check(lib().xt_set_option(opt, float(value)))


def get_option(opt: int) -> float:
    return lib().xt_get_option(opt)


def _argv(args):
    return len(args), (C.c_char_p * max(1, len(args)))(*[a.encode() for a in args])


# -- engine -----------------------------------------------------------------
class Engine:
    """Structs set up by the library from a run folder (the reference main's
    argsParser + initStructures + genSinogram_fromBin)."""

    def __init__(self, path: str, args, read_sinogram: bool = True):
        self.path = path
        self.S, self.O, self.T = Sinogram(), ScannedObject(), TomoInputs()
        n, arr = _argv(args)
        rc = lib().xt_engine_setup(path.encode(), n, arr, C.byref(self.S), C.byref(self.O),
                                   C.byref(self.T))
        if rc != 0:
            lib().xt_engine_free(C.byref(self.S), C.byref(self.O), C.byref(self.T))
            _check(rc)
        if read_sinogram:
            _check(lib().xt_engine_read_sinogram(path.encode(), C.byref(self.S), C.byref(self.O),
                                                 C.byref(self.T)))

    def refs(self):
        return C.byref(self.S), C.byref(self.O), C.byref(self.T)

    def read_sinogram(self) -> None:
        _check(lib().xt_engine_read_sinogram(self.path.encode(), C.byref(self.S),
                                             C.byref(self.O), C.byref(self.T)))

    def close(self):
        if self.S is not None:
            lib().xt_engine_free(*self.refs())
            self.S = self.O = self.T = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # array views
    def _view(self, ptr, levels, ctype, n):
        addr = ptr
        for _ in range(levels):
            addr = C.cast(addr, C.POINTER(C.c_void_p))[0]
        return np.ctypeslib.as_array((ctype * n).from_address(addr))

    def object(self, frame: int) -> np.ndarray:
        """Object[frame] as [N_z+2][N_y][N_x] (halo slices included)."""
        O = self.O
        top = C.cast(O.Object, C.POINTER(C.c_void_p))[frame]
        return self._view(top, 2, REAL, (O.N_z + 2) * O.N_y * O.N_x).reshape(
            O.N_z + 2, O.N_y, O.N_x)

    def volume(self) -> np.ndarray:
        return np.stack([self.object(f)[1:self.O.N_z + 1].copy() for f in range(self.O.N_time)])

    def set_volume(self, vol) -> None:
        for f in range(self.O.N_time):
            v = self.object(f)
            v[:] = 0
            v[1:self.O.N_z + 1] = vol[f]

    def sino(self, which: str) -> np.ndarray:
        S, T = self.S, self.T
        n = S.N_p * S.N_r * S.N_t
        ptr = {"Projection": S.Projection, "Weight": T.Weight}[which]
        return self._view(ptr, 2, REAL, n).reshape(S.N_p, S.N_r, S.N_t)

    def proj_select(self) -> np.ndarray:
        S = self.S
        return self._view(S.ProjSelect, 2, C.c_uint8, S.N_p * S.N_r * S.N_t).reshape(
            S.N_p, S.N_r, S.N_t)

    def offset(self) -> np.ndarray:
        S = self.S
        return self._view(S.ProjOffset, 1, REAL, S.N_r * S.N_t).reshape(S.N_r, S.N_t)


def engine_run(path: str, args) -> None:
    """The whole reference engine on a run folder (object_n0_time_<i>.bin out)."""
    n, arr = _argv(args)
    _check(lib().xt_engine_run(path.encode(), n, arr))


def ICD_BackProject(S, O, T) -> None:
    _check(lib().ICD_BackProject(C.byref(S), C.byref(O), C.byref(T)))


# -- stepwise ---------------------------------------------------------------
class Icd:
    """xt_icd_create / iterate / download over an Engine's structs."""

    def __init__(self, eng: Engine, sino_from: "Icd" = None):
        """sino_from: an open state over the same sinogram whose y and w the new
        state copies on the device (xt_icd_create_from) instead of uploading the
        Engine's Projection and Weight."""
        self.eng = eng
        self.h = C.c_void_p()
        if sino_from is None:
            _check(lib().xt_icd_create(*eng.refs(), C.byref(self.h)))
        else:
            _check(lib().xt_icd_create_from(*eng.refs(), sino_from.h, C.byref(self.h)))

    def iterate(self, n: int = 1) -> bool:
        stop = C.c_int32(0)
        _check(lib().xt_icd_iterate(self.h, n, C.byref(stop)))
        return bool(stop.value)

    def timing(self, on: bool) -> None:
        _check(lib().xt_icd_set_timing(self.h, 1 if on else 0))

    def stats(self) -> IcdStats:
        st = IcdStats()
        _check(lib().xt_icd_get_stats(self.h, C.byref(st)))
        return st

    def cost(self) -> float:
        c = C.c_double()
        _check(lib().xt_icd_cost(self.h, C.byref(c)))
        return c.value

    def resync(self) -> None:
        _check(lib().xt_icd_resync(self.h))

    def fbp(self) -> None:
        """Object := filtered backprojection of y - d; e recomputed."""
        _check(lib().xt_icd_fbp(self.h))

    def synchronize(self) -> None:
        _check(lib().xt_icd_synchronize(self.h))

    def download(self) -> None:
        _check(lib().xt_icd_download(self.h, *self.eng.refs()))

    def upload(self) -> None:
        """The Engine's object and ring offsets into the state; e recomputed."""
        _check(lib().xt_icd_upload(self.h, *self.eng.refs()))

    def close(self) -> None:
        if self.h:
            lib().xt_icd_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class IcdBatch:
    """xt_icd_batch over open Icd objects (not owned): one graph per iteration."""

    def __init__(self, icds):
        self.icds = list(icds)
        arr = (C.c_void_p * len(self.icds))(*[i.h.value for i in self.icds])
        self.h = C.c_void_p()
        _check(lib().xt_icd_batch_create(len(self.icds), arr, C.byref(self.h)))

    def iterate(self, n: int = 1) -> None:
        _check(lib().xt_icd_batch_iterate(self.h, n))

    def close(self) -> None:
        if self.h:
            lib().xt_icd_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# -- learned warm start -----------------------------------------------------
class UNet:
    """xt_unet: the warm-start 3D U-Net for volumes [d0][d1][d2] (d2 fastest).

    Volumes are host float32 arrays, copied through the network's own device
    staging buffers; every call returns after its work is done."""

    def __init__(self, shape, scale=1000.0, seed=0, train=False):
        self.shape = tuple(int(d) for d in shape)
        self.h = C.c_void_p()
        _check(lib().xt_unet_create(*self.shape, float(scale), int(seed), 1 if train else 0,
                                    C.byref(self.h)))
        self.n = int(lib().xt_unet_param_count(self.h))

    @property
    def scale(self) -> float:
        s = C.c_float()
        _check(lib().xt_unet_get_scale(self.h, C.byref(s)))
        return s.value

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n, np.float32)
        _check(lib().xt_unet_get_params(self.h, p.ctypes.data, self.n))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n:
            raise XtError(f"{p.size} parameters given, network has {self.n}")
        _check(lib().xt_unet_set_params(self.h, p.ctypes.data, self.n))

    def _vol(self, x):
        x = np.ascontiguousarray(x, np.float32)
        if x.shape != self.shape:
            raise XtError(f"volume of shape {x.shape}, network expects {self.shape}")
        return x

    def forward(self, x) -> np.ndarray:
        x = self._vol(x)
        out = np.empty_like(x)
        _check(lib().xt_unet_forward_host(self.h, x.ctypes.data, out.ctypes.data))
        return out

    def loss_grad(self, x, tgt, want_grad=False):
        x, tgt = self._vol(x), self._vol(tgt)
        loss = C.c_double()
        g = np.zeros(self.n, np.float32) if want_grad else None
        _check(lib().xt_unet_loss_grad_host(self.h, x.ctypes.data, tgt.ctypes.data,
                                            C.byref(loss), g.ctypes.data if want_grad else None))
        return loss.value, g

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        _check(lib().xt_unet_adam_step(self.h, lr, beta1, beta2, eps))

    def train_step(self, x, tgt, lr) -> float:
        x, tgt = self._vol(x), self._vol(tgt)
        loss = C.c_double()
        _check(lib().xt_unet_train_step_host(self.h, x.ctypes.data, tgt.ctypes.data, lr,
                                             C.byref(loss)))
        return loss.value

    def last_ms(self):
        f, b = C.c_double(), C.c_double()
        _check(lib().xt_unet_last_ms(self.h, C.byref(f), C.byref(b)))
        return f.value, b.value

    def save(self, path) -> None:
        np.savez(path, params=self.get_params(), scale=np.float32(self.scale),
                 arch=np.array("unet3d-4lvl-16-32-64-128"))

    @classmethod
    def load(cls, path, shape, train=False):
        z = np.load(path, allow_pickle=False)
        net = cls(shape, scale=float(z["scale"]), train=train)
        net.set_params(z["params"])
        return net

    def close(self) -> None:
        if self.h:
            lib().xt_unet_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def icd_unet_start(icd: "Icd", net: UNet) -> None:
    """Warm start in place: FBP, the network on the FBP volume, resync of e."""
    _check(lib().xt_icd_unet_start(icd.h, net.h))


# -- seams ------------------------------------------------------------------
def detector_response(eng: Engine) -> np.ndarray:
    H = np.zeros((eng.S.N_p, 65), NP_REAL)
    tree = xt_abi.PtrTree(H)
    _check(lib().xt_detector_response(*eng.refs(), tree.ptr))
    return H


def amatrix_columns(eng: Engine, H, rows, cols, views, max_entries=8):
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    views = np.ascontiguousarray(views, np.int32)
    n = rows.size
    cnt = np.zeros(n, np.uint8)
    idx = np.zeros((n, max_entries), np.int32)
    val = np.zeros((n, max_entries), NP_REAL)
    ht = xt_abi.PtrTree(np.ascontiguousarray(H, NP_REAL))
    _check(lib().xt_amatrix_columns(C.byref(eng.S), C.byref(eng.O), ht.ptr, n,
                                    rows.ctypes.data, cols.ctypes.data, views.ctypes.data,
                                    max_entries, cnt.ctypes.data, idx.ctypes.data,
                                    val.ctypes.data))
    return cnt, idx, val


def counts_to_sinogram(counts, white, dark, N_t: int, N_r: int, bh: float = 0.0):
    """Real-data ingest on the GPU (F/XT_HDFIO.c:160-199): uint16 counts
    [N_p][T_in][R_in] and white/dark frames [n][T_in][R_in] to (projection,
    weight [N_p][N_r][N_t], bright [N_r][N_t]) float32."""
    counts = np.ascontiguousarray(counts, np.uint16)
    white = np.ascontiguousarray(white, np.uint16)
    dark = np.ascontiguousarray(dark, np.uint16)
    n_p, t_in, r_in = counts.shape
    proj = np.zeros((n_p, N_r, N_t), NP_REAL)
    wt = np.zeros_like(proj)
    bright = np.zeros((N_r, N_t), NP_REAL)
    _check(lib().xt_counts_to_sinogram(n_p, t_in, r_in, N_t, N_r, counts.ctypes.data,
                                       white.shape[0], white.ctypes.data, dark.shape[0],
                                       dark.ctypes.data, float(bh), proj.ctypes.data,
                                       wt.ctypes.data, bright.ctypes.data))
    return proj, wt, bright


def upsample_object(init_icd: int, coarse) -> np.ndarray:
    """initObject's upsampling of a coarse [Nzc][Nyc][Nxc] object: initICD 2
    doubles y and x, 3 doubles z, y and x (F/XT_ICD_update.c:316-435)."""
    c = np.ascontiguousarray(coarse, NP_REAL)
    nz, ny, nx = c.shape
    out = np.zeros((2 * nz if init_icd == 3 else nz, 2 * ny, 2 * nx), NP_REAL)
    _check(lib().xt_upsample_object(init_icd, nz, ny, nx, c.ctypes.data, out.ctypes.data))
    return out


def forward_project(eng: Engine, H) -> np.ndarray:
    S = eng.S
    out = np.zeros((S.N_p, S.N_r, S.N_t), np.float64)
    ht = xt_abi.PtrTree(np.ascontiguousarray(H, NP_REAL))
    _check(lib().xt_forward_project(C.byref(eng.S), C.byref(eng.O), ht.ptr, out.ctypes.data))
    return out


def projector(path: str, args):
    """Callable phantom (float64 [N_time][N_z][N_y][N_x]) -> Ax (float64
    [N_p][N_r][N_t]) by the GPU forward projector, for the scan whose
    view_info.txt is in path; the projector synth.generate takes."""
    def project(vol):
        with Engine(path, args, read_sinogram=False) as eng:
            eng.set_volume(vol)
            return forward_project(eng, detector_response(eng))
    return project


def init_error_sino(eng: Engine, H) -> np.ndarray:
    S = eng.S
    E = np.zeros((S.N_p, S.N_r, S.N_t), NP_REAL)
    et, ht = xt_abi.PtrTree(E), xt_abi.PtrTree(np.ascontiguousarray(H, NP_REAL))
    _check(lib().xt_init_error_sino(*eng.refs(), ht.ptr, et.ptr))
    return E


def compute_cost(eng: Engine, E) -> float:
    et = xt_abi.PtrTree(np.ascontiguousarray(E, NP_REAL))
    c = C.c_double()
    _check(lib().xt_compute_cost(*eng.refs(), et.ptr, C.byref(c)))
    return c.value


def offset_update(eng: Engine, E: np.ndarray) -> None:
    et = xt_abi.PtrTree(E)
    _check(lib().xt_offset_update(C.byref(eng.S), C.byref(eng.T), et.ptr))


def variance_update(eng: Engine, E: np.ndarray) -> None:
    et = xt_abi.PtrTree(E)
    _check(lib().xt_variance_update(C.byref(eng.S), C.byref(eng.T), et.ptr))


def vsc_select(eng: Engine, mag: np.ndarray, update_select_num: np.ndarray):
    """mag [N_time][nb][N_y][N_x]; returns per (frame, block) lists of (y, x)."""
    O, T = eng.O, eng.T
    nb = T.num_z_blocks
    usn = eng._view(T.UpdateSelectNum, 1, C.c_int32, O.N_time * nb).reshape(O.N_time, nb)
    usn[:] = update_select_num
    mt = xt_abi.PtrTree(np.ascontiguousarray(mag, NP_REAL))
    _check(lib().xt_vsc_select(C.byref(eng.O), C.byref(eng.T), mt.ptr))
    nxy = O.N_y * O.N_x
    xs = eng._view(T.x_NHICD_select, 2, C.c_int32, O.N_time * nb * nxy).reshape(O.N_time, nb, nxy)
    ys = eng._view(T.y_NHICD_select, 2, C.c_int32, O.N_time * nb * nxy).reshape(O.N_time, nb, nxy)
    num = eng._view(T.NHICDSelectNum, 1, C.c_int32, O.N_time * nb).reshape(O.N_time, nb)
    return [[(ys[i, b, :num[i, b]].copy(), xs[i, b, :num[i, b]].copy()) for b in range(nb)]
            for i in range(O.N_time)]


def update_lines(eng: Engine, E: np.ndarray, frame: int, xs, ys, Iter: int, mask=None):
    """Sequential update of the listed lines; returns (MagUpdate [N_y][N_x], total_vox_mag).
    mask: optional uint8 voxel mask [N_time][N_y][N_x][N_z] (0 = leave the voxel)."""
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    O = eng.O
    mag = np.zeros((O.N_y, O.N_x), NP_REAL)
    tot = C.c_double()
    et = xt_abi.PtrTree(E)
    mp = None
    if mask is not None:
        mask = np.ascontiguousarray(mask, np.uint8)
        if mask.shape != (O.N_time, O.N_y, O.N_x, O.N_z):
            raise ValueError(f"mask shape {mask.shape}")
        mp = mask.ctypes.data
    _check(lib().xt_update_lines(*eng.refs(), et.ptr, frame, xs.size, xs.ctypes.data,
                                 ys.ctypes.data, Iter, mag.ctypes.data, C.byref(tot),
                                 C.c_void_p(mp)))
    return mag, tot.value


def offset_update_noalloc(N_p, N_r, N_t, zinger_T, zinger_delta, zero_mean, e_ptr, w_ptr, d_ptr):
    """xt_offset_update_noalloc on device arrays given by address (E, W: float32
    [N_p][N_r][N_t]; d: float32 [N_r][N_t])."""
    _check(lib().xt_offset_update_noalloc(N_p, N_r, N_t, zinger_T, zinger_delta,
                                          1 if zero_mean else 0, C.c_void_p(e_ptr),
                                          C.c_void_p(w_ptr), C.c_void_p(d_ptr)))


def update_lines_parity(eng: Engine, E: np.ndarray, xs, ys, Iter: int, parity: int):
    """The production form of a voxel-update pass on the listed lines (one z-block
    parity, split sinogram, packed lanes; xt_update_lines_parity); returns
    (MagUpdate [N_y][N_x], total_vox_mag)."""
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    O = eng.O
    mag = np.zeros((O.N_y, O.N_x), NP_REAL)
    tot = C.c_double()
    et = xt_abi.PtrTree(E)
    _check(lib().xt_update_lines_parity(*eng.refs(), et.ptr, xs.size, xs.ctypes.data,
                                        ys.ctypes.data, Iter, parity, mag.ctypes.data,
                                        C.byref(tot)))
    return mag, tot.value


def probe_footprint_atomics(n_lines: int = 16384) -> float:
    """FP32 atomic lane operations per second of the voxel update's write pattern
    alone (xt_probe_footprint_atomics): the ceiling of the line kernel's e update."""
    r = C.c_double()
    _check(lib().xt_probe_footprint_atomics(int(n_lines), C.byref(r)))
    return r.value


def edge_detect(eng: Engine) -> np.ndarray:
    """The edge-only update mask [N_time][N_y][N_x][N_z] (uint8) of the engine's
    current object (xt_edge_detect)."""
    O = eng.O
    mask = np.zeros((O.N_time, O.N_y, O.N_x, O.N_z), np.uint8)
    mt = xt_abi.PtrTree(mask)
    _check(lib().xt_edge_detect(C.byref(O), C.byref(eng.T), mt.ptr))
    return mask
