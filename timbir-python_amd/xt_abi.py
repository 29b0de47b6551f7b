"""ctypes mirror of the engine-API structs and pointer-tree arrays.

The C layout is the one in include/xt_structures.h, which restates
Source_Code_4D_Fast/XT_Structures.h:61-203 (the reference's engine API).
Arrays are pointer trees over one contiguous block, the layout
Source_Code_4D_Fast/allocate.c:91-171 builds, so ``tree.data`` is a plain
row-major numpy array and ``tree.ptr`` is what a ``Real_t ***`` field holds.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

__all__ = ["abi_types", "PtrTree", "np_real"]


def np_real(real_is_double: bool):
    return np.float64 if real_is_double else np.float32


def abi_types(real_is_double: bool = False):
    """Return (Sinogram, ScannedObject, AMatrixCol, TomoInputs) ctypes classes."""
    R = C.c_double if real_is_double else C.c_float
    P = C.c_void_p  # every pointer / pointer-tree field

    class Sinogram(C.Structure):
        _fields_ = [
            ("Projection", P), ("ProjOffset", P), ("ProjSelect", P),
            ("x_start", C.c_int32), ("x_width", C.c_int32), ("N_r", C.c_int32),
            ("N_t", C.c_int32), ("N_p", C.c_int32), ("total_t_slices", C.c_int32),
            ("delta_r", R), ("delta_t", R), ("R0", R), ("RMax", R), ("T0", R), ("TMax", R),
            ("cosine", P), ("sine", P), ("Length_R", R), ("Length_T", R),
            ("OffsetR", R), ("OffsetT", R), ("ViewPtr", P), ("TimePtr", P),
            ("slice_begin", C.c_int32), ("slice_num", C.c_int32), ("z_overlap_num", C.c_int32),
            ("Lap_Kernel", (R * 3) * 3),
        ]

    class ScannedObject(C.Structure):
        _fields_ = [
            ("Object", P), ("Conv_Object", P), ("Length_X", R), ("Length_Y", R), ("Length_Z", R),
            ("N_x", C.c_int32), ("N_z", C.c_int32), ("N_y", C.c_int32), ("N_time", C.c_int32),
            ("x0", R), ("z0", R), ("y0", R), ("delta_xy", R), ("delta_z", R),
            ("mult_xy", R), ("mult_z", R), ("BeamWidth", R), ("Sigma_S", R), ("Sigma_T", R),
            ("MRF_P", R), ("C_S", R), ("C_T", R), ("ProjIdxPtr", P), ("ProjNum", P),
            ("NHICD_Iterations", C.c_int32), ("Rtime0", R), ("delta_Rtime", R),
        ]

    class AMatrixCol(C.Structure):
        _fields_ = [("values", P), ("count", C.c_uint8), ("index", P)]

    class TomoInputs(C.Structure):
        _fields_ = [
            ("NumIter", C.c_int32), ("StopThreshold", R), ("RotCenter", R),
            ("phantom_N_xy", C.c_int32), ("phantom_N_z", C.c_int32), ("radius_obj", R),
            ("Sigma_S_Q", R), ("Sigma_T_Q", R), ("Sigma_S_Q_P", R), ("Sigma_T_Q_P", R),
            ("Spatial_Filter", ((R * 3) * 3) * 3), ("Time_Filter", R * 1),
            ("Weight", P), ("var_est", R), ("alpha", R), ("cost_thresh", R),
            ("time_reg", C.c_uint8), ("sinobin", C.c_uint8), ("initICD", C.c_uint8),
            ("Write2Tiff", C.c_uint8), ("No_Projection_Noise", C.c_uint8),
            ("reconstruct", C.c_uint8), ("num_threads", C.c_int32),
            ("updateProjOffset", C.c_uint8), ("EnforceZeroMeanOffset", C.c_uint8),
            ("no_NHICD", C.c_uint8), ("WritePerIter", C.c_uint8), ("num_z_blocks", C.c_int32),
            ("x_NHICD_select", P), ("y_NHICD_select", P), ("x_rand_select", P),
            ("y_rand_select", P), ("UpdateSelectNum", P), ("NHICDSelectNum", P),
            ("ErrorSinoThresh", R), ("ErrorSinoDelta", R), ("only_Edge_Updates", C.c_uint8),
            ("node_num", C.c_int32), ("node_rank", C.c_int32), ("updateVar", C.c_uint8),
            ("initMagUpMap", C.c_uint8), ("readSino4mHDF", C.c_uint8),
            ("debug_file_ptr", P), ("ErrorSino_Cost", R), ("Forward_Cost", R),
            ("Prior_Cost", R), ("RMSE_converged", C.c_uint8),
        ]

    return Sinogram, ScannedObject, AMatrixCol, TomoInputs


class PtrTree:
    """Pointer tree over a C-contiguous numpy array (numpy keeps ownership).

    For ndim == 1 the pointer is the data pointer itself.  C code must never
    free a PtrTree; use a library allocator for arrays the callee frees.
    """

    def __init__(self, data: np.ndarray):
        if not data.flags["C_CONTIGUOUS"]:
            raise ValueError("PtrTree needs a C-contiguous array")
        self.data = data
        self._levels = []
        shape = data.shape
        base = data.ctypes.data
        if data.ndim == 1:
            self.ptr = C.c_void_p(base)
            return
        # innermost pointer level points into the data block
        n_rows = int(np.prod(shape[:-1]))
        row_bytes = shape[-1] * data.itemsize
        inner = (C.c_void_p * n_rows)(*[base + i * row_bytes for i in range(n_rows)])
        self._levels.append(inner)
        child = inner
        for lvl in range(data.ndim - 2, 0, -1):
            n = int(np.prod(shape[:lvl]))
            stride = shape[lvl] * C.sizeof(C.c_void_p)
            addr = C.addressof(child)
            arr = (C.c_void_p * n)(*[addr + i * stride for i in range(n)])
            self._levels.append(arr)
            child = arr
        self.ptr = C.c_void_p(C.addressof(child))


class ObjectTree:
    """``Real_t ****Object``: [N_time] pointers to [N_z+2][N_y][N_x] trees over one array."""

    def __init__(self, data: np.ndarray):
        self.data = data
        self._frames = [PtrTree(data[i]) for i in range(data.shape[0])]
        self._top = (C.c_void_p * data.shape[0])(*[f.ptr.value for f in self._frames])
        self.ptr = C.c_void_p(C.addressof(self._top))


class RaggedInt32:
    """``int32_t **`` over ragged rows (e.g. ProjIdxPtr[N_time][ProjNum[i]])."""

    def __init__(self, rows):
        self.rows = [np.ascontiguousarray(r, dtype=np.int32) if len(r) else np.zeros(1, np.int32)
                     for r in rows]
        self._top = (C.c_void_p * len(self.rows))(*[r.ctypes.data for r in self.rows])
        self.ptr = C.c_void_p(C.addressof(self._top))
