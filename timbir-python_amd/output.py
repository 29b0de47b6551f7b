"""Result writers of the reference launcher (4D_Recon_Launcher_for_Real_Data/
XT_IOMisc.py:28-51,66-105): the engine's object files as 8-bit TIFF slices on a
Hounsfield scale, or as one HDF5 file.

Host-side format conversion after the GPU path: it reads the run folder's
object_n0_time_<i>.bin ([N_z][N_y][N_x] float32) and proj_offset_n0.bin.
"""
from __future__ import annotations

import os

import numpy as np

# convert_um2HU's linear map (XT_IOMisc.py:28-43): the attenuation of air maps
# to 0 HU and that of water to 1000 HU, with the launcher's mass attenuation
# coefficients and densities in its unit scales
_ATT, _DEN = 100.0 ** 4, 0.001 ** 4
_AIR = 0.496372335005353 * _ATT * 0.001205 * _DEN
_WATER = 0.521225397034623 * _ATT * 1.0 * _DEN
HU_SLOPE = 1000.0 / (_WATER - _AIR)
HU_OFFSET = -HU_SLOPE * _AIR


def um_to_hu(obj):
    """Attenuation per unit length to Hounsfield units (convert_um2HU)."""
    return np.asarray(obj, np.float64) * HU_SLOPE + HU_OFFSET


def hu_to_uint8(hu, max_hu, min_hu):
    """Linear window [min_hu, max_hu] to 0..255, clipped (convert_HU2uint8)."""
    v = (255.0 / (max_hu - min_hu)) * (np.asarray(hu, np.float64) - min_hu)
    v[v <= 0] = 0
    v[v >= 255.0] = 255.0
    return v.astype(np.uint8)


def read_objects(run_folder, N_time, N_z, N_xy):
    """[N_time][N_z][N_xy][N_xy] from the run folder's object files."""
    out = np.zeros((N_time, N_z, N_xy, N_xy), np.float32)
    for j in range(N_time):
        out[j] = np.fromfile(os.path.join(run_folder, f"object_n0_time_{j}.bin"),
                             np.float32).reshape(N_z, N_xy, N_xy)
    return out


def write_tiff_slices(run_folder, out_folder, N_time, N_z, N_xy, max_hu, min_hu):
    """object_time_z_<t>_<z>.tif, 8-bit, per frame and slice
    (write_tiff_from_object_bin_file, XT_IOMisc.py:66-79)."""
    from PIL import Image
    os.makedirs(out_folder, exist_ok=True)
    obj = read_objects(run_folder, N_time, N_z, N_xy)
    names = []
    for j in range(N_time):
        for k in range(N_z):
            name = os.path.join(out_folder, f"object_time_z_{j}_{k}.tif")
            Image.fromarray(hu_to_uint8(um_to_hu(obj[j, k]), max_hu, min_hu)).save(name)
            names.append(name)
    return names


def write_object_hdf5(run_folder, out_file, N_time, N_z, N_xy, N_r, N_t):
    """object.hdf5 with datasets 'object' [N_time][N_z][N_xy][N_xy] (negative
    values set to 0) and 'proj_offset' [N_r][N_t] (write_object2HDF,
    XT_IOMisc.py:81-105).  Needs h5py."""
    try:
        import h5py
    except ImportError as e:
        raise RuntimeError("writing HDF5 needs h5py, which is not installed") from e
    obj = read_objects(run_folder, N_time, N_z, N_xy)
    obj[obj < 0] = 0
    off = np.fromfile(os.path.join(run_folder, "proj_offset_n0.bin"), np.float32).reshape(N_r, N_t)
    with h5py.File(out_file, "w") as f:
        f.create_dataset("object", data=obj, chunks=True)
        f.create_dataset("proj_offset", data=off, chunks=True)
