"""Real-data ingest: counts, white and dark frames to the engine's sinogram files.

Mirrors gen_projection_4m_HDF (Source_Code_4D_Fast/XT_HDFIO.c:14-240) and the
launcher's generate_projections (4D_Recon_Launcher_for_Real_Data/
XT_Projections.py:28-123): the region of interest is the white and dark frames
without the first and the last (:99-121), the slices [slice_begin,
slice_begin + slice_num) and the detector columns [x_start + extras/2,
x_start + extras/2 + true_length) with extras = x_width % N_r; the
arithmetic (flat means, per-bin sums, log ratio, beam-hardening quadratic)
runs on the GPU (xt_counts_to_sinogram).

`roi` works on anything indexable like a [frames][slices][columns] array: numpy
arrays, or the /exchange/data, /exchange/data_white and /exchange/data_dark
datasets of an HDF5 file when h5py is installed (read_hdf5).
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from . import xt_hip


def roi(data, white, dark, N_p: int, proj_start: int, slice_begin: int, slice_num: int,
        x_start: int, x_width: int, N_r: int):
    """The region of interest of the three datasets as uint16 arrays, and the
    true detector length used (a multiple of N_r)."""
    extras = x_width % N_r
    true_len = x_width - extras
    c0 = x_start + extras // 2
    cols = slice(c0, c0 + true_len)
    rows = slice(slice_begin, slice_begin + slice_num)
    counts = np.asarray(data[proj_start:proj_start + N_p, rows, cols], np.uint16)
    w = np.asarray(white[1:white.shape[0] - 1, rows, cols], np.uint16)
    d = np.asarray(dark[1:dark.shape[0] - 1, rows, cols], np.uint16)
    if counts.shape[0] != N_p:
        raise ValueError(f"{counts.shape[0]} projections from {proj_start}, {N_p} requested")
    if w.shape[0] < 1 or d.shape[0] < 1:
        raise ValueError("white and dark datasets need at least 3 frames (first and last dropped)")
    return counts, w, d, true_len


def write_run_folder(path: str, counts, white, dark, N_t: int, N_r: int, bh: float = 0.0,
                     offset: Optional[np.ndarray] = None):
    """Projection and weight [N_p][N_r][N_t] into projection_n0.bin and
    weight_n0.bin (float32, the engine's Real_t), and the bright field into
    bright_n0.bin; returns (projection, weight, bright)."""
    os.makedirs(path, exist_ok=True)
    proj, wt, bright = xt_hip.counts_to_sinogram(counts, white, dark, N_t, N_r, bh)
    proj.tofile(os.path.join(path, "projection_n0.bin"))
    wt.tofile(os.path.join(path, "weight_n0.bin"))
    bright.tofile(os.path.join(path, "bright_n0.bin"))
    if offset is not None:
        np.asarray(offset, np.float32).reshape(N_r, N_t).tofile(
            os.path.join(path, "proj_offset_n0.bin"))
    return proj, wt, bright


def read_hdf5(data_file: str, white_file: str, dark_file: str, **roi_args):
    """roi() on the /exchange datasets of the reference's HDF5 files.  Needs
    h5py; without it this raises (no other HDF5 reader is bundled)."""
    try:
        import h5py
    except ImportError as e:
        raise RuntimeError("reading HDF5 needs h5py, which is not installed; pass arrays to "
                           "roi() instead") from e
    with h5py.File(data_file, "r") as fd, h5py.File(white_file, "r") as fw, \
            h5py.File(dark_file, "r") as fk:
        return roi(fd["/exchange/data"], fw["/exchange/data_white"], fk["/exchange/data_dark"],
                   **roi_args)
