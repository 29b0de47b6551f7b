"""The allocation-free ring-offset update (xt_offset_update_noalloc) against the
allocating entry point (xt_offset_update) on the same values: bit-identical
error sinogram and offsets, on two calls in a row (the second on a smaller
sinogram, so the scratch array holds the first call's values), plus the size
limit of its scratch array.  The device arrays are allocated through the HIP
runtime libxt_hip.so itself links (hipMalloc by ctypes), outside the library."""
import ctypes as C
import dataclasses
import os

import numpy as np
import pytest

from conftest import TINY, make_scan
from timbir_python_amd import xt_hip

pytestmark = pytest.mark.gpu

SMALLER = dataclasses.replace(TINY, name="tiny_half", N_xy=16, N_z=2, N_p=24, seed=9)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime already loaded with libxt_hip.so."""

    def __init__(self):
        xt_hip.lib()
        self.L = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_LAZY)
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, ptr, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _case(cfg, scan_root, seed):
    path, info = make_scan(cfg, scan_root, "noalloc")
    print(f"[gpu] offset_update_noalloc: {cfg.N_p} views x {cfg.N_r} bins x {cfg.N_t} rows",
          flush=True)
    with xt_hip.Engine(path, cfg.cli_args()) as eng:
        S, T = eng.S, eng.T
        T.EnforceZeroMeanOffset = 1
        rng = np.random.default_rng(seed)
        w = eng.sino("Weight")
        e = (rng.normal(0.0, 0.02, w.shape)).astype(np.float32)
        # entries beyond the zinger threshold (|e| sqrt(w) >= T) exercise the
        # generalized-Huber branch
        big = rng.random(w.shape) < 0.01
        e[big] = 50.0 * T.ErrorSinoThresh / np.sqrt(np.maximum(w[big], 1.0))
        d0 = rng.normal(0.0, 0.01, (S.N_r, S.N_t)).astype(np.float32)
        eng.offset()[:] = d0
        # device copies first, then the allocating form on the host arrays
        hip = _Hip()
        try:
            e_p, w_p, d_p = hip.upload(e), hip.upload(w), hip.upload(d0)
            e_ref = e.copy()
            xt_hip.offset_update(eng, e_ref)
            d_ref = eng.offset().copy()
            xt_hip.offset_update_noalloc(S.N_p, S.N_r, S.N_t, T.ErrorSinoThresh,
                                         T.ErrorSinoDelta, 1, e_p, w_p, d_p)
            e_got, d_got = hip.download(e_p, e), hip.download(d_p, d0)
        finally:
            hip.free()
        print(f"  max |d - d_ref| {np.abs(d_got - d_ref).max():.3e}, |e - e_ref| "
              f"{np.abs(e_got - e_ref).max():.3e}, d moved {np.abs(d_ref - d0).max():.3e}",
              flush=True)
        assert np.abs(d_ref - d0).max() > 0
        assert np.array_equal(d_got, d_ref)
        assert np.array_equal(e_got, e_ref)


def test_offset_update_noalloc_matches_allocating_form(scan_root):
    _case(TINY, scan_root, 1)
    _case(SMALLER, scan_root, 2)  # smaller sinogram: the scratch holds the first call's values


def test_offset_update_noalloc_rejects_oversize():
    # 2048 x 128 detector pixels: one view chunk, 4 * 262144 + ... doubles > 8 MiB
    hip = _Hip()
    try:
        e = hip.upload(np.zeros(2 * 2048 * 128, np.float32))
        d = hip.upload(np.zeros(2048 * 128, np.float32))
        with pytest.raises(xt_hip.XtError, match="allocation-free form holds"):
            xt_hip.offset_update_noalloc(2, 2048, 128, 4.0, 0.1, 1, e, e, d)
    finally:
        hip.free()
