"""Multi-resolution ICD: the reference launcher's stage loop on the stepwise API.

The launcher (4D_Recon_Launcher_for_Real_Data/XT_Initialize.py:175-198,
XT_MBIR_Reconstruction.py:113-158) reconstructs the same sinogram in stages of
decreasing voxel size: stage k of n uses voxels 2^(n-1-k) times the final ones
(delta_xy scaled, N_x = N_y = N_r / delta_xy), the first from zero, every later
one from the previous stage's object upsampled in y and x (initICD 2,
Source_Code_4D_Fast/XT_ICD_update.c:484-512, here xt_upsample_object on the
GPU) and the previous stage's ring offsets (updateProjOffset 3; the launcher's
first stage estimates none, its second starts them).

Here the stages pass the object and offsets in memory rather than through the
run folder's files; `Icd.upload` places them in the next stage's state.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import synth, xt_hip


def z_factors(n_stages: int, z_stages: int = 1):
    """delta_z of each stage relative to the final one, the launcher's policy
    (4D_Recon_Launcher_for_Real_Data/XT_Initialize.py:175-179): the z levels
    2^(z_stages-1) .. 1 on the last z_stages stages, the earlier stages at the
    coarsest z."""
    if not 1 <= z_stages <= n_stages:
        raise ValueError(f"{z_stages} z levels for {n_stages} stages")
    zs = [2 ** (z_stages - 1 - k) for k in range(z_stages)]
    return [zs[0]] * (n_stages - z_stages) + zs


def init_icd_of(n_stages: int, z_stages: int = 1):
    """initICD of each stage (XT_Initialize.py:185-188): 0 for the first, 3 where
    the voxels halve in z as well (object upsampled in z, y and x), else 2."""
    zf = z_factors(n_stages, z_stages)
    return [0] + [3 if zf[k] < zf[k - 1] else 2 for k in range(1, n_stages)]


def stage_configs(cfg: synth.ScanConfig, n_stages: int, first_offsets: bool = False,
                  z_stages: int = 1):
    """The n stage configurations, coarsest first (the last is cfg itself): stage
    k's voxels are 2^(n-1-k) times the final ones in x and y and z_factors(...)[k]
    times in z (a slice then spans that many detector rows, z_overlap_num); the
    first stage estimates no ring offsets unless first_offsets (the launcher's
    choice is not to)."""
    out = []
    zf = z_factors(n_stages, z_stages)
    for k in range(n_stages):
        f = 2 ** (n_stages - 1 - k)
        if cfg.N_xy % f or cfg.N_z % zf[k]:
            raise ValueError(f"N_xy={cfg.N_xy}, N_z={cfg.N_z} not divisible by {f}, {zf[k]}")
        c = dataclasses.replace(cfg, N_xy=cfg.N_xy // f, mult_xy=cfg.mult_xy * f,
                                N_z=cfg.N_z // zf[k], mult_z=cfg.mult_z * zf[k],
                                update_offset=0 if (k == 0 and not first_offsets)
                                else cfg.update_offset)
        if c.N_r != cfg.N_r or c.N_t != cfg.N_t:
            raise ValueError("stage sizes must keep the detector")
        out.append(c)
    return out


def coarse_stages(path: str, cfg: synth.ScanConfig, iters, log=None, fbp_first: bool = False,
                  quick: bool = True, z_stages: int = 1, sino_from=None, icd_from=None):
    """Runs the stages before the last one (len(iters) of them, iters[k] ICD
    iterations each) on the run folder `path`; returns the last coarse stage's
    object [N_z][N_y/2][N_x/2] and ring offsets [N_r][N_t].  fbp_first: the
    coarsest stage starts from its filtered backprojection (with the ring-offset
    guess, and offsets estimated from the first stage on) instead of zero.
    quick: the coarse states skip the create-time autotuner and the per-iteration
    FP64 recompute of e (as the reference's float build, which never recomputes
    it): a few iterations of a starting guess gain nothing from either (on the
    512^2 x 32 slab: 0.66 s instead of 0.82 s for 5 + 3 iterations, the same
    start); and unless XT_OPT_LINES is set they update 3/128 of their lines at
    once instead of the default 1/64 (384 instead of 256 lines at 128^2: 0.87 ->
    0.83 s for the slab's 8 + 2 FBP-started stages, the same start within
    run-to-run spread).  z_stages: the number of z
    levels of the launcher's policy (the returned object then has N_z / 2 slices
    when the final stage refines z: start_from_coarse with init_icd 3).
    sino_from: an open Engine of the same run folder (the final stage's) whose
    projection and weights the stages copy instead of reading the files again
    (every stage reconstructs the same sinogram).  icd_from: an open Icd of the
    same sinogram (the final stage's) whose y and w the stages' states copy on the
    device; the stages then neither read nor copy the sinogram on the host."""
    n = len(iters) + 1
    stages = stage_configs(cfg, n, first_offsets=fbp_first, z_stages=z_stages)[:-1]
    inits = init_icd_of(n, z_stages)
    vol = off = None
    saved = {o: xt_hip.get_option(o)
             for o in (xt_hip.OPT_AUTOTUNE, xt_hip.OPT_RESYNC_EVERY, xt_hip.OPT_LINES)}
    try:
        if quick:
            xt_hip.set_option(xt_hip.OPT_AUTOTUNE, 0)
            xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 0)
        for k, (c, n) in enumerate(zip(stages, iters)):
            if quick and saved[xt_hip.OPT_LINES] == 0:
                xt_hip.set_option(xt_hip.OPT_LINES, min(1024, max(16, 3 * c.N_xy * c.N_xy // 128)))
            with xt_hip.Engine(path, c.cli_args(iterations=n),
                               read_sinogram=sino_from is None and icd_from is None) as eng:
                if sino_from is not None and icd_from is None:
                    eng.sino("Projection")[:] = sino_from.sino("Projection")
                    eng.sino("Weight")[:] = sino_from.sino("Weight")
                if vol is not None:
                    eng.set_volume(xt_hip.upsample_object(inits[k], vol)[None])
                    eng.offset()[:] = off
                with xt_hip.Icd(eng, sino_from=icd_from) as icd:
                    if vol is None and fbp_first:
                        icd.fbp()
                    icd.iterate(n)
                    icd.download()
                vol = eng.volume()[0].copy()
                off = eng.offset().copy()
            if log:
                log(f"multi-resolution stage {k + 1}: {c.N_xy}^2 x {c.N_z}, {n} iterations")
    finally:
        for o, v in saved.items():
            xt_hip.set_option(o, v)
    return vol, off


def start_from_coarse(icd: xt_hip.Icd, vol: np.ndarray, off: np.ndarray,
                      init_icd: int = 2) -> None:
    """The final stage's start: the coarse object upsampled (initICD 2, or 3 when
    the final stage also halves the voxels in z) and the coarse stage's offsets,
    uploaded into `icd` (e recomputed)."""
    eng = icd.eng
    eng.set_volume(xt_hip.upsample_object(init_icd, vol)[None])
    eng.offset()[:] = off
    icd.upload()
