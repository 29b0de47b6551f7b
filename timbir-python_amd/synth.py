"""Seeded synthetic scans in the reference's input formats.

Writes, into one directory, what the reference engine reads with ``--sinobin 1``:
``projection_n0.bin`` and ``weight_n0.bin`` (Real_t [N_p][N_r][N_t], raw, no
header; Source_Code_4D_Fast/XT_genSinogram.c:264-300), ``view_info.txt``
(``"<time> - <angle in degrees>,"`` per view; the writer in
4D_Recon_Launcher_for_Real_Data/XT_MBIR_Reconstruction.py:7-13) and, for
reference, the phantom and the true ring offsets.

Geometry convention for every workload here: detector pixel pitch 1 along r and
t (``length_r = N_r``, ``length_t = N_t``), voxel pitch ``mult_xy`` detector
pixels in x-y and ``mult_z`` (1 unless a coarse multi-resolution stage says
otherwise) in z, rotation centre at ``N_r/2``, and an update radius
covering the whole square, so the A-matrix scaling is the same for every
variant of the reference (the ``delta_t`` factor of
Source_Code_4D_Fast/XT_VoxUpdate.c:27 is 1).
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Callable, List, Optional

import numpy as np

# HU conversion of the engine CLI (Source_Code_4D_Fast/XT_Init.c:510-511,543)
_SLOPE_HU = np.float32((1000 - 0) / (0.521225397034623 * 1.0 - 0.496372335005353 * 0.001205)
                       / 0.0001)
_C_HU = np.float32(-_SLOPE_HU * (0.496372335005353 * 0.001205 * 0.0001))


def sigma_to_hu(sigma: float) -> str:
    """CLI string whose HU->attenuation conversion reproduces ``sigma``."""
    return repr(float(sigma) * float(_SLOPE_HU) + float(_C_HU))


@dataclasses.dataclass
class ScanConfig:
    name: str
    N_xy: int
    N_z: int
    N_p: int
    mult_xy: float = 1.5
    mult_z: int = 1               # detector rows per slice (delta_z; z_overlap_num)
    N_time: int = 1
    I0: float = 2e4
    zinger_frac: float = 0.005
    zinger_scale: float = 5.0
    offset_sigma: float = 0.01
    n_ellipsoids: int = 12
    mu_lo: float = 0.5e-3
    mu_hi: float = 2.0e-3
    p: float = 1.2
    c_s: float = 0.01
    sigma_s: float = 1e-4
    c_t: float = 0.01
    sigma_t: float = 2e-4
    iterations: int = 30
    alpha: float = 1.5
    zinger_T: float = 4.0
    zinger_delta: float = 0.1
    update_offset: int = 2
    zero_mean_offset: bool = True
    do_var_est: bool = False
    no_nhicd: bool = False
    num_threads: int = 4
    voxel_thresh: float = 0.0
    interlaced_K: int = 0          # >0: bit-reversed interlaced views (4D workload)
    interlaced_N_theta: int = 0
    motion_per_frame: float = 0.0  # ellipsoid translation, voxels per frame
    seed: int = 1

    @property
    def N_r(self) -> int:
        return int(round(self.N_xy * self.mult_xy))

    @property
    def N_t(self) -> int:
        return self.N_z * self.mult_z

    def views(self):
        """(angles_deg, times) for every view, in acquisition order."""
        if self.interlaced_K:
            return interlaced_views(self.interlaced_K, self.interlaced_N_theta, self.N_p)
        ang = np.arange(self.N_p) * (180.0 / self.N_p)
        return ang, np.arange(self.N_p, dtype=np.float64)

    def rtime(self):
        """(Rtime0, Rtime_delta) as the launcher derives them
        (4D_Recon_Launcher_for_Real_Data/XT_Initialize.py:262-266); 3D uses delta 1."""
        if self.N_time == 1:
            return 0.0, 1.0
        _, t = self.views()
        return float(t[0]), float((t[-1] - t[0]) / self.N_time)

    def cli_args(self, iterations: Optional[int] = None, init_icd: int = 0) -> List[str]:
        r0, rd = self.rtime()
        a = ["--p", repr(self.p), "--sigma_s", sigma_to_hu(self.sigma_s), "--c_s", repr(self.c_s),
             "--sigma_t", sigma_to_hu(self.sigma_t), "--c_t", repr(self.c_t),
             "--delta_xy", repr(self.mult_xy), "--delta_z", str(self.mult_z),
             "--length_r", str(self.N_r), "--length_t", str(self.N_t),
             "--voxel_thresh", repr(self.voxel_thresh), "--cost_thresh", "0",
             "--iter", str(self.iterations if iterations is None else iterations),
             "--rotation_center", repr(self.N_r / 2.0), "--alpha", repr(self.alpha),
             "--sinobin", "1", "--initICD", str(init_icd), "--Rtime0", repr(r0),
             "--Rtime_delta", repr(rd), "--Rtime_num", str(self.N_time),
             "--num_projections", str(self.N_p), "--N_r", str(self.N_r), "--N_t", str(self.N_t),
             "--num_threads", str(self.num_threads), "--radius_obj", str(self.N_r),
             "--updateProjOffset", str(self.update_offset), "--writeTiff", "0",
             "--zingerT", repr(self.zinger_T), "--zingerDel", repr(self.zinger_delta),
             "--Est_of_Var", "1", "--time_reg"]
        if self.zero_mean_offset:
            a.append("--EnforceZeroMeanOffset")
        if self.do_var_est:
            a.append("--do_VarEst")
        if self.no_nhicd:
            a.append("--no_NHICD")
        return a


def phantom(cfg: ScanConfig) -> np.ndarray:
    """Ellipsoid phantom, float32 [N_time][N_z][N_y][N_x]; attenuation per voxel
    uniform in [mu_lo, mu_hi] on a zero background, later ellipsoids painted over
    earlier ones, each translating ``motion_per_frame`` voxels per frame."""
    rng = np.random.default_rng(cfg.seed)
    n, nz = cfg.N_xy, cfg.N_z
    out = np.zeros((cfg.N_time, nz, n, n), np.float32)
    yy, xx = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5, indexing="ij")
    zz = np.arange(nz) + 0.5
    ells = []
    for _ in range(cfg.n_ellipsoids):
        rad = 0.32 * n * math.sqrt(rng.uniform())
        ang = rng.uniform(0, 2 * math.pi)
        cx, cy = n / 2 + rad * math.cos(ang), n / 2 + rad * math.sin(ang)
        cz = rng.uniform(0.3, 0.7) * nz
        a, b = rng.uniform(0.05, 0.2) * n, rng.uniform(0.05, 0.2) * n
        c = rng.uniform(0.3, 0.7) * nz
        phi = rng.uniform(0, math.pi)
        mu = rng.uniform(cfg.mu_lo, cfg.mu_hi)
        direc = rng.uniform(0, 2 * math.pi)
        ells.append((cx, cy, cz, a, b, c, phi, mu, direc))
    for f in range(cfg.N_time):
        for (cx, cy, cz, a, b, c, phi, mu, direc) in ells:
            sx = cx + cfg.motion_per_frame * f * math.cos(direc)
            sy = cy + cfg.motion_per_frame * f * math.sin(direc)
            dx, dy = xx - sx, yy - sy
            u = (dx * math.cos(phi) + dy * math.sin(phi)) / a
            v = (-dx * math.sin(phi) + dy * math.cos(phi)) / b
            r2 = u * u + v * v
            for z in range(nz):
                w = ((zz[z] - cz) / c) ** 2
                m = (r2 + w) <= 1.0
                out[f, z][m] = mu
    return out


def interlaced_views(K: int, N_theta: int, N_p: int):
    """Bit-reversed interlaced acquisition order of 4D_Recon_Launcher_for_Real_Data/
    XT_Interlaced_Angles.py:107-139 (gen_interlaced_views_0_to_Inf): view i has
    index i*K + bitrev_log2(K)(floor(i / L) mod K), L = N_theta / K; returns
    (angles in degrees, index * 180 / N_theta, unwrapped; times, index / K)."""
    L = N_theta // K
    bits = int(round(math.log2(K)))
    i = np.arange(N_p)
    sub = (i // L) % K
    rev = np.zeros_like(sub)
    for b in range(bits):
        rev |= ((sub >> b) & 1) << (bits - 1 - b)
    idx = (i * K + rev).astype(np.float64)
    return idx * (180.0 / N_theta), idx / K


def clip_list_of_views(views, times, min_time_btw_views: float, rotation_speed: float):
    """Views the camera cannot take: 4D_Recon_Launcher_for_Real_Data/
    XT_Interlaced_Angles.py:6-27.  View i (i >= 1) is dropped when it is at most
    rotation_speed * min_time_btw_views degrees past view i - 1 of the *unclipped*
    list, else kept.  Returns (kept views, kept times, deleted views, deleted
    times).  As in the launcher, the first kept time is 0 whatever times[0] is
    (its output array is zero-initialised and entry 0 is never written)."""
    views = np.asarray(views, np.float64)
    times = np.asarray(times, np.float64)
    min_dtheta = rotation_speed * min_time_btw_views
    keep = np.ones(len(views), bool)
    keep[1:] = (views[1:] - views[:-1]) > min_dtheta
    kept_t = times[keep].copy()
    if len(kept_t):
        kept_t[0] = 0.0
    return views[keep], kept_t, views[~keep], times[~keep]


def launcher_views(K: int, N_theta: int, num_cycles: int, min_time_btw_views: float,
                   rotation_speed: float):
    """The launcher's view list (4D_Recon_Launcher_for_Real_Data/XT_Initialize.py:
    73-78): N_theta * num_cycles interlaced views, then clip_list_of_views."""
    ang, t = interlaced_views(K, N_theta, N_theta * num_cycles)
    return clip_list_of_views(ang, t, min_time_btw_views, rotation_speed)


def write_view_info(path: str, angles_deg, times) -> None:
    with open(os.path.join(path, "view_info.txt"), "w") as fh:
        for t, a in zip(times, angles_deg):
            fh.write(f"{float(t)!r} - {float(a)!r},\n")


def measure(cfg: ScanConfig, ax: np.ndarray):
    """Poisson transmission counts for noiseless line integrals ``ax`` (float64
    [N_p][N_r][N_t]).  Returns (y, w, d): y = ln(I0/counts) with zingers,
    w = counts, d the ring offsets that were added to every view."""
    rng = np.random.default_rng(cfg.seed + 7919)
    d = rng.normal(0.0, cfg.offset_sigma, size=(cfg.N_r, cfg.N_t))
    lam = cfg.I0 * np.exp(-(ax + d[None]))
    counts = np.maximum(rng.poisson(lam), 1).astype(np.float64)
    y = np.log(cfg.I0 / counts)
    nz = int(round(cfg.zinger_frac * y.size))
    if nz:
        idx = rng.choice(y.size, size=nz, replace=False)
        y.reshape(-1)[idx] = cfg.zinger_scale * y.max()
    return y, counts, d


def generate(cfg: ScanConfig, path: str, project: Callable[[np.ndarray], np.ndarray],
             real_dtype=np.float32) -> dict:
    """Write a full scan for ``cfg`` into ``path``.  ``project`` maps the phantom
    (float64 [N_time][N_z][N_y][N_x]) to Ax (float64 [N_p][N_r][N_t]) with the
    engine's own system matrix; it is called after view_info.txt exists."""
    os.makedirs(path, exist_ok=True)
    ang, t = cfg.views()
    write_view_info(path, ang, t)
    ph = phantom(cfg)
    ax = project(ph.astype(np.float64))
    y, w, d = measure(cfg, ax)
    y.astype(real_dtype).tofile(os.path.join(path, "projection_n0.bin"))
    w.astype(real_dtype).tofile(os.path.join(path, "weight_n0.bin"))
    np.zeros((cfg.N_r, cfg.N_t), real_dtype).tofile(os.path.join(path, "proj_offset_n0.bin"))
    return {"phantom": ph, "offset": d, "ax": ax}


# The workloads the benchmark and the tests are built on.
C1 = ScanConfig(name="plain3d_128x128x8", N_xy=128, N_z=8, N_p=256, I0=2e4, iterations=30)
C2 = ScanConfig(name="slab_512x512x32", N_xy=512, N_z=32, N_p=1024, I0=1e4, iterations=30,
                num_threads=32)
def batch_configs(n=24, seed=4):
    """The batched workload: n independent slabs with image sizes drawn from
    {128,192,256,384,512}^2 x {4,8,16} slices and view counts from
    {180,360,720,1024} over [0,180), bins 1.5x the width, Poisson I0=1e4 and 0.3%
    zingers (seeded draws)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        out.append(ScanConfig(name=f"batch{i}", N_xy=int(rng.choice([128, 192, 256, 384, 512])),
                              N_z=int(rng.choice([4, 8, 16])),
                              N_p=int(rng.choice([180, 360, 720, 1024])), I0=1e4,
                              zinger_frac=0.003, num_threads=16, seed=100 + i))
    return out


C3 = ScanConfig(name="interlaced4d_256x256x16x16", N_xy=256, N_z=16, N_p=2048, N_time=16,
                I0=1e4, interlaced_K=8, interlaced_N_theta=512, motion_per_frame=0.5,
                sigma_t=2e-4, num_threads=32)
