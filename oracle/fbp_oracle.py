"""CPU restatement of the engine's filtered backprojection (a warm start for
ICD that the reference does not have), in float64 numpy.  Test infrastructure:
the GPU's xt_icd_fbp is checked against it, and it is pinned by the closed-form
Radon transform of a disk (tests/test_oracle_fbp.py).

Steps (the same as the GPU kernels k_offset_guess_*, k_fbp_filter and
k_fbp_backproject); when the ring offsets are estimated, d is first replaced by
offset_guess(y):
  p = y - d per (view, detector row); 3-point median along the bins (isolated
  zingers removed; the end bins kept); Ram-Lak filter of sample spacing delta_r,
  q_j = (p_j / 4 - sum_{n odd} (p_{j-n} + p_{j+n}) / (n pi)^2) / delta_r;
  x(voxel) = sum over the frame's views of w_v q_v(u), u = (r - R0) / delta_r - 1/2
  with r = x_c cos - y_c sin the voxel centre's projection (the A-matrix geometry,
  Source_Code_4D_Fast/XT_AMatrix.c:59-62), linear interpolation (zero outside the
  detector), w_v the view's share of [0, pi) among the frame's angles mod pi.
"""
import numpy as np


def median3(p):
    m = p.copy()
    a, b, c = p[..., :-2], p[..., 1:-1], p[..., 2:]
    m[..., 1:-1] = np.maximum(np.minimum(a, b), np.minimum(np.maximum(a, b), c))
    return m


def ramp(p, delta_r):
    """Ram-Lak filter along the last axis (bins)."""
    n_r = p.shape[-1]
    q = 0.25 * p.astype(np.float64)
    for n in range(1, n_r, 2):
        c = 1.0 / (n * n * np.pi * np.pi)
        q[..., n:] -= c * p[..., :-n]
        q[..., :-n] -= c * p[..., n:]
    return q / delta_r


def angle_weights(angles):
    """Each view's share of [0, pi) among the given angles taken mod pi."""
    t = np.mod(np.asarray(angles, np.float64), np.pi)
    order = np.argsort(t, kind="stable")
    ts = t[order]
    n = len(ts)
    if n == 1:
        return np.array([np.pi])
    lo = np.concatenate([[ts[-1] - np.pi], ts[:-1]])
    hi = np.concatenate([ts[1:], [ts[0] + np.pi]])
    w = np.empty(n)
    w[order] = 0.5 * (hi - lo)
    return w


def backproject(q, cosv, sinv, view_ids, weights, geom):
    """q [N_p][N_r][N_t]; returns [N_t][N_y][N_x] for the listed views."""
    n_x, n_y, x0, y0, dxy, r0, dr = (geom[k] for k in ("N_x", "N_y", "x0", "y0", "delta_xy",
                                                       "R0", "delta_r"))
    n_r, n_t = q.shape[1], q.shape[2]
    xc = (np.float32(x0) + (np.arange(n_x) + 0.5) * np.float64(dxy)).astype(np.float32)
    yc = (np.float32(y0) + (np.arange(n_y) + 0.5) * np.float64(dxy)).astype(np.float32)
    X, Y = np.meshgrid(xc.astype(np.float64), yc.astype(np.float64))
    out = np.zeros((n_t, n_y, n_x))
    for v, w in zip(view_ids, weights):
        r = X * float(cosv[v]) - Y * float(sinv[v])
        u = (r - r0) / dr - 0.5
        i0 = np.floor(u).astype(np.int64)
        fr = u - i0
        qa = np.zeros((n_r + 2, n_t))
        qa[1:-1] = q[v]
        a = qa[np.clip(i0 + 1, 0, n_r + 1)]
        b = qa[np.clip(i0 + 2, 0, n_r + 1)]
        val = a + fr[..., None] * (b - a)
        val[(i0 < -1) | (i0 >= n_r)] = 0.0
        out += w * np.moveaxis(val, -1, 0)
    return out


def offset_guess(y, zero_mean=True):
    """Ring offsets guessed from y [N_p][N_r][N_t]: per detector pixel the mean
    over views of the values within 4 standard deviations of the plain mean
    (zingers excluded), minus its 15-bin running mean along the bins (the end
    bins repeated), optionally shifted to zero mean."""
    y = np.asarray(y, np.float64)
    mu, sd = y.mean(axis=0), y.std(axis=0)
    keep = np.abs(y - mu[None]) <= 4 * sd[None]
    m = (y * keep).sum(axis=0) / np.maximum(keep.sum(axis=0), 1)
    n_r = m.shape[0]
    mp = np.concatenate([np.repeat(m[:1], 7, 0), m, np.repeat(m[-1:], 7, 0)])
    trend = sum(mp[k:k + n_r] for k in range(15)) / 15.0
    d = m - trend
    return d - d.mean() if zero_mean else d


def fbp(y, d, angles, cosv, sinv, frames, geom):
    """y [N_p][N_r][N_t], d [N_r][N_t]; frames: list of view-index lists.
    Returns [N_time][N_t][N_y][N_x]."""
    p = np.moveaxis(np.asarray(y, np.float64) - np.asarray(d, np.float64)[None], 2, 1)
    q = np.moveaxis(ramp(median3(p), geom["delta_r"]), 1, 2)
    return np.stack([backproject(q, cosv, sinv, f, angle_weights(np.asarray(angles)[f]), geom)
                     for f in frames])
