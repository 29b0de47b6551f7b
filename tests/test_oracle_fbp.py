"""The FBP restatement in oracle/fbp_oracle.py, pinned by closed forms: the Radon
transform of a uniform disk of attenuation mu and radius R is
p(r) = 2 mu sqrt(R^2 - r^2), so the reconstruction must return mu inside the
disk and ~0 well outside; the angular weights of uniform views are pi/N."""
import numpy as np

import fbp_oracle


def disk_case(n_p=360, n_r=256, mu=0.7, R=60.0):
    angles = np.arange(n_p) * np.pi / n_p
    r = -n_r / 2 + np.arange(n_r) + 0.5
    p = np.where(np.abs(r) < R, 2 * mu * np.sqrt(np.maximum(R * R - r * r, 0)), 0.0)
    y = np.repeat(p[None, :, None], n_p, axis=0)
    geom = dict(N_x=128, N_y=128, x0=-96.0, y0=-96.0, delta_xy=1.5, R0=-n_r / 2, delta_r=1.0)
    return angles, y, geom


def test_uniform_weights():
    w = fbp_oracle.angle_weights(np.arange(10) * np.pi / 10)
    assert np.allclose(w, np.pi / 10)
    w = fbp_oracle.angle_weights(np.arange(10) * np.pi / 10 + 5 * np.pi)  # mod pi
    assert np.allclose(w, np.pi / 10) and abs(w.sum() - np.pi) < 1e-12


def test_disk_closed_form():
    mu, R = 0.7, 60.0
    angles, y, geom = disk_case(mu=mu, R=R)
    x = fbp_oracle.fbp(y, np.zeros((y.shape[1], 1)), angles, np.cos(angles), np.sin(angles),
                       [list(range(len(angles)))], geom)[0, 0]
    c = geom["x0"] + (np.arange(128) + 0.5) * geom["delta_xy"]
    X, Y = np.meshgrid(c, c)
    rad = np.hypot(X, Y)
    inside, outside = x[rad < 0.8 * R], x[rad > 1.2 * R]
    assert abs(inside.mean() / mu - 1) < 0.01
    assert inside.std() / mu < 0.02
    assert np.abs(outside).mean() < 0.01 * mu


def test_offset_guess_recovers_ring_offsets():
    """A broad object (Gaussian projections, sigma 40 bins) plus per-bin offsets
    and 0.5% zingers at 5x the maximum: the guess returns the offsets (up to the
    zero-mean shift)."""
    n_p, n_r = 180, 256
    rng = np.random.default_rng(3)
    r = np.arange(n_r) - n_r / 2 + 0.5
    p = 2.0 * np.exp(-r * r / (2 * 40.0 ** 2)) * (1 + 0.1 * np.cos(np.arange(n_p)[:, None] * 0.05))
    d = rng.normal(0, 0.05, n_r)
    d -= d.mean()
    y = (p + d[None])[:, :, None].copy()
    z = rng.choice(y.size, y.size // 200, replace=False)
    y.reshape(-1)[z] = 5 * y.max()
    g = fbp_oracle.offset_guess(y, zero_mean=True)[:, 0]
    assert np.corrcoef(g, d)[0, 1] > 0.95
    assert np.sqrt(np.mean((g - d) ** 2)) < 0.35 * d.std()


def test_median_removes_isolated_spikes():
    p = np.ones((1, 9))
    p[0, 4] = 50.0
    assert np.all(fbp_oracle.median3(p)[0, 1:-1] == 1.0)
