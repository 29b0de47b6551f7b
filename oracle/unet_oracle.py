"""CPU restatement of the warm-start 3D U-Net -- test infrastructure only.

Only tests (and the benchmark's CPU leg) may import this module; the product
path runs the network with its own HIP kernels.  The network has no
counterpart in the reference (TIMBIR starts ICD from zero or from a coarser
ICD, Source_Code_4D_Fast/XT_ICD_update.c:318-536); this file restates the
library's network so the GPU forward pass and every parameter gradient can be
compared with an independent computation.

Two modes:
  bf16=True   the GPU's rounding points are reproduced (inputs, weights and
              every stored activation or activation gradient rounded to
              bfloat16 with round-to-nearest-even; sums in float64)
  bf16=False  exact float64 arithmetic, used to pin this restatement by
              finite differences

Pins (tests/test_oracle_unet.py): conv3 against scipy.ndimage.correlate with
zero padding; pooling/upsampling adjointness; the backward pass against
central finite differences of the float64 loss for every layer; bf16 rounding
against hand-worked bit patterns.

Volumes are [d0][d1][d2] (d2 fastest), activations [d0][d1][d2][C].
"""
import numpy as np

# (level, C_in, C_out) of the 14 3x3x3 convs, in parameter order
SPEC = [(0, 1, 16), (0, 16, 16), (1, 16, 32), (1, 32, 32), (2, 32, 64), (2, 64, 64),
        (3, 64, 128), (3, 128, 128), (2, 192, 64), (2, 64, 64), (1, 96, 32), (1, 32, 32),
        (0, 48, 16), (0, 16, 16)]
HEAD_C = 16
TAPS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def param_count():
    return sum(co * 27 * ci + co for _, ci, co in SPEC) + HEAD_C + 1


def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32)
    r = (u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))) & np.uint32(0xFFFF0000)
    return r.view(np.float32)


def unpack(params):
    p = np.asarray(params)
    out, o = [], 0
    for _, ci, co in SPEC:
        w = p[o:o + co * 27 * ci].reshape(co, 27, ci)
        o += co * 27 * ci
        b = p[o:o + co]
        o += co
        out.append((w, b))
    wh = p[o:o + HEAD_C]
    bh = p[o + HEAD_C]
    return out, wh, bh


def shift(x, a):
    """y[v] = x[v + a] with zeros outside (x is [d0][d1][d2][C])."""
    d = x.shape[:3]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    return xp[1 + a[0]:1 + a[0] + d[0], 1 + a[1]:1 + a[1] + d[1], 1 + a[2]:1 + a[2] + d[2]]


def conv3(x, w, b=None):
    """out[v][co] = sum_{t,ci} x[v + a_t][ci] w[co][t][ci] (+ b[co]), float64."""
    x = x.astype(np.float64)
    w = w.astype(np.float64)
    out = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, a in enumerate(TAPS):
        out += shift(x, a) @ w[:, t, :].T
    if b is not None:
        out += b
    return out


def conv3_wgrad(x, g):
    """dW[co][t][ci] = sum_v g[v][co] x[v + a_t][ci]."""
    x = x.astype(np.float64)
    g = g.astype(np.float64)
    co, ci = g.shape[3], x.shape[3]
    dw = np.zeros((co, 27, ci))
    G = g.reshape(-1, co)
    for t, a in enumerate(TAPS):
        dw[:, t, :] = G.T @ shift(x, a).reshape(-1, ci)
    return dw


def conv3_dgrad(g, w):
    """gin[v][ci] = sum_{t,co} g[v - a_t][co] w[co][t][ci]."""
    g = g.astype(np.float64)
    w = w.astype(np.float64)
    out = np.zeros(g.shape[:3] + (w.shape[2],))
    for t, a in enumerate(TAPS):
        out += shift(g, (-a[0], -a[1], -a[2])) @ w[:, t, :]
    return out


def pool(x):
    d0, d1, d2, c = x.shape
    return x.reshape(d0 // 2, 2, d1 // 2, 2, d2 // 2, 2, c).astype(np.float64).sum((1, 3, 5)) * 0.125


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def sum_children(g):
    d0, d1, d2, c = g.shape
    return g.reshape(d0 // 2, 2, d1 // 2, 2, d2 // 2, 2, c).astype(np.float64).sum((1, 3, 5))


class UNet:
    """Forward/backward of the warm-start network on one volume."""

    def __init__(self, params, scale, bf16=True):
        self.params = np.asarray(params, np.float32 if bf16 else np.float64)
        self.scale = scale
        self.bf16 = bf16
        self.layers, self.wh, self.bh = unpack(self.params)
        if bf16:
            self.layers = [(bf16_round(w), b) for w, b in self.layers]

    def R(self, x):
        return bf16_round(np.asarray(x, np.float32)) if self.bf16 else np.asarray(x, np.float64)

    def _conv(self, l, x):
        w, b = self.layers[l]
        return self.R(np.maximum(conv3(x, w, b), 0.0))

    def forward(self, x):
        """Returns the output volume and keeps the activations for backward."""
        R = self.R
        s = self.scale
        xs = np.asarray(x, np.float32 if self.bf16 else np.float64)
        A = {}
        A["t"] = R(xs * np.float32(s) if self.bf16 else xs * s)[..., None]
        A["e1a"] = self._conv(0, A["t"])
        A["e1"] = self._conv(1, A["e1a"])
        A["p1"] = R(pool(A["e1"]))
        A["e2a"] = self._conv(2, A["p1"])
        A["e2"] = self._conv(3, A["e2a"])
        A["p2"] = R(pool(A["e2"]))
        A["e3a"] = self._conv(4, A["p2"])
        A["e3"] = self._conv(5, A["e3a"])
        A["p3"] = R(pool(A["e3"]))
        A["ba"] = self._conv(6, A["p3"])
        A["bb"] = self._conv(7, A["ba"])
        A["u3"] = np.concatenate([up(A["bb"]), A["e3"]], -1)
        A["d3a"] = self._conv(8, A["u3"])
        A["d3"] = self._conv(9, A["d3a"])
        A["u2"] = np.concatenate([up(A["d3"]), A["e2"]], -1)
        A["d2a"] = self._conv(10, A["u2"])
        A["d2"] = self._conv(11, A["d2a"])
        A["u1"] = np.concatenate([up(A["d2"]), A["e1"]], -1)
        A["d1a"] = self._conv(12, A["u1"])
        A["d1"] = self._conv(13, A["d1a"])
        h = A["d1"].astype(np.float64) @ np.asarray(self.wh, np.float64) + float(self.bh)
        A["h"] = h
        A["x"] = xs
        self.A = A
        return (xs.astype(np.float64) * s + h) / s

    def loss_grad(self, x, tgt):
        """loss = mean((x*s + h - tgt*s)^2) and the flat parameter gradient."""
        R = self.R
        s = self.scale
        self.forward(x)
        A = self.A
        n = A["h"].size
        res = np.asarray(x, np.float64) * s + A["h"] - np.asarray(tgt, np.float64) * s
        loss = float(np.mean(res * res))
        g = 2.0 * res / n
        grads = [None] * len(SPEC)
        d1 = A["d1"].astype(np.float64)
        g_wh = np.tensordot(g, d1, axes=([0, 1, 2], [0, 1, 2]))
        g_bh = g.sum()
        g_d1 = R(g[..., None] * np.asarray(self.wh, np.float64)) * (A["d1"] > 0)

        def back(l, xin, gout, relu_in):
            w, _ = self.layers[l]
            grads[l] = (conv3_wgrad(xin, gout), gout.astype(np.float64).sum((0, 1, 2)))
            gin = conv3_dgrad(gout, w)
            if relu_in:
                gin = gin * (xin > 0)
            return R(gin)

        g_d1a = back(13, A["d1a"], g_d1, True)
        g_u1 = back(12, A["u1"], g_d1a, True)
        g_d2 = R(sum_children(g_u1[..., :32]))
        g_d2a = back(11, A["d2a"], g_d2, True)
        g_u2 = back(10, A["u2"], g_d2a, True)
        g_d3 = R(sum_children(g_u2[..., :64]))
        g_d3a = back(9, A["d3a"], g_d3, True)
        g_u3 = back(8, A["u3"], g_d3a, True)
        g_bb = R(sum_children(g_u3[..., :128]))
        g_ba = back(7, A["ba"], g_bb, True)
        g_p3 = back(6, A["p3"], g_ba, False)
        g_e3 = R(g_u3[..., 128:].astype(np.float64) + (A["e3"] > 0) * up(g_p3) * 0.125)
        g_e3a = back(5, A["e3a"], g_e3, True)
        g_p2 = back(4, A["p2"], g_e3a, False)
        g_e2 = R(g_u2[..., 64:].astype(np.float64) + (A["e2"] > 0) * up(g_p2) * 0.125)
        g_e2a = back(3, A["e2a"], g_e2, True)
        g_p1 = back(2, A["p1"], g_e2a, False)
        g_e1 = R(g_u1[..., 32:].astype(np.float64) + (A["e1"] > 0) * up(g_p1) * 0.125)
        g_e1a = back(1, A["e1a"], g_e1, True)
        w0, _ = self.layers[0]
        grads[0] = (conv3_wgrad(A["t"], g_e1a), g_e1a.astype(np.float64).sum((0, 1, 2)))
        flat = []
        for dw, db in grads:
            flat.append(dw.reshape(-1))
            flat.append(db)
        flat.append(g_wh)
        flat.append([g_bh])
        return loss, np.concatenate(flat)


def layer_slices():
    """(name, slice) of each parameter block of the flat vector."""
    out, o = [], 0
    for l, (_, ci, co) in enumerate(SPEC):
        out.append((f"conv{l}.w", slice(o, o + co * 27 * ci)))
        o += co * 27 * ci
        out.append((f"conv{l}.b", slice(o, o + co)))
        o += co
    out.append(("head.w", slice(o, o + HEAD_C)))
    out.append(("head.b", slice(o + HEAD_C, o + HEAD_C + 1)))
    return out
