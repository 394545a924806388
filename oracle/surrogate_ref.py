"""FP64 restatement of the learned template surrogate (TEST INFRASTRUCTURE
ONLY: the checker of the GPU network's loss and gradients; only tests/ may
import it).

Same architecture and flat parameter layout as the product's surrogate
(MLP 4 -> 256 -> 512 -> 13*19*32, SiLU, then three 32-channel and one
1-channel 4x4 stride-2 transposed convolutions, a per-pixel bias map,
crop 200 x 300, mean squared error), written independently: the transposed
convolution as a scatter of every input pixel into a zero-padded output
(out[2y-1+ky, 2x-1+kx] += in[y, x] W[ky, kx]) with hand-derived backward
passes, all in numpy float64.
"""
from __future__ import annotations

import numpy as np

H0, W0, CH, HID0, HID1 = 13, 19, 32, 256, 512
LAT = H0 * W0 * CH
OUT_H, OUT_W, CROP_H, CROP_W = 208, 304, 200, 300

LAYOUT = [("W0", (HID0, 16)), ("b0", (HID0,)), ("W1", (HID1, HID0)), ("b1", (HID1,)),
          ("W2", (LAT, HID1)), ("b2", (LAT,)), ("K1", (16 * CH, CH)), ("c1", (CH,)),
          ("K2", (16 * CH, CH)), ("c2", (CH,)), ("K3", (16 * CH, CH)), ("c3", (CH,)),
          ("w4", (16, CH)), ("bmap", (OUT_H * OUT_W,))]


def unpack(flat):
    out, off = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[off:off + n], dtype=np.float64).reshape(shape)
        off += n
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name, _ in LAYOUT])


def silu(x):
    return x / (1.0 + np.exp(-x))


def dsilu(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def _kernel(K, co):
    """((ky*4 + kx)*co + o, ci) rows (tap-major) -> (co, 4, 4, ci)."""
    return K.reshape(4, 4, co, -1).transpose(2, 0, 1, 3)


def convT(x, K, co):
    """x (B, H, W, ci) -> (B, 2H, 2W, co) by scattering every input pixel."""
    B, H, W, _ = x.shape
    k = _kernel(K, co)
    full = np.zeros((B, 2 * H + 2, 2 * W + 2, co))
    for ky in range(4):
        for kx in range(4):
            full[:, ky:ky + 2 * H:2, kx:kx + 2 * W:2, :] += np.einsum("bhwc,oc->bhwo", x, k[:, ky, kx, :])
    return full[:, 1:2 * H + 1, 1:2 * W + 1, :]


def convT_bwd(x, K, co, dout):
    """Gradients of convT w.r.t. x and K given dout (B, 2H, 2W, co)."""
    B, H, W, ci = x.shape
    k = _kernel(K, co)
    dfull = np.zeros((B, 2 * H + 2, 2 * W + 2, co))
    dfull[:, 1:2 * H + 1, 1:2 * W + 1, :] = dout
    dx = np.zeros_like(x)
    dk = np.zeros_like(k)
    for ky in range(4):
        for kx in range(4):
            g = dfull[:, ky:ky + 2 * H:2, kx:kx + 2 * W:2, :]          # (B, H, W, co)
            dx += np.einsum("bhwo,oc->bhwc", g, k[:, ky, kx, :])
            dk[:, ky, kx, :] = np.einsum("bhwo,bhwc->oc", g, x)
    return dx, dk.transpose(1, 2, 0, 3).reshape(16 * co, ci)


def scale_params(p):
    lo = np.array([6.6, -2.5, 24.0, 0.5])
    hi = np.array([10.1, 0.3, 30.0, 2.0])
    return 2.0 * (np.atleast_2d(np.asarray(p, dtype=np.float64)) - lo) / (hi - lo) - 1.0


def loss_and_grad(flat, x, target, want_grad=True):
    """x: (B, 4) scaled inputs; target (B, 200, 300).  Returns (loss, flat
    gradient or None, output (B, 208, 304))."""
    P = unpack(flat)
    B = x.shape[0]
    x16 = np.zeros((B, 16))
    x16[:, :4] = x
    z0 = x16 @ P["W0"].T + P["b0"]
    a0 = silu(z0)
    z1 = a0 @ P["W1"].T + P["b1"]
    a1 = silu(z1)
    z2 = a1 @ P["W2"].T + P["b2"]
    a2 = silu(z2).reshape(B, H0, W0, CH)
    z3 = convT(a2, P["K1"], CH) + P["c1"]
    a3 = silu(z3)
    z4 = convT(a3, P["K2"], CH) + P["c2"]
    a4 = silu(z4)
    z5 = convT(a4, P["K3"], CH) + P["c3"]
    a5 = silu(z5)
    out = convT(a5, P["w4"], 1)[..., 0] + P["bmap"].reshape(OUT_H, OUT_W)
    diff = out[:, :CROP_H, :CROP_W] - target
    n = B * CROP_H * CROP_W
    loss = float(np.sum(diff * diff) / n)
    if not want_grad:
        return loss, None, out
    G = {}
    dout = np.zeros_like(out)
    dout[:, :CROP_H, :CROP_W] = 2.0 * diff / n
    G["bmap"] = dout.sum(axis=0).ravel()
    da5, G["w4"] = convT_bwd(a5, P["w4"], 1, dout[..., None])
    dz5 = da5 * dsilu(z5)
    G["c3"] = dz5.sum(axis=(0, 1, 2))
    da4, G["K3"] = convT_bwd(a4, P["K3"], CH, dz5)
    dz4 = da4 * dsilu(z4)
    G["c2"] = dz4.sum(axis=(0, 1, 2))
    da3, G["K2"] = convT_bwd(a3, P["K2"], CH, dz4)
    dz3 = da3 * dsilu(z3)
    G["c1"] = dz3.sum(axis=(0, 1, 2))
    da2, G["K1"] = convT_bwd(a2, P["K1"], CH, dz3)
    dz2 = da2.reshape(B, LAT) * dsilu(z2)
    G["b2"] = dz2.sum(axis=0)
    G["W2"] = dz2.T @ a1
    dz1 = (dz2 @ P["W2"]) * dsilu(z1)
    G["b1"] = dz1.sum(axis=0)
    G["W1"] = dz1.T @ a0
    dz0 = (dz1 @ P["W1"]) * dsilu(z0)
    G["b0"] = dz0.sum(axis=0)
    G["W0"] = dz0.T @ x16
    return loss, pack(G), out
