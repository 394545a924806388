"""Learned template surrogate: a network from (logAge, [M/H], distance
modulus, error scale) to a 200 x 300 log-template, trained on the GPU with
the library's own kernels (sfh_nn_*: matrix-core products in BF16 -- or FP32
to check gradients -- with FP32 accumulation, FP32 master weights and Adam).

Architecture (about 4.3M parameters):
    p (4, scaled to [-1, 1]) -> Linear 256 -> SiLU -> Linear 512 -> SiLU
      -> Linear 13*19*32 -> SiLU -> reshape 13 x 19 x 32 (channels last)
      -> ConvT(32->32) -> SiLU -> ConvT(32->32) -> SiLU -> ConvT(32->32) -> SiLU
      -> ConvT(32->1) + per-pixel bias map -> 208 x 304, top-left 200 x 300 kept
    ConvT = transposed convolution, 4x4 kernel, stride 2, padding 1.

No reference counterpart: StarFormationHistories.jl builds every template
with partial_cmd_smooth; here the network fills fine grids cheaply and
partial_cmd_smooth remains the fallback (``fill_templates``).  PyTorch tensors
serve only as device memory; all arithmetic runs in libsfh_hip.so.
"""
from __future__ import annotations

import ctypes as C
import numpy as np

from .core import Context, default_context

# parameter ranges of config 4 (scaled to [-1, 1] on input)
PARAM_LO = np.array([6.6, -2.5, 24.0, 0.5])
PARAM_HI = np.array([10.1, 0.3, 30.0, 2.0])

H0, W0, CH = 13, 19, 32
HID0, HID1 = 256, 512
LAT = H0 * W0 * CH            # 7904
OUT_H, OUT_W = 16 * H0, 16 * W0   # 208 x 304
CROP_H, CROP_W = 200, 300

SILU = 1
IDENT = 0


def _pad16(n):
    return (n + 15) // 16 * 16


def param_layout():
    """(name, shape, fan_in) of every parameter tensor, in the order of the flat vector."""
    return [
        ("W0", (HID0, 16), 4), ("b0", (HID0,), None),
        ("W1", (HID1, HID0), HID0), ("b1", (HID1,), None),
        ("W2", (LAT, HID1), HID1), ("b2", (LAT,), None),
        ("K1", (16 * CH, CH), 4 * CH), ("c1", (CH,), None),
        ("K2", (16 * CH, CH), 4 * CH), ("c2", (CH,), None),
        ("K3", (16 * CH, CH), 4 * CH), ("c3", (CH,), None),
        ("w4", (16, CH), 4 * CH), ("bmap", (OUT_H * OUT_W,), None),
    ]


def init_params(seed=0):
    """Uniform initialisation with variance 1 / fan_in for weights, zeros for
    biases (W0's padding columns 4..15 stay zero).  Returns a flat FP32 array."""
    rng = np.random.default_rng(seed)
    out = []
    for name, shape, fan in param_layout():
        if fan is None:
            out.append(np.zeros(int(np.prod(shape)), dtype=np.float32))
            continue
        a = np.sqrt(3.0 / fan)
        w = rng.uniform(-a, a, size=shape).astype(np.float32)
        if name == "W0":
            w[:, 4:] = 0.0
        out.append(w.ravel())
    return np.concatenate(out)


def scale_params(p):
    """Physical (logAge, [M/H], dmod, error scale) -> [-1, 1]^4."""
    p = np.atleast_2d(np.asarray(p, dtype=np.float64))
    return (2.0 * (p - PARAM_LO) / (PARAM_HI - PARAM_LO) - 1.0).astype(np.float32)


class TemplateSurrogate:
    """The network, its FP32 master parameters, Adam state and work buffers on
    one device.  ``dtype``: "bf16" (training) or "fp32" (gradient checks)."""

    def __init__(self, params=None, dtype="bf16", seed=0, ctx: Context | None = None, max_batch=32):
        import torch
        self.torch = torch
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.dt = {"bf16": 0, "fp32": 1}[dtype]
        self.tdt = torch.bfloat16 if self.dt == 0 else torch.float32
        self.dev = torch.device("cuda")
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        self.flat = torch.from_numpy(p.copy()).to(self.dev)
        self.grad = torch.zeros_like(self.flat)
        self.m = torch.zeros_like(self.flat)
        self.v = torch.zeros_like(self.flat)
        self.step_count = 0
        self.views, self.gviews = {}, {}
        off = 0
        for name, shape, _ in param_layout():
            n = int(np.prod(shape))
            self.views[name] = self.flat[off:off + n].view(*shape)
            self.gviews[name] = self.grad[off:off + n].view(*shape)
            off += n
        self.nparams = off
        self.max_batch = 0
        self._alloc(max_batch)

    # ------------------------------------------------------------ buffers
    def _buf(self, n, dtype=None):
        return self.torch.zeros(int(n), dtype=dtype or self.torch.float32, device=self.dev)

    def _alloc(self, B):
        if B <= self.max_batch:
            return
        self.max_batch = B
        f, t = self.torch.float32, self.tdt
        Bp = _pad16(B)
        P1, P2, P3, P4 = B * H0 * W0, B * 4 * H0 * W0, B * 16 * H0 * W0, B * 64 * H0 * W0
        self.bp = Bp
        self.x0f = self._buf(B * 16)
        self.x0 = self._buf(B * 16, t)
        self.z = self._buf(max(B * LAT, P1 * 16 * CH, P2 * 16 * CH, P3 * 16 * CH))   # GEMM outputs (FP32)
        self.pre0, self.pre1, self.pre2 = self._buf(B * HID0), self._buf(B * HID1), self._buf(B * LAT)
        self.a0, self.a1, self.a2 = self._buf(B * HID0, t), self._buf(B * HID1, t), self._buf(B * LAT, t)
        self.pre3, self.pre4, self.pre5 = self._buf(P2 * CH), self._buf(P3 * CH), self._buf(P4 * CH)
        self.a3, self.a4 = self._buf(P2 * CH, t), self._buf(P3 * CH, t)
        self.a5 = self._buf(P4 * CH)
        self.out = self._buf(B * OUT_H * OUT_W)
        self.dout = self._buf(B * OUT_H * OUT_W)
        self.dA = self._buf(max(P4 * CH, B * LAT))            # gradients w.r.t. activations (FP32)
        self.dB = self._buf(max(P4 * CH, B * LAT))
        self.dcols = self._buf(max(P3 * 16 * CH, B * LAT), t)     # im2col rows / low-precision copies
        self.dcolsT = self._buf(16 * CH * _pad16(P3) + LAT * Bp, t)
        self.actT = self._buf(CH * _pad16(P3) + HID1 * Bp, t)
        # low-precision weight copies and their transposes
        self.w = {k: self._buf(self.views[k].numel(), t) for k in ("W0", "W1", "W2", "K1", "K2", "K3")}
        self.wT = {k: self._buf(self.views[k].numel(), t) for k in ("W1", "W2", "K1", "K2", "K3")}
        self.torch.cuda.synchronize()     # torch filled these on its stream; the library uses its own

    # ---------------------------------------------------------- primitives
    def _c(self, rc):
        self.ctx.check(rc)

    def _gemm(self, A, lda, Bm, ldb, Cm, ldc, M, N, K, beta=0.0):
        self._c(self.lib.sfh_nn_gemm_nt(self.ctx.handle, self.dt, A.data_ptr(), lda, Bm.data_ptr(), ldb,
                                        Cm.data_ptr(), ldc, M, N, K, 1.0, beta))

    def _bias_act(self, X, ldx, bias, rows, cols, act, pre, Y, ldy, dt=None):
        self._c(self.lib.sfh_nn_bias_act(self.ctx.handle, self.dt if dt is None else dt, X.data_ptr(), ldx,
                                         bias.data_ptr() if bias is not None else None, rows, cols, act,
                                         pre.data_ptr() if pre is not None else None, Y.data_ptr(), ldy))

    def _transpose(self, X, ldi, rows, cols, out, ldo, act=IDENT):
        self._c(self.lib.sfh_nn_transpose(self.ctx.handle, self.dt, X.data_ptr(), ldi, rows, cols, out.data_ptr(),
                                          ldo, cols, act))

    def _act_bwd(self, d, pre, n):
        self._c(self.lib.sfh_nn_act_bwd(self.ctx.handle, d.data_ptr(), pre.data_ptr(), n, SILU, d.data_ptr()))

    def _colsum(self, X, ld, rows, cols, out):
        self._c(self.lib.sfh_nn_colsum(self.ctx.handle, X.data_ptr(), ld, rows, cols, 0.0, out.data_ptr()))

    def _refresh_weights(self):
        """Low-precision copies (and transposes) of the master weights."""
        v = self.views
        for k in ("W0", "W1", "W2", "K1", "K2", "K3"):
            r, c = v[k].shape
            self._bias_act(v[k], c, None, r, c, IDENT, None, self.w[k], c)
        for k in ("W1", "W2", "K1", "K2", "K3"):
            r, c = v[k].shape
            self._transpose(v[k], c, r, c, self.wT[k], r)

    # ---------------------------------------------------------- forward
    def forward(self, params, raw=False):
        """Network output for a batch of physical parameters (B x 4): a torch
        FP32 tensor B x 208 x 304 (log-template; rows < 200, columns < 300 are
        the Hess diagram).  raw=True: params are already scaled to [-1, 1]."""
        x = np.asarray(params, dtype=np.float32) if raw else scale_params(params)
        x = np.atleast_2d(x)
        B = x.shape[0]
        self._alloc(B)
        self.B = B
        self._refresh_weights()
        x16 = np.zeros((B, 16), dtype=np.float32)
        x16[:, :4] = x
        self.x0f[:B * 16].copy_(self.torch.from_numpy(x16.ravel()))
        self.torch.cuda.synchronize()
        v = self.views
        self._bias_act(self.x0f, 16, None, B, 16, IDENT, None, self.x0, 16)
        # MLP
        self._gemm(self.x0, 16, self.w["W0"], 16, self.z, HID0, B, HID0, 16)
        self._bias_act(self.z, HID0, v["b0"], B, HID0, SILU, self.pre0, self.a0, HID0)
        self._gemm(self.a0, HID0, self.w["W1"], HID0, self.z, HID1, B, HID1, HID0)
        self._bias_act(self.z, HID1, v["b1"], B, HID1, SILU, self.pre1, self.a1, HID1)
        self._gemm(self.a1, HID1, self.w["W2"], HID1, self.z, LAT, B, LAT, HID1)
        self._bias_act(self.z, LAT, v["b2"], B, LAT, SILU, self.pre2, self.a2, LAT)
        # decoder
        H, W = H0, W0
        ins = (self.a2, self.a3, self.a4)
        pres = (self.pre3, self.pre4, self.pre5)
        outs = (self.a3, self.a4, None)
        for li, (K, cb) in enumerate((("K1", "c1"), ("K2", "c2"), ("K3", "c3"))):
            P = B * H * W
            self._gemm(ins[li], CH, self.w[K], CH, self.z, 16 * CH, P, 16 * CH, CH)
            self._c(self.lib.sfh_nn_col2im(self.ctx.handle, self.z.data_ptr(), 16 * CH, B, H, W, CH,
                                           v[cb].data_ptr(), pres[li].data_ptr()))
            H, W = 2 * H, 2 * W
            if outs[li] is not None:
                self._bias_act(pres[li], CH, None, B * H * W, CH, SILU, None, outs[li], CH)
            else:
                self._bias_act(pres[li], CH, None, B * H * W, CH, SILU, None, self.a5, CH, dt=1)
        self._c(self.lib.sfh_nn_conv1_fwd(self.ctx.handle, self.a5.data_ptr(), B, H, W, CH, v["w4"].data_ptr(),
                                          v["bmap"].data_ptr(), self.out.data_ptr()))
        self.ctx.synchronize()
        return self.out[:B * OUT_H * OUT_W].view(B, OUT_H, OUT_W)

    # ---------------------------------------------------------- backward
    def backward(self, target):
        """Mean squared error of the last forward's crop against target
        (B x 200 x 300, torch FP32 on the device) and the gradient of every
        parameter into self.grad.  Returns the loss."""
        B = self.B
        t = target.contiguous().view(-1)
        self.torch.cuda.synchronize()
        loss = C.c_float()
        self._c(self.lib.sfh_nn_mse(self.ctx.handle, self.out.data_ptr(), t.data_ptr(), B, OUT_H, OUT_W, CROP_H,
                                    CROP_W, self.dout.data_ptr(), C.byref(loss)))
        self._backward_from_dout()
        return float(loss.value)

    def _backward_from_dout(self):
        B, v, g = self.B, self.views, self.gviews
        H, W = 8 * H0, 8 * W0                                   # input grid of the last layer
        # single-channel output layer
        self._c(self.lib.sfh_nn_conv1_bwd(self.ctx.handle, self.dout.data_ptr(), self.a5.data_ptr(), B, H, W, CH,
                                          v["w4"].data_ptr(), self.dA.data_ptr(), g["w4"].data_ptr(),
                                          g["bmap"].data_ptr()))
        d = self.dA                                              # gradient w.r.t. a5
        layers = (("K3", "c3", self.pre5, self.a4, self.pre4),
                  ("K2", "c2", self.pre4, self.a3, self.pre3),
                  ("K1", "c1", self.pre3, self.a2, self.pre2))
        for K, cb, pre_out, a_in, pre_in in layers:
            # d is the gradient w.r.t. act(pre_out) (output grid H x W)
            Hi, Wi = H // 2, W // 2
            Pi, Po = B * Hi * Wi, B * H * W
            self._act_bwd(d, pre_out, Po * CH)                   # -> gradient w.r.t. pre_out
            self._colsum(d, CH, Po, CH, g[cb])
            self._c(self.lib.sfh_nn_im2col(self.ctx.handle, self.dt, d.data_ptr(), B, Hi, Wi, CH,
                                           self.dcols.data_ptr(), 16 * CH))
            # dK = dcols^T a_in: transposed copies (K = pixels, padded)
            Pp = _pad16(Pi)
            self._c(self.lib.sfh_nn_im2col(self.ctx.handle, 1, d.data_ptr(), B, Hi, Wi, CH, self.z.data_ptr(), 16 * CH))
            self._transpose(self.z, 16 * CH, Pi, 16 * CH, self.dcolsT, Pp)
            self._transpose(pre_in, CH, Pi, CH, self.actT, Pp, act=SILU)
            self._gemm(self.dcolsT, Pp, self.actT, Pp, g[K], CH, 16 * CH, CH, Pp)
            # d a_in = dcols K
            dn = self.dB if d is self.dA else self.dA
            self._gemm(self.dcols, 16 * CH, self.wT[K], 16 * CH, dn, CH, Pi, CH, 16 * CH)
            d = dn
            H, W = Hi, Wi
        # MLP; d is the gradient w.r.t. a2 (B x LAT)
        mlp = (("W2", "b2", self.pre2, self.pre1, HID1, LAT),
               ("W1", "b1", self.pre1, self.pre0, HID0, HID1),
               ("W0", "b0", self.pre0, None, 16, HID0))
        Bp = self.bp
        for Wn, bn, pre_out, pre_in, nin, nout in mlp:
            self._act_bwd(d, pre_out, B * nout)
            self._colsum(d, nout, B, nout, g[bn])
            # dW = d^T a_in
            self._transpose(d, nout, B, nout, self.dcolsT, Bp)
            if pre_in is not None:
                self._transpose(pre_in, nin, B, nin, self.actT, Bp, act=SILU)
            else:
                self._transpose(self.x0f, 16, B, 16, self.actT, Bp)
            self._gemm(self.dcolsT, Bp, self.actT, Bp, g[Wn], nin, nout, nin, Bp)
            if pre_in is None:
                break
            # d a_in = d W  (low-precision copy of d times the transposed weights)
            self._bias_act(d, nout, None, B, nout, IDENT, None, self.dcols, nout)
            dn = self.dB if d is self.dA else self.dA
            self._gemm(self.dcols, nout, self.wT[Wn], nout, dn, nin, B, nin, nout)
            d = dn

    # ---------------------------------------------------------- training
    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, weight_decay=0.0):
        self.step_count += 1
        self._c(self.lib.sfh_nn_adam(self.ctx.handle, self.flat.data_ptr(), self.grad.data_ptr(), self.m.data_ptr(),
                                     self.v.data_ptr(), self.nparams, lr, b1, b2, eps, self.step_count,
                                     weight_decay))

    def train_step(self, params, target, lr=1e-3):
        self.forward(params)
        loss = self.backward(target)
        self.adam_step(lr)
        return loss

    def parameters(self) -> np.ndarray:
        self.ctx.synchronize()
        return self.flat.cpu().numpy()

    def gradient(self) -> np.ndarray:
        self.ctx.synchronize()
        return self.grad.cpu().numpy()
