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

The network predicts images of y = (ln(T + LOG_EPS) - LOG_OFFSET) / LOG_SCALE
(T per unit stellar mass, image row = colour bin, column = magnitude bin).

Pipeline: ``sample_parameters`` draws a design over the parameter box,
``exact_templates`` builds those templates with partial_cmd_smooth (host
point preparation and the GPU deposition kernel), ``TemplateSurrogate.train``
fits the network to them and ``fill_templates`` fills a fine
(logAge, [M/H]) grid at one distance and error scale, checking the network
against exact templates on a sparse lattice and computing exactly every cell
whose check fails.  The check is the log-likelihood change of a fit:
fitting N stars drawn from the exact template with the network's template
instead (amplitude free) lowers the best log-likelihood by
dlnL = N KL(exact || network) (``template_delta_lnL``).

No reference counterpart: StarFormationHistories.jl builds every template
with partial_cmd_smooth, which remains the fallback here.  PyTorch tensors
serve only as device memory; all arithmetic runs in libsfh_hip.so.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .core import Context, default_context

# training box of the surrogate: logAge, [M/H], dmod, error scale (scaled to [-1, 1] on input)
PARAM_LO = np.array([6.6, -2.5, 24.0, 0.5])
PARAM_HI = np.array([10.1, 0.3, 30.0, 2.0])

H0, W0, CH = 13, 19, 32
HID0, HID1 = 256, 512
LAT = H0 * W0 * CH            # 7904
OUT_H, OUT_W = 16 * H0, 16 * W0   # 208 x 304
CROP_H, CROP_W = 200, 300

SILU = 1
IDENT = 0

# the log-template transform
LOG_EPS = 1e-12
LOG_OFFSET = -20.0
LOG_SCALE = 5.0
# isochrone wiggles are drawn from this seed for every template, so a template
# is a function of its four parameters alone
ISOCHRONE_SEED = 7


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
        P2, P3, P4 = B * 4 * H0 * W0, B * 16 * H0 * W0, B * 64 * H0 * W0
        self.bp = Bp
        self.x0f = self._buf(B * 16)
        self.x0 = self._buf(B * 16, t)
        self.z = self._buf(B * max(LAT, HID0, HID1))             # MLP product outputs (FP32)
        self.pre0, self.pre1, self.pre2 = self._buf(B * HID0), self._buf(B * HID1), self._buf(B * LAT)
        self.a0, self.a1, self.a2 = self._buf(B * HID0, t), self._buf(B * HID1, t), self._buf(B * LAT, t)
        self.pre3, self.pre4, self.pre5 = self._buf(P2 * CH), self._buf(P3 * CH), self._buf(P4 * CH)
        self.a3, self.a4 = self._buf(P2 * CH, t), self._buf(P3 * CH, t)
        self.out = self._buf(B * OUT_H * OUT_W)
        self.dout = self._buf(B * OUT_H * OUT_W)
        self.dA = self._buf(max(P4 * CH, B * LAT))            # gradients w.r.t. activations (FP32)
        self.dB = self._buf(max(P4 * CH, B * LAT))
        self.dcols = self._buf(B * LAT, t)                       # low-precision copies of MLP gradients
        self.dcolsT = self._buf(LAT * Bp, t)                     # transposed MLP gradients
        self.actT = self._buf(16 * Bp, t)                         # transposed network input
        self.a0T = self._buf(HID0 * Bp, t)                        # transposed MLP activations (from the forward)
        self.a1T = self._buf(HID1 * Bp, t)
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

    def _linear(self, A, lda, Wm, ldw, M, N, K, bias, pre, Y, ldy, YT=None):
        """pre = A Wm^T + bias, Y = SiLU(pre) in the working precision and, with
        YT, its transpose (N x bp, zero padded) for the weight gradient
        (self.z as workspace)."""
        self._c(self.lib.sfh_nn_linear(self.ctx.handle, self.dt, A.data_ptr(), lda, Wm.data_ptr(), ldw, M, N, K,
                                       bias.data_ptr(), SILU, self.z.data_ptr(), pre.data_ptr(), Y.data_ptr(), ldy,
                                       YT.data_ptr() if YT is not None else None, self.bp))

    def _bias_act(self, X, ldx, bias, rows, cols, act, pre, Y, ldy, dt=None):
        self._c(self.lib.sfh_nn_bias_act(self.ctx.handle, self.dt if dt is None else dt, X.data_ptr(), ldx,
                                         bias.data_ptr() if bias is not None else None, rows, cols, act,
                                         pre.data_ptr() if pre is not None else None, Y.data_ptr(), ldy))

    def _transpose(self, X, ldi, rows, cols, out, ldo, act=IDENT):
        self._c(self.lib.sfh_nn_transpose(self.ctx.handle, self.dt, X.data_ptr(), ldi, rows, cols, out.data_ptr(),
                                          ldo, cols, act))

    def _refresh_weights(self):
        """Low-precision copies (and transposes) of the master weights, one launch."""
        v = self.views
        names = ("W0", "W1", "W2", "K1", "K2", "K3")
        mats = (_lib.sfh_nn_wcopy * len(names))()
        for i, k in enumerate(names):
            r, c = v[k].shape
            mats[i] = _lib.sfh_nn_wcopy(v[k].data_ptr(), r, c, self.w[k].data_ptr(),
                                        self.wT[k].data_ptr() if k in self.wT else None)
        self._c(self.lib.sfh_nn_weight_copies(self.ctx.handle, self.dt, len(names), mats))

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
        x16 = np.zeros((B, 16), dtype=np.float32)
        x16[:, :4] = x
        self.x0f[:B * 16].copy_(self.torch.from_numpy(x16.ravel()))
        self.torch.cuda.synchronize()
        self._forward_device(B)
        self.ctx.synchronize()
        return self.out[:B * OUT_H * OUT_W].view(B, OUT_H, OUT_W)

    def _forward_device(self, B):
        """The network on the scaled inputs already in x0f (B x 16), queued on
        the context's stream."""
        v = self.views
        self._refresh_weights()
        self._bias_act(self.x0f, 16, None, B, 16, IDENT, None, self.x0, 16)
        # MLP
        self._linear(self.x0, 16, self.w["W0"], 16, B, HID0, 16, v["b0"], self.pre0, self.a0, HID0, self.a0T)
        self._linear(self.a0, HID0, self.w["W1"], HID0, B, HID1, HID0, v["b1"], self.pre1, self.a1, HID1, self.a1T)
        self._linear(self.a1, HID1, self.w["W2"], HID1, B, LAT, HID1, v["b2"], self.pre2, self.a2, LAT)
        # decoder
        H, W = H0, W0
        ins = (self.a2, self.a3, self.a4)
        pres = (self.pre3, self.pre4, self.pre5)
        outs = (self.a3, self.a4, None)
        for li, (K, cb) in enumerate((("K1", "c1"), ("K2", "c2"), ("K3", "c3"))):
            # pre-activation (+ bias) and, but for the last layer, its SiLU in one kernel
            self._c(self.lib.sfh_nn_conv_t_fwd(self.ctx.handle, self.dt, ins[li].data_ptr(), B, H, W, CH,
                                               self.w[K].data_ptr(), v[cb].data_ptr(), pres[li].data_ptr(),
                                               outs[li].data_ptr() if outs[li] is not None else None))
            H, W = 2 * H, 2 * W
        # single-channel output layer on SiLU(pre5), the activation applied inside
        self._c(self.lib.sfh_nn_conv1_fwd(self.ctx.handle, self.pre5.data_ptr(), B, H, W, CH, v["w4"].data_ptr(),
                                          v["bmap"].data_ptr(), SILU, self.out.data_ptr()))

    # ---------------------------------------------------------- backward
    def backward(self, target):
        """Mean squared error of the last forward's crop against target
        (B x 200 x 300, torch FP32 on the device) and the gradient of every
        parameter into self.grad.  Returns the loss."""
        B = self.B
        t = target.contiguous().view(-1)
        self.torch.cuda.synchronize()
        loss = C.c_float()
        self._c(self.lib.sfh_nn_mse(self.ctx.handle, self.out.data_ptr(), t.data_ptr(), None, B, OUT_H, OUT_W,
                                    CROP_H, CROP_W, self.dout.data_ptr(), C.byref(loss), None))
        self._backward_from_dout()
        return float(loss.value)

    def _backward_from_dout(self):
        B, v, g = self.B, self.views, self.gviews
        H, W = 8 * H0, 8 * W0                                   # input grid of the last layer
        # single-channel output layer
        # (with the channel sums of its input gradient: the last decoder layer's bias gradient)
        self._c(self.lib.sfh_nn_conv1_bwd(self.ctx.handle, self.dout.data_ptr(), self.pre5.data_ptr(), B, H, W, CH,
                                          v["w4"].data_ptr(), SILU, self.dA.data_ptr(), g["c3"].data_ptr(),
                                          g["w4"].data_ptr(), g["bmap"].data_ptr()))
        d = self.dA                                              # gradient w.r.t. pre5
        layers = (("K3", "c3", self.pre5, self.a4, self.pre4),
                  ("K2", "c2", self.pre4, self.a3, self.pre3),
                  ("K1", "c1", self.pre3, self.a2, self.pre2))
        for li, (K, cb, pre_out, a_in, pre_in) in enumerate(layers):
            # d is the gradient w.r.t. act(pre_out) (output grid H x W); the
            # output layer already returned the first one w.r.t. pre5 itself
            Hi, Wi = H // 2, W // 2
            Pi, Po = B * Hi * Wi, B * H * W
            if li > 0:
                # -> gradient w.r.t. pre_out, and its channel sums (the bias gradient)
                self._c(self.lib.sfh_nn_act_bwd_chansum(self.ctx.handle, d.data_ptr(), pre_out.data_ptr(), Po, SILU,
                                                        d.data_ptr(), g[cb].data_ptr()))
            # dK = im2col(d)^T SiLU(pre_in), both operands gathered into LDS by the kernel
            self._c(self.lib.sfh_nn_conv_t_dweight(self.ctx.handle, self.dt, d.data_ptr(), pre_in.data_ptr(), SILU,
                                                   B, Hi, Wi, CH, g[K].data_ptr()))
            # d a_in = im2col(d) K, the column matrix gathered as the product reads it
            dn = self.dB if d is self.dA else self.dA
            self._c(self.lib.sfh_nn_conv_t_dinput(self.ctx.handle, self.dt, d.data_ptr(), B, Hi, Wi, CH,
                                                  self.wT[K].data_ptr(), 16 * CH, CH, dn.data_ptr(), CH))
            d = dn
            H, W = Hi, Wi
        # MLP; d is the gradient w.r.t. a2 (B x LAT)
        mlp = (("W2", "b2", self.pre2, self.pre1, HID1, LAT),
               ("W1", "b1", self.pre1, self.pre0, HID0, HID1),
               ("W0", "b0", self.pre0, None, 16, HID0))
        Bp = self.bp
        for Wn, bn, pre_out, pre_in, nin, nout in mlp:
            # gradient w.r.t. pre_out, its column sums (the bias gradient) and
            # its low-precision copy and transpose, in one launch
            self._c(self.lib.sfh_nn_dense_bwd_prep(self.ctx.handle, self.dt, d.data_ptr(), pre_out.data_ptr(), B, nout,
                                                   SILU, g[bn].data_ptr(), self.dcols.data_ptr(),
                                                   self.dcolsT.data_ptr(), Bp))
            # dW = d^T a_in, a_in transposed by the forward (the network input here)
            if pre_in is None:
                self._transpose(self.x0f, 16, B, 16, self.actT, Bp)
            aT = {"W2": self.a1T, "W1": self.a0T, "W0": self.actT}[Wn]
            self._gemm(self.dcolsT, Bp, aT, Bp, g[Wn], nin, nout, nin, Bp)
            if pre_in is None:
                break
            # d a_in = d W  (low-precision copy of d times the transposed weights)
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

    def images(self, T, nx=CROP_H, ny=CROP_W, out=None):
        """Training images (n x nx x ny, FP32 on the device) of templates T
        (n x nbins device tensor, bin = ix + nx iy)."""
        n = T.shape[0]
        if out is None:
            out = self.torch.empty((n, nx, ny), dtype=self.torch.float32, device=self.dev)
        self.torch.cuda.synchronize()
        for b0 in range(0, n, 65535):
            b1 = min(n, b0 + 65535)
            self._c(self.lib.sfh_nn_log_target(self.ctx.handle, T[b0].data_ptr(), T.stride(0), b1 - b0, nx, ny,
                                               LOG_EPS, LOG_OFFSET, LOG_SCALE, out[b0].data_ptr()))
        self.ctx.synchronize()
        return out

    def train(self, params, images, steps, batch=32, lr=1e-3, lr_min=1e-5, warmup=200, seed=0, log_every=250,
              weight_decay=0.0, kl_weight=1.0, sum_weight=0.0, graph=True, b1=0.9, b2=0.999, eps=1e-8):
        """Adam on shuffled mini-batches of (params: n x 4 physical, images:
        n x 200 x 300 device tensor from ``images``) for ``steps`` steps; the
        learning rate rises linearly over ``warmup`` steps, then follows a
        cosine from lr to lr_min.  The loss is the MSE of the log image plus
        kl_weight times the mean KL(exact || network) per star (the
        log-likelihood a fit loses per star) plus sum_weight times the mean
        squared error of the log template sums.  Each step reads its rows and
        its (lr, bias corrections) from device arrays through a device step
        counter, so after one eager step the step is captured in a HIP graph
        and replayed (graph=False runs every step eagerly; both give the same
        numbers).  The host waits only every ``log_every`` steps, to read the
        means since the last report.  Returns [(step, MSE, KL per star,
        squared log-sum error), ...]."""
        torch = self.torch
        n = images.shape[0]
        if images.shape[1:] != (CROP_H, CROP_W) or len(params) != n:
            raise ValueError("images must be n x 200 x 300 with one parameter row each")
        self._alloc(batch)
        xs = torch.from_numpy(scale_params(params).astype(np.float32)).to(self.dev)
        rng = np.random.default_rng(seed)
        per_epoch = max(1, n // batch)
        nep = -(-steps // per_epoch)
        rows = np.concatenate([rng.permutation(n)[:per_epoch * batch] for _ in range(nep)])
        rows_d = torch.from_numpy(rows.astype(np.int64)).to(self.dev)
        sched = np.empty((steps, 3), dtype=np.float64)
        for k in range(steps):
            if k < warmup:
                rate = lr * (k + 1) / warmup
            else:
                u = (k - warmup) / max(1, steps - warmup)
                rate = lr_min + 0.5 * (lr - lr_min) * (1.0 + np.cos(np.pi * u))
            t = self.step_count + k + 1
            sched[k] = (rate, 1.0 - b1 ** t, 1.0 - b2 ** t)
        sched_d = torch.from_numpy(sched.astype(np.float32).ravel()).to(self.dev)
        counter = torch.zeros(1, dtype=torch.int64, device=self.dev)
        rows_step = torch.zeros(batch, dtype=torch.int64, device=self.dev)
        acc = torch.zeros(3, dtype=torch.float64, device=self.dev)
        self.x0f.zero_()
        torch.cuda.synchronize()
        self.B = batch
        L, h = self.lib, self.ctx.handle

        def body():
            self._c(L.sfh_nn_select_rows(h, rows_d.data_ptr(), counter.data_ptr(), batch, rows_step.data_ptr()))
            self._c(L.sfh_nn_gather_rows(h, xs.data_ptr(), 4, rows_step.data_ptr(), batch, 4, self.x0f.data_ptr(),
                                         16))
            self._forward_device(batch)
            self._c(L.sfh_nn_template_loss(h, self.out.data_ptr(), images.data_ptr(), rows_step.data_ptr(), batch,
                                           OUT_H, OUT_W, CROP_H, CROP_W, LOG_SCALE, LOG_OFFSET, 1.0,
                                           float(kl_weight), float(sum_weight), self.dout.data_ptr(),
                                           acc.data_ptr()))
            self._backward_from_dout()
            self._c(L.sfh_nn_adam_sched(h, self.flat.data_ptr(), self.grad.data_ptr(), self.m.data_ptr(),
                                        self.v.data_ptr(), self.nparams, sched_d.data_ptr(), counter.data_ptr(),
                                        b1, b2, eps, weight_decay))
            self._c(L.sfh_nn_counter_add(h, counter.data_ptr(), 1))

        hist, last = [], 0

        def report(k):
            nonlocal last
            self.ctx.synchronize()
            a = acc.cpu().numpy() / (k - last)
            hist.append((k, float(a[0]), float(a[1]), float(a[2])))
            acc.zero_()
            torch.cuda.synchronize()
            last = k

        body()                                   # also sizes every scratch buffer
        k = 1
        if k % log_every == 0 or k == steps:
            report(k)
        g = None
        if graph and steps > 1:
            self.ctx.synchronize()
            gh = C.c_void_p()
            self._c(L.sfh_ctx_capture_begin(h))
            try:
                body()
            finally:
                rc = L.sfh_ctx_capture_end(h, C.byref(gh))
            self._c(rc)
            g = gh
        try:
            while k < steps:
                nxt = min(steps, (k // log_every + 1) * log_every)
                if g is not None:
                    self._c(L.sfh_graph_launch(g, nxt - k))
                else:
                    for _ in range(nxt - k):
                        body()
                k = nxt
                report(k)
        finally:
            if g is not None:
                L.sfh_graph_destroy(g)
        self.step_count += steps
        self.ctx.synchronize()
        return hist

    def predict(self, params, batch=256, out=None, nx=CROP_H, ny=CROP_W):
        """Templates (n x nbins FP32 device tensor, bin = ix + nx iy, per unit
        stellar mass) for n rows of physical parameters."""
        p = np.atleast_2d(np.asarray(params, dtype=np.float64))
        n = p.shape[0]
        if out is None:
            out = self.torch.empty((n, nx * ny), dtype=self.torch.float32, device=self.dev)
        for b0 in range(0, n, batch):
            b1 = min(n, b0 + batch)
            y = self.forward(p[b0:b1])
            self._c(self.lib.sfh_nn_log_template(self.ctx.handle, y.data_ptr(), b1 - b0, OUT_H, OUT_W, nx, ny,
                                                 LOG_EPS, LOG_OFFSET, LOG_SCALE, out[b0].data_ptr(),
                                                 out.stride(0)))
        self.ctx.synchronize()
        return out

    def save(self, path):
        np.save(path, self.parameters())

    @classmethod
    def load(cls, path, **kws):
        return cls(params=np.load(path), **kws)

    def parameters(self) -> np.ndarray:
        self.ctx.synchronize()
        return self.flat.cpu().numpy()

    def gradient(self) -> np.ndarray:
        self.ctx.synchronize()
        return self.grad.cpu().numpy()


# ---------------------------------------------------------------- data and fills

def sample_parameters(n, seed=0, lo=PARAM_LO, hi=PARAM_HI):
    """n rows of (logAge, [M/H], dmod, error scale): a scrambled stratified
    (Latin hypercube) design over the box lo..hi."""
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
    u = (np.stack([rng.permutation(n) for _ in range(lo.size)], axis=1) + rng.uniform(size=(n, lo.size))) / n
    return lo + u * (hi - lo)


def exact_templates(params, edges=None, npoints=2000, correlation=0.6, mmax_imf=120.0, chunk=1024, out=None,
                    ctx: Context | None = None):
    """partial_cmd_smooth of each (logAge, [M/H], dmod, error scale) row on
    the 200 x 300 grid: synthetic isochrone (seed ISOCHRONE_SEED), point
    preparation and the GPU deposition kernel.  Rows that share their error
    scale with at least ``min_group - 1`` others (a fill grid) are prepared on
    the device (partial_cmd_smooth_batch: the scale s enters the photometric
    error exactly, s (0.01 + e^(0.6 (m - 28))) = 0.01 s + e^(0.6 (m - 28 +
    ln(s) / 0.6))); the others (a sampled design, every scale different) by
    the host preparation.  Returns an n x nbins FP32 device tensor (bin = ix +
    nx iy, per unit stellar mass)."""
    import torch
    from . import synth
    from .templates import PointBatch, deposit_batch, partial_cmd_smooth_batch, ExpPhotErr
    ctx = ctx or default_context()
    edges = edges or synth.edges_200x300()
    p = np.atleast_2d(np.asarray(params, dtype=np.float64))
    n = p.shape[0]
    nbins = (edges[0].length - 1) * (edges[1].length - 1)
    if out is None:
        out = torch.empty((n, nbins), dtype=torch.float32, device="cuda")
    min_group = 8
    scales, inv, counts = np.unique(p[:, 3], return_inverse=True, return_counts=True)
    imf = synth.Kroupa2001(0.08, mmax_imf)
    for k in np.flatnonzero(counts >= min_group):
        s = float(scales[k])
        err = ExpPhotErr(math.e, 0.6, 28.0 - math.log(s) / 0.6, 0.01 * s)
        rows = np.flatnonzero(inv == k)
        for c0 in range(0, rows.size, chunk):
            r = rows[c0:c0 + chunk]
            isos = [synth.isochrone(q[0], q[1], npoints, ISOCHRONE_SEED, mmax_imf=mmax_imf) for q in p[r]]
            d = partial_cmd_smooth_batch(isos, [err, err], 1, (0, 1), imf,
                                         [synth.COMPLETENESS, synth.COMPLETENESS], dmod=[float(v) for v in p[r, 2]],
                                         normalize_value=1.0, mean_mass=imf.mean(),
                                         correlation=0.0 if correlation is None else correlation, edges=edges,
                                         ctx=ctx)
            out.index_copy_(0, torch.from_numpy(r.astype(np.int64)).to(out.device), d)
            del d
    host = np.flatnonzero(counts[inv] < min_group)
    for c0 in range(0, host.size, chunk):
        r = host[c0:c0 + chunk]
        pts = [synth.template_points(synth.TemplateSpec(q[0], q[1], ISOCHRONE_SEED), npoints, edges, mmax_imf,
                                     dmod=q[2], correlation=correlation, err_scale=q[3]) for q in p[r]]
        batch = PointBatch([q for q, _ in pts], edges,
                           cmult_list=[c for _, c in pts] if correlation is not None else None)
        if r.size and r[-1] - r[0] == r.size - 1:
            deposit_batch(batch, out=out[int(r[0]):int(r[-1]) + 1], ctx=ctx)
        else:
            tmp = torch.empty((r.size, nbins), dtype=torch.float32, device=out.device)
            deposit_batch(batch, out=tmp, ctx=ctx)
            out.index_copy_(0, torch.from_numpy(r.astype(np.int64)).to(out.device), tmp)
            del tmp
        ctx.synchronize()
    ctx.synchronize()
    return out


def template_delta_lnL(exact, approx, n_stars=1.0, floor=LOG_EPS, ctx: Context | None = None):
    """Per template (rows of the two n x nbins device tensors): the
    log-likelihood lost by fitting n_stars stars drawn from ``exact`` with
    ``approx`` (floored at ``floor``) instead, amplitude free:
    n_stars KL(exact || approx).  Returns (dlnL, sum exact, sum approx)."""
    import torch
    ctx = ctx or default_context()
    n, nb = exact.shape
    res = torch.empty(3 * n, dtype=torch.float64, device=exact.device)
    torch.cuda.synchronize()
    ctx.check(ctx.lib.sfh_nn_template_kl(ctx.handle, exact.data_ptr(), exact.stride(0), approx.data_ptr(),
                                         approx.stride(0), n, nb, float(floor), res.data_ptr()))
    ctx.synchronize()
    r = res.cpu().numpy().reshape(n, 3)
    return np.asarray(n_stars, dtype=np.float64) * r[:, 0], r[:, 1], r[:, 2]


class FillResult:
    """Templates of a filled grid and how each was obtained."""

    def __init__(self, templates, log_ages, mhs, exact_mask, checked_mask, check_dlnL, seconds):
        self.templates = templates        # (na * nm) x nbins device tensor, row i * nm + j
        self.log_ages, self.mhs = log_ages, mhs
        self.exact_mask = exact_mask      # computed with partial_cmd_smooth (checks and fallbacks)
        self.checked_mask = checked_mask  # lattice points where the network was checked
        self.check_dlnL = check_dlnL      # dlnL of the network at each checked point (nan elsewhere)
        self.seconds = seconds

    @property
    def fallback_rate(self):
        """Fraction of templates computed with partial_cmd_smooth."""
        return float(self.exact_mask.mean())


def _lattice(n, stride):
    idx = list(range(0, n, stride))
    if idx[-1] != n - 1:
        idx.append(n - 1)
    return np.asarray(idx)


def _segments(idx):
    """Consecutive pairs of checked indices along one axis; an axis with a
    single checked index is one degenerate segment."""
    if idx.size == 1:
        return [(int(idx[0]), int(idx[0]))]
    return [(int(idx[k]), int(idx[k + 1])) for k in range(idx.size - 1)]


def fallback_mask(dgrid, ia, im, threshold):
    """Grid points to compute exactly: every point of a lattice cell (a
    rectangle between consecutive checked ages and metallicities, or a segment
    where an axis has a single grid point) with a checked corner whose dlnL
    exceeds ``threshold``."""
    bad = np.zeros(dgrid.shape, dtype=bool)
    for a0, a1 in _segments(np.asarray(ia)):
        for b0, b1 in _segments(np.asarray(im)):
            c = dgrid[np.ix_([a0, a1], [b0, b1])]
            if np.any(c > threshold):
                bad[a0:a1 + 1, b0:b1 + 1] = True
    return bad


def fill_templates(surrogate: TemplateSurrogate, log_ages, mhs, dmod, err_scale, n_stars, threshold=0.5,
                   check_stride=(5, 4), batch=256, exact_kws=None):
    """Templates of the grid log_ages x mhs at distance modulus ``dmod`` and
    error scale ``err_scale``.  The network fills every point; on the check
    lattice (every check_stride-th age and metallicity, ends included) the
    exact template (partial_cmd_smooth) is also built and replaces the
    network's, and dlnL = n_stars KL(exact || network) is measured.  Every
    grid point of a lattice cell with a corner dlnL above ``threshold`` is
    computed with partial_cmd_smooth as well.  Returns a FillResult."""
    import time
    la, mh = np.asarray(log_ages, dtype=np.float64), np.asarray(mhs, dtype=np.float64)
    na, nm = la.size, mh.size
    grid = np.stack(np.meshgrid(la, mh, indexing="ij"), axis=-1).reshape(-1, 2)
    full = np.column_stack([grid, np.full(na * nm, float(dmod)), np.full(na * nm, float(err_scale))])
    kws = exact_kws or {}
    t0 = time.perf_counter()
    T = surrogate.predict(full, batch=batch)
    t_net = time.perf_counter() - t0
    ia, im = _lattice(na, check_stride[0]), _lattice(nm, check_stride[1])
    checked = np.zeros((na, nm), dtype=bool)
    checked[np.ix_(ia, im)] = True
    cidx = np.flatnonzero(checked.ravel())
    E = exact_templates(full[cidx], ctx=surrogate.ctx, **kws)
    dl, _, _ = template_delta_lnL(E, T[cidx], n_stars, ctx=surrogate.ctx)
    dgrid = np.full((na, nm), np.nan)
    dgrid.ravel()[cidx] = dl
    bad = fallback_mask(dgrid, ia, im, threshold)
    T[cidx] = E
    del E
    fb = np.flatnonzero(bad.ravel() & ~checked.ravel())
    if fb.size:
        T[fb] = exact_templates(full[fb], ctx=surrogate.ctx, **kws)
    exact = checked | bad
    surrogate.ctx.synchronize()
    secs = {"network": t_net, "total": time.perf_counter() - t0}
    return FillResult(T, la, mh, exact.ravel(), checked.ravel(), dgrid.ravel(), secs)
