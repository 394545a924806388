"""GPU kernels of the learned template surrogate against the FP64 oracle
network (oracle/surrogate_ref.py): loss and every parameter gradient, with
FP32 operands on the matrix cores (tight) and with BF16 operands (the
training precision, loose), and a short training run that lowers the loss."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import surrogate_ref as R   # noqa: E402

pytestmark = pytest.mark.gpu


def _case(seed=0, B=3):
    from sfh_amd.surrogate import init_params
    rng = np.random.default_rng(seed)
    flat = init_params(seed=seed + 1)
    flat = flat + rng.normal(0, 0.01, flat.size).astype(np.float32)
    x = rng.uniform(-1, 1, (B, 4)).astype(np.float32)
    target = rng.normal(0, 1, (B, R.CROP_H, R.CROP_W)).astype(np.float32)
    return flat, x, target


def _per_tensor_err(g, ref):
    out, off = {}, 0
    for name, shape in R.LAYOUT:
        n = int(np.prod(shape))
        a, b = g[off:off + n], ref[off:off + n]
        out[name] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
        off += n
    return out


@pytest.mark.parametrize("dtype,tol_loss,tol_grad", [("fp32", 1e-5, 1e-5), ("bf16", 2e-2, 2e-2)])
def test_loss_and_gradients_against_oracle(sfh, gpu_ctx, dtype, tol_loss, tol_grad):
    import torch
    from sfh_amd.surrogate import TemplateSurrogate
    flat, x, target = _case()
    net = TemplateSurrogate(params=flat, dtype=dtype, max_batch=3)
    out = net.forward(x, raw=True)
    loss = net.backward(torch.from_numpy(target).cuda())
    g = net.gradient()
    rloss, rg, rout = R.loss_and_grad(flat.astype(np.float64), x.astype(np.float64), target.astype(np.float64))
    oerr = float(np.max(np.abs(out.cpu().numpy() - rout)) / np.max(np.abs(rout)))
    errs = _per_tensor_err(g, rg)
    print(dtype, f"output max err {oerr:.2e}, loss {loss:.6g} vs {rloss:.6g}", {k: f"{v:.1e}" for k, v in errs.items()})
    assert abs(loss - rloss) <= tol_loss * abs(rloss)
    assert max(errs.values()) <= tol_grad


def test_training_lowers_loss(sfh, gpu_ctx):
    import torch
    from sfh_amd.surrogate import TemplateSurrogate
    rng = np.random.default_rng(5)
    B = 16
    x = rng.uniform(-1, 1, (B, 4)).astype(np.float32)
    # smooth targets that depend on the inputs
    yy, xx = np.mgrid[0:R.CROP_H, 0:R.CROP_W] / 100.0
    target = np.stack([np.sin(xx * (1 + p[0])) * np.cos(yy * (1 + p[1])) + p[2] for p in x]).astype(np.float32)
    t = torch.from_numpy(target).cuda()
    net = TemplateSurrogate(dtype="bf16", seed=2, max_batch=B)
    losses = []
    for _ in range(60):
        net.forward(x, raw=True)
        losses.append(net.backward(t))
        net.adam_step(lr=2e-3)
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.3 * losses[0]
