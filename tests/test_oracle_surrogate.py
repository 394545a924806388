"""Pins of the FP64 surrogate network restatement (oracle/surrogate_ref.py):
its gradients against central finite differences of its own loss, on a
random subset of every parameter tensor."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import surrogate_ref as R   # noqa: E402


def test_gradient_finite_differences(sfh):
    from sfh_amd.surrogate import init_params
    rng = np.random.default_rng(0)
    flat = init_params(seed=3).astype(np.float64)
    flat += rng.normal(0, 0.01, flat.size)      # non-zero biases too
    x = rng.uniform(-1, 1, (2, 4))
    target = rng.normal(0, 1, (2, R.CROP_H, R.CROP_W))
    loss, g, _ = R.loss_and_grad(flat, x, target)
    off = 0
    for name, shape in R.LAYOUT:
        n = int(np.prod(shape))
        idx = off + rng.choice(n if name != "W0" else n, size=4, replace=False)
        if name == "W0":   # only the 4 live input columns are parameters
            idx = off + np.array([0, 1 * 16 + 2, 7 * 16 + 3, 100 * 16 + 1])
        for i in idx:
            h = 1e-5 * max(1.0, abs(flat[i]))
            fp, fm = flat.copy(), flat.copy()
            fp[i] += h
            fm[i] -= h
            fd = (R.loss_and_grad(fp, x, target, False)[0] - R.loss_and_grad(fm, x, target, False)[0]) / (2 * h)
            assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), (name, i - off)
        off += n
    assert off == g.size == flat.size


def test_transposed_convolution_shape_and_single_pixel():
    # one input pixel spreads a 4x4 kernel onto the outputs 2y-1..2y+2
    x = np.zeros((1, 3, 3, 1))
    x[0, 1, 1, 0] = 1.0
    K = np.arange(16, dtype=np.float64).reshape(16, 1)
    out = R.convT(x, K, 1)
    assert out.shape == (1, 6, 6, 1)
    assert np.array_equal(out[0, 1:5, 1:5, 0], np.arange(16.0).reshape(4, 4))
    assert out.sum() == K.sum()


def test_transposed_convolution_against_direct_loops():
    """The oracle's scatter form equals the definition written as loops, with
    weight rows (ky*4 + kx)*co + o (tap-major)."""
    rng = np.random.default_rng(0)
    B, H, W, ci, co = 2, 3, 4, 5, 3
    x = rng.normal(size=(B, H, W, ci))
    K = rng.normal(size=(16 * co, ci))
    ref = np.zeros((B, 2 * H, 2 * W, co))
    for b in range(B):
        for y in range(H):
            for xx in range(W):
                for ky in range(4):
                    for kx in range(4):
                        Y, X = 2 * y - 1 + ky, 2 * xx - 1 + kx
                        if 0 <= Y < 2 * H and 0 <= X < 2 * W:
                            for o in range(co):
                                ref[b, Y, X, o] += x[b, y, xx] @ K[(ky * 4 + kx) * co + o]
    assert np.max(np.abs(R.convT(x, K, co) - ref)) < 1e-12
