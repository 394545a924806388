"""GPU primitives of the surrogate pipeline against NumPy restatements
(template <-> log-image transposes, the per-template KL behind dlnL, row
gathers, the indexed MSE, split-K products, column sums) and the bookkeeping
of fill_templates against exact templates built directly; the output layer
against the oracle's transposed convolution."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import surrogate_ref as R   # noqa: E402

pytestmark = pytest.mark.gpu

NX, NY = 200, 300


def _templates(n, seed):
    rng = np.random.default_rng(seed)
    T = rng.lognormal(-14.0, 3.0, (n, NX * NY)).astype(np.float32)
    T[rng.uniform(size=T.shape) < 0.3] = 0.0
    return T


def test_log_image_roundtrip(sfh, gpu_ctx):
    import torch
    from sfh_amd import surrogate as S
    T = _templates(3, 1)
    net = S.TemplateSurrogate(dtype="bf16", max_batch=3)
    img = net.images(torch.from_numpy(T).cuda()).cpu().numpy()
    ref = ((np.log(T.astype(np.float64) + S.LOG_EPS) - S.LOG_OFFSET) / S.LOG_SCALE).reshape(3, NY, NX)
    ref = ref.transpose(0, 2, 1)
    assert np.max(np.abs(img - ref)) <= 2e-6 * np.max(np.abs(ref))
    # network-output layout (208 x 304, junk outside the crop) back to templates
    out = np.random.default_rng(2).normal(0, 1, (3, S.OUT_H, S.OUT_W)).astype(np.float32)
    out[:, :NX, :NY] = img
    Td = torch.empty((3, NX * NY), dtype=torch.float32, device="cuda")
    o = torch.from_numpy(out).cuda()
    torch.cuda.synchronize()
    gpu_ctx.check(gpu_ctx.lib.sfh_nn_log_template(gpu_ctx.handle, o.data_ptr(), 3, S.OUT_H, S.OUT_W, NX, NY,
                                                  S.LOG_EPS, S.LOG_OFFSET, S.LOG_SCALE, Td.data_ptr(), NX * NY))
    gpu_ctx.synchronize()
    back = Td.cpu().numpy().astype(np.float64)
    assert np.all(back >= 0.0)
    assert np.all(np.abs(back - T) <= 2e-5 * T + 1e-16)


def test_template_kl(sfh, gpu_ctx):
    import torch
    from sfh_amd import surrogate as S
    E = _templates(5, 3)
    Sx = E * np.random.default_rng(4).lognormal(0.0, 0.2, E.shape).astype(np.float32)
    Sx[:, ::7] = 0.0                       # zeros in the approximation meet the floor
    E[4] = 0.0                             # an empty exact template gives 0
    dl, se, ss = S.template_delta_lnL(torch.from_numpy(E).cuda(), torch.from_numpy(Sx).cuda(), n_stars=250.0)
    e, s = E.astype(np.float64), np.maximum(Sx.astype(np.float64), S.LOG_EPS)
    for k in range(5):
        sek, ssk = e[k].sum(), s[k].sum()
        if sek == 0:
            assert dl[k] == 0.0
            continue
        pe, ps = e[k] / sek, s[k] / ssk
        m = pe > 0
        kl = np.sum(pe[m] * np.log(pe[m] / ps[m]))
        assert abs(dl[k] - 250.0 * kl) <= 1e-9 * 250.0 * kl + 1e-12
        assert abs(se[k] - sek) <= 1e-12 * sek and abs(ss[k] - ssk) <= 1e-12 * ssk
    # identical templates above the floor lose nothing
    E2 = torch.from_numpy(E[:4] + np.float32(1e-9)).cuda()
    dl0, _, _ = S.template_delta_lnL(E2, E2, 1e4)
    assert np.all(np.abs(dl0) < 1e-9)


def test_gather_and_indexed_mse(sfh, gpu_ctx):
    import ctypes as C
    import torch
    rng = np.random.default_rng(7)
    src = rng.normal(size=(50, 4)).astype(np.float32)
    rows = rng.integers(0, 50, 9).astype(np.int64)
    s_d, r_d = torch.from_numpy(src).cuda(), torch.from_numpy(rows).cuda()
    dst = torch.full((9, 16), 7.0, device="cuda")
    torch.cuda.synchronize()
    gpu_ctx.check(gpu_ctx.lib.sfh_nn_gather_rows(gpu_ctx.handle, s_d.data_ptr(), 4, r_d.data_ptr(), 9, 4,
                                                 dst.data_ptr(), 16))
    gpu_ctx.synchronize()
    d = dst.cpu().numpy()
    assert np.array_equal(d[:, :4], src[rows]) and np.all(d[:, 4:] == 7.0)
    # MSE of 3 outputs (H2 x W2) against target rows picked by index
    B, H2, W2, R, S_ = 3, 12, 10, 9, 7
    out = rng.normal(size=(B, H2, W2)).astype(np.float32)
    tgt = rng.normal(size=(6, R, S_)).astype(np.float32)
    idx = np.array([4, 0, 4], dtype=np.int64)
    o_d, t_d, i_d = torch.from_numpy(out).cuda(), torch.from_numpy(tgt).cuda(), torch.from_numpy(idx).cuda()
    dout = torch.empty_like(o_d)
    acc = torch.full((1,), 0.5, dtype=torch.float64, device="cuda")
    loss = C.c_float()
    torch.cuda.synchronize()
    gpu_ctx.check(gpu_ctx.lib.sfh_nn_mse(gpu_ctx.handle, o_d.data_ptr(), t_d.data_ptr(), i_d.data_ptr(), B, H2, W2,
                                         R, S_, dout.data_ptr(), C.byref(loss), acc.data_ptr()))
    diff = np.zeros_like(out, dtype=np.float64)
    diff[:, :R, :S_] = out[:, :R, :S_].astype(np.float64) - tgt[idx]
    ref = np.sum(diff ** 2) / (B * R * S_)
    assert abs(loss.value - ref) <= 1e-5 * ref
    assert abs(float(acc.item()) - 0.5 - ref) <= 1e-5 * ref
    assert np.allclose(dout.cpu().numpy(), 2 * diff / (B * R * S_), rtol=1e-5, atol=1e-9)


@pytest.mark.parametrize("dtype", [0, 1])
@pytest.mark.parametrize("M,N,K", [(200, 40, 20000), (16, 32, 50016), (300, 130, 64), (1000, 24, 48), (777, 32, 2048),
                                   (70000, 16, 32), (33, 200, 16), (129, 257, 80)])
def test_gemm_split_k(sfh, gpu_ctx, dtype, M, N, K):
    import torch
    rng = np.random.default_rng(M + N + K)
    A = rng.normal(size=(M, K)).astype(np.float32)
    Bm = rng.normal(size=(N, K)).astype(np.float32)
    C0 = rng.normal(size=(M, N)).astype(np.float32)
    tdt = torch.bfloat16 if dtype == 0 else torch.float32
    a_d, b_d = torch.from_numpy(A).to(tdt).cuda(), torch.from_numpy(Bm).to(tdt).cuda()
    c_d = torch.from_numpy(C0).cuda()
    torch.cuda.synchronize()
    gpu_ctx.check(gpu_ctx.lib.sfh_nn_gemm_nt(gpu_ctx.handle, dtype, a_d.data_ptr(), K, b_d.data_ptr(), K,
                                             c_d.data_ptr(), N, M, N, K, 0.5, 2.0))
    gpu_ctx.synchronize()
    Aq = a_d.float().cpu().numpy().astype(np.float64)
    Bq = b_d.float().cpu().numpy().astype(np.float64)
    ref = 0.5 * Aq @ Bq.T + 2.0 * C0
    scale = 0.5 * np.sqrt(K) + 2.0 * np.abs(C0)
    assert np.max(np.abs(c_d.cpu().numpy() - ref) / scale) < 3e-5


@pytest.mark.parametrize("rows,cols", [(70000, 32), (32, 63232), (5, 300)])
def test_colsum(sfh, gpu_ctx, rows, cols):
    import torch
    rng = np.random.default_rng(rows)
    X = rng.normal(size=(rows, cols)).astype(np.float32)
    out0 = rng.normal(size=cols).astype(np.float32)
    x_d, o_d = torch.from_numpy(X).cuda(), torch.from_numpy(out0).cuda()
    torch.cuda.synchronize()
    gpu_ctx.check(gpu_ctx.lib.sfh_nn_colsum(gpu_ctx.handle, x_d.data_ptr(), cols, rows, cols, 0.5, o_d.data_ptr()))
    gpu_ctx.synchronize()
    ref = X.astype(np.float64).sum(0) + 0.5 * out0
    assert np.max(np.abs(o_d.cpu().numpy() - ref)) <= 1e-5 * np.sqrt(rows) + 1e-6


@pytest.mark.parametrize("act", [0, 1])
@pytest.mark.parametrize("B,H,W,Ci", [(2, 13, 37, 32), (1, 8, 32, 32), (3, 5, 7, 32), (2, 6, 9, 16),
                                      (16, 104, 152, 32)])
def test_output_layer_against_oracle(sfh, gpu_ctx, B, H, W, Ci, act):
    """The single-channel output layer (32 input channels: direct pixel-grid
    kernels; other widths: the matrix-product form), on act(in) with act the
    identity or SiLU, against the FP64 oracle's transposed convolution and its
    gradients, on grids that leave partial input tiles and, at 16 x 104 x 152,
    more tiles than workgroups (each workgroup walks several)."""
    import torch
    rng = np.random.default_rng(B * 1000 + H * 10 + W)
    a = rng.normal(size=(B, H, W, Ci)).astype(np.float32)
    w = rng.normal(size=(16, Ci)).astype(np.float32)
    bmap = rng.normal(size=(2 * H * 2 * W,)).astype(np.float32)
    dout = rng.normal(size=(B, 2 * H, 2 * W)).astype(np.float32)
    a_d, w_d, bm_d, do_d = (torch.from_numpy(v).cuda() for v in (a, w, bmap, dout))
    out_d = torch.empty((B, 2 * H, 2 * W), dtype=torch.float32, device="cuda")
    din_d = torch.empty_like(a_d)
    dw_d = torch.empty_like(w_d)
    dbm_d = torch.empty_like(bm_d)
    dsum_d = torch.full((Ci,), np.nan, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    gpu_ctx.check(lib.sfh_nn_conv1_fwd(h, a_d.data_ptr(), B, H, W, Ci, w_d.data_ptr(), bm_d.data_ptr(), act,
                                       out_d.data_ptr()))
    gpu_ctx.check(lib.sfh_nn_conv1_bwd(h, do_d.data_ptr(), a_d.data_ptr(), B, H, W, Ci, w_d.data_ptr(), act,
                                       din_d.data_ptr(), dsum_d.data_ptr(), dw_d.data_ptr(), dbm_d.data_ptr()))
    gpu_ctx.synchronize()
    x64, w64 = a.astype(np.float64), w.astype(np.float64)
    a64 = R.silu(x64) if act else x64
    ref = R.convT(a64, w64, 1)[..., 0] + bmap.reshape(2 * H, 2 * W)
    rdin, rdw = R.convT_bwd(a64, w64, 1, dout.astype(np.float64)[..., None])
    if act:
        rdin = rdin * R.dsilu(x64)
    def rel(x, r):
        return float(np.max(np.abs(x - r)) / np.max(np.abs(r)))
    assert rel(out_d.cpu().numpy(), ref) < 1e-6
    assert rel(din_d.cpu().numpy(), rdin) < 1e-6
    assert rel(dw_d.cpu().numpy(), rdw) < 1e-6
    assert rel(dbm_d.cpu().numpy(), dout.astype(np.float64).sum(0).ravel()) < 1e-6
    # channel sums of din (the previous layer's bias gradient) from the same pass
    rsum = rdin.reshape(-1, Ci).sum(0)
    assert np.all(np.abs(dsum_d.cpu().numpy() - rsum) <= 1e-5 * np.abs(rdin).reshape(-1, Ci).sum(0) + 1e-6)


@pytest.mark.parametrize("dtype", [0, 1])
@pytest.mark.parametrize("B,Hi,Wi,Co,Ci", [(2, 13, 19, 32, 32), (2, 5, 7, 16, 48), (3, 26, 38, 32, 32),
                                           (1, 5, 7, 8, 32), (32, 52, 76, 32, 32)])
def test_conv_t_dinput_equals_im2col_product(sfh, gpu_ctx, dtype, B, Hi, Wi, Co, Ci):
    """The input gradient with the column matrix gathered as it is read is the
    product of the written column matrix (sfh_nn_im2col) with the weights: to
    the bit where both run K in one sequence (256 workgroups or more, like the
    32 x 52 x 76 layer; more than 64 rows and K = 16 Co < 1024, so the written
    product is not split either), and where the gathered product splits the
    taps into four slices (fewer workgroups than CUs) to FP32 rounding of the
    same rounded operands' FP64 product."""
    import torch
    rng = np.random.default_rng(B * 100 + Hi + Wi)
    tdt = torch.bfloat16 if dtype == 0 else torch.float32
    d = torch.from_numpy(rng.normal(size=(B, 2 * Hi, 2 * Wi, Co)).astype(np.float32)).cuda()
    wT = torch.from_numpy(rng.normal(size=(Ci, 16 * Co)).astype(np.float32)).cuda().to(tdt)
    P = B * Hi * Wi
    cols = torch.empty((P, 16 * Co), dtype=tdt, device="cuda")
    ref = torch.empty((P, Ci), dtype=torch.float32, device="cuda")
    out = torch.full((P, Ci), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    gpu_ctx.check(lib.sfh_nn_im2col(h, dtype, d.data_ptr(), B, Hi, Wi, Co, cols.data_ptr(), 16 * Co))
    gpu_ctx.check(lib.sfh_nn_gemm_nt(h, dtype, cols.data_ptr(), 16 * Co, wT.data_ptr(), 16 * Co, ref.data_ptr(), Ci,
                                     P, Ci, 16 * Co, 1.0, 0.0))
    gpu_ctx.check(lib.sfh_nn_conv_t_dinput(h, dtype, d.data_ptr(), B, Hi, Wi, Co, wT.data_ptr(), 16 * Co, Ci,
                                           out.data_ptr(), Ci))
    gpu_ctx.synchronize()
    if -(-P // 256) * -(-Ci // 32) >= 256:
        assert torch.equal(out, ref)
    else:
        ref64 = cols.double() @ wT.double().T
        scale = float(ref64.abs().max())
        assert float((out.double() - ref64).abs().max()) <= 2e-6 * scale
        assert float((ref.double() - ref64).abs().max()) <= 2e-6 * scale


@pytest.mark.parametrize("act", [0, 1])
@pytest.mark.parametrize("B,Hi,Wi", [(2, 13, 19), (1, 5, 7), (3, 26, 38)])
def test_conv_t_dweight(sfh, gpu_ctx, B, Hi, Wi, act):
    """The decoder weight gradient with its operands gathered and transposed in
    LDS: FP32 operands against the oracle's convT_bwd (FP64); BF16 operands
    against the written transposed column matrix times the transposed
    activations (same rounded operands, other summation order)."""
    import torch
    C = 32
    rng = np.random.default_rng(B * 100 + Hi + Wi + act)
    d = rng.normal(size=(B, 2 * Hi, 2 * Wi, C)).astype(np.float32)
    pre = rng.normal(size=(B, Hi, Wi, C)).astype(np.float32)
    d_d, pre_d = torch.from_numpy(d).cuda(), torch.from_numpy(pre).cuda()
    out = {dt: torch.empty((16 * C, C), dtype=torch.float32, device="cuda") for dt in (0, 1)}
    P, Pp = B * Hi * Wi, (B * Hi * Wi + 15) // 16 * 16
    colsT = torch.zeros((16 * C, Pp), dtype=torch.bfloat16, device="cuda")
    actT = torch.zeros((C, Pp), dtype=torch.bfloat16, device="cuda")
    ref0 = torch.empty((16 * C, C), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    for dt in (0, 1):
        gpu_ctx.check(lib.sfh_nn_conv_t_dweight(h, dt, d_d.data_ptr(), pre_d.data_ptr(), act, B, Hi, Wi, C,
                                                out[dt].data_ptr()))
    gpu_ctx.check(lib.sfh_nn_im2col_t(h, 0, d_d.data_ptr(), B, Hi, Wi, C, colsT.data_ptr(), Pp))
    gpu_ctx.check(lib.sfh_nn_transpose(h, 0, pre_d.data_ptr(), C, P, C, actT.data_ptr(), Pp, C, act))
    gpu_ctx.check(lib.sfh_nn_gemm_nt(h, 0, colsT.data_ptr(), Pp, actT.data_ptr(), Pp, ref0.data_ptr(), C, 16 * C, C,
                                     Pp, 1.0, 0.0))
    gpu_ctx.synchronize()
    x64 = pre.astype(np.float64)
    a64 = R.silu(x64) if act else x64
    _, rdk = R.convT_bwd(a64, np.zeros((16 * C, C)), C, d.astype(np.float64))
    def rel(x, r):
        return float(np.max(np.abs(x - r)) / np.max(np.abs(r)))
    assert rel(out[1].cpu().numpy(), rdk) < 2e-5
    assert rel(out[0].cpu().numpy(), ref0.cpu().numpy().astype(np.float64)) < 1e-4
    assert rel(out[0].cpu().numpy(), rdk) < 2e-2


@pytest.mark.parametrize("dtype", [0, 1])
@pytest.mark.parametrize("B,Hi,Wi", [(2, 13, 19), (1, 5, 7), (3, 26, 38)])
def test_conv_t_fwd_against_products(sfh, gpu_ctx, dtype, B, Hi, Wi):
    """The decoder forward by output parity class (gathered operands, bias and
    SiLU in the epilogue) against the column product, col2im and bias/SiLU
    kernels it replaces (same rounded operands), and the FP32 form against the
    oracle's convT."""
    import torch
    C = 32
    rng = np.random.default_rng(B * 10 + Hi + Wi)
    tdt = torch.bfloat16 if dtype == 0 else torch.float32
    x = rng.normal(size=(B, Hi, Wi, C)).astype(np.float32)
    w = (0.2 * rng.normal(size=(16 * C, C))).astype(np.float32)
    bias = rng.normal(size=C).astype(np.float32)
    x_d = torch.from_numpy(x).cuda().to(tdt)
    w_d = torch.from_numpy(w).cuda().to(tdt)
    b_d = torch.from_numpy(bias).cuda()
    P, H2, W2 = B * Hi * Wi, 2 * Hi, 2 * Wi
    z = torch.empty((P, 16 * C), dtype=torch.float32, device="cuda")
    pre_ref = torch.empty((B, H2, W2, C), dtype=torch.float32, device="cuda")
    a_ref = torch.empty((B, H2, W2, C), dtype=tdt, device="cuda")
    pre = torch.full((B, H2, W2, C), float("nan"), dtype=torch.float32, device="cuda")
    a = torch.empty((B, H2, W2, C), dtype=tdt, device="cuda")
    torch.cuda.synchronize()
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    gpu_ctx.check(lib.sfh_nn_gemm_nt(h, dtype, x_d.data_ptr(), C, w_d.data_ptr(), C, z.data_ptr(), 16 * C, P, 16 * C,
                                     C, 1.0, 0.0))
    gpu_ctx.check(lib.sfh_nn_col2im(h, z.data_ptr(), 16 * C, B, Hi, Wi, C, b_d.data_ptr(), pre_ref.data_ptr()))
    gpu_ctx.check(lib.sfh_nn_bias_act(h, dtype, pre_ref.data_ptr(), C, None, B * H2 * W2, C, 1, None,
                                      a_ref.data_ptr(), C))
    gpu_ctx.check(lib.sfh_nn_conv_t_fwd(h, dtype, x_d.data_ptr(), B, Hi, Wi, C, w_d.data_ptr(), b_d.data_ptr(),
                                        pre.data_ptr(), a.data_ptr()))
    gpu_ctx.synchronize()
    def rel(u, v):
        u, v = np.asarray(u, dtype=np.float64), np.asarray(v, dtype=np.float64)
        return float(np.max(np.abs(u - v)) / np.max(np.abs(v)))
    assert rel(pre.cpu().numpy(), pre_ref.cpu().numpy()) < 1e-5
    assert rel(a.float().cpu().numpy(), a_ref.float().cpu().numpy()) < (1e-2 if dtype == 0 else 1e-5)
    if dtype == 1:
        ref = R.convT(x.astype(np.float64), w.astype(np.float64), C) + bias
        assert rel(pre.cpu().numpy(), ref) < 1e-5


def test_fill_templates_bookkeeping(sfh, gpu_ctx):
    """An untrained network fails every check (all templates exact); an
    infinite threshold keeps the network everywhere off the lattice."""
    import torch
    from sfh_amd import surrogate as S
    net = S.TemplateSurrogate(dtype="bf16", seed=3, max_batch=32)
    la = np.array([8.0, 8.1, 8.2, 8.3, 8.4])
    mh = np.array([-1.5, -1.0, -0.5, 0.0])
    full = np.column_stack([np.repeat(la, 4), np.tile(mh, 5), np.full(20, 26.0), np.full(20, 1.2)])
    E = S.exact_templates(full).cpu().numpy()
    r = S.fill_templates(net, la, mh, 26.0, 1.2, n_stars=100.0, check_stride=(2, 2))
    assert r.exact_mask.all() and r.fallback_rate == 1.0
    chk = np.zeros((5, 4), dtype=bool)
    chk[np.ix_([0, 2, 4], [0, 2, 3])] = True
    assert np.array_equal(r.checked_mask, chk.ravel())
    assert np.all(np.isfinite(r.check_dlnL[r.checked_mask])) and np.all(r.check_dlnL[r.checked_mask] > 0.5)
    T = r.templates.cpu().numpy()
    assert np.allclose(T, E, rtol=1e-6, atol=0)
    r2 = S.fill_templates(net, la, mh, 26.0, 1.2, n_stars=100.0, threshold=np.inf, check_stride=(2, 2))
    assert np.array_equal(r2.exact_mask, chk.ravel())
    T2 = r2.templates.cpu().numpy()
    P = net.predict(full).cpu().numpy()
    assert np.allclose(T2[chk.ravel()], E[chk.ravel()], rtol=1e-6, atol=0)
    assert np.array_equal(T2[~chk.ravel()], P[~chk.ravel()])


def test_train_lowers_loss_on_exact_templates(sfh, gpu_ctx):
    from sfh_amd import surrogate as S
    P = S.sample_parameters(96, seed=5, lo=[8.0, -1.5, 25.0, 0.8], hi=[9.0, -0.5, 27.0, 1.5])
    net = S.TemplateSurrogate(dtype="bf16", seed=1, max_batch=16)
    img = net.images(S.exact_templates(P))
    hist = net.train(P, img, steps=300, batch=16, lr=1e-3, warmup=30, log_every=100, kl_weight=0.1)
    assert len(hist) == 3
    assert hist[-1][1] < 0.2 * hist[0][1] and hist[-1][2] < 0.5 * hist[0][2]


def _tl_ref(y, t, sc, off):
    """FP64 restatement: MSE, KL(p_e || p_s) and squared log-sum error of one image pair."""
    sv, ev = np.exp(np.minimum(y, 15.0) * sc + off), np.exp(t * sc + off)
    pe, ps = ev / ev.sum(), sv / sv.sum()
    return (np.mean((y - t) ** 2), np.sum(pe * (np.log(pe) - np.log(ps))),
            (np.log(sv.sum()) - np.log(ev.sum())) ** 2)


def test_template_loss_against_restatement_and_differences(sfh, gpu_ctx):
    import torch
    rng = np.random.default_rng(11)
    B, H2, W2, R, S_ = 2, 14, 12, 11, 9
    sc, off, wm, wk, ws = 5.0, -20.0, 0.7, 0.3, 0.45
    out = rng.normal(0.0, 0.8, (B, H2, W2)).astype(np.float32)
    tgt = rng.normal(0.0, 0.8, (3, R, S_)).astype(np.float32)
    idx = np.array([2, 0], dtype=np.int64)
    o_d, t_d, i_d = torch.from_numpy(out).cuda(), torch.from_numpy(tgt).cuda(), torch.from_numpy(idx).cuda()
    dout = torch.empty_like(o_d)
    acc = torch.zeros(3, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    gpu_ctx.check(gpu_ctx.lib.sfh_nn_template_loss(gpu_ctx.handle, o_d.data_ptr(), t_d.data_ptr(), i_d.data_ptr(), B,
                                                   H2, W2, R, S_, sc, off, wm, wk, ws, dout.data_ptr(),
                                                   acc.data_ptr()))
    gpu_ctx.synchronize()
    y64, t64 = out.astype(np.float64), tgt.astype(np.float64)

    def total(yy):
        return np.mean([_tl_ref(yy[b, :R, :S_], t64[idx[b]], sc, off) for b in range(B)], axis=0)

    mse, kl, sq = total(y64)
    a = acc.cpu().numpy()
    assert abs(a[0] - mse) <= 1e-6 * mse and abs(a[1] - kl) <= 1e-5 * kl and abs(a[2] - sq) <= 1e-5 * sq
    # the gradient of wm MSE + wk KL by central differences of the FP64 restatement
    g = dout.cpu().numpy()
    assert np.all(g[:, R:, :] == 0.0) and np.all(g[:, :, S_:] == 0.0)
    h = 1e-5
    for (b, yy, xx) in [(0, 0, 0), (0, 5, 3), (1, 10, 8), (1, 2, 7)]:
        yp, ym = y64.copy(), y64.copy()
        yp[b, yy, xx] += h
        ym[b, yy, xx] -= h
        fp = np.dot([wm, wk, ws], total(yp))
        fm = np.dot([wm, wk, ws], total(ym))
        num = (fp - fm) / (2 * h)
        assert abs(g[b, yy, xx] - num) <= 1e-4 * abs(num) + 1e-7


def test_captured_training_equals_eager(sfh, gpu_ctx):
    """The training step replayed from a HIP graph gives the same parameters
    and losses as the same steps run eagerly (every replay checked through the
    loss history and the final parameters)."""
    from sfh_amd import surrogate as S
    P = S.sample_parameters(48, seed=6, lo=[8.0, -1.5, 25.0, 0.8], hi=[9.0, -0.5, 27.0, 1.5])
    T = S.exact_templates(P)
    out = []
    for graph in (False, True):
        net = S.TemplateSurrogate(dtype="bf16", seed=4, max_batch=16)
        img = net.images(T)
        hist = net.train(P, img, steps=45, batch=16, warmup=10, log_every=5, graph=graph)
        out.append((hist, net.parameters(), net.step_count))
    (h0, p0, s0), (h1, p1, s1) = out
    assert s0 == s1 == 45 and len(h0) == len(h1) == 9
    assert h0 == h1
    assert np.array_equal(p0, p1)


@pytest.mark.parametrize("dtype", [0, 1])
@pytest.mark.parametrize("rows,cols,act", [(32, 7904, 1), (5, 300, 1), (40, 16, 0), (32, 512, 1)])
def test_dense_bwd_prep_equals_separate_kernels(sfh, gpu_ctx, dtype, rows, cols, act):
    """The one-pass preparation of a dense layer's output gradient equals, bit
    for bit, the four kernels it replaces (activation derivative, column sums,
    low-precision copy, padded transpose), and the NumPy restatement."""
    import torch
    rng = np.random.default_rng(rows * cols + act)
    d = rng.normal(size=(rows, cols)).astype(np.float32)
    pre = rng.normal(scale=2.0, size=(rows, cols)).astype(np.float32)
    ldt = (rows + 15) // 16 * 16
    tdt = torch.bfloat16 if dtype == 0 else torch.float32
    d_d, pre_d = torch.from_numpy(d).cuda(), torch.from_numpy(pre).cuda()
    gb = torch.full((cols,), np.nan, dtype=torch.float32, device="cuda")
    dl = torch.full((rows, cols), np.nan, dtype=tdt, device="cuda")
    dT = torch.full((cols, ldt), np.nan, dtype=tdt, device="cuda")
    d2 = d_d.clone()
    gb2 = torch.empty_like(gb)
    dl2 = torch.empty_like(dl)
    dT2 = torch.empty_like(dT)
    torch.cuda.synchronize()
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    gpu_ctx.check(lib.sfh_nn_dense_bwd_prep(h, dtype, d_d.data_ptr(), pre_d.data_ptr(), rows, cols, act,
                                            gb.data_ptr(), dl.data_ptr(), dT.data_ptr(), ldt))
    gpu_ctx.check(lib.sfh_nn_act_bwd(h, d2.data_ptr(), pre_d.data_ptr(), rows * cols, act, d2.data_ptr()))
    gpu_ctx.check(lib.sfh_nn_colsum(h, d2.data_ptr(), cols, rows, cols, 0.0, gb2.data_ptr()))
    gpu_ctx.check(lib.sfh_nn_bias_act(h, dtype, d2.data_ptr(), cols, None, rows, cols, 0, None, dl2.data_ptr(),
                                      cols))
    gpu_ctx.check(lib.sfh_nn_transpose(h, dtype, d2.data_ptr(), cols, rows, cols, dT2.data_ptr(), ldt, cols, 0))
    gpu_ctx.synchronize()
    assert torch.equal(gb, gb2)
    assert torch.equal(dl, dl2)
    assert torch.equal(dT, dT2)
    x = pre.astype(np.float64)
    s = 1.0 / (1.0 + np.exp(-x))
    ref = d * (s * (1.0 + x * (1.0 - s))) if act else d.astype(np.float64)
    assert np.allclose(gb.cpu().numpy(), ref.sum(axis=0), rtol=1e-5, atol=1e-5)
    assert np.all(dT.float().cpu().numpy()[:, rows:] == 0.0)
    tol = 1e-2 if dtype == 0 else 1e-6
    assert np.allclose(dT.float().cpu().numpy()[:, :rows], ref.T, rtol=tol, atol=tol)


@pytest.mark.parametrize("act", [0, 1])
@pytest.mark.parametrize("P", [1, 7, 4096, 32 * 52 * 76, 32 * 104 * 152])
def test_act_bwd_chansum(sfh, gpu_ctx, P, act):
    """Activation gradient over P pixels of 32 channels with the channel sums
    from the same pass (more pixels than threads at the largest sizes): dX as
    the elementwise kernel gives it, bit for bit, and the sums against FP64."""
    import torch
    rng = np.random.default_rng(P + act)
    dY = torch.from_numpy(rng.normal(size=(P, 32)).astype(np.float32)).cuda()
    pre = torch.from_numpy(rng.normal(scale=2.0, size=(P, 32)).astype(np.float32)).cuda()
    dX = torch.empty_like(dY)
    dX2 = torch.empty_like(dY)
    dsum = torch.full((32,), np.nan, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    gpu_ctx.check(lib.sfh_nn_act_bwd_chansum(h, dY.data_ptr(), pre.data_ptr(), P, act, dX.data_ptr(),
                                             dsum.data_ptr()))
    gpu_ctx.check(lib.sfh_nn_act_bwd(h, dY.data_ptr(), pre.data_ptr(), P * 32, act, dX2.data_ptr()))
    gpu_ctx.synchronize()
    assert torch.equal(dX, dX2)
    ref = dX2.double().cpu().numpy()
    assert np.all(np.abs(dsum.cpu().numpy() - ref.sum(0)) <= 1e-5 * np.abs(ref).sum(0) + 1e-6)


@pytest.mark.parametrize("dtype", [0, 1])
@pytest.mark.parametrize("M,N,K,act,with_pre", [(32, 7904, 512, 1, True), (32, 512, 256, 1, True),
                                                (32, 256, 16, 1, True), (1000, 64, 256, 0, False),
                                                (16, 40, 2048, 1, False)])
def test_linear_equals_product_then_bias_act(sfh, gpu_ctx, dtype, M, N, K, act, with_pre):
    """A dense layer in one call (split products summed by the activation pass)
    equals the product followed by the bias and activation pass, bit for bit,
    and its transposed output the transpose kernel's (zero padding included)."""
    import torch
    rng = np.random.default_rng(M * 7 + N + K)
    tdt = torch.bfloat16 if dtype == 0 else torch.float32
    A = torch.from_numpy(rng.normal(size=(M, K)).astype(np.float32)).cuda().to(tdt)
    W = torch.from_numpy(rng.normal(scale=0.1, size=(N, K)).astype(np.float32)).cuda().to(tdt)
    bias = torch.from_numpy(rng.normal(size=N).astype(np.float32)).cuda()
    ldy = N + 8
    z = torch.empty((M, N), dtype=torch.float32, device="cuda")
    pre = torch.full((M, N), np.nan, dtype=torch.float32, device="cuda")
    Y = torch.full((M, ldy), np.nan, dtype=tdt, device="cuda")
    z2 = torch.empty_like(z)
    pre2 = torch.empty_like(pre)
    Y2 = torch.empty_like(Y)
    torch.cuda.synchronize()
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    ldt = (M + 15) // 16 * 16 + 16
    YT = torch.full((N, ldt), np.nan, dtype=tdt, device="cuda")
    YT2 = torch.empty_like(YT)
    torch.cuda.synchronize()
    gpu_ctx.check(lib.sfh_nn_linear(h, dtype, A.data_ptr(), K, W.data_ptr(), K, M, N, K, bias.data_ptr(), act,
                                    z.data_ptr(), pre.data_ptr() if with_pre else None, Y.data_ptr(), ldy,
                                    YT.data_ptr(), ldt))
    gpu_ctx.check(lib.sfh_nn_gemm_nt(h, dtype, A.data_ptr(), K, W.data_ptr(), K, z2.data_ptr(), N, M, N, K, 1.0, 0.0))
    gpu_ctx.check(lib.sfh_nn_bias_act(h, dtype, z2.data_ptr(), N, bias.data_ptr(), M, N, act, pre2.data_ptr(),
                                      Y2.data_ptr(), ldy))
    # the transposed activations (zero padded to ldt) as the transpose kernel gives them
    gpu_ctx.check(lib.sfh_nn_transpose(h, dtype, pre2.data_ptr(), N, M, N, YT2.data_ptr(), ldt, N, act))
    gpu_ctx.synchronize()
    assert torch.equal(Y, Y2)
    assert torch.equal(YT, YT2)
    if with_pre:
        assert torch.equal(pre, pre2)
    ref = A.double() @ W.double().T + bias.double()
    assert float((pre2.double() - ref).abs().max()) <= 1e-5 * float(ref.abs().max()) + 1e-5
