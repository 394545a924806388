"""What the template surrogate gains over building every template with
partial_cmd_smooth, on a parameter box small enough to train in seconds:

* a 0.01-dex (logAge) grid filled by the network, checked on a sparse lattice
  with partial_cmd_smooth fallback, agrees with the exact templates to
  dlnL <= 0.5 at 100 stars per template almost everywhere, takes most
  templates from the network and is built faster than the exact grid;
* fitting a population (2,000 stars) that lies between the points of the
  coarse grid one would build exactly (0.05 dex, as the 700-template grid)
  reaches a higher likelihood with the filled fine grid, most of which comes
  from the network."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX_LO = [8.45, -1.55, 26.5, 0.9]
BOX_HI = [9.55, -0.45, 27.5, 1.1]
DMOD, ERR = 27.0, 1.0


@pytest.fixture(scope="module")
def trained(sfh, gpu_ctx):
    from sfh_amd import surrogate as S
    P = S.sample_parameters(2000, seed=21, lo=BOX_LO, hi=BOX_HI)
    net = S.TemplateSurrogate(dtype="bf16", seed=2, max_batch=32)
    img = net.images(S.exact_templates(P))
    hist = net.train(P, img, steps=6000, batch=32, lr=1e-3, warmup=200, log_every=1500)
    print("training history (step, MSE, KL/star, log-sum err^2):", hist)
    return net


def test_fill_accuracy_and_speed(trained):
    from sfh_amd import surrogate as S
    la = np.round(8.60 + 0.01 * np.arange(81), 2)
    mh = np.round(-1.4 + 0.1 * np.arange(9), 2)
    n = la.size * mh.size
    full = np.column_stack([np.repeat(la, mh.size), np.tile(mh, la.size), np.full(n, DMOD), np.full(n, ERR)])
    trained.predict(full[:32])                                  # inference shapes warm
    S.exact_templates(full[:8])
    t0 = time.perf_counter()
    E = S.exact_templates(full)
    exact_s = time.perf_counter() - t0
    r = S.fill_templates(trained, la, mh, DMOD, ERR, n_stars=100.0, check_stride=(8, 4))
    dl, _, _ = S.template_delta_lnL(E, r.templates, 100.0)
    print(f"fill {r.seconds['total']:.3f} s (network {r.seconds['network']:.3f} s) vs exact {exact_s:.3f} s; "
          f"exact fraction {r.fallback_rate:.3f}; dlnL q50/q99/max {np.quantile(dl, [0.5, 0.99, 1.0])}")
    assert np.all(dl[r.exact_mask] < 1e-4)      # exact templates: only the floor of empty bins remains
    assert np.mean(dl <= 0.5) >= 0.99
    assert r.fallback_rate <= 0.25
    assert r.seconds["total"] < 0.5 * exact_s


def test_fine_grid_fits_better_than_coarse_grid(trained, gpu_ctx):
    import torch
    from sfh_amd import surrogate as S
    from sfh_amd.multifit import MultiFit
    # a population between the coarse grid points, as expected counts (2,000
    # stars in two components)
    truth = np.array([[9.023, -1.03, DMOD, ERR], [9.027, -0.97, DMOD, ERR]])
    Et = S.exact_templates(truth).cpu().numpy().astype(np.float64)
    lam = Et.sum(0)
    data = 2e3 * lam / lam.sum()
    # coarse exact grid (0.05 dex x 0.2 dex) and the filled fine grid, same span
    la_c, mh_c = np.round(8.80 + 0.05 * np.arange(9), 2), np.round(-1.4 + 0.2 * np.arange(5), 2)
    nc = la_c.size * mh_c.size
    coarse = S.exact_templates(np.column_stack([np.repeat(la_c, mh_c.size), np.tile(mh_c, la_c.size),
                                                np.full(nc, DMOD), np.full(nc, ERR)]))
    la_f, mh_f = np.round(8.80 + 0.01 * np.arange(41), 2), np.round(-1.4 + 0.1 * np.arange(9), 2)
    fine = S.fill_templates(trained, la_f, mh_f, DMOD, ERR, n_stars=1e3, check_stride=(8, 4))
    models = [np.asfortranarray(coarse.cpu().numpy().T), np.asfortranarray(fine.templates.cpu().numpy().T)]
    del coarse
    torch.cuda.empty_cache()
    mf = MultiFit(models, [data, data], ctx=gpu_ctx)
    res = mf.fit_templates()
    # -logL at the MLE of each grid, and of the exact truth (the best possible)
    f_c, f_f = res[0]["mle"].result.minimum, res[1]["mle"].result.minimum
    mt = MultiFit([np.asfortranarray(Et.T.astype(np.float32))], [data], ctx=gpu_ctx)
    f_t = mt.fit_templates()[0]["mle"].result.minimum
    print(f"-logL: truth {f_t:.3f}, coarse exact grid {f_c:.3f} (+{f_c - f_t:.3f}), "
          f"filled fine grid {f_f:.3f} (+{f_f - f_t:.3f}); fine exact fraction {fine.fallback_rate:.3f}")
    assert fine.fallback_rate <= 0.5                 # most of the fine grid comes from the network
    assert f_f < f_c - 0.3
    assert f_f - f_t < 0.5 * (f_c - f_t)
