"""sfh_fg_device_noalloc (the fg! entry point that allocates nothing) against
sfh_fg_device, called through the C ABI on two calls in a row, the second
with a smaller input, plus the oracle's fg! (src/fitting/solvers.jl:20-38)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E64 = float(np.finfo(np.float64).eps)
E32 = float(np.finfo(np.float32).eps)


def _case(N, T, seed):
    rng = np.random.default_rng(seed)
    models = rng.uniform(0.0, 1.0, (N, T)).astype(np.float32)
    models[rng.uniform(size=(N, T)) < 0.3] = 0.0
    x = rng.uniform(0.5, 2.0, T)
    data = rng.poisson(models.astype(np.float64) @ x).astype(np.float64)
    return models, data, x


def _run(lib, fn, ctx, tpl, dat, dx, want_F, want_G, want_comp, N, T):
    import torch
    dnl = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    dG = torch.full((T,), np.nan, dtype=torch.float64, device="cuda") if want_G else None
    dc = torch.full((N,), np.nan, dtype=torch.float64, device="cuda") if want_comp else None
    p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None   # noqa: E731
    rc = getattr(lib, fn)(ctx, tpl, dat, p(dx), 1 if want_F else 0, E64, E32, p(dnl), p(dG), p(dc))
    assert rc == 0, (fn, rc, lib.sfh_last_error(ctx))
    assert lib.sfh_ctx_synchronize(ctx) == 0
    return (dnl.cpu().numpy() if want_F else None, dG.cpu().numpy() if want_G else None,
            dc.cpu().numpy() if want_comp else None)


def test_noalloc_matches_fg_device(sfh, O, gpu_ctx):
    import torch
    lib = sfh._lib.load()
    ctx = C.c_void_p()
    assert lib.sfh_ctx_create(0, None, C.byref(ctx)) == 0
    try:
        for (N, T, seed) in [(60000, 700, 1), (5000, 37, 2)]:
            print(f"noalloc case N={N} T={T}", flush=True)
            models, data, x = _case(N, T, seed)
            tpl, dat = C.c_void_p(), C.c_void_p()
            m = np.asfortranarray(models)
            assert lib.sfh_templates_create(ctx, m.ctypes.data, N, T, C.byref(tpl)) == 0
            assert lib.sfh_data_create(ctx, data.ctypes.data, N, C.byref(dat)) == 0
            dx = torch.tensor(x, dtype=torch.float64, device="cuda")
            for want_F, want_G, want_comp in [(True, True, True), (True, False, False), (False, True, False),
                                              (True, True, False)]:
                a = _run(lib, "sfh_fg_device", ctx, tpl, dat, dx, want_F, want_G, want_comp, N, T)
                b = _run(lib, "sfh_fg_device_noalloc", ctx, tpl, dat, dx, want_F, want_G, want_comp, N, T)
                for u, v in zip(a, b):
                    assert (u is None) == (v is None)
                    if u is not None:
                        assert np.array_equal(u, v), (N, T, want_F, want_G, want_comp)
            rnl, rG, rcomp = O.fg(models, data, x, eps=E64, eps_grad=E32, nthreads=8)
            nl, G, comp = _run(lib, "sfh_fg_device_noalloc", ctx, tpl, dat, dx, True, True, True, N, T)
            assert abs(nl[0] - rnl) <= 1e-12 * abs(rnl)
            assert np.max(np.abs(G - rG)) <= 1e-10 * np.max(np.abs(rG))
            assert np.max(np.abs(comp - rcomp)) <= 1e-12 * np.max(np.abs(rcomp))
            assert lib.sfh_data_destroy(dat) == 0
            assert lib.sfh_templates_destroy(tpl) == 0
        # beyond the single-pass kernel's 704 templates: a clear argument error
        models, data, x = _case(300, 800, 3)
        tpl, dat = C.c_void_p(), C.c_void_p()
        m = np.asfortranarray(models)
        assert lib.sfh_templates_create(ctx, m.ctypes.data, 300, 800, C.byref(tpl)) == 0
        assert lib.sfh_data_create(ctx, data.ctypes.data, 300, C.byref(dat)) == 0
        dx = torch.tensor(x, dtype=torch.float64, device="cuda")
        dnl = torch.empty(1, dtype=torch.float64, device="cuda")
        rc = lib.sfh_fg_device_noalloc(ctx, tpl, dat, C.c_void_p(dx.data_ptr()), 1, E64, E64,
                                       C.c_void_p(dnl.data_ptr()), None, None)
        assert rc == sfh._lib.SFH_ERR_ARGUMENT
        assert b"704" in lib.sfh_last_error(ctx)
        lib.sfh_data_destroy(dat)
        lib.sfh_templates_destroy(tpl)
    finally:
        lib.sfh_ctx_destroy(ctx)
