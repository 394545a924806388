"""Reentrancy of the C ABI: the reference calls fg! concurrently from Julia
threads, one buffer set per task (src/fitting/hmc_sample.jl:127-139,
src/fitting/hierarchical/generic_fitting.jl:617-626).  Here 8 host threads
each with their own context, plus 2 threads sharing one context, call sfh_fg
(host pointers, through ctypes, which releases the GIL) on a 700 x 60,000
template set at the same time; every result must be bitwise equal to the
serial call with the same coefficients."""
import ctypes as C
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E64 = float(np.finfo(np.float64).eps)
N, T = 60000, 700


def _make(lib, ctx, models, data):
    tpl, dat = C.c_void_p(), C.c_void_p()
    assert lib.sfh_templates_create(ctx, models.ctypes.data, N, T, C.byref(tpl)) == 0
    assert lib.sfh_data_create(ctx, data.ctypes.data, N, C.byref(dat)) == 0
    return tpl, dat


def _fg(lib, ctx, tpl, dat, x):
    nl = C.c_double()
    G = np.empty(T)
    rc = lib.sfh_fg(ctx, tpl, dat, x.ctypes.data, 1, E64, E64, C.byref(nl), G.ctypes.data, None)
    assert rc == 0, lib.sfh_last_error(ctx)
    return nl.value, G


def test_concurrent_fg_bitwise_equal_to_serial(sfh, gpu_ctx):
    lib = sfh._lib.load()
    rng = np.random.default_rng(12)
    models = np.asfortranarray(rng.uniform(0, 1, (N, T)).astype(np.float32))
    xs = [rng.uniform(0.5, 2.0, T) for _ in range(10)]
    data = rng.poisson(models[:, :50].astype(np.float64) @ xs[0][:50]).astype(np.float64)
    # serial reference results on one context
    c0 = C.c_void_p()
    assert lib.sfh_ctx_create(0, None, C.byref(c0)) == 0
    t0, d0 = _make(lib, c0, models, data)
    serial = [_fg(lib, c0, t0, d0, x) for x in xs]
    # 8 threads with a context each, 2 threads sharing one context
    ctxs = []
    for _ in range(9):
        c = C.c_void_p()
        assert lib.sfh_ctx_create(0, None, C.byref(c)) == 0
        ctxs.append((c,) + _make(lib, c, models, data))
    plan = [ctxs[k] for k in range(8)] + [ctxs[8], ctxs[8]]
    results = [None] * 10
    errors = []
    start = threading.Barrier(10)

    def work(i):
        try:
            c, tp, dt = plan[i]
            start.wait()
            out = []
            for _ in range(5):
                out.append(_fg(lib, c, tp, dt, xs[i]))
            results[i] = out
        except Exception as e:          # surfaced below
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(10)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=300)
    assert not any(th.is_alive() for th in threads)
    assert not errors, errors
    for i in range(10):
        for nl, G in results[i]:
            assert nl == serial[i][0], i
            assert np.array_equal(G, serial[i][1]), i
    for c, tp, dt in ctxs + [(c0, t0, d0)]:
        lib.sfh_data_destroy(dt)
        lib.sfh_templates_destroy(tp)
        lib.sfh_ctx_destroy(c)
