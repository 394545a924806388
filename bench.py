#!/usr/bin/env python
"""Benchmark of the template-mixture likelihood on one MI355X.

Workload (one "step"): one batched fg! (log-likelihood + gradient) of 4096
coefficient vectors (HMC / ensemble walkers, log-normal 0.3 dex around the
truth) against 700 templates x 60,000 bins (200 x 300 Hess grid) and a
Poisson-drawn observed diagram with ~1e6 stars.  The templates are deposited
on the GPU first (700 synthetic isochrones of 2000 points, covariant kernel
with colour-magnitude correlation 0.6), which is also timed (templates/s).

Prints ONE JSON line.  `value` is walker evaluations per second of the
FP64-MFMA form (log-likelihood at FP64 accuracy); the split-BF16 and FP32
forms, the single-walker ("plain") path, template deposition, errors against
the FP64 oracle and the oracle timed on the host cores are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "bf16": 2500.0, "fp32": 157.3, "hbm": 8000.0}   # TFLOP/s, GB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--walkers", type=int, default=4096)
    ap.add_argument("--precision", default="f64", choices=["f64", "bf16x2", "fp32"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--fits", type=int, default=256, help="independent fits in the uneven multi-fit run")
    ap.add_argument("--no-multifit", action="store_true")
    ap.add_argument("--multifit-cpu-seconds", type=float, default=60.0,
                    help="host time budget of the multi-fit convergence comparison")
    ap.add_argument("--surrogate-steps", type=int, default=40000,
                    help="training steps (batch 32) of the template surrogate on 20,000 templates")
    ap.add_argument("--surrogate-stars", type=float, default=100.0,
                    help="stars per template in the dlnL check of the filled grid")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--mock-mass", type=float, default=1e7,
                    help="birth mass (Msun) of the mock population sampled, observed and binned on the GPU")
    ap.add_argument("--no-mock", action="store_true")
    ap.add_argument("--no-binaries", action="store_true")
    ap.add_argument("--no-config0", action="store_true")
    ap.add_argument("--section", default="all", choices=["all", "templates", "plain", "config0"],
                    help="templates: only the deposition of the 700 templates; plain: deposition, then only the "
                         "single-walker fg! loop (short runs for the profiler's counter passes); config0: only the "
                         "50-template plain-method workload (templates + fit_templates)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (calibrates HBM counters)")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch
    calibrate the counters read for the path's kernels."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4, "dispatches": 3}), flush=True)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Timer:
    """HIP events on the stream the library launches on (torch's current stream)."""

    def __init__(self, torch):
        self.torch = torch

    def __enter__(self):
        self.s = self.torch.cuda.Event(enable_timing=True)
        self.e = self.torch.cuda.Event(enable_timing=True)
        self.torch.cuda.synchronize()
        self.s.record()
        return self

    def __exit__(self, *a):
        self.e.record()
        self.torch.cuda.synchronize()
        self.ms = self.s.elapsed_time(self.e)


def uneven_fits(sfh, torch, ctx, nfits, seed=20261203):
    """Seeded uneven fits built like the 50-template workload: ntemplates ~
    U[50, 700] isochrones picked without repetition from the 70 x 10 (logAge
    6.6-10.05, [M/H] -2.5-0.0) grid of 500-point synthetic isochrones, Kroupa
    IMF over 0.08-2 Msun, the exponential photometric error and logistic
    completeness, deposited by partial_cmd_smooth_batch at the fit's distance
    modulus ~ U[24, 30] on a colour -0.5..2.5 x magnitude dmod-6..dmod+6 grid
    of nbins ~ U[5k, 60k] (aspect 2:3); star total log-uniform in [1e4, 1e6].
    Data: Poisson draws of the composite at log-normal coefficients.  Returns
    [(templates (T, N) FP32 on the device, data (N,) FP64 on the device, T, N,
    dmod, total)]."""
    from sfh_amd import synth
    from sfh_amd.templates import edges_range
    rng = np.random.default_rng(seed)
    specs = synth.specs_70x10()
    imf = synth.Kroupa2001(0.08, 2.0)
    isos = [synth.isochrone(s.log_age, s.mh, 500, s.seed, mmax_imf=2.0) for s in specs]
    errs, comps = [synth.PHOTERR] * 2, [synth.COMPLETENESS] * 2
    fits = []
    for f in range(nfits):
        T = int(rng.integers(50, 701))
        N = int(rng.integers(5000, 60001))
        nx = max(8, int(round(np.sqrt(N * 2.0 / 3.0))))
        ny = max(8, N // nx)
        N = nx * ny
        dmod = float(rng.uniform(24.0, 30.0))
        total = float(10 ** rng.uniform(4.0, 6.0))
        pick = np.sort(rng.choice(len(specs), T, replace=False))
        edges = (edges_range(-0.5, 2.5, nx + 1), edges_range(dmod - 6.0, dmod + 6.0, ny + 1))
        C = sfh.partial_cmd_smooth_batch([isos[k] for k in pick], errs, 1, (0, 1), imf, comps, dmod=dmod,
                                         normalize_value=1.0, mean_mass=imf.mean(), edges=edges, ctx=ctx)
        x = torch.from_numpy(rng.lognormal(0.0, 1.0, T)).cuda()
        comp = x @ C.double()
        comp *= total / comp.sum()
        g = torch.Generator(device="cuda").manual_seed(int(seed + f))
        data = torch.poisson(comp, generator=g)
        fits.append((C, data, T, N, dmod, total))
    return fits


def cpu_fit_templates(O, models, data, nthreads, maxiter=10000):
    """fit_templates (src/fitting/solvers.jl:172-221) on the host: SciPy BFGS
    (inf-norm gradient tolerance 1e-8, as Optim's g_abstol) around the C FP64
    oracle's fg!, MAP then MLE in log-coefficients, at most ``maxiter``
    iterations per stage.  Returns MAP/MLE (mu, fmin, status: "converged",
    "max_iterations" or SciPy's message) and the fg! evaluation count."""
    from scipy.optimize import minimize
    T = models.shape[1]
    x0 = np.ones(T)
    x0 = x0 * (data.sum() / (models.astype(np.float64) @ x0).sum())
    nev = [0]

    def obj(th, mle):
        nev[0] += 1
        x = np.exp(th)
        f, G, _ = O.fg(models, data, x, nthreads=nthreads)
        return (f, G * x) if mle else (f - th.sum(), G * x - 1.0)

    opts = {"gtol": 1e-8, "maxiter": maxiter, "norm": np.inf}
    r1 = minimize(obj, np.log(x0), jac=True, method="BFGS", args=(False,), options=opts)
    r2 = minimize(obj, r1.x, jac=True, method="BFGS", args=(True,), options=opts)

    def status(r):
        return {0: "converged", 1: "max_iterations"}.get(r.status, str(r.message))

    return {"map": (np.exp(r1.x), float(r1.fun), status(r1)), "mle": (np.exp(r2.x), float(r2.fun), status(r2)),
            "evals": nev[0]}


def multifit_bench(sfh, torch, ctx, O, args, ncpu):
    """Uneven batched fits: every fit's fg! in one persistent launch, the
    fit_templates iterations (BFGS, MAP then MLE) on the device in captured
    HIP graphs, launch shape from the autotuner; convergence of a sample of
    fits checked against the same fits done on the host."""
    from sfh_amd.multifit import MultiFit
    t0 = time.time()
    fits = uneven_fits(sfh, torch, ctx, args.fits)
    tpls = [sfh.HIPTemplates(device_ptr=C.data_ptr(), shape=(N, T), ctx=ctx) for C, _, T, N, _, _ in fits]
    datas = [sfh.HIPData(device_ptr=d.data_ptr(), nbins=N, ctx=ctx) for _, d, _, N, _, _ in fits]
    ctx.synchronize()
    sizes = [(T, N) for _, _, T, N, _, _ in fits]
    mf = MultiFit(tpls, datas, ctx=ctx)
    gen_s = time.time() - t0
    elems = float(sum(T * N for T, N in sizes))
    t0 = time.time()
    shape = mf.autotune(reps=3)
    tune_s = time.time() - t0
    log(f"multifit: {len(sizes)} fits, {elems:.3e} template elements ({elems * 4 / 1e9:.1f} GB), "
        f"generated in {gen_s:.1f} s; autotuned shape {shape} in {tune_s:.1f} s")
    # one grouped fg! of every fit: kernel time -> HBM GB/s
    X = [np.ones(T) for T, _ in sizes]
    mf.fg(X)
    ctx.set_timing(True)
    ctx.reset_stats()
    for _ in range(5):
        mf.fg(X)
    kms, kn = ctx.kernel_stats("multifit_fg")
    rms, rn = ctx.kernel_stats("multifit_reduce")
    ctx.set_timing(False)
    fg_ms = kms / max(kn, 1)
    fg_GBps = elems * 4 / (fg_ms * 1e-3) / 1e9
    log(f"multifit grouped fg!: {fg_ms:.3f} ms per launch = {fg_GBps:.0f} GB/s of templates")
    # the fits
    ctx.set_timing(True)
    ctx.reset_stats()
    t0 = time.time()
    res = mf.fit_templates()
    fit_s = time.time() - t0
    gms, gn = ctx.kernel_stats("multifit_graph")
    ctx.set_timing(False)
    st = {}
    evals = 0
    work = 0.0
    for f, r in enumerate(res):
        for key in ("map", "mle"):
            st[r[key].result.status] = st.get(r[key].result.status, 0) + 1
            evals += r[key].result.evaluations
            work += r[key].result.evaluations * sizes[f][0] * sizes[f][1]
    iters = np.array([r["map"].result.iterations + r["mle"].result.iterations for r in res])
    log(f"multifit fit_templates: {fit_s:.2f} s for {len(res)} fits ({len(res) / fit_s:.1f} fits/s), "
        f"{mf.steps} steps, {evals} fg! evaluations, statuses {st}")
    out = {"fits": len(res), "templates_range": [min(t for t, _ in sizes), max(t for t, _ in sizes)],
           "bins_range": [min(n for _, n in sizes), max(n for _, n in sizes)],
           "template_GB": elems * 4 / 1e9, "autotuned_shape": {"item_elems": shape[0], "waves": shape[1],
                                                              "wg_per_cu": shape[2], "tune_s": tune_s},
           "grouped_fg_kernel_ms": fg_ms, "grouped_fg_GBps": fg_GBps, "grouped_fg_hbm_frac": fg_GBps / PEAK["hbm"],
           "reduce_kernel_ms": rms / max(rn, 1),
           "fit_seconds": fit_s, "fits_per_s": len(res) / fit_s, "optimizer_steps": mf.steps,
           "graph_ms_total": gms, "fg_evaluations": evals, "status_counts": st,
           "iterations_median": float(np.median(iters)), "iterations_max": int(iters.max())}
    if not args.no_cpu:
        # host fit_templates, 1000 iterations per stage as on the device, on
        # the three smallest fits and the smallest fits that reached the
        # iteration cap on the device, bounded to about multifit_cpu_seconds
        by_size = [int(f) for f in np.argsort([T * N for T, N in sizes])]
        capped = [f for f in by_size
                  if "max_iterations" in (res[f]["map"].result.status, res[f]["mle"].result.status)]
        order = by_size[:3] + [f for f in capped if f not in by_size[:3]][:3]
        sample, cpu_s, cpu_work = [], 0.0, 0.0
        for f in order:
            if cpu_s > args.multifit_cpu_seconds:
                break
            models = np.asfortranarray(fits[f][0].cpu().numpy().T)
            data = fits[f][1].cpu().numpy()
            t1 = time.time()
            ref = cpu_fit_templates(O, models, data, ncpu, maxiter=1000)
            dt = time.time() - t1
            cpu_s += dt
            cpu_work += ref["evals"] * sizes[f][0] * sizes[f][1]
            row = {"fit": int(f), "templates": sizes[f][0], "bins": sizes[f][1], "dmod": fits[f][4],
                   "stars": fits[f][5], "cpu_s": dt, "cpu_evals": ref["evals"]}
            for key in ("map", "mle"):
                mu_ref, f_ref, st_ref = ref[key]
                r = res[f][key]
                big = mu_ref > 1e-3 * mu_ref.max()
                row[f"{key}_mu_rel_err"] = float(np.max(np.abs(r.mu[big] - mu_ref[big]) / mu_ref[big]))
                row[f"{key}_fmin_rel_diff"] = float((r.result.minimum - f_ref) / abs(f_ref))
                row[f"{key}_status"] = r.result.status
                row[f"{key}_cpu_status"] = st_ref
                row[f"{key}_iterations"] = r.result.iterations
                # coefficients the optimiser drives towards zero in log space
                row[f"{key}_tiny_coeffs"] = int(np.sum(r.mu < 1e-6 * r.mu.max()))
            sample.append(row)
        # the host's time per (template element x fg! evaluation), projected onto the whole group
        proj = cpu_s / max(cpu_work, 1.0) * work
        out["cpu_sample"] = sample
        out["cpu_projected_seconds_all_fits"] = proj
        out["gpu_over_cpu_projected"] = proj / fit_s
        log(f"multifit host sample: {len(sample)} fits ({sum(1 for f in order[:len(sample)] if f in capped)} "
            f"capped on the device) in {cpu_s:.1f} s on {ncpu} threads; projected host time for all fits "
            f"{proj:.0f} s ({proj / fit_s:.0f}x the GPU)")
        for row in sample:
            log(f"  fit {row['fit']} T={row['templates']} N={row['bins']}: device {row['map_status']}/"
                f"{row['mle_status']}, host {row['map_cpu_status']}/{row['mle_cpu_status']}, fmin rel diff "
                f"{row['map_fmin_rel_diff']:.1e}/{row['mle_fmin_rel_diff']:.1e}")
    del fits
    mf.close()
    for t in tpls:
        t.close()
    for d in datas:
        d.close()
    return out


def config0_bench(sfh, torch, ctx, O, args, ncpu):
    """The plain method end to end on the 50-template workload (10 log ages x
    5 [M/H], 500-point isochrones, 75 x 100 grid, IMF 0.08-2 Msun, ~1e5 stars):
    templates by partial_cmd_smooth_batch, then fit_templates (MAP + MLE) by
    the host BFGS around the GPU fg! and by the on-device BFGS, against the
    host path (the oracle's partial_cmd_smooth and SciPy BFGS around its
    fg!)."""
    from sfh_amd import synth
    from sfh_amd.multifit import MultiFit
    edges = synth.edges_75x100()
    specs = synth.specs_10x5()
    imf = synth.Kroupa2001(0.08, 2.0)
    isos = [synth.isochrone(s.log_age, s.mh, 500, s.seed, mmax_imf=2.0) for s in specs]
    kw = dict(dmod=24.0, normalize_value=1.0, mean_mass=imf.mean(), edges=edges, ctx=ctx)
    errs, comps = [synth.PHOTERR] * 2, [synth.COMPLETENESS] * 2
    sfh.partial_cmd_smooth_batch(isos, errs, 1, (0, 1), imf, comps, output="host", **kw)
    t0 = time.time()
    models = sfh.partial_cmd_smooth_batch(isos, errs, 1, (0, 1), imf, comps, output="host", **kw)
    tpl_s = time.time() - t0
    models = np.asfortranarray(models)
    rng = np.random.default_rng(20261203)
    x_true = 10.0 ** (0.3 * rng.standard_normal(len(specs)))
    x_true *= 1e5 / (models.astype(np.float64) @ x_true).sum()
    data = synth.poisson_data(models.astype(np.float64) @ x_true, seed=5)
    sfh.fit_templates(models, data)                       # warm-up (first-call set-up)
    t0 = time.time()
    host_loop = sfh.fit_templates(models, data)
    host_loop_s = time.time() - t0
    mf = MultiFit([models], [data], ctx=ctx)
    mf.fit_templates()
    t0 = time.time()
    dev = mf.fit_templates()[0]
    dev_s = time.time() - t0
    mf.close()
    out = {"templates": len(specs), "bins": models.shape[0], "stars": float(data.sum()),
           "templates_s": tpl_s,
           "fit_host_bfgs_gpu_fg_s": host_loop_s,
           "fit_host_bfgs_evals": int(host_loop["map"].result.nfev + host_loop["mle"].result.nfev),
           "fit_device_bfgs_s": dev_s,
           "device_status": [dev["map"].result.status, dev["mle"].result.status],
           "end_to_end_s": tpl_s + dev_s}
    log(f"config0: 50 templates in {tpl_s * 1e3:.1f} ms; fit_templates host-BFGS/GPU-fg! {host_loop_s:.2f} s "
        f"({out['fit_host_bfgs_evals']} fg!), device BFGS {dev_s:.3f} s ({out['device_status']})")
    if not args.no_cpu:
        f = lambda v: float(synth.photerr(v))           # noqa: E731
        c = lambda v: float(synth.completeness(v))      # noqa: E731
        ex, ey = (O.Range(e.first, e.step, e.length) for e in edges)
        t0 = time.time()
        ref = []
        for m, mags in isos:
            r, _ = O.partial_cmd_smooth(m, mags, [f, f], 1, (0, 1), lambda v: float(imf(v)), [c, c], dmod=24.0,
                                        normalize_value=1.0, mean_mass=imf.mean(), ex=ex, ey=ey, nthreads=ncpu)
            ref.append(r.ravel(order="F"))
        ref = np.asfortranarray(np.stack(ref, axis=1), dtype=np.float32)
        cpu_tpl_s = time.time() - t0
        t0 = time.time()
        cfit = cpu_fit_templates(O, ref, data, ncpu)
        cpu_fit_s = time.time() - t0
        same = cpu_fit_templates(O, models, data, ncpu)
        out["template_max_norm_err_vs_oracle"] = float(max(
            np.max(np.abs(models[:, t] - ref[:, t])) / np.max(np.abs(ref[:, t])) for t in range(len(specs))))
        for key in ("map", "mle"):
            out[f"{key}_fmin_rel_diff_same_templates"] = {
                "host_bfgs_gpu_fg": float((host_loop[key].result.fun - same[key][1]) / abs(same[key][1])),
                "device_bfgs": float((dev[key].result.minimum - same[key][1]) / abs(same[key][1]))}
        out["mle_fmin_rel_diff_vs_oracle_templates"] = float(
            (dev["mle"].result.minimum - cfit["mle"][1]) / abs(cfit["mle"][1]))
        out["cpu_baseline"] = {"value": cpu_tpl_s + cpu_fit_s, "unit": "s end to end", "cores": ncpu,
                               "kind": "port", "templates_s": cpu_tpl_s, "fit_s": cpu_fit_s,
                               "sample": "the whole workload: the oracle's partial_cmd_smooth (Python "
                                         "preparation + C FP64 deposition) for 50 templates, then SciPy BFGS "
                                         "(MAP + MLE) around the C FP64 fg!"}
        log(f"config0 host path: templates {cpu_tpl_s:.2f} s + fit {cpu_fit_s:.2f} s on {ncpu} threads; "
            f"template err {out['template_max_norm_err_vs_oracle']:.2e}")
    return out


def surrogate_bench(sfh, torch, ctx, O, args, ncpu):
    """The learned template surrogate: 20,000 exact templates over (logAge
    6.6-10.1, [M/H] -2.5-0.3, dmod 24-30, error scale 0.5-2) from the
    deposition kernel, BF16 training, held-out accuracy, and the fill of a
    0.01-dex grid (351 ages x 34 metallicities = 11,934 templates at dmod 27,
    error scale 1) with partial_cmd_smooth wherever the lattice check finds
    dlnL > 0.5; an audit of 1,000 filled templates against exact ones; and
    partial_cmd_smooth on the host cores for the same templates."""
    from sfh_amd import surrogate as S
    from sfh_amd import synth
    out = {}
    t0 = time.time()
    P = S.sample_parameters(20000, seed=20261204)
    T = S.exact_templates(P, ctx=ctx)
    gen_s = time.time() - t0
    net = S.TemplateSurrogate(dtype="bf16", seed=0, ctx=ctx, max_batch=32)
    img = net.images(T)
    del T
    torch.cuda.empty_cache()
    log(f"surrogate: 20,000 exact training templates in {gen_s:.1f} s ({20000 / gen_s:.0f} templates/s, "
        f"host preparation + GPU deposition)")
    steps, batch = args.surrogate_steps, 32
    t0 = time.time()
    hist = net.train(P, img, steps=steps, batch=batch, lr=1e-3, log_every=max(1, steps // 5))
    train_s = time.time() - t0
    log(f"surrogate training: {steps} steps x {batch} in {train_s:.1f} s = {steps * batch / train_s:.0f} "
        f"templates/s; (step, MSE, KL/star, log-sum err^2): {[tuple(round(v, 6) for v in h) for h in hist]}")
    del img
    torch.cuda.empty_cache()
    Pv = S.sample_parameters(500, seed=20261205)
    Ev = S.exact_templates(Pv, ctx=ctx)
    dl, se, ss = S.template_delta_lnL(Ev, net.predict(Pv), 1.0, ctx=ctx)
    del Ev
    q = [float(v) for v in np.quantile(dl, [0.5, 0.9, 0.99, 1.0])]
    out.update({"parameters": int(net.nparams), "train_templates": 20000, "train_seconds": train_s,
                "train_steps": steps, "batch": batch, "train_templates_per_s": steps * batch / train_s,
                "exact_templates_per_s": 20000 / gen_s, "loss_history": hist,
                "heldout_KL_per_star_q50_q90_q99_max": q,
                "heldout_sum_ratio_q01_q50_q99": [float(v) for v in np.quantile(ss / se, [0.01, 0.5, 0.99])],
                "heldout_fail_frac": {str(n): float((dl * n > 0.5).mean())
                                      for n in (10, 100, 1000, 10000, 100000, 1000000)}})
    log(f"surrogate held-out (500 over the whole box): KL/star q50/q90/q99/max {q}; "
        f"dlnL > 0.5 at 100 stars: {out['heldout_fail_frac']['100']:.3f}")
    # the fill
    la = np.round(6.6 + 0.01 * np.arange(351), 2)
    mh = np.linspace(-2.5, 0.3, 34)
    net.predict(np.array([[8.0, -1.0, 27.0, 1.0]]))        # warm-up of the inference shapes
    # the fallback rate rises with the star count (dlnL = N KL): fills at
    # 1e2 .. 1e5 stars
    for nstars in (args.surrogate_stars, 10.0 * args.surrogate_stars, 100.0 * args.surrogate_stars,
                   1000.0 * args.surrogate_stars):
        out[f"fill_{nstars:g}_stars"] = _fill_bench(S, torch, ctx, O, args, ncpu, net, la, mh, nstars,
                                                    20000 / gen_s)
    return out


def binaries_bench(sfh, ctx, args):
    """Unresolved-binary templates (RandomBinaryPairs(0.3)) of the 700
    2000-point deposition-benchmark isochrones (200 x 300 grid), downsampled 4x as partial_cmd_smooth's binary branch
    does: all pair histograms in one launch and all templates in one
    deposition (binary_hess_batch), against binary_hess one template at a
    time on a sample."""
    from sfh_amd import synth
    edges = synth.edges_200x300()
    imf = synth.Kroupa2001(0.08, 120.0)
    pts = [synth.template_points(s, 2000, edges, 120.0)[0] for s in synth.specs_70x10()]
    isos = [(np.asarray(sp.new_mini)[::4], [np.asarray(m)[::4] for m in sp.new_iso_mags]) for sp in pts]
    kw = dict(normalize_value=1.0, mean_mass=imf.mean(), ctx=ctx)
    model = sfh.RandomBinaryPairs(0.3)
    errs, comp = [synth.photerr, synth.photerr], [synth.completeness, synth.completeness]
    sfh.binary_hess_batch(model, isos[:4], errs, 1, (0, 1), imf, comp, None, edges, **kw)        # warm-up
    ctx.set_timing(True)
    ctx.reset_stats()
    t0 = time.time()
    T = sfh.binary_hess_batch(model, isos, errs, 1, (0, 1), imf, comp, None, edges, **kw)
    batch_s = time.time() - t0
    pk, _ = ctx.kernel_stats("binary_pairs")
    dk, _ = ctx.kernel_stats("deposit_cov")
    ctx.set_timing(False)
    n = len(isos)
    t0 = time.time()
    k = 8
    for m, g in isos[:k]:
        sfh.binary_hess(model, m, g, errs, 1, (0, 1), imf, comp, None, edges, **kw)
    single_s = (time.time() - t0) / k
    out = {"templates": n, "points_per_isochrone": int(np.mean([m.size for m, _ in isos])),
           "batch_seconds": batch_s, "batch_templates_per_s": n / batch_s, "pairs_kernel_ms": pk,
           "deposit_kernel_ms": dk, "single_seconds_per_template": single_s,
           "batch_over_single": single_s * n / batch_s, "sum_check": float(T.sum().item())}
    log(f"binaries: {n} templates in {batch_s:.2f} s batched ({n / batch_s:.0f} templates/s; pair kernel "
        f"{pk:.2f} ms, deposition {dk:.2f} ms) vs {single_s * 1e3:.1f} ms per template one at a time "
        f"({out['batch_over_single']:.1f}x)")
    return out


def mock_bench(sfh, ctx, args):
    """The mock-observation path: a population of one 2000-point isochrone
    (logAge 9, [M/H] -1, 2000 points) with a Kroupa IMF and random binary
    pairs (fraction 0.3) sampled to a birth mass on the GPU, observed with
    the deposition benchmark's error and completeness model, and binned on the 200 x 300
    grid; kernel times from HIP events.  The NumPy restatement of the same
    sampling rule times the host on a bounded sample (one thread)."""
    import importlib
    from sfh_amd import mock, synth
    m, (B, V) = synth.isochrone(9.0, -1.0, 2000, 7, mmax_imf=120.0)
    g = np.stack([B, V], axis=1)
    imf = mock.Kroupa2001(0.08, 120.0)
    edges = synth.edges_200x300()
    limit = args.mock_mass
    kw = dict(dist_mod=24.0, binary_model=sfh.RandomBinaryPairs(0.3), catalog=True, ctx=ctx)
    sfh.generate_stars_mass(m, g, ["B", "V"], 1e4, imf, seed=1, **kw).close()          # warm-up
    ctx.set_timing(True)
    ctx.reset_stats()
    t0 = time.time()
    cat = sfh.generate_stars_mass(m, g, ["B", "V"], limit, imf, seed=2, **kw)
    gen_s = time.time() - t0
    t0 = time.time()
    obs = sfh.model_cmd(cat, [synth.photerr, synth.photerr], [synth.completeness, synth.completeness], seed=3)
    obs_s = time.time() - t0
    t0 = time.time()
    H = mock.bin_cmd_catalog(obs, (0, 1), 1, edges).weights
    bin_s = time.time() - t0
    gk, gn = ctx.kernel_stats("mock_generate")
    ok, _ = ctx.kernel_stats("mock_observe")
    bk, _ = ctx.kernel_stats("mock_bin")
    ctx.set_timing(False)
    drawn, kept, observed = cat.sampled, cat.n, obs.n
    # algorithmic bytes of the sampling kernel per system: 8 (contribution) +
    # 2 x 8 (masses) + 2 x 8 (magnitudes) + 1 (kept flag) written
    gen_bytes = drawn * 41.0
    out = {"birth_mass": limit, "systems_drawn": drawn, "systems_kept": kept, "stars_observed": observed,
           "binned": float(H.sum()), "generate_kernel_ms": gk, "generate_launches": gn,
           "generate_systems_per_s_kernel": drawn / (gk * 1e-3), "generate_wall_s": gen_s,
           "generate_systems_per_s_wall": drawn / gen_s, "generate_GBps_algorithmic": gen_bytes / (gk * 1e-3) / 1e9,
           "observe_kernel_ms": ok, "observe_stars_per_s_kernel": kept / (ok * 1e-3), "observe_wall_s": obs_s,
           "bin_kernel_ms": bk, "bin_stars_per_s_kernel": observed / (bk * 1e-3), "bin_wall_s": bin_s}
    log(f"mock: {limit:.3g} Msun -> {drawn} systems drawn ({gn} batches), {kept} kept, {observed} observed; "
        f"generate {gk:.2f} ms kernel ({drawn / (gk * 1e-3):.3g} systems/s), {gen_s:.2f} s wall; "
        f"model_cmd {ok:.2f} ms; bin_cmd {bk:.2f} ms")
    cat.close()
    obs.close()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle"))
        R = importlib.import_module("mock_ref")
        gi = g + 24.0
        oimf = R.PowerLaw(imf.breaks, imf.slopes, imf.coefs)
        small = max(1e3, min(limit, 2e5))
        t0 = time.time()
        _, _, _, nd = R.generate(m, gi, 0, small, oimf, 1, 0.3, seed=2, mmin_cut=m[0])
        dt = time.time() - t0
        out["cpu_numpy_systems_per_s"] = nd / dt
        out["cpu_sample"] = f"{small:.3g} Msun ({nd} systems) with the NumPy restatement, one thread, {dt:.2f} s"
        out["gpu_over_cpu_kernel"] = out["generate_systems_per_s_kernel"] / (nd / dt)
        log(f"mock host (NumPy): {nd / dt:.3g} systems/s; GPU kernel / host = {out['gpu_over_cpu_kernel']:.0f}x")
    return out


def _fill_bench(S, torch, ctx, O, args, ncpu, net, la, mh, nstars, exact_rate):
    from sfh_amd import synth
    out = {}
    fill = S.fill_templates(net, la, mh, 27.0, 1.0, n_stars=nstars, check_stride=(5, 4))
    n = la.size * mh.size
    chk = fill.check_dlnL[fill.checked_mask]
    out["fill"] = {"grid": [int(la.size), int(mh.size)], "templates": n, "dmod": 27.0, "err_scale": 1.0,
                   "n_stars": nstars, "network_seconds": fill.seconds["network"],
                   "network_templates_per_s": n / fill.seconds["network"], "total_seconds": fill.seconds["total"],
                   "templates_per_s": n / fill.seconds["total"], "checked": int(fill.checked_mask.sum()),
                   "fallback_rate": fill.fallback_rate,
                   "check_dlnL_q50_q90_max": [float(v) for v in np.quantile(chk, [0.5, 0.9, 1.0])]}
    log(f"surrogate fill: {n} templates in {fill.seconds['total']:.2f} s (network {fill.seconds['network']:.2f} s); "
        f"{int(fill.checked_mask.sum())} checked, exact fraction {fill.fallback_rate:.3f}")
    # audit: 1,000 filled templates off the lattice against exact templates
    grid = np.stack(np.meshgrid(la, mh, indexing="ij"), axis=-1).reshape(-1, 2)
    off = np.flatnonzero(~fill.exact_mask)
    rng = np.random.default_rng(3)
    pick = rng.choice(off, size=min(1000, off.size), replace=False) if off.size else off
    if pick.size:
        full = np.column_stack([grid[pick], np.full(pick.size, 27.0), np.full(pick.size, 1.0)])
        t0 = time.time()
        Ea = S.exact_templates(full, ctx=ctx)
        ex_s = time.time() - t0
        idx = torch.from_numpy(pick.astype(np.int64)).to(fill.templates.device)
        da, _, _ = S.template_delta_lnL(Ea, fill.templates.index_select(0, idx), nstars, ctx=ctx)
        del Ea
        out["fill"]["audit"] = {"templates": int(pick.size), "dlnL_q50_q90_q99_max":
                                [float(v) for v in np.quantile(da, [0.5, 0.9, 0.99, 1.0])],
                                "frac_within_0.5": float((da <= 0.5).mean())}
        exact_gpu_rate = pick.size / ex_s
        log(f"surrogate audit: {pick.size} network templates, dlnL at {nstars:g} stars "
            f"q50/q90/q99/max {out['fill']['audit']['dlnL_q50_q90_q99_max']}, "
            f"{out['fill']['audit']['frac_within_0.5']:.3f} within 0.5")
    else:
        exact_gpu_rate = exact_rate
    out["fill"]["exact_gpu_seconds_projected"] = n / exact_gpu_rate
    out["fill"]["speedup_vs_exact_gpu"] = n / exact_gpu_rate / fill.seconds["total"]
    if not args.no_cpu and nstars == args.surrogate_stars:
        # partial_cmd_smooth on the host: the same host preparation, then the C
        # FP64 restatement of bin_cmd_smooth on all host cores; bounded sample
        edges = synth.edges_200x300()
        ex_, ey_ = (O.Range(e.first, e.step, e.length) for e in edges)
        rng = np.random.default_rng(4)
        cpu_s, k = 0.0, 0
        while cpu_s < args.cpu_seconds and k < 64:
            r = grid[rng.integers(0, n)]
            t1 = time.time()
            sp, cm = synth.template_points(synth.TemplateSpec(r[0], r[1], S.ISOCHRONE_SEED), 2000, edges, 120.0,
                                           dmod=27.0, correlation=0.6)
            O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, sp.cov_mult, sp.weights, ex_, ey_,
                             cmult=cm, nthreads=ncpu)
            cpu_s += time.time() - t1
            k += 1
        rate = k / cpu_s
        out["fill"]["cpu_partial_cmd_smooth"] = {"templates_per_s": rate, "cores": ncpu, "sample": k,
                                                 "projected_seconds": n / rate}
        out["fill"]["speedup_vs_cpu"] = n / rate / fill.seconds["total"]
        log(f"surrogate: host partial_cmd_smooth {rate:.3g} templates/s on {ncpu} threads; the fill is "
            f"{out['fill']['speedup_vs_cpu']:.0f}x the host and {out['fill']['speedup_vs_exact_gpu']:.1f}x "
            f"exact templates on the GPU")
    del fill
    torch.cuda.empty_cache()
    return out["fill"]


# VALU peak of the MI355X in lane-instructions/s: 256 CUs x 4 SIMDs x 2.4 GHz x
# 32 lanes per cycle.  A SIMD issues a wave64 FP32 instruction over 4 cycles,
# and two of them, from two waves, in the same 4 cycles (measured on gfx950,
# SQ_ACTIVE_INST_VALU2); this is the published 157.3 TFLOP/s FP32 vector rate
# counted in FMAs.  Transcendentals never pair, and two instructions that read
# scalar registers do not pair with each other.
VALU_PEAK = 256 * 4 * 32 * 2.4e9


def templates_bench(sfh, torch, ctx, O, args, ncpu):
    """The 700-template workload (70 ages x 10 metallicities, 2000-point
    isochrones, 200 x 300 grid, IMF 0.08-120 Msun, colour-magnitude
    correlation 0.6) built by partial_cmd_smooth_batch: the host preparation
    and the deposition on the device.  Returns the report and the device
    templates."""
    from sfh_amd import synth
    edges = synth.edges_200x300()
    nbins = (edges[0].length - 1) * (edges[1].length - 1)
    specs = synth.specs_70x10()
    T = len(specs)
    imf = synth.Kroupa2001(0.08, 120.0)
    t0 = time.time()
    isos = sfh.IsochroneBatch.from_list([synth.isochrone(s.log_age, s.mh, 2000, s.seed, mmax_imf=120.0)
                                         for s in specs])
    gen_s = time.time() - t0               # synthetic inputs (packed host arrays), not part of the path
    kw = dict(dmod=24.0, normalize_value=1.0, mean_mass=imf.mean(), correlation=0.6, edges=edges)
    errs, comps = [synth.PHOTERR] * 2, [synth.COMPLETENESS] * 2
    out, (npts, pairs) = sfh.partial_cmd_smooth_batch(isos, errs, 1, (0, 1), imf, comps, return_info=True,
                                                      ctx=ctx, **kw)
    nrep = 5
    reps = []
    for _ in range(nrep):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = sfh.partial_cmd_smooth_batch(isos, errs, 1, (0, 1), imf, comps, ctx=ctx, **kw)
        torch.cuda.synchronize()
        reps.append(time.perf_counter() - t0)
    e2e_s = float(np.median(reps))
    ctx.set_timing(True)
    ctx.reset_stats()
    for _ in range(nrep):
        out = sfh.partial_cmd_smooth_batch(isos, errs, 1, (0, 1), imf, comps, ctx=ctx, **kw)
    ctx.synchronize()
    kern = {k: ctx.kernel_stats(k) for k in ("pcs_count", "pcs_points", "pcs_desc", "deposit_order", "deposit_cov")}
    ctx.set_timing(False)
    kms = {k: v[0] / max(v[1], 1) for k, v in kern.items()}
    dep_ms = kms["deposit_cov"]
    prep_ms = kms["pcs_count"] + kms["pcs_points"] + kms["pcs_desc"] + kms["deposit_order"]
    pairs_per_s = pairs / (dep_ms * 1e-3)
    # VALU instructions per pair of the gather's dominant path (the Gaussian
    # form over a full-width row strip of 16 bins, 94% of the pairs): per GL
    # node and bin one FMA and one recurrence multiply (6 per pair), plus the
    # strip's set-up amortised over its 16 bins (139 VALU instructions per
    # strip in the compiled kernel).  The published peak counts two wave64
    # instructions per 4 cycles per SIMD; what FP32 code of the strip's mix
    # (an FMA reading a scalar register, a two-register multiply) reaches is
    # measured in the same run by the library's issue-rate probe.
    inst_per_strip = 139.0
    ops_per_pair = inst_per_strip / 16.0
    achieved = pairs_per_s * ops_per_pair
    issue = ctx.valu_issue_rate()
    roof = {"bound": "valu", "kernel": "k_dep_gather<false>", "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12,
            "unit": "T VALU lane-instructions/s", "frac": achieved / VALU_PEAK, "traffic": None,
            "measured_issue_rate": issue / 1e12, "frac_of_measured_issue_rate": achieved / issue,
            "peak_note": "peak: the published FP32 vector rate (157.3 TFLOP/s = 78.6 T FMA lane-ops/s), two VALU "
                         "instructions per 4 cycles per SIMD from two waves (no packed FP32 in this build); "
                         "measured_issue_rate: the rate independent streams of the strip's instruction mix reach "
                         "on this device in this run (sfh_probe_valu_issue, four waves per SIMD)",
            "algorithmic": f"{ops_per_pair:g} VALU instructions per (point, bin) pair (the Gaussian-form strip: "
                           f"3 GL nodes x (accumulate FMA + recurrence multiply) per bin, plus the strip's set-up "
                           f"of 5 exp2 and 44 other instructions over 16 bins) x {pairs:.4g} pairs per launch; "
                           f"peak = 256 CUs x 4 SIMDs x 32 lanes x 2.4 GHz lane-instructions/s"}
    # the same templates through the host (NumPy) preparation, for comparison
    t0 = time.time()
    pc = [synth.template_points(s, 2000, edges, 120.0, correlation=0.6) for s in specs]
    host_prep_s = time.time() - t0
    rep = {"templates": T, "bins": nbins, "kernel_points": npts, "pairs": pairs,
           "end_to_end_s": e2e_s, "end_to_end_s_each": reps, "templates_per_s_end_to_end": T / e2e_s,
           "kernel_ms": kms, "device_prep_ms": prep_ms, "deposit_ms": dep_ms,
           "templates_per_s_kernels": T / ((prep_ms + dep_ms) * 1e-3), "pairs_per_s": pairs_per_s,
           "hbm_GBps_algorithmic": (nbins * T * 4 + npts * 112) / (dep_ms * 1e-3) / 1e9,
           "host_numpy_prep_s": host_prep_s, "synthetic_isochrones_s": gen_s, "roofline": roof,
           "note": "end-to-end (the median of 5 calls) = host isochrone arrays in (packed: one array per quantity with offsets, "
                   "IsochroneBatch), templates resident on the device out (upload, device preparation, "
                   "deposition); the isochrones themselves are synthetic inputs generated beforehand"}
    log(f"templates: {T} in {e2e_s * 1e3:.1f} ms end to end ({T / e2e_s:.4g}/s); kernels: prep {prep_ms:.2f} ms, "
        f"deposit {dep_ms:.2f} ms ({pairs_per_s:.3g} pairs/s, VALU frac {roof['frac']:.2f}); host NumPy prep "
        f"would take {host_prep_s:.2f} s")
    # parity of the bench's own templates on three of them
    ex, ey = (O.Range(e.first, e.step, e.length) for e in edges)
    errs_t, errs_l1, errs_bin = [], [], []
    for t in (0, T // 2, T - 1):
        sp, cm = pc[t]
        ref, _ = O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, 1, sp.weights, ex, ey, cmult=cm,
                                  nthreads=ncpu)
        got = out[t].cpu().numpy().reshape(ref.shape, order="F").astype(np.float64)
        d = np.abs(got - ref)
        errs_t.append(float(np.max(d) / np.max(np.abs(ref))))
        errs_l1.append(float(d.sum() / np.abs(ref).sum()))
        big = np.abs(ref) > 1e-6 * np.max(np.abs(ref))
        errs_bin.append(float(np.max(d[big] / np.abs(ref[big]))))
    rep["max_norm_err_vs_oracle"] = max(errs_t)
    rep["l1_rel_err_vs_oracle"] = max(errs_l1)
    rep["per_bin_rel_err_vs_oracle"] = {"max": max(errs_bin), "bins": "above 1e-6 of the template maximum",
                                        "templates_checked": 3}
    if not args.no_cpu:
        # CPU baseline: the NumPy preparation + the C FP64 bin_cmd_smooth of the
        # oracle on all host cores, bounded to ~cpu_seconds
        k, cpu_s, cpu_pairs = 0, 0.0, 0
        while cpu_s < args.cpu_seconds and k < T:
            s = specs[(k * 37) % T]
            t1 = time.time()
            sp, cm = synth.template_points(s, 2000, edges, 120.0, correlation=0.6)
            _, pr = O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, 1, sp.weights, ex, ey, cmult=cm,
                                     nthreads=ncpu)
            cpu_s += time.time() - t1
            cpu_pairs += pr
            k += 1
        rep["cpu_baseline"] = {"value": k / cpu_s, "unit": "templates/s", "cores": ncpu, "kind": "port",
                               "pairs_per_s": cpu_pairs / cpu_s,
                               "sample": f"{k} of the {T} templates: NumPy preparation + the C FP64 restatement of "
                                         f"bin_cmd_smooth (OpenMP), {cpu_s:.1f} s"}
        rep["gpu_over_cpu_end_to_end"] = (T / e2e_s) / (k / cpu_s)
        log(f"templates cpu oracle: {k / cpu_s:.3g} templates/s on {ncpu} threads")
    return rep, out


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU sharding of "
              "these paths is a design, not something this benchmark runs.", file=sys.stderr)
        sys.exit(2)
    import torch
    import __graft_entry__ as ge
    sfh = ge.load_package()
    from sfh_amd import synth
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.init()
    # A dedicated (default-priority) stream shared by torch and the library, so
    # torch's events bracket exactly the library's kernels.  (torch's default
    # stream has handle 0, which the library would replace by its own stream.)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx = sfh.Context(0, stream=stream.cuda_stream)
    T_, B = 700, args.walkers
    res = {}
    O = ge.load_oracle()
    ncpu = os.cpu_count() or 1
    if args.section == "config0":
        print(json.dumps({"section": "config0", **config0_bench(sfh, torch, ctx, O, args, ncpu)}), flush=True)
        return

    # ---------------------------------------------------- templates (deposition)
    edges = synth.edges_200x300()
    nbins = (edges[0].length - 1) * (edges[1].length - 1)
    res["templates"], out = templates_bench(sfh, torch, ctx, O, args, ncpu)
    if args.section == "templates":
        print(json.dumps({"section": "templates", **res["templates"]}), flush=True)
        return

    tpl = sfh.templates_from_device(out, nbins, ctx=ctx)
    models = np.asfortranarray(out.cpu().numpy().T)        # nbins x 700 FP32 (host copy for the oracle)

    # ------------------------------------------------------------- data, walkers
    rng = np.random.default_rng(20261202)
    x_true = rng.lognormal(0.0, 1.0, T_)
    comp = np.empty(nbins)
    sfh.composite_(comp, x_true, tpl)
    x_true *= 1e6 / comp.sum()
    sfh.composite_(comp, x_true, tpl)
    data = synth.poisson_data(comp, seed=7)
    hd = sfh.HIPData(data, ctx=ctx)
    X = synth.walkers(x_true, B, 0.3, seed=11)
    dX = torch.from_numpy(np.asfortranarray(X).ravel(order="F")).cuda()
    dnl = torch.empty(B, dtype=torch.float64, device="cuda")
    dG = torch.empty(T_ * B, dtype=torch.float64, device="cuda")

    def batched(prec, steps, warmup):
        for _ in range(warmup):
            sfh.fg_batched_device(tpl, hd, dX, B, dnl, dG, precision=prec)
        ctx.set_timing(True)
        ctx.reset_stats()
        with Timer(torch) as tm:
            for _ in range(steps):
                sfh.fg_batched_device(tpl, hd, dX, B, dnl, dG, precision=prec)
        ctx.synchronize()
        stats = {k: ctx.kernel_stats(k) for k in ("batched_prep_x", "batched_composite_ll",
                                                    "batched_gradient", "batched_finish")}
        ctx.set_timing(False)
        return tm.ms / steps, stats

    # oracle on a sample of walkers (errors of every form)
    nsample = 8
    rnl, rG = O.fg_batched(models, data, X[:, :nsample], nthreads=ncpu)

    def errors():
        nl = dnl[:nsample].cpu().numpy()
        G = dG.view(B, T_)[:nsample].cpu().numpy().T
        rel = np.abs(nl - rnl) / np.abs(rnl)
        return {"ll_rel_err_max": float(np.max(rel)), "ll_rel_err_median": float(np.median(rel)),
                "grad_norm_rel_err_max": float(np.max(np.linalg.norm(G - rG, axis=0) / np.linalg.norm(rG, axis=0))),
                "grad_maxabs_rel_err_max": float(np.max(np.max(np.abs(G - rG), axis=0) / np.max(np.abs(rG), axis=0))),
                "walkers_checked": nsample}

    if args.section == "plain":
        d1 = torch.from_numpy(X[:, 0].copy()).cuda()
        n1 = torch.empty(1, dtype=torch.float64, device="cuda")
        g1 = torch.empty(T_, dtype=torch.float64, device="cuda")
        for _ in range(5):
            sfh.fg_device(tpl, hd, d1, n1, g1)
        with Timer(torch) as tm:
            for _ in range(50):
                sfh.fg_device(tpl, hd, d1, n1, g1)
        x1, Gh = X[:, 0].copy(), np.empty(T_)
        for _ in range(5):
            sfh.fg_(True, Gh, x1, tpl, hd)
        t0 = time.perf_counter()
        for _ in range(200):
            sfh.fg_(True, Gh, x1, tpl, hd)
        host_ms = (time.perf_counter() - t0) / 200 * 1e3
        print(json.dumps({"section": "plain", "ms_per_eval": tm.ms / 50, "host_abi_ms_per_eval": host_ms}),
              flush=True)
        return
    ms, stats = batched(args.precision, args.steps, args.warmup)
    err = errors()
    value = B / (ms * 1e-3)
    comp_ms = stats["batched_composite_ll"][0] / max(stats["batched_composite_ll"][1], 1)
    flops = 2.0 * T_ * nbins * B
    peak_key = {"f64": "fp64", "bf16x2": "bf16", "fp32": "fp32"}[args.precision]
    achieved = flops / (comp_ms * 1e-3) / 1e12
    log(f"batched {args.precision}: {ms:.3f} ms/step, {value:.4g} evals/s, composite kernel {comp_ms:.3f} ms "
        f"({achieved:.1f} TFLOP/s), LL rel-err {err['ll_rel_err_max']:.2e}")
    kernels = {k: (v[0] / max(v[1], 1)) for k, v in stats.items()}
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if args.precision == "f64" and os.path.exists(tfile):
        # calibrated FETCH_SIZE + WRITE_SIZE of this kernel from the committed rocprofv3 --pmc runs
        t = json.load(open(tfile)).get("k_gemm_f64_ll")
        if t:
            traffic = t["read_bytes"] + t["write_bytes"]
    roofline = {"bound": "mfma", "kernel": "k_gemm_f64_ll" if args.precision == "f64" else "k_gemm<LL>",
                "achieved": achieved, "peak": PEAK[peak_key], "unit": "TFLOP/s",
                "frac": achieved / PEAK[peak_key], "traffic": traffic,
                "traffic_source": "not measured in this run: calibrated FETCH_SIZE + WRITE_SIZE bytes per dispatch "
                                  "of this kernel on this workload from rocprofv3 --pmc runs "
                                  "(profiles/hbm_traffic.json)" if traffic else None,
                "algorithmic": f"2*T*N*B = {flops:.3e} FLOP per launch (composite), {PEAK[peak_key]} TFLOP/s "
                               f"dense {peak_key} peak"}

    extras = {}
    if not args.no_extras:
        for prec in ("bf16x2", "fp32", "f64"):
            if prec == args.precision:
                continue
            m2, st2 = batched(prec, max(3, args.steps // 4), 1)
            extras[f"batched_{prec}"] = {"evals_per_s": B / (m2 * 1e-3), "ms_per_step": m2,
                                         "kernel_ms": {k: v[0] / max(v[1], 1) for k, v in st2.items()},
                                         **errors()}
            log(f"batched {prec}: {m2:.3f} ms/step, {B / (m2 * 1e-3):.4g} evals/s, {extras[f'batched_{prec}']}")
        # the plain (single coefficient vector) path
        d1 = torch.from_numpy(X[:, 0].copy()).cuda()
        n1 = torch.empty(1, dtype=torch.float64, device="cuda")
        g1 = torch.empty(T_, dtype=torch.float64, device="cuda")
        for _ in range(5):
            sfh.fg_device(tpl, hd, d1, n1, g1)
        nplain = 200
        with Timer(torch) as tm:
            for _ in range(nplain):
                sfh.fg_device(tpl, hd, d1, n1, g1)
        pms = tm.ms / nplain
        # per-kernel times in a separate run (event timing adds host work per launch)
        ctx.set_timing(True)
        ctx.reset_stats()
        for _ in range(50):
            sfh.fg_device(tpl, hd, d1, n1, g1)
        pk = {k: ctx.kernel_stats(k) for k in ("single_fg", "single_reduce")}
        ctx.set_timing(False)
        # one pass: every template element read once (FP32), the FP64 data once
        plain_bytes = T_ * nbins * 4 + nbins * 8
        fgk = pk["single_fg"][0] / max(pk["single_fg"][1], 1)
        bw = plain_bytes / (fgk * 1e-3) / 1e9
        extras["plain"] = {
            "evals_per_s": 1e3 / pms, "ms_per_eval": pms,
            "kernel_ms": {k: v[0] / max(v[1], 1) for k, v in pk.items()},
            "roofline": {"bound": "hbm", "kernel": "k_single_fg", "achieved": bw, "peak": PEAK["hbm"],
                         "unit": "GB/s", "frac": bw / PEAK["hbm"],
                         "algorithmic": f"{plain_bytes / 1e6:.1f} MB per evaluation (templates read once, FP32, "
                                        f"+ FP64 data); the 168 MB template set fits the 256 MB Infinity Cache, "
                                        f"so repeated evaluations can be served above the HBM rate"},
            "ll_rel_err": float(abs(n1.item() - rnl[0]) / abs(rnl[0]))}
        # the drop-in fg! through the host-pointer C ABI (sfh_fg): coefficients
        # in, F and G out, with and without downloading the N-element composite
        x1 = X[:, 0].copy()
        Gh = np.empty(T_)
        comp_h = np.empty(nbins, dtype=np.float32)
        host = {}
        for label, comp_arg in (("no_composite", None), ("with_composite_f32", comp_h)):
            for _ in range(5):
                sfh.fg_(True, Gh, x1, tpl, hd, comp_arg)
            t0 = time.perf_counter()
            for _ in range(nplain):
                sfh.fg_(True, Gh, x1, tpl, hd, comp_arg)
            host[label] = (time.perf_counter() - t0) / nplain * 1e3
        extras["plain"]["host_abi_ms_per_eval"] = host
        log(f"plain through the host C ABI: {host['no_composite']:.4f} ms/eval, {host['with_composite_f32']:.4f} ms "
            f"with the composite downloaded")
        extras["batched_over_plain"] = value / extras["plain"]["evals_per_s"]
        log(f"plain: {pms:.4f} ms/eval ({1e3 / pms:.4g} evals/s); batched/plain = {extras['batched_over_plain']:.1f}x")

    cpu = None
    if not args.no_cpu:
        # oracle fg! over walkers, all host cores, bounded to ~cpu_seconds
        t0 = time.time()
        O.fg_batched(models, data, X[:, :2], nthreads=ncpu)
        per = (time.time() - t0) / 2
        n = int(max(2, min(B, args.cpu_seconds / max(per, 1e-6))))
        t0 = time.time()
        O.fg_batched(models, data, X[:, :n], nthreads=ncpu)
        dt = time.time() - t0
        cpu = {"value": n / dt, "unit": "evals/s", "cores": ncpu, "kind": "port",
               "sample": f"{n} of the {B} walkers, same 700 x 60,000 problem: the C FP64 restatement of "
                         f"fg! (OpenMP over bins and templates), {dt:.1f} s"}
        log(f"cpu oracle: {n / dt:.3g} evals/s on {ncpu} threads")

    if not args.no_extras and not args.no_config0:
        extras["config0"] = config0_bench(sfh, torch, ctx, O, args, ncpu)
    if not args.no_extras and not args.no_multifit:
        extras["multifit"] = multifit_bench(sfh, torch, ctx, O, args, ncpu)
    if not args.no_extras and not args.no_surrogate:
        extras["surrogate"] = surrogate_bench(sfh, torch, ctx, O, args, ncpu)
    if not args.no_extras and not args.no_mock:
        extras["mock"] = mock_bench(sfh, ctx, args)
    if not args.no_extras and not args.no_binaries:
        extras["binaries"] = binaries_bench(sfh, ctx, args)

    line = {
        "metric": "loglike+grad evals/s (700 tpl×60k bins) batched vs plain; templates/s; HBM GB/s",
        "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"f64": "f64/bf16x2", "bf16x2": "bf16x2", "fp32": "f32"}[args.precision],
        "data": "synthetic (seeded isochrones deposited on the GPU; Poisson data; log-normal walkers)",
        "config": {"workload": f"batched fg! of {B} walkers, 700 templates x 60,000 bins "
                               f"({args.precision}: composite on the matrix cores, FP64 log-likelihood)",
                   "templates": T_, "bins": nbins, "walkers": B, "precision": args.precision,
                   "precision_detail": {"f64": "composite on the FP64 matrix cores, FP64 log-likelihood, gradient "
                                               "in split-BF16 hi+lo products with FP32 accumulation",
                                        "bf16x2": "split-BF16 hi+lo products, FP32 accumulation, FP64 LL",
                                        "fp32": "FP32 matrix-core products, FP64 LL"}[args.precision]},
        "roofline": roofline, "cpu_baseline": cpu, "errors": err, "kernel_ms": kernels,
        "templates": res["templates"], **extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
