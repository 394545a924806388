"""GPU parity of template deposition (bin_cmd_smooth / partial_cmd_smooth,
single and batched) against the oracle and the mpmath golden."""
import numpy as np
import pytest

from helpers import golden, max_norm_err, oracle_range

pytestmark = pytest.mark.gpu

TOL = 1e-5   # template max-normalised error bound


def test_golden_template(sfh, gpu_ctx):
    for case in golden("template.json"):
        ex = sfh.Range(*case["xedges"][:2], case["xedges"][2])
        ey = sfh.Range(*case["yedges"][:2], case["yedges"][2])
        h = sfh.bin_cmd_smooth(case["colors"], case["mags"], case["color_err"], case["mag_err"],
                               case["cov_mult"], weights=case["weights"], edges=(ex, ey))
        ref = np.array(case["matrix_colmajor"]).reshape(h.weights.shape, order="F")
        assert max_norm_err(h.weights, ref) < 2e-6, case["cov_mult"]
        assert h.weights.shape == (ex.length - 1, ey.length - 1)


@pytest.mark.parametrize("y_index,cidx", [(1, (0, 1)), (0, (0, 1))])
def test_partial_cmd_smooth_75x100(sfh, O, gpu_ctx, y_index, cidx):
    from sfh_amd import synth
    edges = synth.edges_75x100()
    imf = synth.Kroupa2001(0.08, 2.0)
    m, mags = synth.isochrone(9.2, -0.8, 500, 21, mmax_imf=2.0)
    h = sfh.partial_cmd_smooth(m, mags, [synth.photerr] * 2, y_index, cidx, imf, [synth.completeness] * 2,
                               dmod=24.0, normalize_value=1e4, mean_mass=imf.mean(), edges=edges)
    ref, _ = O.partial_cmd_smooth(m, mags, [lambda v: float(synth.photerr(v))] * 2, y_index, cidx,
                                  lambda v: float(imf(v)), [lambda v: float(synth.completeness(v))] * 2,
                                  dmod=24.0, normalize_value=1e4, mean_mass=imf.mean(),
                                  ex=oracle_range(O, edges[0]), ey=oracle_range(O, edges[1]), nthreads=8)
    assert max_norm_err(h.weights, ref) < TOL


def test_asymmetric_kernel(sfh, O, gpu_ctx):
    from sfh_amd import synth
    edges = synth.edges_75x100()
    sp, _ = synth.template_points(synth.TemplateSpec(9.8, -1.5, 4), 500, edges, 2.0)
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    ref, _ = O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, 0, sp.weights, ex, ey, nthreads=8)
    h = sfh.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, 0, weights=sp.weights, edges=edges)
    assert max_norm_err(h.weights, ref) < TOL


def test_batch_200x300_shape(sfh, O, gpu_ctx):
    """A few templates at the 200 x 300 / 2000-point config, one launch."""
    from sfh_amd import synth
    edges = synth.edges_200x300()
    specs = synth.specs_70x10()[::97][:6]
    pts = [synth.template_points(s, 2000, edges, 120.0)[0] for s in specs]
    batch = sfh.PointBatch(pts, edges)
    out, pairs = sfh.deposit_batch(batch, return_pairs=True)
    out = out.cpu().numpy()
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    total = 0
    for t, sp in enumerate(pts[:3]):
        ref, pr = O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, 1, sp.weights, ex, ey, nthreads=8)
        total += pr
        assert max_norm_err(out[t].reshape(ref.shape, order="F"), ref) < TOL, t
    counts = sum(O.count_pairs(sp.colors, sp.mags, sp.color_err, sp.mag_err, 1, ex, ey) for sp in pts)
    assert pairs == counts


@pytest.mark.parametrize("cov_mult", [1.0, 0.0])
def test_batch_launch_order_does_not_change_templates(sfh, O, gpu_ctx, cov_mult):
    """Batches of up to 16,384 (template, row band) workgroups run in the
    order of their estimated cost, larger ones in the plain rotation; every
    workgroup writes its own band, so both give the same templates bit for
    bit: 6 templates alone (ordered) and the same 6 repeated to 1,030 templates
    (16,480 workgroups on the 200 x 300 grid, unordered); the covariant and the
    asymmetric kernel."""
    import torch
    from sfh_amd import synth
    edges = synth.edges_200x300()
    specs = synth.specs_70x10()[::131][:6]
    pts = [synth.template_points(s, 2000, edges, 120.0)[0] for s in specs]
    small = sfh.deposit_batch(sfh.PointBatch(pts, edges, cov_mult=cov_mult))
    big = sfh.deposit_batch(sfh.PointBatch([pts[i % 6] for i in range(1030)], edges, cov_mult=cov_mult))
    torch.cuda.synchronize()
    for i in range(1030):
        assert torch.equal(big[i], small[i % 6]), i
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    sp = pts[0]
    ref, _ = O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, cov_mult, sp.weights, ex, ey, nthreads=8)
    assert max_norm_err(small[0].cpu().numpy().reshape(ref.shape, order="F"), ref) < TOL


def test_batch_config1_exact_combination(sfh, O, gpu_ctx):
    """The 700-template workload's exact combination: 200 x 300 grid,
    2000-point isochrones, IMF 0.08-120 Msun, colour-magnitude correlation 0.6
    (per-point covariance multipliers), on 3 templates (youngest, middle,
    oldest) in one launch; max-normalised error <= 1e-5 and per-bin relative
    error <= 2e-5 above 1e-6 of the maximum, pair counts equal."""
    from sfh_amd import synth
    edges = synth.edges_200x300()
    specs = synth.specs_70x10()
    specs = [specs[0], specs[345], specs[-1]]
    pc = [synth.template_points(s, 2000, edges, 120.0, correlation=0.6) for s in specs]
    batch = sfh.PointBatch([p for p, _ in pc], edges, cmult_list=[c for _, c in pc])
    out, pairs = sfh.deposit_batch(batch, return_pairs=True)
    out = out.cpu().numpy()
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    total = 0
    for t, (sp, cm) in enumerate(pc):
        ref, pr = O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, 1, sp.weights, ex, ey,
                                   cmult=cm, nthreads=8)
        total += pr
        got = out[t].reshape(ref.shape, order="F")
        assert max_norm_err(got, ref) <= TOL, t
        big = ref > 1e-6 * ref.max()
        assert np.max(np.abs(got[big] - ref[big]) / ref[big]) <= 2e-5, t
    assert pairs == total


def test_batch_real_valued_covariance_and_empty(sfh, O, gpu_ctx):
    from sfh_amd import synth
    edges = synth.edges_75x100()
    specs = [synth.TemplateSpec(8.5, -1.0, 1), synth.TemplateSpec(9.9, -0.2, 2)]
    pc = [synth.template_points(s, 500, edges, 2.0, correlation=0.6) for s in specs]
    pts = [p for p, _ in pc]
    from sfh_amd.templates import SmoothPoints
    empty = SmoothPoints(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), 1)
    batch = sfh.PointBatch([pts[0], empty, pts[1]], edges, cmult_list=[pc[0][1], np.zeros(0), pc[1][1]])
    out = sfh.deposit_batch(batch).cpu().numpy()
    ex, ey = oracle_range(O, edges[0]), oracle_range(O, edges[1])
    assert not out[1].any()
    for t, k in ((0, 0), (2, 1)):
        sp, cm = pc[k]
        ref, _ = O.bin_cmd_smooth(sp.colors, sp.mags, sp.color_err, sp.mag_err, 1, sp.weights, ex, ey,
                                  cmult=cm, nthreads=8)
        assert max_norm_err(out[t].reshape(ref.shape, order="F"), ref) < TOL


def test_deposited_templates_feed_likelihood(sfh, O, gpu_ctx):
    from sfh_amd import synth
    edges = synth.edges_75x100()
    pts = [synth.template_points(s, 500, edges, 2.0, normalize_value=1e3)[0] for s in synth.specs_10x5()[:8]]
    out = sfh.deposit_batch(sfh.PointBatch(pts, edges))
    nbins = out.shape[1]
    tpl = sfh.templates_from_device(out, nbins)
    models = np.asfortranarray(out.cpu().numpy().T)
    x = np.linspace(1.0, 3.0, 8)
    data = synth.poisson_data(models.astype(np.float64) @ x, seed=1)
    G = np.empty(8)
    nl = sfh.fg_(True, G, x, tpl, data)
    rnl, rG, _ = O.fg(models, data, x)
    assert nl == pytest.approx(rnl, rel=1e-12)
    assert np.allclose(G, rG, rtol=1e-9, atol=1e-12 * np.abs(rG).max())


def test_invalid_inputs(sfh, gpu_ctx):
    edges = (sfh.edges_range(0, 1, 11), sfh.edges_range(0, 1, 11))
    with pytest.raises(sfh.ArgumentError):
        sfh.bin_cmd_smooth([0.5], [0.5], [0.1], [0.1], 2, weights=[1.0], edges=edges)
    with pytest.raises(sfh.InexactError):
        sfh.bin_cmd_smooth([np.nan], [0.5], [0.1], [0.1], 1, weights=[1.0], edges=edges)
    h = sfh.bin_cmd_smooth([], [], [], [], 1, weights=[], edges=edges)
    assert not h.weights.any()
