"""Generate golden vectors for the oracle with mpmath (40 significant digits).

Independent of the oracle's C code: the kernel formulas are evaluated here
in arbitrary precision straight from the reference's definitions, the
footprint rules restated from the reference's lines in plain Python.

  kernels.json   gaussian_psf_covariant (src/StarFormationHistories.jl:315-334)
                 and gaussian_int_general (:198-205) on 7x7 pixel stencils,
                 sigma in {0.2, 1, 5, 50} pixels, c in {-1, 0, 0.6, 1}.
  template.json  bin_cmd_smooth (:574-621) of 20 seeded points on a 24 x 32
                 grid for cov_mult in {-1, 0, 1}, kernel values in mpmath.
  likelihood.json loglikelihood (fitting_base.jl:84-96) and ∇loglikelihood!
                 (:265-285) on a 200-bin, 6-template seeded case.

Run:  python tests/golden/gen_golden.py   (writes next to this script)
"""
import json
import math
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))

GLX = [mp.mpf(-0.7745966692414834), mp.mpf(0.0), mp.mpf(0.7745966692414834)]   # :303 (Float64 literals)
GLW = [mp.mpf(0.5555555555555556), mp.mpf(0.8888888888888888), mp.mpf(0.5555555555555556)]  # :304


def cov_kernel(x, y, hx, hy, x0, y0, sx, sy, c, A):
    x, y, hx, hy, x0, y0, sx, sy, c, A = map(mp.mpf, (x, y, hx, hy, x0, y0, sx, sy, c, A))
    dx = x - x0
    pref = A / 2 / mp.sqrt(2 * mp.pi) / sy
    s2 = mp.sqrt(2)
    r = mp.mpf(0)
    for k in range(3):
        yv = GLX[k] * hy + y
        yw = GLW[k] * hy
        Dy = yv - y0
        Dx = dx + Dy * c
        r += yw * mp.exp(-(Dy / sy) ** 2 / 2) * (mp.erf((Dx + hx) / s2 / sx) + mp.erf((-Dx + hx) / s2 / sx))
    return r * pref


def asym_kernel(dx, dy, hx, hy, sx, sy, A):
    dx, dy, hx, hy, sx, sy, A = map(mp.mpf, (dx, dy, hx, hy, sx, sy, A))
    s2 = mp.sqrt(2)
    return A / 4 * (mp.erf((dx - hx) / (s2 * sx)) - mp.erf((dx + hx) / (s2 * sx))) * \
        (mp.erf((dy - hy) / (s2 * sy)) - mp.erf((dy + hy) / (s2 * sy)))


def kernels():
    out = []
    x0, y0 = 0.3, -0.2
    for s in (0.2, 1.0, 5.0, 50.0):
        for c in (-1.0, 0.0, 0.6, 1.0):
            vals = [[float(cov_kernel(i, j, 0.5, 0.5, x0, y0, s, 0.7 * s, c, 2.5))
                     for j in range(-3, 4)] for i in range(-3, 4)]
            out.append({"kind": "covariant", "x0": x0, "y0": y0, "sx": s, "sy": 0.7 * s, "c": c,
                        "A": 2.5, "h": 0.5, "grid": list(range(-3, 4)), "values": vals})
        vals = [[float(asym_kernel(i - x0, j - y0, 0.5, 0.5, s, 0.7 * s, 2.5))
                 for j in range(-3, 4)] for i in range(-3, 4)]
        out.append({"kind": "asymmetric", "x0": x0, "y0": y0, "sx": s, "sy": 0.7 * s, "A": 2.5,
                    "h": 0.5, "grid": list(range(-3, 4)), "values": vals})
    return out


def jl_round(x):
    return int(np.rint(x))


def template_case(cov_mult, seed):
    rng = np.random.default_rng(seed)
    ex = (-0.2, 0.05, 25)        # first, step, length (edges)
    ey = (20.0, 0.25, 33)
    nx, ny = ex[2] - 1, ey[2] - 1
    P = 20
    colors = rng.uniform(-0.3, 1.1, P)
    mags = rng.uniform(19.5, 28.5, P)
    cerr = 10 ** rng.uniform(-2.2, -0.3, P)
    merr = 10 ** rng.uniform(-2.2, 0.0, P)
    w = rng.uniform(0.5, 3.0, P)
    mat = [[mp.mpf(0)] * ny for _ in range(nx)]
    for p in range(P):
        if cov_mult == 0:
            # :598-609 and addstar! pixel space :348-364
            x0 = (colors[p] - ex[0]) / ex[1] + 1.0
            y0 = (mags[p] - ey[0]) / ey[1] + 1.0
            sx, sy = cerr[p] / ex[1], merr[p] / ey[1]
            cx, cy = math.ceil(sx * 10.0), math.ceil(sy * 10.0)
            x, y = jl_round(x0), jl_round(y0)
            xo, yo = max(1, cx // 2), max(1, cy // 2)
            xa, xb = max(1, x - xo), min(nx, x + xo)
            ya, yb = max(1, y - yo), min(ny, y + yo)
            if xb - xa + 1 <= 1 or yb - ya + 1 <= 1:
                continue
            for i in range(xa, xb + 1):
                for j in range(ya, yb + 1):
                    mat[i - 1][j - 1] += asym_kernel(i + 0.5 - x0, j + 0.5 - y0, 0.5, 0.5, sx, sy, w[p])
        else:
            # addstar! real space :366-408
            xp = jl_round((colors[p] - ex[0]) / ex[1] + 1.0)
            yp = jl_round((mags[p] - ey[0]) / ey[1] + 1.0)
            xo = max(1, jl_round(15.0 * cerr[p] / ex[1] / 2.0))
            yo = max(1, jl_round(10.0 * merr[p] / ey[1] / 2.0))
            xa, xb = max(1, xp - xo), min(nx, xp + xo)
            ya, yb = max(1, yp - yo), min(ny, yp + yo)
            lx, ly = xb - xa + 1, yb - ya + 1
            if lx <= 1 or ly <= 1:
                continue
            xs = mp.mpf((xa + 0.5 - 1.0) * ex[1] + ex[0])
            xe = mp.mpf((xb + 0.5 - 1.0) * ex[1] + ex[0])
            ys = mp.mpf((ya + 0.5 - 1.0) * ey[1] + ey[0])
            ye = mp.mpf((yb + 0.5 - 1.0) * ey[1] + ey[0])
            xst, yst = (xe - xs) / (lx - 1), (ye - ys) / (ly - 1)
            for i in range(lx):
                for j in range(ly):
                    mat[xa - 1 + i][ya - 1 + j] += cov_kernel(xs + i * xst, ys + j * yst, xst / 2, yst / 2,
                                                              colors[p], mags[p], cerr[p], merr[p],
                                                              cov_mult, w[p])
    return {"cov_mult": cov_mult, "seed": seed, "xedges": ex, "yedges": ey,
            "colors": colors.tolist(), "mags": mags.tolist(), "color_err": cerr.tolist(),
            "mag_err": merr.tolist(), "weights": w.tolist(),
            "matrix_colmajor": [float(mat[i][j]) for j in range(ny) for i in range(nx)]}


def likelihood_case(seed):
    rng = np.random.default_rng(seed)
    N, T = 200, 6
    models = rng.uniform(0, 1, (N, T)).astype(np.float32)
    models[rng.uniform(size=(N, T)) < 0.2] = 0.0
    coeffs = rng.uniform(0.5, 4.0, T)
    data = rng.poisson(3.0, N).astype(np.float64)
    data[:10] = 0.0
    eps = mp.mpf(np.finfo(np.float64).eps)
    m = [sum(mp.mpf(float(models[i, j])) * mp.mpf(coeffs[j]) for j in range(T)) for i in range(N)]
    ll = mp.mpf(0)
    R = []
    for i in range(N):
        ci = max(m[i], eps)
        ni = mp.mpf(data[i])
        ll += (ni - ci - ni * mp.log(ni / ci)) if ni > 0 else -ci
        R.append(1 - ni / ci)
    G = [-sum(mp.mpf(float(models[i, j])) * R[i] for i in range(N)) for j in range(T)]
    return {"models_colmajor": models.ravel(order="F").astype(float).tolist(), "nbins": N, "ntemplates": T,
            "coeffs": coeffs.tolist(), "data": data.tolist(), "composite": [float(v) for v in m],
            "loglikelihood": float(ll), "grad": [float(g) for g in G]}


def main():
    with open(os.path.join(HERE, "kernels.json"), "w") as f:
        json.dump(kernels(), f)
    with open(os.path.join(HERE, "template.json"), "w") as f:
        json.dump([template_case(c, 11 + k) for k, c in enumerate((-1, 0, 1))], f)
    with open(os.path.join(HERE, "likelihood.json"), "w") as f:
        json.dump(likelihood_case(5), f)


if __name__ == "__main__":
    main()
