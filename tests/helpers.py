"""Shared helpers for the test-suite."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def max_norm_err(a, ref):
    a, ref = np.asarray(a, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


def oracle_range(O, r):
    return O.Range(r.first, r.step, r.length)


def ref_matrices():
    """The reference's 3x3 known-answer inputs (test/fitting/fitting_core_test.jl:132-137)."""
    models = [np.array([[1, 1, 1], [0, 0, 0], [0, 0, 0]], dtype=np.float32),
              np.array([[0, 0, 0], [1, 1, 1], [0, 0, 0]], dtype=np.float32),
              np.array([[0, 0, 0], [0, 0, 0], [1, 1, 1]], dtype=np.float32)]
    data = np.array([[1, 1, 1], [2, 2, 2], [2, 2, 2]], dtype=np.float64)
    coeffs = np.array([1.5, 3.0, 3.0])
    return models, data, coeffs


def stack(models):
    return np.asfortranarray(np.stack([m.ravel(order="F") for m in models], axis=1))


def cpu_fit_templates(O, models, data, x0=None, gtol=1e-8, transform="log"):
    """fit_templates (src/fitting/solvers.jl:172-221; transform="log": MAP then
    MLE in log-coefficients) or fit_templates_fast (:248-275; transform="sqrt")
    on the CPU: SciPy BFGS (inf-norm gradient tolerance, as Optim's g_abstol)
    around the FP64 oracle's fg!.  Returns {"map": (mu, fmin), "mle": (mu, fmin)}
    ("map" absent for "sqrt")."""
    from scipy.optimize import minimize
    models = np.asfortranarray(models, dtype=np.float32)
    T = models.shape[1]
    x0 = np.ones(T) if x0 is None else np.asarray(x0, dtype=np.float64)
    comp, _, _ = O.fg(models, data, x0, want_F=False, want_G=False)
    s = np.sum(models.astype(np.float64) @ x0)
    x0 = x0 * (np.sum(data) / s) if s != 0 else x0        # renormalize_x0
    opts = {"gtol": gtol, "maxiter": 10000, "norm": np.inf}

    def obj(th, mode):
        x = np.exp(th) if transform == "log" else th * th
        f, G, _ = O.fg(models, data, x, nthreads=4)
        if transform == "sqrt":
            return f, G * 2.0 * th
        if mode == "map":
            return f - np.sum(th), G * x - 1.0
        return f, G * x

    out = {}
    if transform == "log":
        r1 = minimize(obj, np.log(x0), jac=True, method="BFGS", args=("map",), options=opts)
        r2 = minimize(obj, r1.x, jac=True, method="BFGS", args=("mle",), options=opts)
        out["map"] = (np.exp(r1.x), float(r1.fun))
        out["mle"] = (np.exp(r2.x), float(r2.fun))
    else:
        r = minimize(obj, np.sqrt(x0), jac=True, method="BFGS", args=("mle",), options=opts)
        out["mle"] = (r.x ** 2, float(r.fun))
    return out
