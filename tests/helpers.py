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
