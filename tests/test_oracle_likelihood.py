"""Pins the oracle's likelihood functions: the reference's own known answers
(test/fitting/fitting_core_test.jl), an mpmath golden (tests/golden/likelihood.json)
and properties the mathematics fixes."""
import numpy as np
import pytest

from helpers import golden, ref_matrices, stack

EPS = np.finfo(np.float64).eps


def test_composite_known_answer(O):
    # fitting_core_test.jl:14-28
    A = np.array([[0, 0, 0], [1, 1, 1], [0, 0, 0]], dtype=np.float64)
    B = np.array([[0, 0, 0], [0, 0, 0], [1, 1, 1]], dtype=np.float64)
    m = stack([A, B])
    assert np.array_equal(m, np.array([[0, 0], [1, 0], [0, 1]] * 3, dtype=np.float64))
    C = O.composite(m, [1.0, 2.0])
    assert np.array_equal(C, [0, 1, 2, 0, 1, 2, 0, 1, 2])


def test_loglikelihood_known_answers(O):
    C = np.array([[1, 1, 1], [2, 2, 2], [3, 3, 3]], dtype=np.float64)
    data = np.array([[1, 1, 1], [2, 2, 2], [2, 2, 2]], dtype=np.float64)
    assert O.loglikelihood(C, data) == pytest.approx(-0.5672093513510137, rel=1e-14)   # :39
    Cz = np.array([[1.5] * 3, [3] * 3, [3] * 3])
    dz = np.array([[0] * 3, [2] * 3, [2] * 3], dtype=np.float64)
    assert O.loglikelihood(Cz, dz) == pytest.approx(-5.6344187027020260, rel=1e-14)    # :66


def test_loglikelihood_zero_is_minus_inf(O):
    # fitting_base.jl:95: a sum of exactly 0 is penalised as -typemax(T)
    assert O.loglikelihood(np.array([2.0, 3.0]), np.array([2.0, 3.0])) == -np.inf


def test_loglikelihood_clamps_to_eps(O):
    # a zero model bin with data is clamped to eps, not a division by zero
    v = O.loglikelihood(np.array([0.0, 1.0]), np.array([2.0, 1.0]))
    ref = (2 - EPS - 2 * np.log(2 / EPS)) + 0.0
    assert v == pytest.approx(ref, rel=1e-14)


def test_gradient_known_answers(O):
    models, data, coeffs = ref_matrices()
    m = stack(models)
    d = data.ravel(order="F")
    # ∇loglikelihood(model, C, data) = -1 (:77-81)
    C = np.array([[1, 1, 1], [2, 2, 2], [3, 3, 3]], dtype=np.float64).ravel(order="F")
    model = np.array([[0, 0, 0], [0, 0, 0], [1, 1, 1]], dtype=np.float64).ravel(order="F")
    assert O.grad_single(model, C, d) == pytest.approx(-1.0, rel=1e-14)
    # zero-data bins (:92-96) -> -3
    mz = np.array([[1, 1, 1], [0, 0, 0], [0, 0, 0]], dtype=np.float64).ravel(order="F")
    Cz = np.array([[1.5] * 3, [3] * 3, [3] * 3]).ravel(order="F")
    dz = np.array([[0] * 3, [2] * 3, [2] * 3], dtype=np.float64).ravel(order="F")
    assert O.grad_single(mz, Cz, dz) == pytest.approx(-3.0, rel=1e-14)
    # ∇loglikelihood!(G, C, models, data) (:132-159)
    comp = O.composite(m, coeffs)
    G = np.empty(3)
    O.grad_inplace(G, comp, m, d)
    assert np.allclose(G, [-1, -1, -1], rtol=1e-14)
    # the reference overwrites composite with 1 - n/c (fitting_base.jl:219, :279)
    assert np.allclose(comp, 1 - d / O.composite(m, coeffs), rtol=1e-15)
    comp3 = O.composite(m, coeffs)
    O.grad_inplace(G, comp3, m, dz)
    assert np.allclose(G, [-3, -1, -1], rtol=1e-14)


def test_fg_known_answer(O):
    # fg! (:163-192): -logL = 1.4180233783775342, -grad = [-1,-1,-1]
    models, data, coeffs = ref_matrices()
    nl, G, comp = O.fg(stack(models), data.ravel(order="F"), coeffs)
    assert nl == pytest.approx(1.4180233783775342, rel=1e-14)
    assert np.allclose(-G, [-1, -1, -1], rtol=1e-14)
    nl2, G2, _ = O.fg(stack(models), data.ravel(order="F"), coeffs, want_G=False)
    assert nl2 == pytest.approx(1.4180233783775342, rel=1e-14) and G2 is None


def test_likelihood_mpmath_golden(O):
    g = golden("likelihood.json")
    N, T = g["nbins"], g["ntemplates"]
    m = np.array(g["models_colmajor"], dtype=np.float32).reshape((N, T), order="F")
    comp = O.composite(m, g["coeffs"])
    assert np.allclose(comp, g["composite"], rtol=1e-14, atol=0)
    ll = O.loglikelihood(comp, g["data"])
    assert ll == pytest.approx(g["loglikelihood"], rel=1e-13)
    G = np.empty(T)
    O.grad_inplace(G, comp, m, np.array(g["data"]))
    assert np.allclose(G, g["grad"], rtol=1e-12)


def test_batched_matches_single(O):
    rng = np.random.default_rng(3)
    N, T, B = 500, 7, 5
    m = rng.uniform(0, 1, (N, T)).astype(np.float32)
    data = rng.poisson(4.0, N).astype(np.float64)
    X = rng.uniform(0.5, 2, (T, B))
    nl, G = O.fg_batched(m, data, X)
    for b in range(B):
        nl1, G1, _ = O.fg(m, data, X[:, b])
        assert nl[b] == nl1 and np.array_equal(G[:, b], G1)


def test_gradient_is_derivative(O):
    # central finite difference of loglikelihood(coeffs) matches -G of fg!
    rng = np.random.default_rng(4)
    N, T = 300, 4
    m = rng.uniform(0, 1, (N, T)).astype(np.float32)
    data = rng.poisson(5.0, N).astype(np.float64)
    x = rng.uniform(1, 3, T)
    _, G, _ = O.fg(m, data, x)
    for j in range(T):
        h = 1e-6 * x[j]
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        fd = (O.loglikelihood(O.composite(m, xp), data) - O.loglikelihood(O.composite(m, xm), data)) / (2 * h)
        assert -G[j] == pytest.approx(fd, rel=1e-6)
