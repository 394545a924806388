"""CPU checks of host-side fitting helpers that do not touch the GPU."""
import numpy as np
import pytest


def test_construct_x0_doc_example(sfh):
    # src/fitting/utilities.jl:24-28: sum close to normalize_value, constant SFR
    x0 = sfh.construct_x0(np.tile([7.0, 8.0, 9.0], 3), 10.0, normalize_value=5.0)
    assert x0.shape == (9,)
    assert 4.9 < x0.sum() <= 5.0
    # constant SFR: mass per unique age proportional to its time span
    la = np.array([7.0, 8.0, 9.0])
    dt = np.diff(10.0 ** np.concatenate([la, [10.0]]))
    x1 = sfh.construct_x0(la, 10.0, normalize_value=1.0)
    assert np.allclose(x1 / x1.sum(), dt / dt.sum(), rtol=1e-12)


def test_construct_x0_checks(sfh):
    with pytest.raises(sfh.ArgumentError):
        sfh.construct_x0(np.array([7.0, 10.5]), 10.0)
