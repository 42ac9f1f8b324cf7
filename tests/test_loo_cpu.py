"""Leave-one-out kriging: the FP64 oracle against closed forms, invariants and
an independent torch implementation (CPU only)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))


def _points(n=150, seed=0):
    rng = np.random.default_rng(seed)
    xyz = (rng.random((n, 3)) * np.array([60.0, 60.0, 15.0])).astype(np.float32)
    vals = (np.sin(xyz[:, 0] / 7.0) + np.cos(xyz[:, 1] / 5.0) + 0.1 * xyz[:, 2]).astype(np.float32)
    params = np.array([[1.0, 0.1, 12.0, 9.0, 4.0], [1.3, 0.0, 25.0, 25.0, 6.0], [0.8, 0.25, 6.0, 14.0, 3.0]])
    return xyz, vals, params


def test_knn_is_sorted_and_excludes_self():
    from oracle import loo_oracle as O
    xyz, _, _ = _points()
    nbr = O.knn(xyz, 8)
    x = xyz.astype(np.float64)
    for i in range(len(x)):
        assert i not in nbr[i]
        d = ((x[nbr[i]] - x[i]) ** 2).sum(1)
        assert np.all(np.diff(d) >= 0)
        others = np.delete(np.arange(len(x)), np.r_[nbr[i], i])
        assert d[-1] <= ((x[others] - x[i]) ** 2).sum(1).min()


@pytest.mark.parametrize("model", ["spherical", "exponential", "gaussian"])
def test_oracle_matches_torch_reference(model):
    import torch
    import loo_torch_ref as T
    from oracle import loo_oracle as O
    xyz, vals, params = _points()
    nbr = O.knn(xyz, 12)
    a = O.estimates(params, xyz, vals, nbr, model)
    b = T.estimates(torch.tensor(params), xyz, vals, nbr, model).numpy()
    assert np.allclose(a, b, rtol=1e-11, atol=1e-12)


def test_single_neighbour_copies_its_value():
    from oracle import loo_oracle as O
    xyz, vals, params = _points()
    nbr = O.knn(xyz, 1)
    est = O.estimates(params, xyz, vals, nbr)
    assert np.allclose(est, vals.astype(np.float64)[nbr[:, 0]][None, :], rtol=0, atol=1e-14)


def test_invariant_to_covariance_scale_and_constant_field():
    from oracle import loo_oracle as O
    xyz, vals, params = _points()
    nbr = O.knn(xyz, 10)
    scaled = params.copy()
    scaled[:, :2] *= 3.7
    assert np.allclose(O.estimates(params, xyz, vals, nbr), O.estimates(scaled, xyz, vals, nbr), rtol=1e-10,
                       atol=1e-12)
    const = np.full_like(vals, 0.42)
    assert np.allclose(O.estimates(params, xyz, const, nbr), float(np.float32(0.42)), rtol=0, atol=1e-12)


@pytest.mark.parametrize("model", ["spherical", "gaussian"])
def test_torch_reference_derivative_matches_finite_differences(model):
    """Pins the reference Jacobian (used against the GPU adjoint) on the
    oracle: central differences, step 1e-6 relative."""
    import loo_torch_ref as T
    from oracle import loo_oracle as O
    xyz, vals, params = _points(60)
    nbr = O.knn(xyz, 8)
    jac = T.jacobian(params, xyz, vals, nbr, model).numpy()
    for c in range(5):
        h = 1e-6 * np.maximum(np.abs(params[:, c]), 1e-2)
        up, dn = params.copy(), params.copy()
        up[:, c] += h
        dn[:, c] -= h
        fd = (O.estimates(up, xyz, vals, nbr, model) - O.estimates(dn, xyz, vals, nbr, model)) / (2 * h[:, None])
        assert np.allclose(jac[..., c], fd, rtol=1e-4, atol=1e-7), c
