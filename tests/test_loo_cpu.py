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


# ----------------------------------------------------------------------------
# The leave-one-out oracle's covariance and estimate against the grid oracle,
# whose covariance (oracle/hpgl_oracle.cpp, cov_model.h:17-77, 146-183) is
# pinned bitwise to the reference build.  Ranges differ on every axis and the
# offsets are asymmetric, so a swapped axis or a range that multiplies where
# it should divide changes the values.
_MODELS = [(0, "spherical"), (1, "exponential"), (2, "gaussian")]


@pytest.mark.parametrize("code,model", _MODELS)
def test_loo_covariance_equals_grid_oracle(code, model):
    from oracle import loo_oracle as LO
    from oracle import oracle as orc
    rng = np.random.default_rng(11 + code)
    for prm in ([1.0, 0.1, 12.0, 7.0, 3.0], [2.5, 0.0, 5.0, 21.0, 9.5], [0.7, 0.3, 30.0, 30.0, 2.0]):
        sill, nug, rx, ry, rz = prm
        cov = orc.make_cov(code, (rx, ry, rz), (0.0, 0.0, 0.0), sill, nug)
        # general offsets (inside and beyond the ranges), tiny ones on the
        # nugget branch h < 1e-4 and ones straddling the spherical cut-off
        d = rng.uniform(-1.6, 1.6, (1500, 3)) * np.array([rx, ry, rz])
        # h = |(dx, dy rx/ry, dz rx/rz)| < 1e-4 on the nugget branch
        tiny = rng.uniform(-5e-5, 5e-5, (100, 3)) / np.sqrt(3.0) / np.array([1.0, rx / ry, rx / rz])
        u = rng.normal(size=(200, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        cut = u * np.array([rx, ry, rz]) * rng.uniform(0.999, 1.001, (200, 1))
        axes = np.array([[rx * 0.5, 0, 0], [0, ry * 0.5, 0], [0, 0, rz * 0.5], [rx, 0, 0], [0, ry, 0], [0, 0, rz]])
        pts = np.concatenate([d, tiny, cut, axes, -d[:50]])
        want = np.array([orc.cov_value(cov, *p) for p in pts])
        got = LO.cov(model, prm, pts)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-15), (prm, np.abs(got - want).max())
        assert np.all(got[1500:1600] == sill)


@pytest.mark.parametrize("code,model", _MODELS)
def test_loo_estimate_equals_grid_oracle_ordinary_kriging(code, model):
    """Points on integer coordinates: the oracle's leave-one-out estimate of a
    point equals the grid oracle's FP64 ordinary-kriging estimate of that cell
    from the same neighbours (my_kriging_weights.h:198-361)."""
    from oracle import loo_oracle as LO
    from oracle import oracle as orc
    rng = np.random.default_rng(21 + code)
    shape = (25, 19, 9)
    nx, ny, nz = shape
    rad = (7, 5, 2)
    prm = [1.0, 0.08, 13.0, 8.0, 3.5]
    cov = orc.make_cov(code, prm[2:], (0.0, 0.0, 0.0), prm[0], prm[1])
    for trial in range(4):
        mask = np.zeros(shape[::-1], np.uint8)  # (z, y, x): x fastest when flat
        cz, cy, cx = 4, 9, 12
        box = [(z, y, x) for z in range(cz - rad[2], cz + rad[2] + 1) for y in range(cy - rad[1], cy + rad[1] + 1)
               for x in range(cx - rad[0], cx + rad[0] + 1) if (z, y, x) != (cz, cy, cx)]
        pick = rng.choice(len(box), 14, replace=False)
        for i in pick:
            mask[box[i]] = 1
        values = rng.normal(0.2, 0.05, shape[::-1]).astype(np.float32)
        mask_f, values_f = mask.reshape(-1), values.reshape(-1)
        target = (cz * ny + cy) * nx + cx
        kmax = 9 + trial  # the grid oracle's template order picks the subset
        o = orc.krige(0, cov, rad, kmax, shape, mask_f, values=values_f)
        assert o["status"][target] == orc.ST_KRIGED
        nodes = o["nb"][target][: o["count"][target]]
        assert len(nodes) == kmax
        # point set: the neighbours, then the target cell
        allnodes = np.concatenate([nodes, [target]])
        xyz = np.stack([allnodes % nx, (allnodes // nx) % ny, allnodes // (nx * ny)], 1).astype(np.float32)
        nbr = np.full((len(allnodes), kmax), 0, np.int64)
        nbr[-1] = np.arange(kmax)
        est = LO.estimates(np.array([prm]), xyz, values_f[allnodes], nbr, model)[0, -1]
        want = o["est64"][target]
        assert est == pytest.approx(want, rel=1e-10, abs=1e-12), (trial, est, want)


def test_neighbour_check_runs_once_per_content():
    """The index check of a neighbour tensor (two host reads) runs once per
    tensor content: again after an in-place change, which a bad index then
    fails."""
    import torch
    from geo_bsd import loo as L
    nbr = torch.tensor([[1, 2], [0, 2], [0, 1]], dtype=torch.int32)
    calls = []
    orig_min = torch.Tensor.min

    def counting_min(self, *a, **k):
        calls.append(1)
        return orig_min(self, *a, **k)

    torch.Tensor.min = counting_min
    try:
        L._check_nbr(nbr, 3)
        L._check_nbr(nbr, 3)
        assert len(calls) == 1
        nbr[0, 0] = 7  # in place: the version changes
        with pytest.raises(ValueError, match=r"\[0, n\)"):
            L._check_nbr(nbr, 3)
        assert len(calls) == 2
        with pytest.raises(ValueError):
            L._check_nbr(nbr[:2], 3)
    finally:
        torch.Tensor.min = orig_min
