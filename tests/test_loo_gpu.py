"""Leave-one-out kriging on the GPU (through the C ABI) against the FP64
oracle, the FP64 torch reference and central finite differences."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
pytestmark = pytest.mark.gpu


def _data(n=400, seed=3, nsets=4):
    rng = np.random.default_rng(seed)
    xyz = (rng.random((n, 3)) * np.array([80.0, 80.0, 20.0])).astype(np.float32)
    vals = (np.sin(xyz[:, 0] / 9.0) * np.cos(xyz[:, 1] / 6.0) + 0.05 * xyz[:, 2]).astype(np.float32)
    params = np.column_stack([rng.uniform(0.8, 1.2, nsets), rng.uniform(0.0, 0.3, nsets),
                              rng.uniform(5.0, 40.0, (nsets, 3))])
    return xyz, vals, params


def _dev(*arrays):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def test_knn_matches_oracle():
    from geo_bsd import loo as L
    from oracle import loo_oracle as O
    xyz, _, _ = _data(2500)
    (dx,) = _dev(xyz)
    for k in (1, 8, 24, 32):
        print(f"[gpu] knn n=2500 k={k}", flush=True)
        got = L.knn(dx, k).cpu().numpy()
        assert np.array_equal(got, O.knn(xyz, k)), k


@pytest.mark.parametrize("model", ["spherical", "exponential", "gaussian"])
@pytest.mark.parametrize("k", [6, 16, 24, 32])
def test_estimates_match_oracle(model, k):
    import torch
    from geo_bsd import loo as L
    from oracle import loo_oracle as O
    xyz, vals, params = _data()
    nbr = O.knn(xyz, k)
    dx, dv, dn, dp = _dev(xyz, vals, nbr.astype(np.int32), params)
    print(f"[gpu] loo estimates model={model} k={k}", flush=True)
    with torch.no_grad():
        est = L.loo_estimates(dp, dx, dv, dn, model).cpu().numpy()
    assert np.allclose(est, O.estimates(params, xyz, vals, nbr, model), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("model", ["spherical", "exponential", "gaussian"])
@pytest.mark.parametrize("k", [8, 24])
def test_adjoint_jacobian_matches_torch_reference(model, k):
    import loo_torch_ref as T
    from geo_bsd import loo as L
    from oracle import loo_oracle as O
    xyz, vals, params = _data(200)
    nbr = O.knn(xyz, k)
    dx, dv, dn, dp = _dev(xyz, vals, nbr.astype(np.int32), params)
    print(f"[gpu] loo jacobian model={model} k={k}", flush=True)
    _, jac = L.loo_jacobian(dp, dx, dv, dn, model)
    ref = T.jacobian(params, xyz, vals, nbr, model).numpy()
    assert np.allclose(jac.cpu().numpy(), ref, rtol=1e-8, atol=1e-11)


@pytest.mark.parametrize("model", ["spherical", "gaussian"])
def test_adjoint_jacobian_matches_finite_differences(model):
    """Central differences of the GPU forward itself, within 1e-4 relative."""
    import torch
    from geo_bsd import loo as L
    from oracle import loo_oracle as O
    xyz, vals, params = _data(300)
    nbr = O.knn(xyz, 24)
    dx, dv, dn, dp = _dev(xyz, vals, nbr.astype(np.int32), params)
    print(f"[gpu] loo finite differences model={model}", flush=True)
    _, jac = L.loo_jacobian(dp, dx, dv, dn, model)
    jac = jac.cpu().numpy()
    for c in range(5):
        h = 1e-6 * np.maximum(np.abs(params[:, c]), 1e-2)
        up, dn_ = params.copy(), params.copy()
        up[:, c] += h
        dn_[:, c] -= h
        with torch.no_grad():
            fu = L.loo_estimates(_dev(up)[0], dx, dv, dn, model).cpu().numpy()
            fd_ = L.loo_estimates(_dev(dn_)[0], dx, dv, dn, model).cpu().numpy()
        fd = (fu - fd_) / (2 * h[:, None])
        scale = np.maximum(np.abs(fd).max(axis=1, keepdims=True), 1e-6)
        assert np.all(np.abs(jac[..., c] - fd) <= 1e-4 * scale), c


def test_loss_gradient_through_autograd_matches_torch_reference():
    import torch
    import loo_torch_ref as T
    from geo_bsd import loo as L
    from oracle import loo_oracle as O
    xyz, vals, params = _data(250, nsets=5)
    nbr = O.knn(xyz, 16)
    dx, dv, dn, dp = _dev(xyz, vals, nbr.astype(np.int32), params)
    print("[gpu] loo loss gradient", flush=True)
    p = dp.clone().requires_grad_(True)
    loss = L.loo_loss(p, dx, dv, dn).sum()
    loss.backward()
    pr = torch.tensor(params, requires_grad=True)
    est = T.estimates(pr, xyz, vals, nbr)
    ref = ((est - torch.tensor(vals, dtype=torch.float32).to(torch.float64)[None]) ** 2).mean(1).sum()
    ref.backward()
    assert abs(loss.item() - ref.item()) <= 1e-12 * max(1.0, abs(ref.item()))
    assert np.allclose(p.grad.cpu().numpy(), pr.grad.numpy(), rtol=1e-8, atol=1e-13)


def test_duplicate_points_fall_back_to_the_neighbour_mean():
    import torch
    from geo_bsd import loo as L
    from oracle import loo_oracle as O
    xyz, vals, params = _data(120)
    xyz[5] = xyz[6]  # two coincident points make the systems that hold both singular
    nbr = O.knn(xyz, 6)
    dx, dv, dn, dp = _dev(xyz, vals, nbr.astype(np.int32), params)
    print("[gpu] loo duplicate points", flush=True)
    with torch.no_grad():
        out = L.LooKriging.apply(dp, dx, dv, dn, "spherical")
    est = out.cpu().numpy()
    both = np.array([(5 in nbr[i]) and (6 in nbr[i]) for i in range(len(xyz))])
    assert both.any()
    mean = vals.astype(np.float64)[nbr].mean(1)
    assert np.allclose(est[:, both], mean[None, both], rtol=0, atol=1e-12)
    ok = ~both
    assert np.allclose(est[:, ok], O.estimates(params, xyz, vals, nbr)[:, ok], rtol=1e-10, atol=1e-12)


def test_fit_decreases_the_loss():
    from geo_bsd import loo as L, synthetic as S
    wl = S.loo_c5(n=3000, nsets=6)
    dx, dv, dp = _dev(wl.xyz, wl.values, wl.params)
    nbr = L.knn(dx, wl.k)
    print("[gpu] loo fit", flush=True)
    fitted, hist = L.fit(dp, dx, dv, nbr, steps=30)
    h = hist.cpu().numpy()
    assert np.all(np.isfinite(h))
    assert np.all(h[-1] < h[0])
