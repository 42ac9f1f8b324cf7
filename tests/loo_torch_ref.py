"""FP64 torch reference of the leave-one-out kriging estimates, differentiable
through torch autograd (CPU); used by the tests to check the hand-written
adjoint.  Same covariance model and OK system as oracle/loo_oracle.py, with
the derivatives taken by torch instead of by formula."""
import torch


def _rho(model, x):
    if model in (0, "spherical"):
        return torch.where(x < 1.0, 1.0 - 1.5 * x + 0.5 * x ** 3, torch.zeros_like(x))
    if model in (1, "exponential"):
        return torch.exp(-3.0 * x)
    return torch.exp(-3.0 * x * x)


def _cov(model, params, d, diag=None):
    """params (P, 5); d (..., 3) offsets broadcast against (P, 1, ..., 1)."""
    shape = (params.shape[0],) + (1,) * (d.dim() - 1)
    sill = params[:, 0].reshape(shape)
    nug = params[:, 1].reshape(shape)
    r = params[:, 2:]
    u2 = torch.zeros((params.shape[0],) + d.shape[:-1], dtype=torch.float64)
    for c in range(3):
        u2 = u2 + (d[..., c][None] / r[:, c].reshape(shape)) ** 2
    if diag is not None:  # keep sqrt away from 0 where the sill branch is taken
        u2 = u2 + diag.to(torch.float64)[None]
    x = torch.sqrt(u2)
    c = (sill - nug) * _rho(model, x)
    if diag is not None:
        c = torch.where(diag[None], sill.expand_as(c), c)
    return c


def estimates(params, xyz, values, nbr, model="spherical"):
    """(P, n) estimates; params (P, 5) float64 CPU tensor (may require grad)."""
    xyz = torch.as_tensor(xyz, dtype=torch.float32).to(torch.float64)
    v = torch.as_tensor(values, dtype=torch.float32).to(torch.float64)
    nbr = torch.as_tensor(nbr, dtype=torch.int64)
    n, k = nbr.shape
    pts = xyz[nbr]                                     # (n, k, 3)
    dA = pts[:, :, None, :] - pts[:, None, :, :]       # (n, k, k, 3)
    eye = torch.eye(k, dtype=torch.bool).expand(n, k, k)
    A = _cov(model, params, dA, diag=eye)             # (P, n, k, k)
    b = _cov(model, params, pts - xyz[:, None, :])    # (P, n, k)
    P = params.shape[0]
    K = torch.zeros((P, n, k + 1, k + 1), dtype=torch.float64)
    K[..., :k, :k] = A
    K[..., :k, k] = 1.0
    K[..., k, :k] = 1.0
    rhs = torch.cat([b, torch.ones((P, n, 1), dtype=torch.float64)], dim=-1)
    y = torch.linalg.solve(K, rhs.unsqueeze(-1)).squeeze(-1)
    return (y[..., :k] * v[nbr][None]).sum(-1)


def jacobian(params, xyz, values, nbr, model="spherical"):
    """(P, n, 5) d est[p, i] / d params[p] by forward-mode differentiation."""
    params = torch.as_tensor(params, dtype=torch.float64)
    cols = []
    for c in range(5):
        t = torch.zeros_like(params)
        t[:, c] = 1.0
        _, jv = torch.func.jvp(lambda q: estimates(q, xyz, values, nbr, model), (params,), (t,))
        cols.append(jv)
    return torch.stack(cols, dim=-1)
