"""FP64 CPU oracle of the leave-one-out kriging path (numpy).

TEST INFRASTRUCTURE ONLY: tests/ and bench.py's checks import this module;
the product path (geo_bsd.loo, csrc/loo*.hip) never does.

Restates, per point i and parameter set p (sill, nugget, rx, ry, rz):
  covariance  C(d) = sill if h < 1e-4 else (sill - nugget) * rho(h / r0),
              h / r0 = sqrt((dx/rx)^2 + (dy/ry)^2 + (dz/rz)^2)
              (src/geo_bsd/hpgl/cov_model.h:17-77 with zero angles, :146-183)
  OK system   [A 1; 1' 0] [w; mu] = [b; 1], est = w . v
              (src/geo_bsd/hpgl/my_kriging_weights.h:198-361), solved here
              with numpy.linalg.solve on the bordered matrix
  neighbours  the k nearest other points by Euclidean distance, ties by the
              lower index.
"""
import numpy as np


def rho(model, x):
    x = np.asarray(x, dtype=np.float64)
    if model in (0, "spherical"):
        return np.where(x < 1.0, 1.0 - 1.5 * x + 0.5 * x ** 3, 0.0)
    if model in (1, "exponential"):
        return np.exp(-3.0 * x)
    return np.exp(-3.0 * x * x)


def cov(model, prm, d):
    """Covariance for offsets d (..., 3) under prm (sill, nugget, rx, ry, rz)."""
    sill, nug, rx, ry, rz = [float(v) for v in prm]
    u = d / np.array([rx, ry, rz])
    x = np.sqrt((u * u).sum(-1))
    return np.where(x * rx < 1e-4, sill, (sill - nug) * rho(model, x))


def knn(xyz, k):
    xyz = np.asarray(xyz, dtype=np.float32).astype(np.float64)
    n = xyz.shape[0]
    out = np.empty((n, k), dtype=np.int64)
    idx = np.arange(n)
    for i in range(n):
        d = ((xyz - xyz[i]) ** 2).sum(1)
        d[i] = np.inf
        order = np.lexsort((idx, d))  # by distance, then index
        out[i] = order[:k]
    return out


def estimates(params, xyz, values, nbr, model="spherical"):
    """(P, n) leave-one-out estimates in FP64 (NaN where the system is
    singular)."""
    xyz = np.asarray(xyz, dtype=np.float32).astype(np.float64)
    v = np.asarray(values, dtype=np.float32).astype(np.float64)
    nbr = np.asarray(nbr)
    params = np.atleast_2d(np.asarray(params, dtype=np.float64))
    n, k = nbr.shape
    est = np.empty((params.shape[0], n))
    for p, prm in enumerate(params):
        for i in range(n):
            pts = xyz[nbr[i]]
            A = cov(model, prm, pts[:, None, :] - pts[None, :, :])
            b = cov(model, prm, pts - xyz[i])
            K = np.ones((k + 1, k + 1))
            K[:k, :k] = A
            K[k, k] = 0.0
            try:
                y = np.linalg.solve(K, np.concatenate([b, [1.0]]))
            except np.linalg.LinAlgError:  # singular (coincident neighbours)
                est[p, i] = np.nan
                continue
            est[p, i] = y[:k] @ v[nbr[i]]
    return est
