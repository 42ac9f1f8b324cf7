"""Kriging workloads with search radiuses far beyond the covariance's reach
(the reference accepts radiuses up to 10^6, validation.py:37-38): the search
template is still the +-radius box's entries above C(0)/100, the clusters are
r cells wide (so sparse data take the cluster path over the whole grid), and
the simple-kriging table is clipped to the +-radius box.
"""
import numpy as np

from geo_bsd import synthetic as S

# name: kind, shape, data fraction (or a data count), model (type, ranges, angles, nugget), radiuses, k, seed
CASES = {
    # dense enough that every cluster is over its limit: template walk
    "ok_exp_dense": (0, (48, 40, 8), 0.04, (1, (10.0, 8.0, 3.0), (0.0, 0.0, 0.0), 0.1), (300, 280, 12), 16, 31),
    # eight data in a grid of few, wide clusters (limit |template| / 250 = 11):
    # the cluster path everywhere
    "ok_sph_sparse": (0, (40, 40, 6), 8, (0, (15.0, 15.0, 4.0), (0.0, 0.0, 0.0), 0.05), (300, 300, 10), 8, 32),
    # rotated model (no octant symmetry, full covariance table), simple kriging
    "sk_gauss_rotated": (1, (32, 32, 6), 0.05, (2, (12.0, 6.0, 3.0), (30.0, 0.0, 0.0), 0.1), (260, 260, 8), 12, 33),
    # radius large in x only; median IK over two indicators
    "mik_sph_x": (2, (40, 24, 6), 0.05, (0, (9.0, 6.0, 3.0), (0.0, 0.0, 0.0), 0.02), (400, 6, 3), 12, 34),
}


def workload(name):
    kind, shape, frac, (t, ranges, angles, nugget), radiuses, k, seed = CASES[name]
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    if frac >= 1:
        mask = np.zeros(n, np.uint8)
        mask[rng.choice(n, size=int(frac), replace=False)] = 1
    else:
        mask = S.random_mask(rng, n, frac)
    field = S.moving_average_field(rng, shape, (5, 5, 3), 0.3, 0.1)
    wl = S.Workload(name=name, kind={0: "ok", 1: "sk", 2: "mik"}[kind], shape=shape, mask=mask,
                    values=S.with_undefined(field, mask),
                    cov=dict(type=t, ranges=ranges, angles=angles, sill=1.0, nugget=nugget), radiuses=radiuses,
                    k=k, mean=0.3)
    if kind == 2:
        wl.thresholds = np.quantile(field[mask == 1], [0.5]).astype(np.float32)
        wl.levels = np.array([0.5], np.float64)
    return wl
