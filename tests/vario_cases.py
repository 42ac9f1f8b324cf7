"""Variogram test cases: search templates (the reference's ellipsoid with
several azimuth / dip / rotation settings, overlapping and separated lags)
and seeded grids and point sets."""
import numpy as np

TEMPLATES = [
    # azimuth, dip, rotation, R1, R2, R3, lag_width, lag_sep, tol, num_lags, first
    (0.0, 0.0, 0.0, 20.0, 2.0, 1.0, 1.0, 1.0, 1.0, 10, 0.0),
    (45.0, 0.0, 0.0, 20.0, 3.0, 1.5, 2.0, 1.5, 0.5, 12, 0.5),
    (30.0, 10.0, 5.0, 20.0, 2.5, 2.0, 3.0, 2.0, 2.0, 8, 1.0),
    (90.0, 0.0, 0.0, 20.0, 1.0, 1.0, 0.5, 1.0, 1.0, 15, 0.0),
]


def template(orc, row):
    az, dip, rot, R1, R2, R3, w, sep, tol, nl, first = row
    return dict(dirs=orc.vario_directions(az, dip, rot), R2=R2, R3=R3, lag_width=w, lag_sep=sep, tol=tol,
                num_lags=nl, first=first, angles=(az, dip, rot), R1=R1)


def grid(seed=21, shape=(40, 30, 8), p=0.3):
    from geo_bsd import synthetic as S
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    f = S.moving_average_field(rng, shape, (5, 5, 3), 0.2, 0.05)
    m = (rng.random(n) < p).astype(np.uint8)
    return f, m


def points(seed=22, n=1500):
    rng = np.random.default_rng(seed)
    xyz = (rng.random((n, 3)) * [40, 40, 10]).astype(np.float32)
    v = (np.sin(xyz[:, 0] / 6) + 0.3 * rng.standard_normal(n)).astype(np.float32)
    return xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy(), v
