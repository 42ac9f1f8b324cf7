"""Seeded synthetic workloads for the kriging benchmarks and tests.

Recipe (one numpy Generator per workload, ``default_rng(seed)``):
  * field: i.i.d. N(0,1) on the grid, box moving average of the given window
    (scipy.ndimage.uniform_filter, mode 'reflect'), standardised, then
    rescaled to the target mean / sd and cast to float32;
  * hard data: exactly round(p * N) cells drawn without replacement;
  * uninformed cells hold -99.
Arrays are flat in x-fastest order (= the reference's F-order (nx, ny, nz)).
"""
import numpy

UNDEFINED = -99.0


def moving_average_field(rng, shape, window, mean=0.0, sd=1.0):
    from scipy.ndimage import uniform_filter
    nx, ny, nz = shape
    g = rng.standard_normal((nx, ny, nz))
    g = uniform_filter(g, size=window, mode="reflect")
    g = (g - g.mean()) / g.std()
    return (mean + sd * g).astype(numpy.float32).reshape(-1, order="F")


def random_mask(rng, n, p):
    k = int(round(p * n))
    idx = rng.choice(n, size=k, replace=False)
    m = numpy.zeros(n, numpy.uint8)
    m[idx] = 1
    return m


def with_undefined(values, mask):
    v = values.copy()
    v[mask == 0] = UNDEFINED
    return v


class Workload:
    """Attributes: name, kind, shape, values, mask, cov (dict), radiuses, k, and per workload
    mean / thresholds / levels."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    def get(self, key, default=None):
        return self.__dict__.get(key, default)


def sk_c1(scale=1.0, seed=1):
    """Simple kriging, 50x50x10, 2% hard data, spherical (10,10,5), k=12."""
    shape = (max(4, int(50 * scale)), max(4, int(50 * scale)), max(2, int(10 * scale)))
    rng = numpy.random.default_rng(seed)
    n = shape[0] * shape[1] * shape[2]
    f = moving_average_field(rng, shape, (7, 7, 3), 0.20, 0.05)
    m = random_mask(rng, n, 0.02)
    return Workload(name="sk_50x50x10", kind="sk", shape=shape, values=with_undefined(f, m), mask=m,
                    cov=dict(type=0, ranges=(10.0, 10.0, 5.0), angles=(0.0, 0.0, 0.0), sill=1.0, nugget=0.1),
                    radiuses=(10, 10, 5), k=12, mean=0.20)


def ok_c2(scale=1.0, seed=2):
    """Ordinary kriging, 100x100x50 (500k cells), 5% hard data, exponential (20,20,4), k=24."""
    shape = (max(4, int(100 * scale)), max(4, int(100 * scale)), max(2, int(50 * scale)))
    rng = numpy.random.default_rng(seed)
    n = shape[0] * shape[1] * shape[2]
    f = moving_average_field(rng, shape, (9, 9, 3), 0.2, 0.05)
    m = random_mask(rng, n, 0.05)
    return Workload(name="ok_100x100x50", kind="ok", shape=shape, values=with_undefined(f, m), mask=m,
                    cov=dict(type=1, ranges=(20.0, 20.0, 4.0), angles=(0.0, 0.0, 0.0), sill=1.0, nugget=0.05),
                    radiuses=(20, 20, 4), k=24)


def mik_c3(scale=1.0, seed=3):
    """Median IK, 200x200x100, 1% hard data of a lognormal field, 5 thresholds, spherical (15,15,3), k=16."""
    shape = (max(4, int(200 * scale)), max(4, int(200 * scale)), max(2, int(100 * scale)))
    rng = numpy.random.default_rng(seed)
    n = shape[0] * shape[1] * shape[2]
    g = moving_average_field(rng, shape, (9, 9, 3), 0.0, 1.0).astype(numpy.float64)
    z = numpy.exp(0.5 * g).astype(numpy.float32)
    m = random_mask(rng, n, 0.01)
    levels = numpy.array([0.1, 0.3, 0.5, 0.7, 0.9])
    thr = numpy.quantile(z[m == 1], levels).astype(numpy.float32)
    return Workload(name="mik_200x200x100", kind="mik", shape=shape, values=with_undefined(z, m), mask=m,
                    cov=dict(type=0, ranges=(15.0, 15.0, 3.0), angles=(0.0, 0.0, 0.0), sill=0.25, nugget=0.02),
                    radiuses=(15, 15, 3), k=16, thresholds=thr, levels=levels)


def ok_wells_c4(scale=1.0, seed=4):
    """Ordinary kriging from 40 vertical wells, 256x256x64, Gaussian (30,30,6), nugget 0.01, k=32."""
    shape = (max(8, int(256 * scale)), max(8, int(256 * scale)), max(4, int(64 * scale)))
    nx, ny, nz = shape
    rng = numpy.random.default_rng(seed)
    f = moving_average_field(rng, shape, (9, 9, 3), 0.2, 0.05)
    cols = set()
    while len(cols) < min(40, nx * ny):
        xy = rng.integers(0, [nx, ny])
        cols.add((int(xy[0]), int(xy[1])))
    m = numpy.zeros((nx, ny, nz), numpy.uint8)
    for (x, y) in sorted(cols):
        m[x, y, :] = 1
    m = m.reshape(-1, order="F")
    return Workload(name="ok_wells_256x256x64", kind="ok", shape=shape, values=with_undefined(f, m), mask=m,
                    cov=dict(type=2, ranges=(30.0, 30.0, 6.0), angles=(0.0, 0.0, 0.0), sill=1.0, nugget=0.01),
                    radiuses=(30, 30, 6), k=32)


def lvm_c1(scale=1.0, seed=6):
    """Kriging with a locally varying mean, 50x50x10, 2% hard data,
    spherical (10,10,5), k=12: the mean trend
    0.20 + 0.05 sin(x/8) cos(y/11) + 0.004 z plus a 7x7x3 moving-average
    residual of sd 0.03."""
    shape = (max(4, int(50 * scale)), max(4, int(50 * scale)), max(2, int(10 * scale)))
    nx, ny, nz = shape
    rng = numpy.random.default_rng(seed)
    n = nx * ny * nz
    x, y, z = numpy.meshgrid(numpy.arange(nx), numpy.arange(ny), numpy.arange(nz), indexing="ij")
    trend = (0.20 + 0.05 * numpy.sin(x / 8.0) * numpy.cos(y / 11.0) + 0.004 * z).astype(numpy.float32)
    trend = trend.reshape(-1, order="F")
    f = trend + moving_average_field(rng, shape, (7, 7, 3), 0.0, 0.03)
    m = random_mask(rng, n, 0.02)
    return Workload(name=f"lvm_{nx}x{ny}x{nz}", kind="sk", shape=shape, values=with_undefined(f, m), mask=m,
                    lvm_mean=trend, cov=dict(type=0, ranges=(10.0, 10.0, 5.0), angles=(0.0, 0.0, 0.0), sill=1.0,
                                             nugget=0.1),
                    radiuses=(10, 10, 5), k=12)


def ik_c3(scale=1.0, seed=7):
    """Indicator kriging over 4 categories, 200x200x100, 1% hard data: the
    categories are the quartile classes of a 9x9x3 moving-average field;
    spherical models of ranges (8,8,2), (12,12,3), (6,6,2) and again
    (12,12,3) (categories 1 and 3 share a model), sill 0.25, nugget 0.02,
    radiuses equal to the ranges, k = 16; marginals are the hard-data
    proportions."""
    shape = (max(4, int(200 * scale)), max(4, int(200 * scale)), max(2, int(100 * scale)))
    rng = numpy.random.default_rng(seed)
    n = shape[0] * shape[1] * shape[2]
    g = moving_average_field(rng, shape, (9, 9, 3), 0.0, 1.0)
    m = random_mask(rng, n, 0.01)
    cuts = numpy.quantile(g[m == 1], [0.25, 0.5, 0.75])
    data = numpy.searchsorted(cuts, g, side="right").astype(numpy.uint8)
    data[m == 0] = 0
    ranges = [(8.0, 8.0, 2.0), (12.0, 12.0, 3.0), (6.0, 6.0, 2.0), (12.0, 12.0, 3.0)]
    cats = [dict(cov=dict(type=0, ranges=r, angles=(0.0, 0.0, 0.0), sill=0.25, nugget=0.02),
                 radiuses=tuple(int(v) for v in r), k=16) for r in ranges]
    marg = [float((data[m == 1] == c).mean()) for c in range(4)]
    return Workload(name="ik4_200x200x100", kind="ik", shape=shape, values=data, mask=m, categories=cats,
                    marginal=marg)


def loo_c5(n=20000, nsets=64, seed=5):
    """Leave-one-out kriging of scattered points: n float32 points uniform in
    [0,200) x [0,200) x [0,50), values of the 7x7x3 moving-average field
    (mean 0.20, sd 0.05) on a 200x200x50 grid sampled at floor(coordinate);
    spherical model, k = 24; nsets parameter sets with ranges U[5,40] (each
    axis), nugget U[0,0.3], sill 1."""
    rng = numpy.random.default_rng(seed)
    shape = (200, 200, 50)
    field = moving_average_field(rng, shape, (7, 7, 3), 0.20, 0.05)
    xyz = (rng.random((n, 3)) * numpy.array(shape, dtype=numpy.float64)).astype(numpy.float32)
    ix = numpy.minimum(numpy.floor(xyz).astype(numpy.int64), numpy.array(shape) - 1)
    vals = field[ix[:, 0] + shape[0] * (ix[:, 1] + shape[1] * ix[:, 2])]
    ranges = rng.uniform(5.0, 40.0, size=(nsets, 3))
    nugget = rng.uniform(0.0, 0.3, size=nsets)
    params = numpy.column_stack([numpy.ones(nsets), nugget, ranges]).astype(numpy.float64)
    return Workload(name=f"loo_{n}pts_{nsets}sets", kind="loo", xyz=xyz, values=vals.astype(numpy.float32),
                    params=params, k=24, model="spherical")
