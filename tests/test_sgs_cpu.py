"""Pins the oracle's sequential Gaussian simulation (CPU only).

The reference's simulation cannot be built here (its path generator includes
a prime table the reference does not ship), so each piece is pinned on its
own: the path by a hand-computed cycle and the full-period property, the
uniform draws by the C++ standard's known answer for std::mt19937 and
numpy's independent MT19937 with libstdc++'s canonical formula, the Gaussian
approximations by their published error bounds against scipy, and the walk
(skips, one draw per simulated cell, path order) by a closed-form case.
"""
import numpy as np
import pytest


def test_path_hand_cycle(orc):
    # size 10: the primes dividing 10 are 2 and 5 -> A = 2*5 + 1 = 11; the
    # largest prime <= 10/3 is 3, which does not divide 10 -> C = 3;
    # R0 = 7 % 10.  R <- (11 R + 3) mod 10.
    assert orc.sgs_path(10, 7).tolist() == [7, 0, 3, 6, 9, 2, 5, 8, 1, 4]


@pytest.mark.parametrize("size", [1, 3, 5, 7, 9, 16, 30, 97, 100, 210, 4096, 2 * 3 * 5 * 7 * 11 * 13, 500000,
                                  1000003])
def test_path_full_period(orc, size):
    """Every such size is visited node by node exactly once (Hull-Dobell:
    C coprime to size, A - 1 divisible by every prime factor of size and by
    4 when 4 divides it)."""
    for seed in (0, 1, 12345):
        p = orc.sgs_path(size, seed)
        assert p[0] == seed % size
        assert np.array_equal(np.sort(p), np.arange(size))


def test_path_short_cycles(orc):
    """Where every prime up to size / 3 divides the size the walk's C shares
    a factor with it and the walk cycles early: size 4 (A = 5, C = 2) and
    size 6 (A = 7, C = 2)."""
    assert orc.sgs_path(4, 0).tolist() == [0, 2, 0, 2]
    assert orc.sgs_path(6, 1).tolist() == [1, 3, 5, 1, 3, 5]
    assert orc.sgs_path(2, 1).tolist() == [1, 1]


def test_path_rejects_out_of_range(orc):
    with pytest.raises(ValueError):
        orc.sgs_path(10000001, 1, n=1)


def test_mt19937_known_answer():
    """numpy's legacy MT19937 with an integer seed is std::mt19937 seeded the
    same way: the C++ standard fixes the 10000th output of the default seed
    5489 at 4123659995."""
    rs = np.random.RandomState(5489)
    g = rs.randint(0, 2 ** 32, size=10000, dtype=np.uint64)
    assert int(g[-1]) == 4123659995


@pytest.mark.parametrize("seed", [0, 1, 42, 5489, 2 ** 31 - 1, 2 ** 32 + 7])
def test_uniform_draws(orc, seed):
    """std::uniform_real_distribution<double>(0, 1) over std::mt19937 takes
    two 32-bit outputs per draw: (g1 + g2 2^32) / 2^64 (libstdc++'s
    generate_canonical); the seed is reduced to 32 bits."""
    n = 257
    u = orc.mt_uniform(seed, n)
    g = np.random.RandomState(seed % 2 ** 32).randint(0, 2 ** 32, size=2 * n, dtype=np.uint64)
    ref = [(float(g[2 * i]) + float(g[2 * i + 1]) * 4294967296.0) / 2.0 ** 64 for i in range(n)]
    assert u.tolist() == ref


def test_gauss_quantile_error_bound(orc):
    """The rational approximation's published bound is 1.5e-8 (absolute)."""
    from scipy.special import ndtri
    p = np.concatenate([np.linspace(1e-12, 1e-6, 2001), np.linspace(1e-6, 1 - 1e-6, 20001),
                        1 - np.linspace(1e-12, 1e-6, 2001)])
    err = np.abs(orc.gauss(p, inverse=True) - ndtri(p))
    assert err.max() < 1.6e-8
    # scaled by sqrt(var) and shifted by the mean
    x = orc.gauss(np.array([0.3, 0.9]), mean=2.0, var=4.0, inverse=True)
    assert np.allclose(x, 2.0 + 2.0 * ndtri(np.array([0.3, 0.9])), atol=4e-8)


def test_gauss_quantile_outside_unit_interval(orc):
    """inverse() returns the mean for p <= 0 and p >= 1."""
    assert orc.gauss(np.array([0.0, 1.0, -0.5, 1.5]), mean=3.5, var=2.0, inverse=True).tolist() == [3.5] * 4


def test_gauss_cdf_error_bound(orc):
    """Abramowitz-Stegun 26.2.17: |error| < 7.5e-8."""
    from scipy.special import ndtr
    x = np.linspace(-8, 8, 40001)
    assert np.abs(orc.gauss(x) - ndtr(x)).max() < 7.5e-8


def _sgs_case(seed=3, shape=(9, 7, 4), frac=0.2):
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    mask = (rng.random(n) < frac).astype(np.uint8)
    values = np.where(mask == 1, rng.normal(size=n), -99).astype(np.float32)
    return shape, values, mask


@pytest.mark.parametrize("kind", [0, 1])
def test_sgs_pure_nugget_is_the_path_of_draws(orc, kind):
    """With a pure nugget model the template is the zero offset alone, no
    cell has neighbours, and every simulated cell takes the standard deviate
    of its own draw: the draws are consumed one per simulated cell, in path
    order, skipping data and cells outside the simulation mask."""
    shape, values, mask = _sgs_case()
    n = values.size
    simmask = np.ones(n, np.uint8)
    simmask[::5] = 0
    simmask[3] = 2  # only exactly 1 is simulated
    seed = 20240611
    cov = orc.make_cov(0, (4, 4, 2), sill=1.0, nugget=1.0)
    v, m, order = orc.sgs(kind, cov, (3, 3, 1), 8, shape, values, mask, seed, simmask=simmask)
    path = orc.sgs_path(n, seed)
    expect_order = [int(c) for c in path if mask[c] == 0 and simmask[c] == 1]
    assert order.tolist() == expect_order
    u = orc.mt_uniform(seed, len(expect_order))
    z = orc.gauss(u, inverse=True).astype(np.float32)
    assert np.array_equal(v[expect_order], z)
    assert np.array_equal(v[mask == 1], values[mask == 1])
    assert np.array_equal(m, np.where((mask == 1) | (simmask == 1), 1, 0).astype(np.uint8))


def test_sgs_first_cell_is_kriging(orc):
    """The first simulated cell is kriged on the data alone: its value is
    inverse(u1) of N(estimate, variance) with the estimate and variance of
    simple kriging (mean 0) over the radius-box covariance table."""
    shape, values, mask = _sgs_case(seed=5, shape=(12, 10, 3), frac=0.3)
    seed = 77
    cov = orc.make_cov(1, (6, 6, 2), sill=1.0, nugget=0.1)
    r = (4, 4, 1)
    v, m, order = orc.sgs(1, cov, r, 6, shape, values, mask, seed)
    first = order[0]
    k = orc.krige(1, cov, r, 6, shape, mask, values=values, means=[0.0])
    assert k["status"][first] == orc.ST_KRIGED
    u = orc.mt_uniform(seed, 1)
    want = orc.gauss(u, mean=float(k["value"][first]), var=float(k["var"][first]), inverse=True).astype(np.float32)
    assert v[first] == want[0]


def test_calc_cdf_hand_case():
    from geo_bsd import ContProperty, calc_cdf
    data = np.array([3.0, 1.0, 2.0, 1.0, 9.0, 2.0], np.float32).reshape(6, 1, 1, order="F")
    mask = np.array([1, 1, 1, 1, 0, 1], np.uint8).reshape(6, 1, 1, order="F")
    cdf = calc_cdf(ContProperty(data, mask))
    # distinct informed values 1, 2, 3 with counts 2, 2, 1 of 5; the last dropped
    assert cdf.values.tolist() == [1.0, 2.0]
    assert cdf.probs.tolist() == [np.float32(0.4), np.float32(0.8)]


def test_sgs_lvm_constant_mean_field_is_simple_kriging(orc):
    """A mean field of zeros is the simple-kriging mean 0: the realization is
    the plain simulation's bit for bit; a nonzero field changes it."""
    shape, values, mask = _sgs_case(seed=9, shape=(10, 9, 3), frac=0.25)
    cov = orc.make_cov(1, (5, 5, 2), sill=1.0, nugget=0.1)
    base, bm, _ = orc.sgs(1, cov, (3, 3, 1), 8, shape, values, mask, 5)
    z, zm, _ = orc.sgs(1, cov, (3, 3, 1), 8, shape, values, mask, 5, lvm_mean=np.zeros(values.size, np.float32))
    assert np.array_equal(base, z) and np.array_equal(bm, zm)
    lv, _, _ = orc.sgs(1, cov, (3, 3, 1), 8, shape, values, mask, 5, lvm_mean=np.full(values.size, 0.5, np.float32))
    assert not np.array_equal(lv[mask == 0], base[mask == 0])
    with pytest.raises(ValueError):
        orc.sgs(0, cov, (3, 3, 1), 8, shape, values, mask, 5, lvm_mean=np.zeros(values.size, np.float32))


@pytest.mark.parametrize("shape,r,seed", [((12, 10, 4), (3, 3, 1), 7), ((9, 7, 5), (2, 4, 2), 123),
                                          ((30, 1, 1), (4, 0, 0), 5), ((16, 16, 6), (5, 5, 1), 20240611)])
def test_sgs_schedule_levels_keep_the_walk(orc, shape, r, seed):
    """The product's host schedule (what the GPU simulation runs): the cells
    and deviates of the oracle's walk, and levels under which running a
    level's cells at once gives the one-by-one walk -- two cells inside each
    other's +-r box never share a level and the earlier on the path comes
    first; cells of one level keep path order."""
    from geo_bsd.sgs import sgs_schedule
    n = int(np.prod(shape))
    rng = np.random.default_rng(seed)
    informed = (rng.random(n) < 0.15).astype(np.uint8)
    simmask = (rng.random(n) < 0.8).astype(np.uint8)
    cells, dev, lev = sgs_schedule(shape, r, informed, seed, simulation_mask=simmask)
    path = orc.sgs_path(n, seed)
    seen, expect = set(), []
    for c in path:
        c = int(c)
        if informed[c] or simmask[c] != 1 or c in seen:
            continue
        seen.add(c)
        expect.append(c)
    pos = {c: i for i, c in enumerate(expect)}
    assert sorted(cells.tolist()) == sorted(expect)
    # deviates of the draws in path order
    z = orc.gauss(orc.mt_uniform(seed, len(expect)), inverse=True)
    assert np.array_equal(dev, z[[pos[int(c)] for c in cells]])
    # levels ascend through the output, path order within a level
    assert np.all(np.diff(lev) >= 0)
    for l in np.unique(lev):
        p = [pos[int(c)] for c in cells[lev == l]]
        assert p == sorted(p)
    nx, ny = shape[0], shape[1]
    x, y, zc = cells % nx, (cells // nx) % ny, cells // (nx * ny)
    near = ((np.abs(x[:, None] - x[None, :]) <= r[0]) & (np.abs(y[:, None] - y[None, :]) <= r[1]) &
            (np.abs(zc[:, None] - zc[None, :]) <= r[2]))
    np.fill_diagonal(near, False)
    order = np.array([pos[int(c)] for c in cells])
    earlier = order[:, None] < order[None, :]
    # every pair inside a common box: the earlier cell's level is lower
    i, j = np.nonzero(near & earlier)
    assert np.all(lev[i] < lev[j])
    assert lev.max() + 1 < len(cells) or len(cells) <= 1  # some cells do share levels


def test_sgs_schedule_rejects_bad_input():
    from geo_bsd.sgs import sgs_schedule
    with pytest.raises(ValueError):
        sgs_schedule((4, 4, 2), (1, 1, 1), np.zeros(32, np.uint8), -1)
    with pytest.raises(ValueError):
        sgs_schedule((4, 4, 2), (1, 1, 1), np.zeros(31, np.uint8), 1)


def test_sgs_zero_radius_is_the_path_of_draws(orc):
    """Search radiuses (0, 0, 0): the template is the cell itself, never
    informed when it is simulated, so every simulated cell takes the deviate
    of its own draw, as with a pure nugget; kriging still rejects a zero
    radius (the reference's clusters divide by it)."""
    shape, values, mask = _sgs_case(seed=4)
    seed = 99
    cov = orc.make_cov(0, (1.0, 1.0, 1.0), sill=1.0, nugget=0.0)
    v, m, order = orc.sgs(1, cov, (0, 0, 0), 1, shape, values, mask, seed)
    u = orc.mt_uniform(seed, order.size)
    assert np.array_equal(v[order], orc.gauss(u, inverse=True).astype(np.float32))
    assert m.all()
    with pytest.raises(ValueError):
        orc.krige(0, cov, (0, 0, 0), 4, shape, mask, values=values)


def test_sgs_invalid_kriging_type_raises_key_error():
    """sgs.py:122 looks the kind up in a dict: an unknown kriging_type is a
    KeyError, raised before anything runs."""
    from geo_bsd import ContProperty, CovarianceModel, SugarboxGrid, sgs_simulation
    prop = ContProperty(np.zeros(8, np.float32), np.zeros(8, np.uint8))
    model = CovarianceModel(0, (2.0, 2.0, 1.0), (0.0, 0.0, 0.0), 1.0, 0.0)
    with pytest.raises(KeyError):
        sgs_simulation(prop, SugarboxGrid(2, 2, 2), None, (1, 1, 1), 4, model, 1, kriging_type="invalid")
