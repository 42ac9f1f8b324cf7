"""Edge cases the reference's own tests exercise (tests/python/
test_edge_cases.py), run through geo_bsd on the GPU and checked against the
oracle where it covers them: tiny and flat grids, a zero search radius,
max_neighbours of 0 and 1, simulation masks that select nothing or
everything, and CDF tables whose probabilities reach 0 and 1.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ok(prop, grid, radiuses, k, model):
    import geo_bsd
    return geo_bsd.ordinary_kriging(prop=prop, grid=grid, radiuses=radiuses, max_neighbours=k, cov_model=model)


def _model(ranges, nugget=0.1, type_=0):
    import geo_bsd
    return geo_bsd.CovarianceModel(type=type_, ranges=ranges, angles=(0.0, 0.0, 0.0), sill=1.0, nugget=nugget)


@pytest.mark.parametrize("dims,radiuses,ranges", [((1, 1, 1), (1, 1, 1), (1.0, 1.0, 1.0)),
                                                  ((1, 10, 10), (1, 5, 5), (1.0, 5.0, 5.0)),
                                                  ((10, 10, 1), (5, 5, 1), (5.0, 5.0, 1.0)),
                                                  ((5, 10, 20), (2, 5, 10), (3.0, 5.0, 10.0)),
                                                  ((2, 3, 5), (1, 1, 1), (1.0, 1.0, 1.0))])
def test_small_and_flat_grids_match_oracle(orc, dims, radiuses, ranges):
    """test_edge_cases.py:43-222: 1x1x1, flat and small grids, OK."""
    import geo_bsd
    n = int(np.prod(dims))
    rng = np.random.default_rng(n)
    values = (rng.random(n) * 100).astype(np.float32)
    mask = np.ones(n, np.uint8)
    if n > 1:
        mask[::5] = 0
    print(f"[gpu] edge grid {dims} r={radiuses}", flush=True)
    k = min(12, n)
    out = _ok(geo_bsd.ContProperty(values, mask), geo_bsd.SugarboxGrid(*dims), radiuses, k, _model(ranges))
    cov = orc.make_cov(0, ranges, sill=1.0, nugget=0.1)
    o = orc.krige(0, cov, radiuses, k, dims, mask, values=values)
    got = out.data.reshape(-1, order="F")
    assert np.array_equal(out.mask.reshape(-1, order="F"), o["mask"])
    assert np.all(np.abs(got - o["value"]) <= 1e-5 * np.maximum(1.0, np.abs(o["value"])))
    if n == 1:
        assert got[0] == values[0]


@pytest.mark.parametrize("k", [0, 1])
def test_max_neighbours_zero_and_one(orc, k):
    """test_edge_cases.py:603-664: max_neighbours 1 kriges from the nearest
    datum; 0 either raises or leaves the uninformed cells as they were."""
    import geo_bsd
    dims = (10, 10, 5)
    rng = np.random.default_rng(3)
    values = (rng.random(500) * 100).astype(np.float32)
    mask = np.ones(500, np.uint8)
    mask[::10] = 0
    prop = geo_bsd.ContProperty(values, mask)
    print(f"[gpu] edge max_neighbours={k}", flush=True)
    if k == 0:
        from geo_bsd.validation import ValidationError
        try:
            out = _ok(prop, geo_bsd.SugarboxGrid(*dims), (5, 5, 3), 0, _model((5.0, 5.0, 3.0)))
        except (ValidationError, ValueError, RuntimeError):
            return
        assert np.array_equal(out.mask.reshape(-1, order="F"), mask)
        return
    out = _ok(prop, geo_bsd.SugarboxGrid(*dims), (5, 5, 3), 1, _model((5.0, 5.0, 3.0)))
    cov = orc.make_cov(0, (5.0, 5.0, 3.0), sill=1.0, nugget=0.1)
    o = orc.krige(0, cov, (5, 5, 3), 1, dims, mask, values=values)
    got = out.data.reshape(-1, order="F")
    assert np.array_equal(out.mask.reshape(-1, order="F"), o["mask"])
    assert np.all(np.abs(got - o["value"]) <= 1e-5 * np.maximum(1.0, np.abs(o["value"])))


def test_sgs_zero_radius_matches_oracle(orc):
    """test_edge_cases.py:1256-1301: search radiuses (0, 0, 0) find no
    neighbours, so every simulated cell draws from the CDF alone; the CDF's
    probabilities run from 0 to 1 (deviates of the ends are the mean)."""
    import geo_bsd
    dims = (5, 5, 5)
    rng = np.random.default_rng(42)
    values = (rng.random(125) * 100).astype(np.float32)
    mask = np.ones(125, np.uint8)
    mask[::2] = 0
    cdf = geo_bsd.CdfData(values=np.linspace(0, 100, 50).astype(np.float32),
                          probs=np.linspace(0, 1, 50).astype(np.float32))
    print("[gpu] edge sgs radius 0", flush=True)
    out = geo_bsd.sgs_simulation(prop=geo_bsd.ContProperty(values, mask), grid=geo_bsd.SugarboxGrid(*dims),
                                 cdf_data=cdf, radiuses=(0, 0, 0), max_neighbours=1,
                                 cov_model=_model((1.0, 1.0, 1.0), nugget=0.0), seed=42)
    got = out.data.reshape(-1, order="F")
    assert np.all(np.isfinite(got))
    cov = orc.make_cov(0, (1.0, 1.0, 1.0), sill=1.0, nugget=0.0)
    ov, om, _ = orc.sgs(1, cov, (0, 0, 0), 1, dims, values, mask, 42, cdf=(cdf.values, cdf.probs))
    assert np.array_equal(out.mask.reshape(-1, order="F"), om)
    assert np.abs(got.astype(np.float64) - ov).max() <= 1e-3
    # hard data come back through the normal-score round trip, as the oracle's
    assert np.array_equal(got[mask == 1], ov[mask == 1].astype(np.float32))


def test_sgs_simulation_mask_selecting_nothing_and_everything(orc):
    """test_edge_cases.py:1118-1209: a simulation mask of zeros simulates
    nothing (the data come back as they were); a mask of ones without hard
    data simulates every cell."""
    import geo_bsd
    dims = (10, 10, 5)
    rng = np.random.default_rng(42)
    values = (rng.random(500) * 100).astype(np.float32)
    model = _model((5.0, 5.0, 3.0))
    inf = np.sort(values)
    cdf = geo_bsd.CdfData(values=inf, probs=np.linspace(0, 1, inf.size).astype(np.float32))
    print("[gpu] edge sgs simulation masks", flush=True)
    out = geo_bsd.sgs_simulation(prop=geo_bsd.ContProperty(values, np.ones(500, np.uint8)),
                                 grid=geo_bsd.SugarboxGrid(*dims), cdf_data=cdf, radiuses=(5, 5, 3),
                                 max_neighbours=12, cov_model=model, seed=42, mask=np.zeros(500, np.uint8))
    got = out.data.reshape(-1, order="F")
    assert not np.any(np.isnan(got))
    cov = orc.make_cov(0, (5.0, 5.0, 3.0), sill=1.0, nugget=0.1)
    ov, om, order = orc.sgs(1, cov, (5, 5, 3), 12, dims, values, np.ones(500, np.uint8), 42,
                            simmask=np.zeros(500, np.uint8), cdf=(cdf.values, cdf.probs))
    assert order.size == 0
    assert np.abs(got.astype(np.float64) - ov).max() <= 1e-3
    cdf3 = geo_bsd.CdfData(values=np.array([0.0, 50.0, 100.0], np.float32),
                           probs=np.array([0.33, 0.66, 1.0], np.float32))
    out = geo_bsd.sgs_simulation(prop=geo_bsd.ContProperty(values, np.zeros(500, np.uint8)),
                                 grid=geo_bsd.SugarboxGrid(*dims), cdf_data=cdf3, radiuses=(5, 5, 3),
                                 max_neighbours=12, cov_model=model, seed=42, mask=np.ones(500, np.uint8),
                                 use_harddata=False)
    got = out.data.reshape(-1, order="F")
    assert out.mask.all() and np.all(np.isfinite(got))
    ov, om, order = orc.sgs(1, cov, (5, 5, 3), 12, dims, np.zeros(500, np.float32), np.zeros(500, np.uint8), 42,
                            simmask=np.ones(500, np.uint8), cdf=(cdf3.values, cdf3.probs))
    assert order.size == 500
    assert np.abs(got.astype(np.float64) - ov).max() <= 1e-3


def test_kriging_rejects_zero_radius():
    """Kriging clusters the grid by the search radiuses
    (clusterizer.cpp:87-89), so a zero radius is refused with an error,
    where the reference would divide by zero; simulation accepts it (see
    test_sgs_zero_radius_matches_oracle)."""
    import geo_bsd
    values = np.arange(125, dtype=np.float32)
    mask = np.ones(125, np.uint8)
    mask[::3] = 0
    print("[gpu] edge kriging radius 0", flush=True)
    with pytest.raises(RuntimeError, match="radiuses"):
        _ok(geo_bsd.ContProperty(values, mask), geo_bsd.SugarboxGrid(5, 5, 5), (0, 2, 2), 4, _model((2.0, 2.0, 2.0)))
