"""The GPU experimental variogram against the oracle (pinned to the
reference's _cvariogram build by the CPU tests).  The GPU sums the pairs in
a different order than the reference's single loop, so lags are compared to
within 2 float32 ulps (in practice they are equal)."""
import numpy as np
import pytest

import vario_cases as V

pytestmark = pytest.mark.gpu


def _close(got, exp):
    tol = 2 * np.spacing(np.abs(exp).astype(np.float32))
    bad = np.abs(got.astype(np.float64) - exp.astype(np.float64)) > tol
    assert not bad.any(), list(zip(np.nonzero(bad)[0][:5], got[bad][:5], exp[bad][:5]))


def _templ(row):
    from geo_bsd import cvariogram as CV
    az, dip, rot, R1, R2, R3, w, sep, tol, nl, first = row
    return CV.VariogramSearchTemplate(w, sep, tol, nl, first, CV.Ellipsoid(R1, R2, R3, az, dip, rot))


@pytest.mark.parametrize("row", V.TEMPLATES)
@pytest.mark.parametrize("order", ["F", "C"])
def test_calc_variograms_matches_oracle(orc, row, order):
    from geo_bsd import cvariogram as CV
    shape = (40, 30, 8)
    f, m = V.grid(shape=shape)
    data = np.array(f.reshape(shape, order="F"), order=order)
    mask = np.array(m.reshape(shape, order="F"), order=order)
    print(f"[gpu] variogram grid {row[:3]} order={order}", flush=True)
    borders, g = CV.CalcVariograms(_templ(row), (data, mask))
    exp = orc.vario_grid(V.template(orc, row), f, m, shape)
    _close(g, exp)
    assert np.allclose(borders, np.arange(row[9]) * row[7])


@pytest.mark.parametrize("row", V.TEMPLATES)
def test_point_set_matches_oracle(orc, row):
    from geo_bsd import cvariogram as CV
    xs, ys, zs, vs = V.points(n=3000)
    print(f"[gpu] variogram points {row[:3]}", flush=True)
    _, g = CV.CalcVariogramsFromPointSet(_templ(row), {"X": xs, "Y": ys, "Z": zs, "Property": vs}, None)
    _close(g, orc.vario_points(V.template(orc, row), xs, ys, zs, vs))


def test_many_lags_point_set_passes(orc):
    """More lags than one pass of the point-set kernel holds (64)."""
    from geo_bsd import cvariogram as CV
    row = (20.0, 0.0, 0.0, 50.0, 3.0, 2.0, 0.7, 0.5, 1.0, 150, 0.0)
    xs, ys, zs, vs = V.points(n=1200)
    _, g = CV.CalcVariogramsFromPointSet(_templ(row), {"X": xs, "Y": ys, "Z": zs, "Property": vs})
    _close(g, orc.vario_points(V.template(orc, row), xs, ys, zs, vs))


def test_device_entry_points_and_sampling(orc):
    import torch
    from geo_bsd import cvariogram as CV
    shape = (60, 50, 10)
    f, m = V.grid(shape=shape, p=0.2)
    dev = torch.device("cuda:0")
    t = _templ(V.TEMPLATES[1])
    df, dm = torch.from_numpy(f).to(dev), torch.from_numpy(m).to(dev)
    g = CV.calc_variograms_device(t, df, dm, shape).cpu().numpy()
    _close(g, orc.vario_grid(V.template(orc, V.TEMPLATES[1]), f, m, shape))
    g50a = CV.calc_variograms_device(t, df, dm, shape, percent=50).cpu().numpy()
    g50b = CV.calc_variograms_device(t, df, dm, shape, percent=50).cpu().numpy()
    assert np.array_equal(g50a, g50b)  # deterministic
    ok = g > 0
    assert np.all(np.abs(g50a[ok] - g[ok]) < 0.25 * g[ok])  # half the pairs: same semivariogram, roughly
    xs, ys, zs, vs = V.points(n=2000)
    pts = torch.from_numpy(np.stack([xs, ys, zs, vs], 1).astype(np.float32)).to(dev)
    gp = CV.calc_variograms_points_device(t, pts).cpu().numpy()
    _close(gp, orc.vario_points(V.template(orc, V.TEMPLATES[1]), xs, ys, zs, vs))


def test_empty_and_degenerate():
    from geo_bsd import cvariogram as CV
    t = _templ(V.TEMPLATES[0])
    shape = (10, 10, 2)
    data = np.zeros(shape, np.float32, order="F")
    mask = np.zeros(shape, np.uint8, order="F")
    _, g = CV.CalcVariograms(t, (data, mask))
    assert not g.any()
    _, g = CV.CalcVariogramsFromPointSet(t, {k: np.zeros(0, np.float32) for k in ("X", "Y", "Z", "Property")})
    assert not g.any()
