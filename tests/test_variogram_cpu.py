"""Pins the variogram oracle: against the reference's own _cvariogram build
(oracle/_ref) bit for bit, against brute-force pair enumeration, and the
library's host-side ellipsoid against the reference's."""
import numpy as np
import pytest

import vario_cases as V


@pytest.mark.parametrize("row", V.TEMPLATES)
def test_directions_match_reference(ref, row):
    az, dip, rot = row[:3]
    assert np.array_equal(ref.vario_directions(az, dip, rot), ref.ref_vario_directions(az, dip, rot))


def test_library_directions_match_reference(ref):
    """fill_ellipsoid_directions of libhpgl.so (host code, no GPU needed)."""
    from geo_bsd import cvariogram as CV
    for row in V.TEMPLATES:
        az, dip, rot = row[:3]
        e = CV.Ellipsoid(20, 2, 1, az, dip, rot).ell
        got = np.array([list(e.direction1.data), list(e.direction2.data), list(e.direction3.data)])
        assert np.array_equal(got, ref.ref_vario_directions(az, dip, rot))


@pytest.mark.parametrize("row", V.TEMPLATES)
def test_grid_matches_reference(ref, row):
    t = V.template(ref, row)
    shape = (24, 18, 6)
    f, m = V.grid(shape=shape)
    assert np.array_equal(ref.vario_grid(t, f, m, shape), ref.ref_vario_grid(t, f, m, shape))


@pytest.mark.parametrize("row", V.TEMPLATES)
def test_points_match_reference(ref, row):
    t = V.template(ref, row)
    xs, ys, zs, vs = V.points(n=600)
    assert np.array_equal(ref.vario_points(t, xs, ys, zs, vs), ref.ref_vario_points(t, xs, ys, zs, vs))


def test_grid_against_brute_force(orc):
    """Every informed pair enumerated directly in numpy (azimuth 0: the tunnel
    is along +x; R2 = R3 = 0.5 keeps only offsets on the x axis)."""
    t = V.template(orc, (0.0, 0.0, 0.0, 20.0, 0.5, 0.5, 1.0, 1.0, 1.0, 6, 0.0))
    shape = (16, 5, 3)
    f, m = V.grid(shape=shape, p=0.6)
    F = f.reshape(shape, order="F").astype(np.float64)
    M = m.reshape(shape, order="F").astype(bool)
    exp = np.zeros(6)
    for h in range(0, 6):  # offsets (h, 0, 0), lag h covers [h - 0.5, h + 0.5]
        a, b = F[: shape[0] - h], F[h:]
        ok = M[: shape[0] - h] & M[h:]
        d = (a - b)[ok]
        exp[h] = (d * d).sum() / ok.sum() / 2 if ok.sum() else 0
    assert np.allclose(orc.vario_grid(t, f, m, shape), exp.astype(np.float32), rtol=1e-6, atol=0)


def test_constant_field_is_zero(orc):
    t = V.template(orc, V.TEMPLATES[1])
    shape = (20, 20, 4)
    m = np.ones(int(np.prod(shape)), np.uint8)
    assert not orc.vario_grid(t, np.full(m.size, 3.5, np.float32), m, shape).any()
