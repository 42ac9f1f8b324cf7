"""The oracle against the reference's own build for search radiuses far
beyond the covariance's reach: neighbour sets and float estimates bit for
bit.  And the library's plan: its template equals the oracle's, and the
radius box is still scanned whole (the template is the box's entries above
C(0)/100, whatever the radius)."""
import numpy as np
import pytest

import large_radius_cases as LR


def _cov(o, wl):
    c = wl.cov
    return o.make_cov(c["type"], c["ranges"], c["angles"], c["sill"], c["nugget"])


@pytest.mark.parametrize("name", ["ok_exp_dense", "ok_sph_sparse", "sk_gauss_rotated"])
def test_large_radius_oracle_matches_reference(ref, name):
    wl = LR.workload(name)
    kind = {"ok": 0, "sk": 1}[wl.kind]
    cov = _cov(ref, wl)
    o = ref.krige(kind, cov, wl.radiuses, wl.k, wl.shape, wl.mask, values=wl.values, means=[wl.mean])
    cnt, idx = ref.ref_neighbours(kind, cov, wl.radiuses, wl.k, wl.shape, wl.mask)
    searched = o["status"] != ref.ST_INFORMED
    assert np.array_equal(o["count"][searched], cnt[searched])
    assert np.array_equal(o["nb"][searched], idx[searched])
    if kind == 0:
        rv, rm = ref.ref_ordinary_kriging(cov, wl.radiuses, wl.k, wl.shape, wl.values, wl.mask, threads=4)
    else:
        rv, rm = ref.ref_simple_kriging(cov, wl.radiuses, wl.k, wl.shape, wl.values, wl.mask, mean=wl.mean,
                                        threads=4)
    assert np.array_equal(o["mask"], rm)
    assert np.array_equal(o["value"].view(np.uint32), rv.view(np.uint32))


@pytest.mark.parametrize("name", list(LR.CASES))
def test_large_radius_library_template(orc, name):
    from geo_bsd import CovarianceModel, device as D
    wl = LR.workload(name)
    lib = D.search_template(CovarianceModel(**wl.cov), wl.radiuses)
    t, _ = orc.template(_cov(orc, wl), wl.radiuses)
    assert np.array_equal(np.asarray(lib), np.asarray(t))
