"""The multi-threshold CDF with order-relation correction (the median-IK
workload as its output is defined) and the FP64 re-solve of the mixed
solver's fallbacks, through the C ABI against the oracle.

* CDF: per threshold k the oracle's median-IK probability (SK of
  I_k = z <= z_k with mean p_k, median_ik.cpp:24-90), then the reference's
  correct_order_relations (indicator_kriging.h:34-59) over the K values of a
  cell.  The device's corrected CDF equals the oracle's corrected CDF within
  float tolerance, and equals the correction applied to its own uncorrected
  output bit for bit.
* Fallback: a cell whose FP32 factor fails or whose smallest FP32 pivot is
  below pivot_tol * C(0) is re-solved in FP64 (the reference's failure
  semantics, dpotrf info != 0, my_kriging_weights.h:147-157, 284-294, stay
  those of the FP64 form).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _wl(name, scale):
    from geo_bsd import synthetic as S
    return {"ok": S.ok_c2, "mik": S.mik_c3, "wells": S.ok_wells_c4}[name](scale)


def _model(wl):
    from geo_bsd import CovarianceModel
    return CovarianceModel(**wl.cov)


def _cov(orc, wl):
    c = wl.cov
    return orc.make_cov(c["type"], c["ranges"], c["angles"], c["sill"], c["nugget"])


def _cdf(wl, order_relations, precision):
    import torch
    from geo_bsd import device as D
    dev = torch.device("cuda:0")
    print(f"[gpu] cdf {wl.shape} k={wl.k} K={len(wl.thresholds)} order_relations={order_relations} "
          f"precision={precision}", flush=True)
    out = D.indicator_cdf(torch.from_numpy(wl.values).to(dev), torch.from_numpy(wl.mask).to(dev), wl.shape,
                          wl.thresholds, wl.levels, wl.radiuses, wl.k, _model(wl), precision=precision,
                          order_relations=order_relations, status=True)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.parametrize("k,scale", [(16, 0.25), (40, 0.08)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_indicator_cdf_order_relations_matches_oracle(orc, k, scale, precision):
    """Five thresholds at the data quantiles 0.1/0.3/0.5/0.7/0.9, one
    spherical model; k = 16 runs the register solver, k = 40 the wide one."""
    wl = _wl("mik", scale)
    wl.k = k
    assert len(wl.thresholds) == 5
    ind = np.stack([(wl.values <= t).astype(np.uint8) for t in wl.thresholds])
    o = orc.krige(2, _cov(orc, wl), wl.radiuses, wl.k, wl.shape, wl.mask, indicators=ind, means=wl.levels,
                  want_neighbours=False)
    # the reference's per-threshold probabilities are float (median IK's
    # double sum cast once), then corrected cell by cell
    want = orc.order_relations(o["prob"].T.astype(np.float32)).T
    raw = _cdf(wl, False, precision)
    got = _cdf(wl, True, precision)
    assert np.array_equal(got["status"], o["status"].astype(np.uint8))
    assert got["mask"].all()
    # the correction, bit for bit, on the device's own probabilities
    assert np.array_equal(got["prob"], orc.order_relations(raw["prob"].T).T)
    # and the corrected CDF against the oracle's
    assert np.allclose(got["prob"], want, rtol=0, atol=2e-6), np.abs(got["prob"] - want).max()
    # a corrected CDF is monotone and inside [0, 1]
    assert (got["prob"] >= 0).all() and (got["prob"] <= 1).all()
    changed = (got["prob"] != raw["prob"]).any(0)
    assert changed.sum() > 0, "no cell needed a correction: the case does not exercise it"
    print(f"[gpu] cells corrected: {int(changed.sum())} of {changed.size}", flush=True)


def _krige(wl, precision, pivot_tol=1e-3):
    import torch
    from geo_bsd import device as D
    dev = torch.device("cuda:0")
    print(f"[gpu] {wl.name} {wl.shape} k={wl.k} precision={precision} pivot_tol={pivot_tol}", flush=True)
    out = D.krige(0, wl.shape, wl.radiuses, wl.k, _model(wl), torch.from_numpy(wl.mask).to(dev),
                  values=torch.from_numpy(wl.values).to(dev), precision=precision, est64=True, status=True,
                  pivot_tol=pivot_tol)
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["fallbacks"] = D.last_call_stats()["fp64_fallbacks"]
    return res


@pytest.mark.parametrize("name,scale", [("ok", 0.3), ("wells", 0.2)])
def test_every_cell_falls_back_to_fp64(name, scale):
    """pivot_tol so large that every kriged cell is queued: the mixed call
    then returns the FP64 form's results bit for bit."""
    wl = _wl(name, scale)
    ref = _krige(wl, "fp64")
    got = _krige(wl, "mixed", pivot_tol=1e30)
    kriged = int((ref["status"] == 0).sum())
    assert kriged > 0
    assert got["fallbacks"] == kriged, (got["fallbacks"], kriged)
    for key in ("value", "mask", "variance", "est64", "status"):
        assert np.array_equal(got[key].view(np.uint8), ref[key].view(np.uint8)), key


def test_wells_mixed_fallbacks_within_bound(orc):
    """The Gaussian model with nugget 0.01 (the wells workload): with a
    threshold that queues part of the cells, the queued cells carry the FP64
    bits and every estimate stays within 1e-6 of the FP64 oracle."""
    wl = _wl("wells", 0.2)
    o = orc.krige(0, _cov(orc, wl), wl.radiuses, wl.k, wl.shape, wl.mask, values=wl.values, want_neighbours=False)
    kr = o["status"] == orc.ST_KRIGED
    ref = _krige(wl, "fp64")
    dflt = _krige(wl, "mixed")
    err = np.abs(dflt["est64"][kr] - o["est64"][kr]).max()
    print(f"[gpu] default pivot_tol: {dflt['fallbacks']} fallbacks, max err {err:.3g}", flush=True)
    assert err <= 1e-6
    # A = (sill - nugget) R + nugget I has every eigenvalue >= nugget, so no
    # pivot falls below 0.01 C(0) and the default threshold never fires here.
    # A threshold between the smallest and largest FP32 pivots queues some
    # cells; the others keep the refined FP32 solution.
    for tol in (0.011, 0.013, 0.016, 0.02, 0.03, 0.05, 0.08, 0.12):
        part = _krige(wl, "mixed", pivot_tol=tol)
        if 0 < part["fallbacks"] < kr.sum():
            break
    same = (part["est64"].view(np.uint64) == ref["est64"].view(np.uint64)) & kr
    print(f"[gpu] pivot_tol {tol}: {part['fallbacks']} fallbacks, {int(same.sum())} cells with FP64 bits", flush=True)
    assert 0 < part["fallbacks"] < kr.sum()
    assert same.sum() >= part["fallbacks"]
    assert np.array_equal(part["status"], o["status"].astype(np.uint8))
    assert np.abs(part["est64"][kr] - o["est64"][kr]).max() <= 1e-6
