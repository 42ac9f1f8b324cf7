"""Search radiuses far beyond the covariance's reach (up to 400 cells) on the
GPU: neighbour sets and search paths equal the oracle's cell by cell, FP64
estimates within 1e-9, float outputs within 1e-5 relative, the mixed form
within 1e-6 of the FP64 oracle.  The device keeps only what the covariance
reaches: a padded mask with the template's reach as halo, covariance tables
over twice the neighbour reach (candidates beyond them are below C(0)/100)."""
import numpy as np
import pytest

import large_radius_cases as LR
from test_gpu_parity import KIND, _cov, _device, _node_lists, _oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(LR.CASES))
def test_large_radius_matches_oracle(orc, name):
    wl = LR.workload(name)
    o = _oracle(orc, wl)
    g = _device(wl)
    searched = o["status"] != orc.ST_INFORMED
    assert np.array_equal(g["count"][searched], o["count"][searched])
    assert np.array_equal(_node_lists(wl, g["count"], g["nboff"])[searched], o["nb"][searched].astype(np.int64))
    assert np.array_equal(g["cluster_path"][searched], o["path"][searched].astype(bool))
    assert np.array_equal(g["status"], o["status"].astype(np.uint8))
    kr = o["status"] == orc.ST_KRIGED
    assert kr.sum() > 0
    if wl.kind == "mik":
        assert np.allclose(g["est64"][:, kr], o["est64"][:, kr], rtol=0, atol=1e-9)
        assert np.allclose(g["prob"][:, kr], o["prob"][:, kr].astype(np.float32), rtol=1e-5, atol=1e-6)
    else:
        assert np.all(np.abs(g["value"] - o["value"]) <= 1e-5 * np.maximum(1.0, np.abs(o["value"])))
        assert np.allclose(g["est64"][kr], o["est64"][kr], rtol=1e-9, atol=1e-12)
    gm = _device(wl, "mixed")
    assert np.array_equal(gm["status"], o["status"].astype(np.uint8))
    assert np.abs(gm["est64"][..., kr] - o["est64"][..., kr]).max() <= 1e-6


def test_large_radius_cluster_path_taken(orc):
    """The sparse case runs the cluster path over candidates from the whole grid."""
    wl = LR.workload("ok_sph_sparse")
    g = _device(wl)
    searched = wl.mask == 0
    assert g["cluster_path"][searched].all()


def test_large_radius_sgs_matches_oracle(orc):
    """Sequential simulation with radiuses (300, 300, 6) against the oracle's walk."""
    import torch
    from geo_bsd import CovarianceModel, device as D
    shape, r, k, seed = (24, 22, 5), (300, 300, 6), 12, 77
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    mask = (rng.random(n) < 0.08).astype(np.uint8)
    values = np.where(mask == 1, rng.normal(size=n), 0.0).astype(np.float32)
    cov = orc.make_cov(1, (6.0, 6.0, 2.0), sill=1.0, nugget=0.1)
    model = CovarianceModel(1, (6.0, 6.0, 2.0), (0.0, 0.0, 0.0), 1.0, 0.1)
    ov, om, order = orc.sgs(1, cov, r, k, shape, values, mask, seed)
    print(f"[gpu] sgs large radius shape={shape} r={r} k={k}", flush=True)
    dev = torch.device("cuda:0")
    v = torch.from_numpy(values.copy()).to(dev)
    m = torch.from_numpy(mask.copy()).to(dev)
    D.sgs_simulation(v, m, shape, r, k, model, seed, "sk")
    torch.cuda.synchronize()
    gv, gm = v.cpu().numpy(), m.cpu().numpy()
    assert np.array_equal(gm, om)
    assert np.abs(gv.astype(np.float64) - ov.astype(np.float64)).max() <= 1e-4
