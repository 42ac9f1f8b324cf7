"""Sequential Gaussian simulation on the GPU against the oracle's one-by-one
walk with the same path and the same draws.  The GPU runs, level by level,
cells that lie outside each other's search box, which changes nothing in
what each cell sees.  The informed masks and the simulated cells agree
exactly.  Values agree to float rounding: the GPU's FP64 kriging variance
differs from the oracle's in the last bits, which can move a drawn value by
a float32 ulp and, through later estimates, others by about as much
(tolerance 1e-4 on values of order 1).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    dict(name="sk_exponential", kind=1, shape=(24, 20, 6), frac=0.08, cov=(1, (8, 8, 3), 0.1), r=(5, 5, 2), k=12,
         seed=11),
    dict(name="ok_spherical", kind=0, shape=(20, 18, 5), frac=0.10, cov=(0, (7, 7, 3), 0.05), r=(4, 4, 2), k=16,
         seed=3),
    dict(name="sk_gaussian_k24", kind=1, shape=(40, 36, 8), frac=0.05, cov=(2, (10, 10, 3), 0.2), r=(6, 6, 2),
         k=24, seed=2024),
    dict(name="ok_k32", kind=0, shape=(30, 28, 6), frac=0.06, cov=(1, (9, 9, 3), 0.1), r=(5, 5, 2), k=32, seed=9),
    # above 32 neighbours: a wavefront per cell
    dict(name="sk_k40", kind=1, shape=(28, 26, 6), frac=0.08, cov=(1, (9, 9, 3), 0.1), r=(6, 6, 2), k=40, seed=41),
    dict(name="ok_k64", kind=0, shape=(30, 28, 6), frac=0.10, cov=(0, (10, 10, 3), 0.05), r=(7, 7, 3), k=64,
         seed=64),
]


def _data(shape, frac, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    mask = (rng.random(n) < frac).astype(np.uint8)
    values = np.where(mask == 1, rng.normal(size=n), 0.0).astype(np.float32)
    return values, mask


def _models(orc, c):
    from geo_bsd import CovarianceModel
    t, ranges, nugget = c["cov"]
    return orc.make_cov(t, ranges, sill=1.0, nugget=nugget), CovarianceModel(t, ranges, (0.0, 0.0, 0.0), 1.0, nugget)


def _device_sgs(values, mask, c, model, simmask=None, cdf=None, lvm=None):
    import torch
    from geo_bsd import device as D
    print(f"[gpu] sgs {c['name']} shape={c['shape']} k={c['k']} kind={c['kind']}", flush=True)
    dev = torch.device("cuda:0")
    v = torch.from_numpy(values.copy()).to(dev)
    m = torch.from_numpy(mask.copy()).to(dev)
    sm = None if simmask is None else torch.from_numpy(simmask).to(dev)
    lm = None if lvm is None else torch.from_numpy(lvm).to(dev)
    D.sgs_simulation(v, m, c["shape"], c["r"], c["k"], model, c["seed"], "sk" if c["kind"] == 1 else "ok", cdf=cdf,
                     simulation_mask=sm, mean_field=lm)
    torch.cuda.synchronize()
    return v.cpu().numpy(), m.cpu().numpy()


def _compare(gv, gm, ov, om):
    assert np.array_equal(gm, om)
    d = np.abs(gv.astype(np.float64) - ov.astype(np.float64))
    assert d.max() <= 1e-4, (d.max(), int((d > 0).sum()))
    return int((d > 0).sum())


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_sgs_matches_oracle(orc, c):
    values, mask = _data(c["shape"], c["frac"], c["seed"] + 100)
    cov, model = _models(orc, c)
    ov, om, order = orc.sgs(c["kind"], cov, c["r"], c["k"], c["shape"], values, mask, c["seed"])
    assert order.size == int((mask == 0).sum())
    gv, gm = _device_sgs(values, mask, c, model)
    diff = _compare(gv, gm, ov, om)
    print(f"[gpu] {c['name']}: {order.size} cells simulated, {diff} differ by rounding", flush=True)
    # data are kept bit for bit
    assert np.array_equal(gv[mask == 1], values[mask == 1])


def test_sgs_simulation_mask_and_transform(orc):
    c = CASES[0]
    values, mask = _data(c["shape"], c["frac"], 7)
    values = np.where(mask == 1, np.exp(values), 0.0).astype(np.float32)  # skewed data
    cov, model = _models(orc, c)
    rng = np.random.default_rng(8)
    simmask = (rng.random(values.size) < 0.6).astype(np.uint8)
    simmask[:5] = 2  # only exactly 1 is simulated
    vals = np.unique(values[mask == 1])
    cnt = np.array([(values[mask == 1] == x).sum() for x in vals], np.float64)
    cdf_v = vals[:-1].astype(np.float32)
    cdf_p = np.add.accumulate(cnt[:-1] / mask.sum()).astype(np.float32)
    ov, om, order = orc.sgs(c["kind"], cov, c["r"], c["k"], c["shape"], values, mask, c["seed"], simmask=simmask,
                            cdf=(cdf_v, cdf_p))
    gv, gm = _device_sgs(values, mask, c, model, simmask=simmask, cdf=(cdf_v, cdf_p))
    _compare(gv, gm, ov, om)
    assert np.array_equal(gm, ((mask == 1) | (simmask == 1)).astype(np.uint8))


def test_geo_sgs_simulation_matches_oracle(orc):
    """geo_bsd.sgs_simulation (the reference's Python entry point over the
    host ABI) with calc_cdf, a simulation mask and hard data."""
    from geo_bsd import CdfData, ContProperty, CovarianceModel, SugarboxGrid, calc_cdf, sgs_simulation
    c = CASES[1]
    values, mask = _data(c["shape"], c["frac"], 21)
    values = np.where(mask == 1, 0.3 + 0.05 * values, -99).astype(np.float32)
    nx, ny, nz = c["shape"]
    prop = ContProperty(values.reshape(c["shape"], order="F"), mask.reshape(c["shape"], order="F"))
    t, ranges, nugget = c["cov"]
    model = CovarianceModel(t, ranges, (0.0, 0.0, 0.0), 2.0, 2.0 * nugget)  # scaled to unit sill inside
    cdf = calc_cdf(prop)
    assert isinstance(cdf, CdfData)
    out = sgs_simulation(prop, SugarboxGrid(nx, ny, nz), cdf, c["r"], c["k"], model, c["seed"], kriging_type="ok")
    cov = orc.make_cov(t, ranges, sill=1.0, nugget=nugget)
    ov, om, order = orc.sgs(0, cov, c["r"], c["k"], c["shape"], values, mask, c["seed"],
                            cdf=(cdf.values, cdf.probs))
    _compare(out.data.reshape(-1, order="F"), out.mask.reshape(-1, order="F"), ov, om)
    assert out.mask.all()
    # the input property is untouched
    assert np.array_equal(prop.data.reshape(-1, order="F"), values)


def test_geo_sgs_without_hard_data(orc):
    from geo_bsd import ContProperty, CovarianceModel, SugarboxGrid, sgs_simulation
    c = CASES[0]
    values, mask = _data(c["shape"], c["frac"], 5)
    prop = ContProperty(values.reshape(c["shape"], order="F"), mask.reshape(c["shape"], order="F"))
    t, ranges, nugget = c["cov"]
    model = CovarianceModel(t, ranges, (0.0, 0.0, 0.0), 1.0, nugget)
    out = sgs_simulation(prop, SugarboxGrid(*c["shape"]), None, c["r"], c["k"], model, 4, use_harddata=False)
    cov = orc.make_cov(t, ranges, sill=1.0, nugget=nugget)
    zero = np.zeros_like(mask)
    ov, om, _ = orc.sgs(1, cov, c["r"], c["k"], c["shape"], np.zeros_like(values), zero, 4)
    _compare(out.data.reshape(-1, order="F"), out.mask.reshape(-1, order="F"), ov, om)


def test_sgs_invalid_parameters_raise():
    import torch
    from geo_bsd import CovarianceModel
    from geo_bsd import device as D
    dev = torch.device("cuda:0")
    v = torch.zeros(8 * 8 * 2, dtype=torch.float32, device=dev)
    m = torch.zeros(8 * 8 * 2, dtype=torch.uint8, device=dev)
    model = CovarianceModel(0, (4, 4, 2), (0.0, 0.0, 0.0), 1.0, 0.0)
    with pytest.raises(RuntimeError):
        D.sgs_simulation(v, m, (8, 8, 2), (2, 2, 1), 80, model, 1)  # max_neighbours above 64
    with pytest.raises(RuntimeError):
        D.sgs_simulation(v, m, (8, 8, 2), (2, 2, 1), 8, model, -3)  # negative seed


def test_sgs_locally_varying_mean(orc):
    """Simple kriging with one mean per cell, with and without the
    normal-score transform (which also takes the means)."""
    c = CASES[2]
    values, mask = _data(c["shape"], c["frac"], 31)
    cov, model = _models(orc, c)
    rng = np.random.default_rng(32)
    lvm = (0.3 * rng.standard_normal(values.size)).astype(np.float32)
    ov, om, _ = orc.sgs(1, cov, c["r"], c["k"], c["shape"], values, mask, c["seed"], lvm_mean=lvm)
    gv, gm = _device_sgs(values, mask, c, model, lvm=lvm)
    _compare(gv, gm, ov, om)
    pos = np.where(mask == 1, np.exp(values), 0.0).astype(np.float32)
    vals = np.unique(pos[mask == 1])
    cnt = np.array([(pos[mask == 1] == x).sum() for x in vals], np.float64)
    cdf = (vals[:-1].astype(np.float32), np.add.accumulate(cnt[:-1] / mask.sum()).astype(np.float32))
    lvm2 = np.exp(lvm).astype(np.float32)
    ov, om, _ = orc.sgs(1, cov, c["r"], c["k"], c["shape"], pos, mask, c["seed"], cdf=cdf, lvm_mean=lvm2)
    gv, gm = _device_sgs(pos, mask, c, model, cdf=cdf, lvm=lvm2)
    _compare(gv, gm, ov, om)


def test_geo_sgs_lvm_matches_oracle(orc):
    from geo_bsd import ContProperty, CovarianceModel, SugarboxGrid, sgs_simulation
    c = CASES[0]
    values, mask = _data(c["shape"], c["frac"], 41)
    prop = ContProperty(values.reshape(c["shape"], order="F"), mask.reshape(c["shape"], order="F"))
    t, ranges, nugget = c["cov"]
    model = CovarianceModel(t, ranges, (0.0, 0.0, 0.0), 1.0, nugget)
    lvm = np.linspace(-0.5, 0.5, values.size).astype(np.float32)
    out = sgs_simulation(prop, SugarboxGrid(*c["shape"]), None, c["r"], c["k"], model, 17,
                         mean=lvm.reshape(c["shape"], order="F"))
    cov = orc.make_cov(t, ranges, sill=1.0, nugget=nugget)
    ov, om, _ = orc.sgs(1, cov, c["r"], c["k"], c["shape"], values, mask, 17, lvm_mean=lvm)
    _compare(out.data.reshape(-1, order="F"), out.mask.reshape(-1, order="F"), ov, om)


def _realizations(values, mask, c, model, seeds, **kw):
    import torch
    from geo_bsd import device as D
    print(f"[gpu] sgs_realizations {c['name']} shape={c['shape']} seeds={seeds}", flush=True)
    dev = torch.device("cuda:0")
    v = torch.from_numpy(values.copy()).to(dev)
    m = torch.from_numpy(mask.copy()).to(dev)
    ov, om = D.sgs_realizations(v, m, c["shape"], c["r"], c["k"], model, seeds,
                                "sk" if c["kind"] == 1 else "ok", **kw)
    torch.cuda.synchronize()
    # the conditioning data are left as they were
    assert np.array_equal(v.cpu().numpy(), values) and np.array_equal(m.cpu().numpy(), mask)
    return ov.cpu().numpy(), om.cpu().numpy()


@pytest.mark.parametrize("ci", [0, 3])
def test_sgs_realizations_match_oracle(orc, ci):
    """Realizations run together equal the oracle's one-by-one walk for each
    seed, and the one-realization device call for the same seed bit for bit."""
    c = CASES[ci]
    values, mask = _data(c["shape"], c["frac"], c["seed"] + 200)
    cov, model = _models(orc, c)
    seeds = [c["seed"], 1, 77, 123456, 0]
    rv, rm = _realizations(values, mask, c, model, seeds)
    assert rv.shape == (len(seeds), values.size)
    for i, s in enumerate(seeds):
        ov, om, _ = orc.sgs(c["kind"], cov, c["r"], c["k"], c["shape"], values, mask, s)
        _compare(rv[i], rm[i], ov, om)
    one = dict(c, seed=seeds[2])
    gv, gm = _device_sgs(values, mask, one, model)
    assert np.array_equal(rv[2], gv) and np.array_equal(rm[2], gm)
    # different seeds give different realizations
    assert not np.array_equal(rv[0], rv[1])


def test_sgs_realizations_transform_mask_and_lvm(orc):
    """The batched call with a simulation mask, a CDF and a mean field: each
    row equals the single call with its seed."""
    c = CASES[2]
    values, mask = _data(c["shape"], c["frac"], 51)
    pos = np.where(mask == 1, np.exp(values), 0.0).astype(np.float32)
    _, model = _models(orc, c)
    rng = np.random.default_rng(52)
    simmask = (rng.random(values.size) < 0.7).astype(np.uint8)
    vals = np.unique(pos[mask == 1])
    cnt = np.array([(pos[mask == 1] == x).sum() for x in vals], np.float64)
    cdf = (vals[:-1].astype(np.float32), np.add.accumulate(cnt[:-1] / mask.sum()).astype(np.float32))
    lvm = np.exp(0.2 * rng.standard_normal(values.size)).astype(np.float32)
    import torch
    dev = torch.device("cuda:0")
    seeds = [5, 6, 7]
    rv, rm = _realizations(pos, mask, c, model, seeds, cdf=cdf, simulation_mask=torch.from_numpy(simmask).to(dev),
                           mean_field=torch.from_numpy(lvm).to(dev))
    for i, s in enumerate(seeds):
        gv, gm = _device_sgs(pos, mask, dict(c, seed=s), model, simmask=simmask, cdf=cdf, lvm=lvm)
        assert np.array_equal(rv[i], gv) and np.array_equal(rm[i], gm)


def test_sgs_realizations_invalid():
    import torch
    from geo_bsd import CovarianceModel
    from geo_bsd import device as D
    dev = torch.device("cuda:0")
    v = torch.zeros(8 * 8 * 2, dtype=torch.float32, device=dev)
    m = torch.zeros(8 * 8 * 2, dtype=torch.uint8, device=dev)
    model = CovarianceModel(0, (4, 4, 2), (0.0, 0.0, 0.0), 1.0, 0.0)
    with pytest.raises(RuntimeError):
        D.sgs_realizations(v, m, (8, 8, 2), (2, 2, 1), 8, model, [1, -2])  # a negative seed
    with pytest.raises(ValueError):
        D.sgs_realizations(v, m, (8, 8, 2), (2, 2, 1), 8, model, [])
