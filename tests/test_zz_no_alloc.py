"""hpglx_krige_noalloc (every buffer from the library's static scratch) against
hpglx_krige (workspace allocated by the library), through the C ABI: two
calls in a row, the second on a smaller grid, each bit for bit equal to the
allocating entry point's results; a call too large for the scratch is
refused with a message before any kernel runs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc_call(q, stream):
    from geo_bsd import device as D
    rc = D._hpgl_so.hpglx_krige_noalloc(C.byref(q), C.c_void_p(stream))
    if rc != 0:
        raise RuntimeError("hpglx_krige_noalloc: " + D._hpgl_so.hpglx_last_error().decode())


def _run(wl, kind, precision, noalloc, monkeypatch):
    import torch
    from geo_bsd import CovarianceModel, device as D
    dev = torch.device("cuda:0")
    print(f"[gpu] {wl.name} {wl.shape} k={wl.k} {precision} noalloc={noalloc}", flush=True)
    with monkeypatch.context() as m:
        if noalloc:
            m.setattr(D, "_call", _noalloc_call)
        mask = torch.from_numpy(wl.mask).to(dev)
        model = CovarianceModel(**wl.cov)
        if kind == 2:
            out = D.krige(2, wl.shape, wl.radiuses, wl.k, model, mask, zfield=torch.from_numpy(wl.values).to(dev),
                          thresholds=wl.thresholds, marginal=wl.levels, precision=precision, est64=True, status=True,
                          neighbours=True, order_relations=True)
        else:
            out = D.krige(kind, wl.shape, wl.radiuses, wl.k, model, mask, values=torch.from_numpy(wl.values).to(dev),
                          mean=wl.get("mean"), precision=precision, est64=True, status=True, neighbours=True)
        torch.cuda.synchronize()
        stats = D.last_call_stats()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    valid = np.arange(wl.k)[None, :] < np.maximum(res["count"], 0)[:, None]
    res["nboff"] = np.where(valid, res["nboff"], 0)
    return res, stats


def _same(a, b):
    assert a.keys() == b.keys()
    for key in a:
        assert a[key].dtype == b[key].dtype and a[key].shape == b[key].shape, key
        assert np.array_equal(a[key].view(np.uint8), b[key].view(np.uint8)), key


CASES = [
    # (first call, second call on a smaller grid)
    (("ok", 1.0, 0, "mixed"), ("ok", 0.4, 0, "mixed")),
    (("wells", 0.35, 0, "mixed"), ("sk", 0.7, 1, "fp64")),
    (("mik", 0.3, 2, "fp64"), ("mik", 0.15, 2, "mixed")),
]


@pytest.mark.parametrize("first,second", CASES)
def test_noalloc_equals_allocating_entry_point(monkeypatch, first, second):
    from geo_bsd import synthetic as S
    gen = {"ok": S.ok_c2, "sk": S.sk_c1, "mik": S.mik_c3, "wells": S.ok_wells_c4}
    for name, scale, kind, precision in (first, second):
        wl = gen[name](scale)
        ref, ref_stats = _run(wl, kind, precision, False, monkeypatch)
        got, got_stats = _run(wl, kind, precision, True, monkeypatch)
        _same(got, ref)
        assert got_stats == ref_stats


def test_noalloc_refuses_a_call_beyond_its_scratch(monkeypatch):
    """Neighbour lists of 256^3 cells x 24 need 1.6 GB: refused with a
    message naming the bytes, and the next call still works."""
    import torch
    from geo_bsd import CovarianceModel, device as D
    from geo_bsd import synthetic as S
    dev = torch.device("cuda:0")
    shape = (256, 256, 256)
    n = int(np.prod(shape))
    mask = torch.zeros(n, dtype=torch.uint8, device=dev)
    mask[:: 97] = 1
    values = torch.zeros(n, dtype=torch.float32, device=dev)
    print("[gpu] noalloc oversized request", flush=True)
    with monkeypatch.context() as m:
        m.setattr(D, "_call", _noalloc_call)
        with pytest.raises(RuntimeError, match="allocation-free scratch is too small.*bytes"):
            D.krige(0, shape, (20, 20, 4), 24, CovarianceModel(type=1, ranges=(20, 20, 4), sill=1.0, nugget=0.05),
                    mask, values=values)
    torch.cuda.synchronize()
    wl = S.ok_c2(0.3)
    ref, _ = _run(wl, 0, "mixed", False, monkeypatch)
    got, _ = _run(wl, 0, "mixed", True, monkeypatch)
    _same(got, ref)
