"""Graph replay of repeated grid calls (the library captures a call equal to
the previous one into a hipGraph and replays it): every replay equals the
eager call bit for bit, with the program's own eager work and
synchronisations between replays; a replay reads the buffers' current
contents; a call with another geometry in between (which rewrites the shared
padded mask) is not answered by a stale graph."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def graphs_on():
    from geo_bsd import device as D
    D.set_graph_mode(True)
    yield
    D.set_graph_mode(False)


def _request(wl, precision, kind=0):
    """The request geo_bsd.device builds for a call, with its output tensors
    (kept alive by the caller)."""
    import torch
    from geo_bsd import CovarianceModel, device as D
    dev = torch.device("cuda:0")
    held = {}

    def grab(q, stream):
        held["q"] = q

    inputs = dict(mask=torch.from_numpy(wl.mask).to(dev))
    if kind == 2:
        # the request points at these: the test keeps them alive
        inputs["zfield"] = torch.from_numpy(wl.values).to(dev)
        inputs["thresholds"] = torch.tensor(np.asarray(wl.thresholds, np.float32), device=dev)
        inputs["marginal"] = np.ascontiguousarray(wl.levels, np.float64)
    else:
        inputs["values"] = torch.from_numpy(wl.values).to(dev)
    orig = D._call
    D._call = grab
    try:
        if kind == 2:
            out = D.krige(2, wl.shape, wl.radiuses, wl.k, CovarianceModel(**wl.cov), inputs["mask"],
                          zfield=inputs["zfield"], thresholds=inputs["thresholds"], marginal=inputs["marginal"],
                          precision=precision, est64=True, status=True, order_relations=True)
        else:
            out = D.krige(kind, wl.shape, wl.radiuses, wl.k, CovarianceModel(**wl.cov), inputs["mask"],
                          values=inputs["values"], precision=precision, est64=True, status=True)
    finally:
        D._call = orig
    return held["q"], out, inputs


def _call(q):
    import torch
    from geo_bsd import device as D
    rc = D._hpgl_so.hpglx_krige(C.byref(q), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, D._hpgl_so.hpglx_last_error().decode()


def _host(out):
    import torch
    torch.cuda.synchronize()
    return {k: v.cpu().numpy().copy() for k, v in out.items()}


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


@pytest.mark.parametrize("name,scale,kind,precision", [("ok", 0.4, 0, "mixed"), ("ok", 0.4, 0, "fp64"),
                                                       ("wells", 0.25, 0, "mixed"), ("mik", 0.2, 2, "fp64")])
def test_replays_equal_eager(name, scale, kind, precision):
    import torch
    from geo_bsd import synthetic as S
    wl = {"ok": S.ok_c2, "wells": S.ok_wells_c4, "mik": S.mik_c3}[name](scale)
    print(f"[gpu] graph replays {wl.name} {wl.shape} {precision}", flush=True)
    q, out, inputs = _request(wl, precision, kind)
    _call(q)  # eager
    first = _host(out)
    scratch = torch.randn(1 << 20, device="cuda:0")
    for i in range(5):  # the second call is captured, the rest replay
        for v in out.values():
            v.zero_()
        _call(q)
        scratch.mul_(1.0001).add_(float(i))  # eager work between replays
        _same(first, _host(out))
        assert torch.isfinite(scratch).all()


def test_replay_reads_current_inputs_and_survives_other_geometry():
    import torch
    from geo_bsd import synthetic as S
    wl = S.ok_c2(0.4)
    q, out, inputs = _request(wl, "mixed")
    _call(q)
    _call(q)  # captured
    base = _host(out)
    # new contents in the same buffers: the replay must see them
    inputs["values"].mul_(2.0)
    _call(q)
    doubled = _host(out)
    q2, out2, inputs2 = _request(wl, "mixed")  # other output buffers: an eager call
    inputs2["values"].mul_(2.0)
    _call(q2)
    _same(doubled, _host(out2))
    assert not np.array_equal(doubled["value"], base["value"])
    # a call with another geometry rewrites the shared padded mask; the
    # first call's graph must not be replayed over it
    inputs["values"].div_(2.0)
    other = S.ok_c2(0.3)
    qo, outo, _ = _request(other, "mixed")
    _call(qo)
    _call(q)
    _same(base, _host(out))
    _call(q)
    _same(base, _host(out))
