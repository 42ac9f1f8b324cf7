"""The launch autotuner (hpglx_autotune): every candidate only changes how
work is dealt to the GPU, so results after tuning equal the untuned results
bit for bit, and the choice is one of the candidates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,scale,precision", [("ok", 0.4, "mixed"), ("wells", 0.25, "fp64"), ("sk", 1.0, "mixed")])
def test_tuned_calls_equal_untuned(name, scale, precision):
    import torch
    from geo_bsd import CovarianceModel, device as D
    from geo_bsd import synthetic as S
    wl = {"ok": S.ok_c2, "wells": S.ok_wells_c4, "sk": S.sk_c1}[name](scale)
    dev = torch.device("cuda:0")
    kind = 1 if wl.kind == "sk" else 0
    args = (kind, wl.shape, wl.radiuses, wl.k, CovarianceModel(**wl.cov), torch.from_numpy(wl.mask).to(dev))
    kw = dict(values=torch.from_numpy(wl.values).to(dev), mean=wl.get("mean"), precision=precision, est64=True,
              status=True, neighbours=True)
    print(f"[gpu] tuning {wl.name} {wl.shape} {precision}", flush=True)
    before = {k: v.cpu().numpy() for k, v in D.krige(*args, **kw).items()}
    tuned_out, choice = D.autotune(*args, reps=2, **kw)
    print(f"[gpu] choice {choice}", flush=True)
    assert choice["search_head"] in (256, 512, 1024) and choice["blocks_per_cu"] in (0, 1) and choice["ms"] > 0
    after = {k: v.cpu().numpy() for k, v in D.krige(*args, **kw).items()}
    tuned = {k: v.cpu().numpy() for k, v in tuned_out.items()}
    for res in (after, tuned):
        for key in before:
            a, b = before[key], res[key]
            if key == "nboff":  # slots past a cell's count are never written
                valid = np.arange(wl.k)[None, :] < np.maximum(before["count"], 0)[:, None]
                a, b = np.where(valid, a, 0), np.where(valid, b, 0)
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), key
