"""Leave-one-out ordinary kriging of scattered points, differentiable in the
variogram parameters, on the GPU.

There is no counterpart in the reference; the path reuses its covariance
model (src/geo_bsd/hpgl/cov_model.h:146-183) and its ordinary-kriging system
(src/geo_bsd/hpgl/my_kriging_weights.h:198-361).  Each point is estimated
from its k nearest other points (Euclidean, fixed for every parameter set)
for P parameter sets at once; parameters are rows (sill, nugget, range_x,
range_y, range_z).  The backward pass is a hand-written adjoint (one more
solve with the forward Cholesky factor; covariance derivatives recomputed),
wrapped in a torch.autograd.Function.

    xyz   (n, 3) float32 device tensor     values (n,) float32
    nbr = knn(xyz, k)                       (n, k) int32
    est = loo_estimates(params, xyz, values, nbr, "spherical")   (P, n) float64
    loss = loo_loss(params, xyz, values, nbr)                    (P,)
"""
import ctypes as C
import weakref

import torch

from .hpgl_wrap import _hpgl_so

MODELS = {"spherical": 0, "exponential": 1, "gaussian": 2, 0: 0, 1: 1, 2: 2}


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream(dev):
    return C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {_hpgl_so.hpglx_last_error().decode()}")


def _points(xyz, values):
    if not (xyz.is_cuda and values.is_cuda):
        raise ValueError("xyz and values must be device tensors")
    if xyz.dim() != 2 or xyz.shape[1] != 3 or values.shape != (xyz.shape[0],):
        raise ValueError("xyz must be (n, 3) and values (n,)")
    return xyz.to(torch.float32).contiguous(), values.to(torch.float32).contiguous()


# Neighbour tensors whose indices were checked, by id: a weak reference to the
# tensor and its version counter when checked.  Reading min and max waits for
# the device, so a tensor is checked once per content, not once per forward
# (the fitting loop passes the same tensor every step).
_nbr_checked = {}


def _check_nbr(nbr, n):
    if nbr.dim() != 2 or nbr.shape[0] != n:
        raise ValueError("nbr must be (n, k)")
    seen = _nbr_checked.get(id(nbr))
    if seen is not None and seen[0]() is nbr and seen[1] == nbr._version and seen[2] == n:
        return
    lo, hi = int(nbr.min()), int(nbr.max())
    if lo < 0 or hi >= n:
        raise ValueError("neighbour indices must lie in [0, n)")
    if len(_nbr_checked) > 64:
        _nbr_checked.clear()
    _nbr_checked[id(nbr)] = (weakref.ref(nbr), nbr._version, n)


# Free forward-state buffers by (device, elements).  The state of 64 parameter
# sets x 20,000 points is about 4 GB; the backward hands its buffer back here
# and the next forward on the same stream reuses it instead of allocating
# one of that size again.
_state_pool = {}


def _take_state(numel, dev):
    free = _state_pool.get((dev, numel))
    if free:
        return free.pop()
    return torch.empty(numel, dtype=torch.float64, device=dev)


def _give_state(t):
    free = _state_pool.setdefault((t.device, t.numel()), [])
    if len(free) < 2:
        free.append(t)


def knn(xyz, k):
    """Indices of the k nearest other points (Euclidean; ties by index)."""
    if not xyz.is_cuda or xyz.dim() != 2 or xyz.shape[1] != 3:
        raise ValueError("xyz must be an (n, 3) device tensor")
    xyz = xyz.to(torch.float32).contiguous()
    n = xyz.shape[0]
    nbr = torch.empty((n, int(k)), dtype=torch.int32, device=xyz.device)
    _check(_hpgl_so.hpglx_knn(_ptr(xyz), n, int(k), _ptr(nbr), _stream(xyz.device)), "knn")
    return nbr


class LooKriging(torch.autograd.Function):
    """est[p, i] = leave-one-out OK estimate of point i under params[p]."""

    @staticmethod
    def forward(ctx, params, xyz, values, nbr, model):
        xyz, values = _points(xyz, values)
        params = params.detach().to(torch.float64).contiguous()
        nbr = nbr.to(torch.int32).contiguous()
        _check_nbr(nbr, xyz.shape[0])
        n, k = nbr.shape
        p = params.shape[0]
        if params.dim() != 2 or params.shape[1] != 5:
            raise ValueError("params must be (P, 5): sill, nugget, range_x, range_y, range_z")
        if not params.is_cuda:
            raise ValueError("params must be a device tensor")
        dev = xyz.device
        est = torch.empty((p, n), dtype=torch.float64, device=dev)
        status = torch.empty((p, n), dtype=torch.int32, device=dev)
        nbytes = _hpgl_so.hpglx_loo_state_size(n, k, p)
        if nbytes == 0:
            raise ValueError("invalid sizes for leave-one-out kriging")
        state = _take_state(nbytes // 8, dev)
        m = MODELS[model]
        _check(_hpgl_so.hpglx_loo_forward(_ptr(xyz), _ptr(values), n, _ptr(nbr), k, _ptr(params), p, m, _ptr(est),
                                          _ptr(status), _ptr(state), _stream(dev)), "loo forward")
        ctx.save_for_backward(params, xyz, values, nbr, state)
        ctx.model = m
        ctx.status = status
        return est

    @staticmethod
    def backward(ctx, grad_est):
        params, xyz, values, nbr, state = ctx.saved_tensors
        n, k = nbr.shape
        p = params.shape[0]
        grad_est = grad_est.to(torch.float64).contiguous()
        grad = torch.empty((p, 5), dtype=torch.float64, device=xyz.device)
        _check(_hpgl_so.hpglx_loo_backward(_ptr(xyz), _ptr(values), n, _ptr(nbr), k, _ptr(params), p, ctx.model,
                                           _ptr(state), _ptr(grad_est), _ptr(grad), None, _stream(xyz.device)),
               "loo backward")
        _give_state(state)  # later work on this stream is ordered after the kernel that reads it
        return grad, None, None, None, None


def loo_estimates(params, xyz, values, nbr, model="spherical"):
    """(P, n) leave-one-out estimates, differentiable w.r.t. params."""
    return LooKriging.apply(params, xyz, values, nbr, model)


def loo_jacobian(params, xyz, values, nbr, model="spherical"):
    """(P, n) estimates and (P, n, 5) d est / d params (no autograd)."""
    xyz, values = _points(xyz, values)
    params = params.detach().to(torch.float64).contiguous()
    nbr = nbr.to(torch.int32).contiguous()
    _check_nbr(nbr, xyz.shape[0])
    n, k = nbr.shape
    p = params.shape[0]
    dev = xyz.device
    est = torch.empty((p, n), dtype=torch.float64, device=dev)
    state = torch.empty(_hpgl_so.hpglx_loo_state_size(n, k, p) // 8, dtype=torch.float64, device=dev)
    jac = torch.empty((p, n, 5), dtype=torch.float64, device=dev)
    m = MODELS[model]
    _check(_hpgl_so.hpglx_loo_forward(_ptr(xyz), _ptr(values), n, _ptr(nbr), k, _ptr(params), p, m, _ptr(est), None,
                                      _ptr(state), _stream(dev)), "loo forward")
    _check(_hpgl_so.hpglx_loo_backward(_ptr(xyz), _ptr(values), n, _ptr(nbr), k, _ptr(params), p, m, _ptr(state),
                                       None, None, _ptr(jac), _stream(dev)), "loo backward")
    return est, jac


def loo_loss(params, xyz, values, nbr, model="spherical"):
    """(P,) mean squared leave-one-out error per parameter set."""
    est = loo_estimates(params, xyz, values, nbr, model)
    v = values.to(torch.float64)
    return ((est - v[None, :]) ** 2).mean(dim=1)


def fit(params0, xyz, values, nbr, model="spherical", steps=200, lr=0.05, min_range=0.5):
    """Adam on log(sill), logit(nugget / sill) and log(ranges) minimising the
    leave-one-out loss of every parameter set independently.  Returns the
    fitted (P, 5) parameters and the (steps + 1, P) loss history."""
    p0 = params0.detach().to(torch.float64)
    sill0 = p0[:, 0].clamp_min(1e-12)
    frac = (p0[:, 1] / sill0).clamp(1e-6, 1 - 1e-6)
    u = torch.cat([sill0.log()[:, None], torch.log(frac / (1 - frac))[:, None],
                   (p0[:, 2:] - min_range).clamp_min(1e-6).log()], dim=1).detach().requires_grad_(True)

    def params_of(u):
        sill = u[:, 0].exp()
        nug = sill * torch.sigmoid(u[:, 1])
        ranges = u[:, 2:].exp() + min_range
        return torch.cat([sill[:, None], nug[:, None], ranges], dim=1)

    # one int32 tensor for every step, so its indices are checked once
    nbr = nbr.to(torch.int32).contiguous()
    opt = torch.optim.Adam([u], lr=lr)
    hist = []
    for _ in range(steps):
        opt.zero_grad()
        loss = loo_loss(params_of(u), xyz, values, nbr, model)
        hist.append(loss.detach())
        loss.sum().backward()
        opt.step()
    with torch.no_grad():
        final = params_of(u)
        hist.append(loo_loss(final, xyz, values, nbr, model).detach())
    return final.detach(), torch.stack(hist)
