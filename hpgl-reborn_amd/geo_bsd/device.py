"""Device-tensor entry points (PyTorch-ROCm tensors in, tensors out).

These enqueue the HIP pipeline of libhpgl.so on torch's current stream and
return without synchronising.  Tensors are flat (or any contiguous shape)
over the grid in x-fastest order -- the order of the reference's F-order
(nx, ny, nz) arrays.

Beyond the reference's host entry points they return the kriging variance
(computed but discarded by the reference, my_kriging_weights.h:175-194,
338-358), offer the mixed-precision solver (FP32 Cholesky with FP64 residual
refinement and an FP64 fallback) and multi-threshold indicator CDFs.
"""
import ctypes as C

import numpy

from .hpgl_wrap import _HPGL_IK_PARAMS, _HPGLX_GRID_REQUEST, _hpgl_so

FP64 = 0
MIXED = 1
KIND_OK, KIND_SK, KIND_MIK = 0, 1, 2
_PREC = {"fp64": FP64, "mixed": MIXED, "fp32+refine": MIXED, FP64: FP64, MIXED: MIXED}


def _torch():
    import torch
    return torch


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _check_dev(t, dtype, n, name):
    # Only plain attributes reach _require, so an error report never has to
    # format a device tensor.
    _require(t is not None, f"{name} is required")
    _require(t.device.type == "cuda", f"{name} must be a device tensor")
    _require(t.dtype == dtype, f"{name} must be {dtype}, got {t.dtype}")
    _require(t.is_contiguous(), f"{name} must be contiguous")
    _require(t.numel() == n, f"{name} has {t.numel()} elements, the grid has {n}")


def _require(ok, msg):
    if not ok:
        raise ValueError(msg)


def _call(q, stream):
    rc = _hpgl_so.hpglx_krige(C.byref(q), C.c_void_p(stream))
    if rc != 0:
        raise RuntimeError("hpglx_krige: " + _hpgl_so.hpglx_last_error().decode())


def autotune(*args, reps=5, **kw):
    """Times the launch variants of one krige() request (same arguments) and
    keeps the fastest for later calls with the same parameters, grid and
    precision.  Returns (outputs, choice) with choice = dict(search_head,
    blocks_per_cu, dynamic, ms)."""
    global _call
    chosen = (C.c_int * 3)()
    best = C.c_float(0)

    def tune(q, stream):
        rc = _hpgl_so.hpglx_autotune(C.byref(q), C.c_void_p(stream), int(reps), chosen, C.byref(best))
        if rc != 0:
            raise RuntimeError("hpglx_autotune: " + _hpgl_so.hpglx_last_error().decode())

    saved = _call
    _call = tune
    try:
        out = krige(*args, **kw)
    finally:
        _call = saved
    return out, dict(search_head=chosen[0] or 1024, blocks_per_cu=chosen[1], dynamic=bool(chosen[2]),
                     ms=float(best.value))


def _dims(grid):
    if hasattr(grid, "x"):
        return (int(grid.x), int(grid.y), int(grid.z))
    return tuple(int(v) for v in grid)


def krige(kind, grid, radiuses, max_neighbours, cov_model, mask, values=None, indicators=None, zfield=None,
          thresholds=None, marginal=None, mean=None, precision="fp64", variance=True, est64=False, status=False,
          neighbours=False, order_relations=False, pivot_tol=1e-3, refine_steps=-1, stream=None, lvm_mean=None):
    """Grid kriging on device tensors; returns a dict of result tensors.

    kind: KIND_OK, KIND_SK or KIND_MIK.  For MIK give ``indicators`` (uint8,
    nind x N, value 1 = indicator set) or ``zfield`` + ``thresholds``
    (indicator_k = z <= thresholds[k]), and ``marginal`` (nind floats).
    For SK, ``lvm_mean`` (float32, N) gives every cell its own mean.
    """
    torch = _torch()
    nx, ny, nz = _dims(grid)
    n = nx * ny * nz
    _check_dev(mask, torch.uint8, n, "mask")
    dev = mask.device
    q = _HPGLX_GRID_REQUEST()
    q.kind = int(kind)
    q.precision = _PREC[precision]
    q.dims[:] = [nx, ny, nz]
    q.radiuses[:] = [int(r) for r in radiuses]
    q.max_neighbours = int(max_neighbours)
    q.cov.covariance_type = int(cov_model.type)
    q.cov.ranges[:] = [float(r) for r in cov_model.ranges]
    q.cov.angles[:] = [float(a) for a in cov_model.angles]
    q.cov.sill = float(cov_model.sill)
    q.cov.nugget = float(cov_model.nugget)
    q.mask = _ptr(mask)
    out = {}
    keep = []
    if kind in (KIND_OK, KIND_SK):
        _check_dev(values, torch.float32, n, "values")
        q.values = _ptr(values)
        out["value"] = torch.empty(n, dtype=torch.float32, device=dev)
        q.out_val = _ptr(out["value"])
        nout = 1
    else:
        if zfield is not None:
            _check_dev(zfield, torch.float32, n, "zfield")
            thr = thresholds if torch.is_tensor(thresholds) else torch.tensor(
                numpy.asarray(thresholds, numpy.float32), device=dev)
            thr = thr.to(device=dev, dtype=torch.float32).contiguous()
            keep.append(thr)
            q.zfield = _ptr(zfield)
            q.thresholds = _ptr(thr)
            nind = thr.numel()
        else:
            nind = indicators.numel() // n
            _check_dev(indicators, torch.uint8, nind * n, "indicators")
            q.ind = _ptr(indicators)
        marg = numpy.ascontiguousarray(numpy.asarray(marginal, numpy.float64).reshape(-1))
        if marg.size != nind:
            raise ValueError("one marginal probability per indicator is required")
        keep.append(marg)
        q.marginal = marg.ctypes.data_as(C.POINTER(C.c_double))
        q.nind = nind
        q.order_relations = 1 if order_relations else 0
        out["prob"] = torch.empty((nind, n), dtype=torch.float32, device=dev)
        out["class"] = torch.empty((nind, n), dtype=torch.uint8, device=dev)
        q.out_prob = _ptr(out["prob"])
        q.out_class = _ptr(out["class"])
        nout = nind
    if kind == KIND_SK:
        q.auto_mean = 1 if mean is None else 0
        q.mean = 0.0 if mean is None else float(mean)
    if lvm_mean is not None:
        _require(kind == KIND_SK, "a locally varying mean applies to simple kriging only")
        _check_dev(lvm_mean, torch.float32, n, "lvm_mean")
        q.lvm_mean = _ptr(lvm_mean)
    out["mask"] = torch.empty(n, dtype=torch.uint8, device=dev)
    q.out_mask = _ptr(out["mask"])
    if variance:
        out["variance"] = torch.empty(n, dtype=torch.float64, device=dev)
        q.out_var = _ptr(out["variance"])
    if est64:
        out["est64"] = torch.empty((nout, n) if kind == KIND_MIK else n, dtype=torch.float64, device=dev)
        q.out_est64 = _ptr(out["est64"])
    if status:
        out["status"] = torch.empty(n, dtype=torch.uint8, device=dev)
        q.out_status = _ptr(out["status"])
    if neighbours:
        out["count"] = torch.empty(n, dtype=torch.int32, device=dev)
        out["nboff"] = torch.empty((n, int(max_neighbours)), dtype=torch.int32, device=dev)
        q.out_count = _ptr(out["count"])
        q.out_nboff = _ptr(out["nboff"])
    q.pivot_tol = float(pivot_tol)
    q.refine_steps = int(refine_steps)
    if stream is None:
        stream = torch.cuda.current_stream(dev).cuda_stream
    _call(q, stream)
    if neighbours:
        # Bit 16 of a non-negative count marks a cluster-path cell.
        raw = out["count"]
        out["cluster_path"] = (raw >= 0) & ((raw & 0x10000) != 0)
        out["count"] = torch.where(raw >= 0, raw & 0xFFFF, raw)
    return out


def ordinary_kriging(values, mask, grid, radiuses, max_neighbours, cov_model, precision="fp64", **kw):
    """OK on device tensors (work of geo.ordinary_kriging, geo.py:572-607)."""
    return krige(KIND_OK, grid, radiuses, max_neighbours, cov_model, mask, values=values, precision=precision, **kw)


def simple_kriging(values, mask, grid, radiuses, max_neighbours, cov_model, mean=None, precision="fp64", **kw):
    """SK on device tensors (work of geo.simple_kriging, geo.py:609-649)."""
    return krige(KIND_SK, grid, radiuses, max_neighbours, cov_model, mask, values=values, mean=mean,
                 precision=precision, **kw)


def lvm_kriging(values, mask, mean_data, grid, radiuses, max_neighbours, cov_model, precision="fp64", **kw):
    """SK with a locally varying mean on device tensors (work of
    geo.lvm_kriging, geo.py:651-696; lvm_kriging.cpp:16-39)."""
    return krige(KIND_SK, grid, radiuses, max_neighbours, cov_model, mask, values=values, lvm_mean=mean_data,
                 precision=precision, **kw)


def indicator_kriging(data, mask, grid, categories, marginal_probs, precision="fp64", stream=None):
    """Indicator kriging over categories on device tensors (work of
    hpgl_indicator_kriging, api.cpp:457-482; indicator_kriging.h:88-161).

    ``data``: uint8 category per cell; ``categories``: one dict per category
    with cov_model, radiuses and max_neighbours (as geo.indicator_kriging).
    Returns dict(category, mask, prob) where prob (ncat x N) holds the
    probabilities after the order-relation correction.  Unlike
    geo.indicator_kriging, two categories are kriged the same way as more
    (the reference sends them to median_ik).
    """
    torch = _torch()
    nx, ny, nz = _dims(grid)
    n = nx * ny * nz
    ncat = len(categories)
    _require(1 <= ncat <= 256, "1..256 categories are supported")
    _require(len(marginal_probs) == ncat, "one marginal probability per category is required")
    _check_dev(mask, torch.uint8, n, "mask")
    _check_dev(data, torch.uint8, n, "data")
    dev = mask.device
    params = (_HPGL_IK_PARAMS * ncat)()
    for i, c in enumerate(categories):
        cm = c["cov_model"]
        params[i] = _HPGL_IK_PARAMS(covariance_type=int(cm.type), ranges=tuple(float(r) for r in cm.ranges),
                                    angles=tuple(float(a) for a in cm.angles), sill=float(cm.sill),
                                    nugget=float(cm.nugget), radiuses=tuple(int(r) for r in c["radiuses"]),
                                    max_neighbours=int(c["max_neighbours"]), marginal_prob=float(marginal_probs[i]))
    out = {"category": torch.empty(n, dtype=torch.uint8, device=dev),
           "mask": torch.empty(n, dtype=torch.uint8, device=dev),
           "prob": torch.empty((ncat, n), dtype=torch.float32, device=dev)}
    if stream is None:
        stream = torch.cuda.current_stream(dev).cuda_stream
    rc = _hpgl_so.hpglx_indicator_kriging(params, ncat, (C.c_int * 3)(nx, ny, nz), _ptr(data), _ptr(mask),
                                          _ptr(out["category"]), _ptr(out["mask"]), _ptr(out["prob"]),
                                          _PREC[precision], C.c_void_p(stream))
    if rc != 0:
        raise RuntimeError("hpglx_indicator_kriging: " + _hpgl_so.hpglx_last_error().decode())
    return out


def _sgs_args(values, mask, grid, radiuses, max_neighbours, cov_model, seed, kriging_type, cdf, simulation_mask,
              mean_field):
    torch = _torch()
    import numpy as np
    from .hpgl_wrap import _HPGL_SGS_PARAMS, hpgl_non_parametric_cdf_t
    nx, ny, nz = _dims(grid)
    n = nx * ny * nz
    _check_dev(values, torch.float32, n, "values")
    _check_dev(mask, torch.uint8, n, "mask")
    if simulation_mask is not None:
        _check_dev(simulation_mask, torch.uint8, n, "simulation_mask")
    if mean_field is not None:
        _check_dev(mean_field, torch.float32, n, "mean_field")
    _require(kriging_type in ("sk", "ok"), 'kriging_type must be "sk" or "ok"')
    p = _HPGL_SGS_PARAMS(covariance_type=int(cov_model.type), ranges=tuple(float(r) for r in cov_model.ranges),
                         angles=tuple(float(a) for a in cov_model.angles), sill=float(cov_model.sill),
                         nugget=float(cov_model.nugget), radiuses=tuple(int(r) for r in radiuses),
                         max_neighbours=int(max_neighbours), kriging_kind=1 if kriging_type == "sk" else 0,
                         seed=int(seed), min_neighbours=0)
    cdf_arg, keep = None, ()
    if cdf is not None:
        cv = np.ascontiguousarray(cdf[0], np.float32)
        cp = np.ascontiguousarray(cdf[1], np.float32)
        keep = (cv, cp)
        cdf_arg = C.byref(hpgl_non_parametric_cdf_t(values=cv.ctypes.data_as(C.POINTER(C.c_float)),
                                                    probs=cp.ctypes.data_as(C.POINTER(C.c_float)), size=cv.size))
    return (nx, ny, nz), p, cdf_arg, keep


def sgs_simulation(values, mask, grid, radiuses, max_neighbours, cov_model, seed, kriging_type="sk", cdf=None,
                   simulation_mask=None, mean_field=None, stream=None):
    """Sequential Gaussian simulation in place on device tensors (work of
    hpgl_sgs_simulation, api.cpp:485-516): float32 values and uint8 mask of
    N cells are overwritten with the realization; the model is used as
    given (geo_bsd.sgs_simulation scales it to unit sill first).  cdf =
    (values, probs) host arrays of the normal-score transform, or None;
    simulation_mask: uint8 tensor (1 = simulate) or None; mean_field:
    float32 tensor of one mean per cell (simple kriging with a locally
    varying mean) or None.  Returns the number of cells informed
    afterwards."""
    torch = _torch()
    dims, p, cdf_arg, keep = _sgs_args(values, mask, grid, radiuses, max_neighbours, cov_model, seed, kriging_type,
                                       cdf, simulation_mask, mean_field)
    if stream is None:
        stream = torch.cuda.current_stream(values.device).cuda_stream
    rc = _hpgl_so.hpglx_sgs_simulation(_ptr(values), _ptr(mask), (C.c_int * 3)(*dims), C.byref(p), cdf_arg,
                                       _ptr(simulation_mask) if simulation_mask is not None else None,
                                       _ptr(mean_field) if mean_field is not None else None, stream)
    del keep
    if rc != 0:
        raise RuntimeError("sgs_simulation: " + _hpgl_so.hpglx_last_error().decode())
    return int(mask.sum().item())


def sgs_realizations(values, mask, grid, radiuses, max_neighbours, cov_model, seeds, kriging_type="sk", cdf=None,
                     simulation_mask=None, mean_field=None, stream=None):
    """Several realizations from one conditioning data set, one per seed:
    returns (values, mask) tensors of shape (len(seeds), N), row r equal to
    what sgs_simulation gives with seed seeds[r].  The inputs are left as
    they are.  The realizations advance together on the GPU, one launch per
    level index, so a few cells per level from each of them fill the chip."""
    torch = _torch()
    seeds = [int(s) for s in seeds]
    _require(1 <= len(seeds) <= 4096, "seeds must hold 1..4096 values")
    dims, p, cdf_arg, keep = _sgs_args(values, mask, grid, radiuses, max_neighbours, cov_model, seeds[0],
                                       kriging_type, cdf, simulation_mask, mean_field)
    n = dims[0] * dims[1] * dims[2]
    out_v = torch.empty((len(seeds), n), dtype=torch.float32, device=values.device)
    out_m = torch.empty((len(seeds), n), dtype=torch.uint8, device=values.device)
    sarr = (C.c_long * len(seeds))(*seeds)
    if stream is None:
        stream = torch.cuda.current_stream(values.device).cuda_stream
    rc = _hpgl_so.hpglx_sgs_realizations(_ptr(values), _ptr(mask), (C.c_int * 3)(*dims), C.byref(p), sarr,
                                         len(seeds), cdf_arg,
                                         _ptr(simulation_mask) if simulation_mask is not None else None,
                                         _ptr(mean_field) if mean_field is not None else None, _ptr(out_v),
                                         _ptr(out_m), stream)
    del keep
    if rc != 0:
        raise RuntimeError("sgs_realizations: " + _hpgl_so.hpglx_last_error().decode())
    return out_v, out_m


def indicator_cdf(zfield, mask, grid, thresholds, marginal, radiuses, max_neighbours, cov_model, precision="fp64",
                  order_relations=True, **kw):
    """Median IK with several thresholds sharing one model: one solve per cell,
    one CDF value per threshold, order relations corrected."""
    return krige(KIND_MIK, grid, radiuses, max_neighbours, cov_model, mask, zfield=zfield, thresholds=thresholds,
                 marginal=marginal, precision=precision, order_relations=order_relations, **kw)


def unpack_offsets(nboff):
    """Decodes packed neighbour offsets into (dx, dy, dz) integer arrays."""
    p = numpy.asarray(nboff).astype(numpy.int64)
    dx = (p & 2047) - 1024
    dy = ((p >> 11) & 2047) - 1024
    dz = ((p >> 22) & 1023) - 512
    return dx, dy, dz


def set_profiling(on):
    _hpgl_so.hpglx_set_profiling(1 if on else 0)


def last_timing():
    """Milliseconds (search kernel, main solve kernel) of the last profiled call."""
    ms = (C.c_float * 2)()
    if _hpgl_so.hpglx_last_timing(ms) != 0:
        raise RuntimeError(_hpgl_so.hpglx_last_error().decode())
    return tuple(float(v) for v in ms)


def profile_summary():
    """Mean milliseconds of the search kernel and of the main solve kernel
    over the calls profiled since set_profiling(True) (start/stop events
    recorded by the kernel dispatches), and the number of calls."""
    ms = (C.c_float * 2)()
    nc = C.c_int(0)
    if _hpgl_so.hpglx_profile_summary(ms, C.byref(nc)) != 0:
        raise RuntimeError(_hpgl_so.hpglx_last_error().decode())
    return {"search": float(ms[0]), "solve_main": float(ms[1]), "calls": nc.value}


def last_call_stats():
    fb = C.c_uint(0)
    cl = C.c_int(0)
    if _hpgl_so.hpglx_last_call_stats(C.byref(fb), C.byref(cl)) != 0:
        raise RuntimeError(_hpgl_so.hpglx_last_error().decode())
    return {"fp64_fallbacks": fb.value, "cluster_path_cells": cl.value}


def set_graph_mode(on):
    """Capture repeated identical grid calls into a hipGraph and replay them."""
    _hpgl_so.hpglx_set_graph_mode(1 if on else 0)


def set_default_precision(precision):
    """Precision of the host-array entry points (ordinary_kriging etc.)."""
    _hpgl_so.hpglx_set_default_precision(_PREC[precision])


def search_template(cov_model, radiuses):
    """The ordered search template as an (n, 3) int array."""
    from .hpgl_wrap import _HPGL_COV_PARAMS
    cp = _HPGL_COV_PARAMS(covariance_type=int(cov_model.type), ranges=tuple(float(r) for r in cov_model.ranges),
                          angles=tuple(float(a) for a in cov_model.angles), sill=float(cov_model.sill),
                          nugget=float(cov_model.nugget))
    r = (C.c_int * 3)(*[int(v) for v in radiuses])
    n = _hpgl_so.hpglx_template(C.byref(cp), r, None, 0)
    if n < 0:
        raise RuntimeError(_hpgl_so.hpglx_last_error().decode())
    out = numpy.zeros((n, 3), numpy.int32)
    _hpgl_so.hpglx_template(C.byref(cp), r, out.ctypes.data_as(C.c_void_p), n)
    return out


def host_introsort_perm(keys):
    """Host run of the device introsort (slot -> input position)."""
    keys = numpy.ascontiguousarray(keys, numpy.float64)
    perm = numpy.zeros(keys.size, numpy.int32)
    _hpgl_so.hpglx_host_introsort_perm(keys.ctypes.data_as(C.c_void_p), keys.size, perm.ctypes.data_as(C.c_void_p))
    return perm


def host_introsort_first(keys, kneed):
    """Host run of the device introsort when only the first kneed slots are
    wanted (slot -> input position; slots from kneed on are unspecified)."""
    keys = numpy.ascontiguousarray(keys, numpy.float64)
    perm = numpy.zeros(keys.size, numpy.int32)
    _hpgl_so.hpglx_host_introsort_first(keys.ctypes.data_as(C.c_void_p), keys.size, int(kneed),
                                        perm.ctypes.data_as(C.c_void_p))
    return perm


# ------------------------------------------------------------ property files
def read_inc_float(filename, undefined_value, size, device="cuda:0"):
    """INC file -> (data float32, mask uint8) device tensors; the text is
    parsed on the GPU (work of hpgl_read_inc_file_float, api.cpp:53-75)."""
    torch = _torch()
    data = torch.empty(int(size), dtype=torch.float32, device=device)
    mask = torch.empty(int(size), dtype=torch.uint8, device=device)
    stream = torch.cuda.current_stream(data.device).cuda_stream
    rc = _hpgl_so.hpglx_read_inc_float(str(filename).encode(), float(undefined_value), int(size), _ptr(data),
                                       _ptr(mask), C.c_void_p(stream))
    if rc != 0:
        raise RuntimeError("hpglx_read_inc_float: " + _hpgl_so.hpglx_last_error().decode())
    return data, mask


def read_inc_byte(filename, undefined_value, size, indicator_values=(), device="cuda:0"):
    """INC file of integers -> (data uint8, mask uint8) device tensors, values
    remapped to their index in indicator_values (api.cpp:77-116)."""
    torch = _torch()
    data = torch.empty(int(size), dtype=torch.uint8, device=device)
    mask = torch.empty(int(size), dtype=torch.uint8, device=device)
    vals = numpy.ascontiguousarray(numpy.asarray(indicator_values, numpy.uint8).reshape(-1))
    stream = torch.cuda.current_stream(data.device).cuda_stream
    rc = _hpgl_so.hpglx_read_inc_byte(str(filename).encode(), int(undefined_value), int(size), _ptr(data), _ptr(mask),
                                      vals.ctypes.data_as(C.c_void_p) if vals.size else None, int(vals.size),
                                      C.c_void_p(stream))
    if rc != 0:
        raise RuntimeError("hpglx_read_inc_byte: " + _hpgl_so.hpglx_last_error().decode())
    return data, mask


def write_inc(filename, data, mask, undefined_value, name, gslib=False, indicator_values=None):
    """Device property -> INC (or GSLIB) file, printed on the GPU
    (property_writer.cpp).  float32 data are written with "%E", uint8 data as
    integers through indicator_values."""
    torch = _torch()
    n = data.numel()
    _check_dev(mask, torch.uint8, n, "mask")
    stream = torch.cuda.current_stream(data.device).cuda_stream
    if data.dtype == torch.float32:
        _require(data.is_contiguous(), "data must be contiguous")
        rc = _hpgl_so.hpglx_write_inc_float(str(filename).encode(), _ptr(data), _ptr(mask), int(n),
                                            float(undefined_value), name.encode(), 1 if gslib else 0,
                                            C.c_void_p(stream))
        fn = "hpglx_write_inc_float"
    else:
        _check_dev(data, torch.uint8, n, "data")
        vals = numpy.ascontiguousarray(numpy.asarray(indicator_values if indicator_values is not None else [],
                                                     numpy.uint8).reshape(-1))
        rc = _hpgl_so.hpglx_write_inc_byte(str(filename).encode(), _ptr(data), _ptr(mask), int(n),
                                           int(undefined_value), name.encode(),
                                           vals.ctypes.data_as(C.c_void_p) if vals.size else None, int(vals.size),
                                           1 if gslib else 0, C.c_void_p(stream))
        fn = "hpglx_write_inc_byte"
    if rc != 0:
        raise RuntimeError(fn + ": " + _hpgl_so.hpglx_last_error().decode())
