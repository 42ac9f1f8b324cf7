"""ctypes front-end to the CPU checkers (TEST INFRASTRUCTURE ONLY).

* ``libhpgl_oracle.so`` -- the clean-room FP64 restatement in hpgl_oracle.cpp.
* ``_ref/libhpgl_ref.so`` -- the reference's own kriging core compiled from its
  sources by ``make ref`` (present only where the reference tree was available
  at build time; tests skip what needs it otherwise).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
import this module.  The product path never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libhpgl_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhpgl_ref.so")

ST_KRIGED, ST_NO_NEIGHBOURS, ST_SINGULAR, ST_INFORMED = 0, 1, 2, 3
OK, SK, MIK = 0, 1, 2


class CovT(C.Structure):
    _fields_ = [("type", C.c_int), ("ranges", C.c_double * 3), ("angles", C.c_double * 3),
                ("sill", C.c_double), ("nugget", C.c_double)]


def make_cov(type_, ranges, angles=(0.0, 0.0, 0.0), sill=1.0, nugget=0.0):
    return CovT(int(type_), (C.c_double * 3)(*ranges), (C.c_double * 3)(*angles), float(sill), float(nugget))


def build(ref=False):
    """Compile the oracle (and the reference checker when ``ref``)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_orc = None
_ref = None


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _orc = C.CDLL(ORACLE_SO)
        _orc.orc_template.restype = C.c_int
        _orc.orc_cov.restype = C.c_double
        _orc.orc_cov.argtypes = [C.POINTER(CovT), C.c_double, C.c_double, C.c_double]
        _orc.orc_krige.restype = C.c_int
        _orc.orc_sk_weights.restype = C.c_int
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
        _ref.ref_quiet()
    return _ref


# ------------------------------------------------------------------ oracle
def template(cov, radiuses):
    r = (C.c_int * 3)(*radiuses)
    n = lib().orc_template(C.byref(cov), r, None, None, 0)
    if n < 0:
        raise ValueError("invalid covariance parameters")
    out = np.zeros((n, 3), np.int32)
    keys = np.zeros(n, np.float64)
    lib().orc_template(C.byref(cov), r, _p(out, C.c_int), _p(keys, C.c_double), n)
    return out, keys


def cov_value(cov, dx, dy, dz):
    return lib().orc_cov(C.byref(cov), dx, dy, dz)


def krige(kind, cov, radiuses, max_nb, shape, mask, values=None, indicators=None, means=None,
          auto_mean=False, want_neighbours=True, lvm_mean=None):
    """Run the oracle over a whole grid.  Arrays are flat, x fastest.
    ``lvm_mean`` (SK only): one mean per cell, as lvm_kriging."""
    nx, ny, nz = shape
    n = nx * ny * nz
    mask = np.ascontiguousarray(mask, np.uint8).reshape(-1)
    assert mask.size == n
    if values is None:
        values = np.zeros(n, np.float32)
    values = np.ascontiguousarray(values, np.float32).reshape(-1)
    nind = 0
    ind = None
    if kind == MIK:
        ind = np.ascontiguousarray(indicators, np.uint8).reshape(-1, n)
        nind = ind.shape[0]
    m = np.zeros(max(nind, 1), np.float64)
    if means is not None:
        m[:len(np.atleast_1d(means))] = np.atleast_1d(means)
    mean_used = C.c_double(0)
    lvm = None if lvm_mean is None else np.ascontiguousarray(lvm_mean, np.float32).reshape(-1)
    assert lvm is None or (kind == SK and lvm.size == n)
    out = dict(
        value=np.zeros(n, np.float32), mask=np.zeros(n, np.uint8),
        var=np.zeros(n, np.float64), count=np.zeros(n, np.int32),
        path=np.zeros(n, np.uint8), status=np.zeros(n, np.int32),
        est64=np.zeros(max(nind, 1) * n, np.float64))
    if kind == MIK:
        out["prob"] = np.zeros(nind * n, np.float64)
    if want_neighbours:
        out["nb"] = np.zeros(n * max_nb, np.int32)
    rc = lib().orc_krige(
        C.c_int(kind), C.byref(cov), (C.c_int * 3)(*radiuses), C.c_int(max_nb), C.c_int(nx), C.c_int(ny),
        C.c_int(nz), _p(values, C.c_float), _p(ind, C.c_ubyte), C.c_int(nind), _p(mask, C.c_ubyte),
        _p(m, C.c_double), C.c_int(1 if auto_mean else 0), C.byref(mean_used), _p(out["value"], C.c_float),
        _p(out["mask"], C.c_ubyte), _p(out.get("prob"), C.c_double), _p(out["est64"], C.c_double),
        _p(out["var"], C.c_double), _p(out["count"], C.c_int), _p(out.get("nb"), C.c_int),
        _p(out["path"], C.c_ubyte), _p(out["status"], C.c_int), _p(lvm, C.c_float))
    if rc != 0:
        raise ValueError("oracle rejected the parameters")
    if want_neighbours:
        out["nb"] = out["nb"].reshape(n, max_nb)
    if kind == MIK:
        out["prob"] = out["prob"].reshape(nind, n)
        out["est64"] = out["est64"].reshape(nind, n)
    out["mean"] = mean_used.value
    return out


def indicator_kriging(shape, data, mask, cats):
    """Indicator kriging over categories (indicator_kriging.h:88-161).
    ``cats``: one dict per category with keys cov (CovT), radiuses, max_nb,
    marginal.  Returns dict(cat, mask, raw, prob, status); raw/prob/status are
    (ncat, N)."""
    n = int(np.prod(shape))
    k = len(cats)
    d = np.ascontiguousarray(data, np.uint8).reshape(-1)
    m = np.ascontiguousarray(mask, np.uint8).reshape(-1)
    assert d.size == n and m.size == n
    covs = (CovT * k)(*[c["cov"] for c in cats])
    rad = np.ascontiguousarray([r for c in cats for r in c["radiuses"]], np.int32)
    kmax = np.ascontiguousarray([c["max_nb"] for c in cats], np.int32)
    marg = np.ascontiguousarray([c["marginal"] for c in cats], np.float64)
    out = dict(cat=np.zeros(n, np.uint8), mask=np.zeros(n, np.uint8), raw=np.zeros(k * n, np.float32),
               prob=np.zeros(k * n, np.float32), status=np.zeros(k * n, np.int32))
    rc = lib().orc_indicator_kriging(
        _p(d, C.c_ubyte), _p(m, C.c_ubyte), *[C.c_int(s) for s in shape], C.c_int(k), covs, _p(rad, C.c_int),
        _p(kmax, C.c_int), _p(marg, C.c_double), _p(out["cat"], C.c_ubyte), _p(out["mask"], C.c_ubyte),
        _p(out["raw"], C.c_float), _p(out["prob"], C.c_float), _p(out["status"], C.c_int))
    if rc != 0:
        raise ValueError("oracle rejected the parameters")
    for key in ("raw", "prob", "status"):
        out[key] = out[key].reshape(k, n)
    return out


def order_relations(probs):
    """indicator_kriging.h:34-59 on an (cells, K) float32 array, in place copy."""
    p = np.ascontiguousarray(probs, np.float32).copy()
    cells, k = p.shape
    lib().orc_order_relations(_p(p, C.c_float), C.c_int(k), C.c_longlong(cells))
    return p


def sk_weights(cov, center, xs, ys, zs):
    n = len(xs)
    w = np.zeros(n, np.float64)
    var = C.c_double(0)
    rc = lib().orc_sk_weights(C.byref(cov), (C.c_float * 3)(*center), _p(np.ascontiguousarray(xs, np.float32), C.c_float),
                              _p(np.ascontiguousarray(ys, np.float32), C.c_float), _p(np.ascontiguousarray(zs, np.float32), C.c_float),
                              C.c_int(n), _p(w, C.c_double), C.byref(var))
    return rc, w, var.value


def std_sort_perm(keys):
    keys = np.ascontiguousarray(keys, np.float64)
    perm = np.zeros(keys.size, np.int32)
    lib().orc_std_sort_perm(_p(keys, C.c_double), C.c_int(keys.size), _p(perm, C.c_int))
    return perm


# --------------------------------------------------------------- reference
def _cov_args(cov):
    return (C.c_int(cov.type), (C.c_double * 3)(*cov.ranges), (C.c_double * 3)(*cov.angles),
            C.c_double(cov.sill), C.c_double(cov.nugget))


def ref_ordinary_kriging(cov, radiuses, max_nb, shape, values, mask, threads=1):
    r = ref_lib()
    r.ref_set_threads(C.c_int(threads))
    v = np.ascontiguousarray(values, np.float32).reshape(-1).copy()
    m = np.ascontiguousarray(mask, np.uint8).reshape(-1).copy()
    ov, om = v.copy(), m.copy()
    rc = r.ref_ordinary_kriging(_p(v, C.c_float), _p(m, C.c_ubyte), *[C.c_int(s) for s in shape], *_cov_args(cov),
                                (C.c_int * 3)(*radiuses), C.c_int(max_nb), _p(ov, C.c_float), _p(om, C.c_ubyte))
    if rc != 0:
        raise RuntimeError("reference raised")
    return ov, om


def ref_simple_kriging(cov, radiuses, max_nb, shape, values, mask, mean=None, threads=1):
    r = ref_lib()
    r.ref_set_threads(C.c_int(threads))
    v = np.ascontiguousarray(values, np.float32).reshape(-1).copy()
    m = np.ascontiguousarray(mask, np.uint8).reshape(-1).copy()
    ov, om = v.copy(), m.copy()
    rc = r.ref_simple_kriging(_p(v, C.c_float), _p(m, C.c_ubyte), *[C.c_int(s) for s in shape], *_cov_args(cov),
                              (C.c_int * 3)(*radiuses), C.c_int(max_nb), C.c_int(1 if mean is None else 0),
                              C.c_double(0.0 if mean is None else mean), _p(ov, C.c_float), _p(om, C.c_ubyte))
    if rc != 0:
        raise RuntimeError("reference raised")
    return ov, om


def ref_median_ik(cov, radiuses, max_nb, shape, data, mask, marginal_probs, threads=1):
    r = ref_lib()
    r.ref_set_threads(C.c_int(threads))
    d = np.ascontiguousarray(data, np.uint8).reshape(-1).copy()
    m = np.ascontiguousarray(mask, np.uint8).reshape(-1).copy()
    od, om = d.copy(), m.copy()
    rc = r.ref_median_ik(_p(d, C.c_ubyte), _p(m, C.c_ubyte), *[C.c_int(s) for s in shape], *_cov_args(cov),
                         (C.c_int * 3)(*radiuses), C.c_int(max_nb), C.c_double(marginal_probs[0]),
                         C.c_double(marginal_probs[1]), _p(od, C.c_ubyte), _p(om, C.c_ubyte))
    if rc != 0:
        raise RuntimeError("reference raised")
    return od, om


def ref_lvm_kriging(cov, radiuses, max_nb, shape, values, mask, mean_data, threads=1):
    r = ref_lib()
    r.ref_set_threads(C.c_int(threads))
    v = np.ascontiguousarray(values, np.float32).reshape(-1).copy()
    m = np.ascontiguousarray(mask, np.uint8).reshape(-1).copy()
    md = np.ascontiguousarray(mean_data, np.float32).reshape(-1).copy()
    ov, om = v.copy(), m.copy()
    rc = r.ref_lvm_kriging(_p(v, C.c_float), _p(m, C.c_ubyte), _p(md, C.c_float), *[C.c_int(s) for s in shape],
                           *_cov_args(cov), (C.c_int * 3)(*radiuses), C.c_int(max_nb), _p(ov, C.c_float),
                           _p(om, C.c_ubyte))
    if rc != 0:
        raise RuntimeError("reference raised")
    return ov, om


def ref_indicator_kriging(shape, data, mask, cats, threads=1):
    """The reference's hpgl::indicator_kriging; ``cats`` as for indicator_kriging()."""
    r = ref_lib()
    r.ref_set_threads(C.c_int(threads))
    k = len(cats)
    d = np.ascontiguousarray(data, np.uint8).reshape(-1).copy()
    m = np.ascontiguousarray(mask, np.uint8).reshape(-1).copy()
    od, om = d.copy(), m.copy()
    types = np.ascontiguousarray([c["cov"].type for c in cats], np.int32)
    ranges = np.ascontiguousarray([x for c in cats for x in c["cov"].ranges], np.float64)
    angles = np.ascontiguousarray([x for c in cats for x in c["cov"].angles], np.float64)
    sills = np.ascontiguousarray([c["cov"].sill for c in cats], np.float64)
    nuggets = np.ascontiguousarray([c["cov"].nugget for c in cats], np.float64)
    rad = np.ascontiguousarray([x for c in cats for x in c["radiuses"]], np.int32)
    kmax = np.ascontiguousarray([c["max_nb"] for c in cats], np.int32)
    marg = np.ascontiguousarray([c["marginal"] for c in cats], np.float64)
    rc = r.ref_indicator_kriging(_p(d, C.c_ubyte), _p(m, C.c_ubyte), *[C.c_int(s) for s in shape], C.c_int(k),
                                 _p(types, C.c_int), _p(ranges, C.c_double), _p(angles, C.c_double),
                                 _p(sills, C.c_double), _p(nuggets, C.c_double), _p(rad, C.c_int),
                                 _p(kmax, C.c_int), _p(marg, C.c_double), _p(od, C.c_ubyte), _p(om, C.c_ubyte))
    if rc != 0:
        raise RuntimeError("reference raised")
    return od, om


def ref_neighbours(kind, cov, radiuses, max_nb, shape, mask, threads=8):
    r = ref_lib()
    r.ref_set_threads(C.c_int(threads))
    n = int(np.prod(shape))
    m = np.ascontiguousarray(mask, np.uint8).reshape(-1).copy()
    counts = np.zeros(n, np.int32)
    idx = np.zeros(n * max_nb, np.int32)
    rc = r.ref_neighbours(C.c_int(kind), _p(m, C.c_ubyte), *[C.c_int(s) for s in shape], *_cov_args(cov),
                          (C.c_int * 3)(*radiuses), C.c_int(max_nb), _p(counts, C.c_int), _p(idx, C.c_int))
    if rc != 0:
        raise RuntimeError("reference raised")
    return counts, idx.reshape(n, max_nb)


def ref_template(cov, radiuses):
    r = ref_lib()
    cap = 1
    for x in radiuses:
        cap *= 2 * x + 1
    out = np.zeros((cap, 3), np.int32)
    n = r.ref_template(*_cov_args(cov), (C.c_int * 3)(*radiuses), _p(out, C.c_int), C.c_int(cap))
    if n < 0:
        raise RuntimeError("reference raised")
    return out[:n]


def ref_sk_weights(cov, center, xs, ys, zs):
    r = ref_lib()
    n = len(xs)
    w = np.zeros(n, np.float64)
    var = C.c_double(0)
    got = r.ref_sk_weights(*_cov_args(cov), (C.c_float * 3)(*center), _p(np.ascontiguousarray(xs, np.float32), C.c_float),
                           _p(np.ascontiguousarray(ys, np.float32), C.c_float), _p(np.ascontiguousarray(zs, np.float32), C.c_float),
                           C.c_int(n), _p(w, C.c_double), C.byref(var))
    return got, w, var.value


# ---------------------------------------------------------- property files
def read_inc(text, size, as_int=False, undefined=-99):
    """The INC reader (read_inc_file.cpp) over file bytes ``text``.
    Returns (data, mask); raises ValueError with the reference's message."""
    buf = C.create_string_buffer(bytes(text), len(text) + 1)
    data_f = np.zeros(max(size, 1), np.float32)
    data_b = np.zeros(max(size, 1), np.uint8)
    mask = np.zeros(max(size, 1), np.uint8)
    err = C.create_string_buffer(512)
    rc = lib().orc_read_inc(buf, C.c_longlong(len(text)), C.c_int(1 if as_int else 0), C.c_float(undefined),
                            C.c_int(int(undefined)), C.c_int(size), _p(data_f, C.c_float), _p(data_b, C.c_ubyte),
                            _p(mask, C.c_ubyte), err, C.c_int(512))
    if rc != 0:
        raise ValueError(err.value.decode())
    return (data_b if as_int else data_f)[:size], mask[:size]


def remap_bytes(data, mask, values):
    d = np.ascontiguousarray(data, np.uint8).copy()
    v = np.ascontiguousarray(values, np.uint8)
    lib().orc_remap_bytes(_p(d, C.c_ubyte), _p(np.ascontiguousarray(mask, np.uint8), C.c_ubyte), C.c_int(d.size),
                          _p(v, C.c_ubyte), C.c_int(v.size))
    return d


def format_values(mask, floats=None, bytes_=None, undefined=-99.0, remap=None):
    """The value lines a property writer prints (property_writer.cpp:18-85)."""
    m = np.ascontiguousarray(mask, np.uint8)
    f = None if floats is None else np.ascontiguousarray(floats, np.float32)
    b = None if bytes_ is None else np.ascontiguousarray(bytes_, np.uint8)
    r = np.zeros(1, np.uint8) if remap is None else np.ascontiguousarray(remap, np.uint8)
    args = (_p(f, C.c_float), _p(b, C.c_ubyte), _p(m, C.c_ubyte), C.c_int(m.size), C.c_float(float(undefined)),
            C.c_ubyte(int(undefined) & 0xFF), _p(r, C.c_ubyte), C.c_int(0 if remap is None else r.size))
    lib().orc_format_values.restype = C.c_longlong
    n = lib().orc_format_values(*args, None, C.c_longlong(0))
    out = C.create_string_buffer(int(n) + 1)
    lib().orc_format_values(*args, out, C.c_longlong(n))
    return out.raw[:n]


def ref_read_inc(path, size, as_int=False, undefined=-99):
    """The reference's read_inc_file_float / _byte on a file."""
    r = ref_lib()
    data = np.zeros(max(size, 1), np.uint8 if as_int else np.float32)
    mask = np.zeros(max(size, 1), np.uint8)
    if as_int:
        rc = r.ref_read_inc_byte(str(path).encode(), C.c_int(int(undefined)), C.c_int(size), _p(data, C.c_ubyte),
                                 _p(mask, C.c_ubyte))
    else:
        rc = r.ref_read_inc_float(str(path).encode(), C.c_float(undefined), C.c_int(size), _p(data, C.c_float),
                                  _p(mask, C.c_ubyte))
    if rc != 0:
        raise ValueError("reference raised")
    return data[:size], mask[:size]


def ref_write(path, name, mask, floats=None, bytes_=None, undefined=-99.0, remap=None, indicator_count=None,
              gslib=False):
    """The reference's property_writer_t on a float or byte property."""
    r = ref_lib()
    m = np.ascontiguousarray(mask, np.uint8).copy()
    if floats is not None:
        f = np.ascontiguousarray(floats, np.float32).copy()
        rc = r.ref_write_float(str(path).encode(), name.encode(), _p(f, C.c_float), _p(m, C.c_ubyte), C.c_int(m.size),
                               C.c_double(float(undefined)), C.c_int(1 if gslib else 0))
    else:
        b = np.ascontiguousarray(bytes_, np.uint8).copy()
        t = np.ascontiguousarray(remap, np.uint8)
        rc = r.ref_write_byte(str(path).encode(), name.encode(), _p(b, C.c_ubyte), _p(m, C.c_ubyte), C.c_int(m.size),
                              C.c_int(indicator_count or t.size), C.c_ubyte(int(undefined) & 0xFF), _p(t, C.c_ubyte),
                              C.c_int(t.size), C.c_int(1 if gslib else 0))
    if rc != 0:
        raise RuntimeError("reference raised")


# ------------------------------------------------------ experimental variogram
def vario_directions(azimuth, dip, rotation):
    d = np.zeros(9, np.float64)
    lib().orc_vario_directions(C.c_double(azimuth), C.c_double(dip), C.c_double(rotation), _p(d, C.c_double))
    return d.reshape(3, 3)


def vario_grid(t, data, mask, shape, lag_count=None):
    """``t``: dict(dirs (3x3, rows = directions), R2, R3, lag_width, lag_sep,
    tol, num_lags, first).  data / mask flat, x fastest."""
    n = t["num_lags"] if lag_count is None else lag_count
    out = np.zeros(n, np.float32)
    dirs = np.ascontiguousarray(t["dirs"], np.float64).reshape(-1)
    lib().orc_vario_grid(_p(dirs, C.c_double), C.c_double(t["R2"]), C.c_double(t["R3"]),
                         C.c_double(t["lag_width"]), C.c_double(t["lag_sep"]), C.c_double(t["tol"]),
                         C.c_int(t["num_lags"]), C.c_double(t["first"]),
                         _p(np.ascontiguousarray(data, np.float32), C.c_float),
                         _p(np.ascontiguousarray(mask, np.uint8), C.c_ubyte), *[C.c_int(s) for s in shape],
                         C.c_int(n), _p(out, C.c_float))
    return out


def vario_points(t, xs, ys, zs, vs, lag_count=None):
    n = t["num_lags"] if lag_count is None else lag_count
    out = np.zeros(n, np.float32)
    dirs = np.ascontiguousarray(t["dirs"], np.float64).reshape(-1)
    f = lambda a: _p(np.ascontiguousarray(a, np.float32), C.c_float)
    lib().orc_vario_points(_p(dirs, C.c_double), C.c_double(t["R2"]), C.c_double(t["R3"]),
                           C.c_double(t["lag_width"]), C.c_double(t["lag_sep"]), C.c_double(t["tol"]),
                           C.c_double(t["first"]), f(xs), f(ys), f(zs), f(vs), C.c_int(len(xs)), C.c_int(n),
                           _p(out, C.c_float))
    return out


class _VecT(C.Structure):
    _fields_ = [("data", C.c_double * 3)]


class _EllT(C.Structure):
    _fields_ = [("d1", _VecT), ("d2", _VecT), ("d3", _VecT), ("R1", C.c_double), ("R2", C.c_double),
                ("R3", C.c_double)]


class _TemplT(C.Structure):
    _fields_ = [("lag_width", C.c_double), ("lag_separation", C.c_double), ("tol_distance", C.c_double),
                ("num_lags", C.c_int), ("first_lag_distance", C.c_double), ("ell", _EllT)]


class _HardT(C.Structure):
    _fields_ = [("data", C.c_void_p), ("mask", C.c_void_p), ("data_shape", C.c_int * 3),
                ("data_strides", C.c_int * 3), ("mask_shape", C.c_int * 3), ("mask_strides", C.c_int * 3)]


class _PtsT(C.Structure):
    _fields_ = [("xs", C.c_void_p), ("ys", C.c_void_p), ("zs", C.c_void_p), ("values", C.c_void_p),
                ("size", C.c_int)]


def _ref_template(t, R1=1.0):
    e = _EllT()
    for k, name in enumerate(("d1", "d2", "d3")):
        getattr(e, name).data[:] = list(np.asarray(t["dirs"])[k])
    e.R1, e.R2, e.R3 = R1, t["R2"], t["R3"]
    return _TemplT(t["lag_width"], t["lag_sep"], t["tol"], t["num_lags"], t["first"], e)


def ref_vario_directions(azimuth, dip, rotation):
    e = _EllT()
    ref_lib().fill_ellipsoid_directions(C.byref(e), C.c_double(azimuth), C.c_double(dip), C.c_double(rotation))
    return np.array([list(e.d1.data), list(e.d2.data), list(e.d3.data)])


def ref_vario_grid(t, data, mask, shape, lag_count=None):
    """The reference's calc_variograms (percentToUse 100) on dense x-fastest arrays."""
    n = t["num_lags"] if lag_count is None else lag_count
    d = np.ascontiguousarray(data, np.float32)
    m = np.ascontiguousarray(mask, np.uint8)
    nx, ny, nz = shape
    st = (C.c_int * 3)(1, nx, nx * ny)
    hd = _HardT(d.ctypes.data, m.ctypes.data, (C.c_int * 3)(*shape), st, (C.c_int * 3)(*shape), st)
    out = np.zeros(n, np.float32)
    tp = _ref_template(t)
    ref_lib().calc_variograms(C.byref(tp), C.byref(hd), _p(out, C.c_float), C.c_int(n), C.c_int(100))
    return out


def ref_vario_points(t, xs, ys, zs, vs, lag_count=None):
    n = t["num_lags"] if lag_count is None else lag_count
    arrs = [np.ascontiguousarray(a, np.float32) for a in (xs, ys, zs, vs)]
    ps = _PtsT(arrs[0].ctypes.data, arrs[1].ctypes.data, arrs[2].ctypes.data, arrs[3].ctypes.data, len(xs))
    out = np.zeros(n, np.float32)
    tp = _ref_template(t)
    ref_lib().calc_variograms_from_point_set(C.byref(tp), C.byref(ps), _p(out, C.c_float), C.c_int(n))
    return out


# ------------------------------------------------------------------ SGS
def mt_uniform(seed, n):
    """First ``n`` draws of std::uniform_real_distribution<double>(0, 1) over
    std::mt19937 seeded with (uint32) seed, as bs_random_generator.cpp takes them."""
    out = np.zeros(n, np.float64)
    lib().orc_mt_uniform(C.c_long(seed), C.c_int(n), _p(out, C.c_double))
    return out


def sgs_path(size, seed, n=None):
    """First ``n`` nodes (default all) of the reference's random path."""
    n = size if n is None else n
    out = np.zeros(n, np.int64)
    rc = lib().orc_sgs_path(C.c_longlong(size), C.c_int(seed), C.c_int(n), _p(out, C.c_longlong))
    if rc != 0:
        raise ValueError("path size out of range")
    return out


def gauss(x, mean=0.0, var=1.0, inverse=False):
    """gaussian_cdf_t: the CDF approximation, or its quantile when ``inverse``."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    out = np.zeros_like(x)
    lib().orc_gauss(_p(x, C.c_double), C.c_int(x.size), C.c_double(mean), C.c_double(var),
                    C.c_int(1 if inverse else 0), _p(out, C.c_double))
    return out


def sgs(kind, cov, radiuses, max_nb, shape, values, mask, seed, simmask=None, cdf=None, lvm_mean=None):
    """Sequential Gaussian simulation along the reference's path.  Returns
    (values, mask, order): the realization, the final informed flags and the
    simulated nodes in path order.  ``cdf`` = (values, probs) or None;
    ``lvm_mean`` (simple kriging): one mean per cell, or None."""
    nx, ny, nz = shape
    n = nx * ny * nz
    v = np.array(values, np.float32).reshape(-1)
    m = np.array(mask, np.uint8).reshape(-1)
    sm = None if simmask is None else np.ascontiguousarray(simmask, np.uint8).reshape(-1)
    cv = cp = None
    csize = 0
    if cdf is not None:
        cv = np.ascontiguousarray(cdf[0], np.float32)
        cp = np.ascontiguousarray(cdf[1], np.float32)
        csize = cv.size
    order = np.zeros(n, np.int32)
    lm = None if lvm_mean is None else np.ascontiguousarray(lvm_mean, np.float32).reshape(-1)
    f = lib().orc_sgs
    f.restype = C.c_longlong
    k = f(C.c_int(kind), C.byref(cov), (C.c_int * 3)(*radiuses), C.c_int(max_nb), C.c_int(nx), C.c_int(ny),
          C.c_int(nz), _p(v, C.c_float), _p(m, C.c_ubyte), _p(sm, C.c_ubyte), C.c_long(seed),
          _p(cv, C.c_float), _p(cp, C.c_float), C.c_longlong(csize), _p(order, C.c_int), _p(lm, C.c_float))
    if k < 0:
        raise ValueError("oracle rejected the parameters")
    return v, m, order[:k]
