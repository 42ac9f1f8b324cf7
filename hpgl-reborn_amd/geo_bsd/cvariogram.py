"""Experimental variogram on the MI355X: the reference's geo_bsd.cvariogram
API (cvariogram.py:38-215) -- Ellipsoid, VariogramSearchTemplate,
CalcVariograms, CalcVariogramsFromPointSet -- over the same C structs
(_cvariogram/api.h:24-89), computed by libhpgl.so's GPU kernels.

Difference: with percent < 100, CalcVariograms keeps a deterministic
pseudo-random share of the pairs (the reference draws the C library's
rand(), so its subsample depends on the process's earlier rand() calls).
"""
import ctypes as C

import numpy

from .hpgl_wrap import _hpgl_so as cvar


class vector_t(C.Structure):
    _fields_ = [("data", C.c_double * 3)]


class ellipsoid_t(C.Structure):
    _fields_ = [("direction1", vector_t), ("direction2", vector_t), ("direction3", vector_t), ("R1", C.c_double),
                ("R2", C.c_double), ("R3", C.c_double)]


class variogram_search_template_t(C.Structure):
    _fields_ = [("lag_width", C.c_double), ("lag_separation", C.c_double), ("tol_distance", C.c_double),
                ("num_lags", C.c_int), ("first_lag_distance", C.c_double), ("ellipsoid", ellipsoid_t)]


class hard_data_t(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("mask", C.POINTER(C.c_ubyte)), ("data_shape", C.c_int * 3),
                ("data_strides", C.c_int * 3), ("mask_shape", C.c_int * 3), ("mask_strides", C.c_int * 3)]


class cont_point_set_t(C.Structure):
    _fields_ = [("xs", C.POINTER(C.c_float)), ("ys", C.POINTER(C.c_float)), ("zs", C.POINTER(C.c_float)),
                ("values", C.POINTER(C.c_float)), ("values_count", C.c_int)]


cvar.fill_ellipsoid_directions.restype = None
cvar.fill_ellipsoid_directions.argtypes = [C.POINTER(ellipsoid_t), C.c_double, C.c_double, C.c_double]
cvar.calc_variograms.restype = None
cvar.calc_variograms.argtypes = [C.POINTER(variogram_search_template_t), C.POINTER(hard_data_t),
                                 numpy.ctypeslib.ndpointer(dtype=numpy.float32), C.c_int, C.c_int]
cvar.calc_variograms_from_point_set.restype = None
cvar.calc_variograms_from_point_set.argtypes = [C.POINTER(variogram_search_template_t), C.POINTER(cont_point_set_t),
                                                numpy.ctypeslib.ndpointer(dtype=numpy.float32), C.c_int]
cvar.hpglx_calc_variograms.restype = C.c_int
cvar.hpglx_calc_variograms.argtypes = [C.POINTER(variogram_search_template_t), C.c_void_p, C.c_void_p, C.c_int * 3,
                                       C.c_int, C.c_int, C.c_void_p, C.c_void_p]
cvar.hpglx_calc_variograms_points.restype = C.c_int
cvar.hpglx_calc_variograms_points.argtypes = [C.POINTER(variogram_search_template_t), C.c_void_p, C.c_int, C.c_int,
                                              C.c_void_p, C.c_void_p]


def checked_create(T, **kargs):
    """cvariogram.py:104-112: every field must be given."""
    fields = [f for f, _ in T._fields_]
    for k in kargs:
        if k in fields:
            fields.remove(k)
    assert len(fields) == 0, "No values for parameters: %s" % fields
    return T(**kargs)


def _strides(array):
    return tuple(s // array.itemsize for s in array.strides)


def _c_array(t, size, values):
    return (t * size)(*values)


def _check_native(name):
    msg = cvar.hpgl_get_last_exception_message()
    if msg:
        raise RuntimeError(f"{name}: {msg.decode('utf-8', errors='replace')}")


class Ellipsoid:
    """cvariogram.py:123-137."""

    def __init__(self, R1, R2, R3, azimuth, dip, rotation):
        vec = checked_create(vector_t, data=_c_array(C.c_double, 3, (0, 0, 0)))
        self.ell = checked_create(ellipsoid_t, direction1=vec, direction2=vec, direction3=vec, R1=R1, R2=R2, R3=R3)
        cvar.fill_ellipsoid_directions(C.byref(self.ell), azimuth, dip, rotation)


class VariogramSearchTemplate:
    """cvariogram.py:139-154."""

    def __init__(self, lag_width, lag_separation, tol_distance, num_lags, first_lag_distance, ellipsoid):
        self.templ = checked_create(variogram_search_template_t, lag_width=lag_width, lag_separation=lag_separation,
                                    tol_distance=tol_distance, num_lags=num_lags,
                                    first_lag_distance=first_lag_distance, ellipsoid=ellipsoid.ell)
        self.num_lags = num_lags
        self.lag_separation = lag_separation


def _borders(templ):
    b = numpy.zeros(templ.num_lags)
    for k in range(templ.num_lags):
        b[k] = k * templ.lag_separation
    return b


def CalcVariograms(templ, hard_data, percent=100):
    """cvariogram.py:157-186: (lag borders, semivariogram) of gridded hard
    data, hard_data = (data, mask) 3-D arrays of any memory order."""
    if templ.num_lags <= 0:
        raise ValueError("CalcVariograms: num_lags must be positive")
    if percent < 1 or percent > 100:
        raise ValueError(f"CalcVariograms: percent must be in [1, 100], got {percent}")
    variogram = numpy.array([0] * templ.num_lags, dtype="float32")
    data, mask = hard_data[0], hard_data[1]
    hd = checked_create(hard_data_t, data=data.ctypes.data_as(C.POINTER(C.c_float)),
                        mask=mask.ctypes.data_as(C.POINTER(C.c_ubyte)),
                        data_shape=_c_array(C.c_int, 3, data.shape), mask_shape=_c_array(C.c_int, 3, mask.shape),
                        data_strides=_c_array(C.c_int, 3, _strides(data)),
                        mask_strides=_c_array(C.c_int, 3, _strides(mask)))
    cvar.calc_variograms(C.byref(templ.templ), C.byref(hd), variogram, variogram.size, int(percent))
    _check_native("CalcVariograms")
    return (_borders(templ), variogram)


def CalcVariogramsFromPointSet(templ, point_set, variogram=None):
    """cvariogram.py:188-215: (lag borders, semivariogram) of a point set
    given as a dict of float32 arrays X, Y, Z, Property."""
    if templ.num_lags <= 0:
        raise ValueError("CalcVariogramsFromPointSet: num_lags must be positive")
    for key in ("X", "Y", "Z", "Property"):
        if key not in point_set:
            raise ValueError(f"CalcVariogramsFromPointSet: point_set missing required key '{key}'")
    variogram = numpy.array([0] * templ.num_lags, dtype="float32")
    arrs = {k: numpy.ascontiguousarray(point_set[k], dtype=numpy.float32) for k in ("X", "Y", "Z", "Property")}
    ps = checked_create(cont_point_set_t, xs=arrs["X"].ctypes.data_as(C.POINTER(C.c_float)),
                        ys=arrs["Y"].ctypes.data_as(C.POINTER(C.c_float)),
                        zs=arrs["Z"].ctypes.data_as(C.POINTER(C.c_float)),
                        values=arrs["Property"].ctypes.data_as(C.POINTER(C.c_float)),
                        values_count=arrs["Property"].size)
    cvar.calc_variograms_from_point_set(C.byref(templ.templ), C.byref(ps), variogram, variogram.size)
    _check_native("CalcVariogramsFromPointSet")
    return (_borders(templ), variogram)


def calc_variograms_device(templ, data, mask, dims, lag_count=None, percent=100):
    """Semivariogram of a dense device grid (torch tensors, x fastest);
    returns a float32 device tensor of the lags."""
    import torch
    n = templ.num_lags if lag_count is None else int(lag_count)
    out = torch.empty(n, dtype=torch.float32, device=data.device)
    stream = torch.cuda.current_stream(data.device).cuda_stream
    rc = cvar.hpglx_calc_variograms(C.byref(templ.templ), C.c_void_p(data.data_ptr()), C.c_void_p(mask.data_ptr()),
                                    (C.c_int * 3)(*[int(d) for d in dims]), n, int(percent),
                                    C.c_void_p(out.data_ptr()), C.c_void_p(stream))
    if rc != 0:
        raise RuntimeError("hpglx_calc_variograms: " + cvar.hpglx_last_error().decode())
    return out


def calc_variograms_points_device(templ, xyzv, lag_count=None):
    """Semivariogram of a device point set (n x 4 float32 tensor: x, y, z,
    value)."""
    import torch
    n = templ.num_lags if lag_count is None else int(lag_count)
    out = torch.empty(n, dtype=torch.float32, device=xyzv.device)
    stream = torch.cuda.current_stream(xyzv.device).cuda_stream
    rc = cvar.hpglx_calc_variograms_points(C.byref(templ.templ), C.c_void_p(xyzv.data_ptr()), int(xyzv.shape[0]), n,
                                           C.c_void_p(out.data_ptr()), C.c_void_p(stream))
    if rc != 0:
        raise RuntimeError("hpglx_calc_variograms_points: " + cvar.hpglx_last_error().decode())
    return out
