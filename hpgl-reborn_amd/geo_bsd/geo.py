"""geo_bsd kriging API on the MI355X library.

Drop-in for the reference's grid-kriging functions (src/geo_bsd/geo.py):
same names, arguments, property containers and results.  Arrays are numpy
in Fortran order (x fastest), exactly as the reference expects; the native
calls go through the reference's C ABI, now served by libhpgl.so (HIP).

Differences from the reference, all deliberate:
  * a native error (for example a zero covariance range) raises RuntimeError
    instead of being stored silently (the reference's hpgl_wrap.py never
    reads hpgl_get_last_exception_message for these calls);
  * median_ik returns the same indicators, with cells that cannot be kriged
    classified from the marginal probability (the reference reads an
    uninitialised value there, median_ik.cpp:65-77).
"""
import ctypes as C

import numpy

from . import validation
from .hpgl_wrap import (_HPGL_CONT_MASKED_ARRAY, _HPGL_IND_MASKED_ARRAY, _HPGL_IK_PARAMS, _HPGL_MEDIAN_IK_PARAMS,
                        _HPGL_FLOAT_ARRAY, _HPGL_OK_PARAMS, _HPGL_SHAPE, _HPGL_SK_PARAMS, _HPGL_COV_PARAMS,
                        _HPGL_UBYTE_ARRAY, _hpgl_so,
                        hpgl_output_handler, hpgl_progress_handler, last_error)
from .validation import (CriticalValidationError, GridValidator, ParameterValidator, PathValidator, ValidationError,
                         ValidationWarning)

__all__ = ["ContProperty", "IndProperty", "SugarboxGrid", "CovarianceModel", "covariance", "ordinary_kriging",
           "simple_kriging", "lvm_kriging", "median_ik", "indicator_kriging", "simple_kriging_weights", "calc_mean", "set_output_handler",
           "set_progress_handler", "set_thread_num", "get_thread_num", "append_mask", "load_cont_property",
           "load_ind_property", "read_inc_file_float", "read_inc_file_byte", "write_property", "write_gslib_property",
           "ValidationError",
           "CriticalValidationError", "ValidationWarning"]


def checkFWA(a):
    """geo.py:140-144: Fortran-order, writable, aligned."""
    assert a.flags["F"] and a.flags["W"] and a.flags["A"]


class ContProperty:
    """geo.py:215-239."""

    def __init__(self, data, mask):
        self.data = numpy.require(data, "float32", "F")
        self.mask = numpy.require(mask, "uint8", "F")

    def validate(self):
        checkFWA(self.data)
        checkFWA(self.mask)
        assert self.data.shape == self.mask.shape

    def fix_shape(self, grid):
        if self.data.ndim != 3 and self.data.size == grid.x * grid.y * grid.z:
            self.data = self.data.reshape((grid.x, grid.y, grid.z), order="F")
        if self.mask.ndim != 3 and self.mask.size == grid.x * grid.y * grid.z:
            self.mask = self.mask.reshape((grid.x, grid.y, grid.z), order="F")

    def __getitem__(self, idx):
        if idx == 0:
            return self.data
        if idx == 1:
            return self.mask
        raise RuntimeError("Index out of range.")


class IndProperty:
    """geo.py:241-260."""

    def __init__(self, data, mask, indicator_count):
        self.data = numpy.require(data, "uint8", "F")
        self.mask = numpy.require(mask, "uint8", "F")
        self.indicator_count = indicator_count
        if numpy.sum(numpy.bitwise_and((numpy.asarray(mask) > 0), (numpy.asarray(data) >= indicator_count))) > 0:
            raise RuntimeError("Property contains some indicators outside of [0..%s] range." % (indicator_count - 1))
        assert numpy.asarray(data).shape == numpy.asarray(mask).shape

    def validate(self):
        checkFWA(self.data)
        checkFWA(self.mask)
        assert self.data.shape == self.mask.shape

    def __getitem__(self, idx):
        if idx == 0:
            return self.data
        if idx == 1:
            return self.mask
        if idx == 2:
            return self.indicator_count
        raise RuntimeError("Index out of range.")


def append_mask(prop, mask):
    """geo.py:270-273."""
    infs = prop[1]
    infs &= mask
    infs.choose(-99, prop[0], out=prop[0])


class covariance:
    """geo.py:275-278."""
    spherical = 0
    exponential = 1
    gaussian = 2


class SugarboxGrid:
    """geo.py:281-287."""

    def __init__(self, x, y, z):
        GridValidator.validate_grid_dimensions(x, y, z)
        self.x = x
        self.y = y
        self.z = z


class CovarianceModel:
    """geo.py:289-298."""

    def __init__(self, type=0, ranges=(0, 0, 0), angles=(0.0, 0.0, 0.0), sill=1.0, nugget=0.0):
        self.type = type
        self.ranges = ranges
        self.angles = angles
        self.sill = sill
        self.nugget = nugget
        ParameterValidator.validate_covariance_parameters(sill, nugget, ranges, angles)


def _empty_clone(prop):
    """geo.py:360-370: a property of the same kind and shape, zero data,
    nothing informed (the indicator count kept)."""
    data2 = prop.data.copy("F")
    data2.fill(0)
    mask2 = prop.mask.copy("F")
    mask2.fill(0)
    if isinstance(prop, IndProperty):
        return IndProperty(data2, mask2, prop.indicator_count)
    return ContProperty(data2, mask2)


def _clone_prop(prop):
    """geo.py:372-380: F-order copies of data and mask."""
    data2 = prop.data.copy("F")
    mask2 = prop.mask.copy("F")
    if isinstance(prop, IndProperty):
        return IndProperty(data2, mask2, prop.indicator_count)
    return ContProperty(data2, mask2)


def accepts_tuple(arg_name, arg_pos):
    """geo.py:390-410: (data, mask) or (data, mask, count) tuples become properties."""
    def tuple_to_prop(t):
        if isinstance(t, tuple):
            if len(t) == 3:
                return IndProperty(*t)
            if len(t) == 2:
                return ContProperty(*t)
            assert False
        assert isinstance(t, (ContProperty, IndProperty))
        return t

    def decorator(f):
        def new_f(*args, **kargs):
            if arg_name in kargs:
                kargs[arg_name] = tuple_to_prop(kargs[arg_name])
            elif len(args) > arg_pos:
                args = args[:arg_pos] + (tuple_to_prop(args[arg_pos]),) + args[arg_pos + 1:]
            else:
                assert False
            return f(*args, **kargs)
        new_f.__name__ = f.__name__
        new_f.__doc__ = f.__doc__
        return new_f
    return decorator


def _shape(dims):
    sh = _HPGL_SHAPE()
    for i in range(3):
        sh.data[i] = int(dims[i])
    sh.strides[0] = 1
    sh.strides[1] = int(dims[0])
    sh.strides[2] = int(dims[0]) * int(dims[1])
    return sh


def _require_cont_data(data):
    """geo.py:382-385: float32, Fortran order (None passes through)."""
    return None if data is None else numpy.require(data, dtype="float32", requirements="F")


def _requite_ind_data(data):
    """geo.py:387-390 (the reference's spelling): uint8, Fortran order."""
    return None if data is None else numpy.require(data, dtype="uint8", requirements="F")


def _array_shape(array, grid):
    if grid is not None:
        return _shape((grid.x, grid.y, grid.z))
    dims = tuple(array.shape) + (1,) * (3 - array.ndim)
    return _shape(dims)


def _create_hpgl_float_array(array, grid):
    """geo.py:203-214: an hpgl_float_array_t over a Fortran-order float32
    array (the array is kept alive by the struct)."""
    checkFWA(array)
    res = _HPGL_FLOAT_ARRAY(data=array.ctypes.data_as(C.POINTER(C.c_float)), shape=_array_shape(array, grid))
    res._array_ref = array
    return res


def _create_hpgl_ubyte_array(array, grid):
    """geo.py:191-201: an hpgl_ubyte_array_t over a Fortran-order uint8
    array."""
    checkFWA(array)
    res = _HPGL_UBYTE_ARRAY(data=array.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=_array_shape(array, grid))
    res._array_ref = array
    return res


def _grid_arrays(prop, grid):
    n = grid.x * grid.y * grid.z
    if prop.data.size != n:
        raise RuntimeError("Invalid data size. Size of data = %s. Size of grid = %s" % (prop.data.size, n))
    return _shape((grid.x, grid.y, grid.z))


def _check_native(name):
    msg = last_error()
    if msg:
        raise RuntimeError(f"{name}: {msg}")


def _validate_common(grid, radiuses, max_neighbours, cov_model):
    GridValidator.validate_grid_dimensions(grid.x, grid.y, grid.z)
    valid_radiuses = ParameterValidator.validate_radius(radiuses, "radiuses")
    ParameterValidator.validate_max_neighbors(max_neighbours)
    ParameterValidator.validate_covariance_parameters(cov_model.sill, cov_model.nugget, cov_model.ranges,
                                                      cov_model.angles)
    return tuple(int(r) for r in valid_radiuses)


def calc_mean(prop):
    """geo.py:560-570."""
    informed = numpy.asarray(prop.mask).reshape(-1, order="F") == 1
    if not informed.any():
        raise ValueError("calc_mean: no informed values (all masked)")
    vals = numpy.asarray(prop.data).reshape(-1, order="F")[informed]
    s = 0
    for v in vals:
        s += v
    return s / int(informed.sum())


@accepts_tuple("prop", 0)
def ordinary_kriging(prop, grid, radiuses, max_neighbours, cov_model):
    """Ordinary kriging of every uninformed cell (geo.py:572-607)."""
    valid_radiuses = _validate_common(grid, radiuses, max_neighbours, cov_model)
    out_prop = _clone_prop(prop)
    okp = _HPGL_OK_PARAMS(covariance_type=cov_model.type, ranges=tuple(cov_model.ranges),
                          angles=tuple(cov_model.angles), sill=cov_model.sill, nugget=cov_model.nugget,
                          radiuses=valid_radiuses, max_neighbours=max_neighbours)
    sh = _grid_arrays(prop, grid)
    inp = _HPGL_CONT_MASKED_ARRAY(data=prop.data.ctypes.data_as(C.POINTER(C.c_float)),
                                  mask=prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh)
    outp = _HPGL_CONT_MASKED_ARRAY(data=out_prop.data.ctypes.data_as(C.POINTER(C.c_float)),
                                   mask=out_prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh)
    _hpgl_so.hpgl_ordinary_kriging(C.byref(inp), C.byref(okp), C.byref(outp))
    _check_native("ordinary_kriging")
    return out_prop


@accepts_tuple("prop", 0)
def simple_kriging(prop, grid, radiuses, max_neighbours, cov_model, mean=None):
    """Simple kriging of every uninformed cell (geo.py:609-649); mean=None
    uses the mean of the informed values."""
    valid_radiuses = _validate_common(grid, radiuses, max_neighbours, cov_model)
    out_prop = _clone_prop(prop)
    skp = _HPGL_SK_PARAMS(covariance_type=cov_model.type, ranges=tuple(cov_model.ranges),
                          angles=tuple(cov_model.angles), sill=cov_model.sill, nugget=cov_model.nugget,
                          radiuses=valid_radiuses, max_neighbours=max_neighbours, automatic_mean=(mean is None),
                          mean=(mean if mean is not None else 0))
    sh = _grid_arrays(prop, grid)
    _hpgl_so.hpgl_simple_kriging(prop.data, prop.mask, C.byref(sh), C.byref(skp), out_prop.data, out_prop.mask,
                                 C.byref(sh))
    _check_native("simple_kriging")
    return out_prop


@accepts_tuple("prop", 0)
def median_ik(prop, grid, marginal_probs, radiuses, max_neighbours, cov_model):
    """Median indicator kriging with two indicators (geo.py:698-738)."""
    valid_radiuses = _validate_common(grid, radiuses, max_neighbours, cov_model)
    if len(marginal_probs) != 2:
        raise ValueError("median_ik: marginal_probs must have exactly 2 elements")
    for i, p in enumerate(marginal_probs):
        ParameterValidator.validate_probability(p, f"marginal_probs[{i}]")
    out_prop = _clone_prop(prop)
    mp = _HPGL_MEDIAN_IK_PARAMS(covariance_type=cov_model.type, ranges=tuple(cov_model.ranges),
                                angles=tuple(cov_model.angles), sill=cov_model.sill, nugget=cov_model.nugget,
                                radiuses=valid_radiuses, max_neighbours=max_neighbours,
                                marginal_probs=tuple(marginal_probs))
    sh = _shape((grid.x, grid.y, grid.z))
    inp = _HPGL_IND_MASKED_ARRAY(data=prop.data.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                 mask=prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh,
                                 indicator_count=prop.indicator_count)
    outp = _HPGL_IND_MASKED_ARRAY(data=out_prop.data.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                  mask=out_prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh,
                                  indicator_count=out_prop.indicator_count)
    _hpgl_so.hpgl_median_ik(C.byref(inp), C.byref(mp), C.byref(outp))
    _check_native("median_ik")
    return out_prop


@accepts_tuple("prop", 0)
def lvm_kriging(prop, grid, mean_data, radiuses, max_neighbours, cov_model):
    """Simple kriging with a locally varying mean (geo.py:651-696):
    mean_data holds one mean per cell, x fastest."""
    valid_radiuses = _validate_common(grid, radiuses, max_neighbours, cov_model)
    if not isinstance(mean_data, numpy.ndarray):
        raise ValueError("lvm_kriging: mean_data must be a numpy array")
    mean_data = numpy.require(mean_data, "float32", ["F", "W", "A"])
    if mean_data.size != grid.x * grid.y * grid.z:
        raise ValueError("lvm_kriging: mean_data must hold one value per grid cell")
    out_prop = _clone_prop(prop)
    okp = _HPGL_OK_PARAMS(covariance_type=cov_model.type, ranges=tuple(cov_model.ranges),
                          angles=tuple(cov_model.angles), sill=cov_model.sill, nugget=cov_model.nugget,
                          radiuses=valid_radiuses, max_neighbours=max_neighbours)
    sh = _grid_arrays(prop, grid)
    _hpgl_so.hpgl_lvm_kriging(prop.data, prop.mask, C.byref(sh), mean_data, C.byref(sh), C.byref(okp),
                              out_prop.data, out_prop.mask, C.byref(sh))
    _check_native("lvm_kriging")
    return out_prop


def _ik_params(data, marginal_probs):
    """geo.py:740-757: one _HPGL_IK_PARAMS per category."""
    n = len(data)
    arr = (_HPGL_IK_PARAMS * n)()
    for i, ikd in enumerate(data):
        cm = ikd["cov_model"]
        arr[i] = _HPGL_IK_PARAMS(covariance_type=cm.type, ranges=tuple(cm.ranges), angles=tuple(cm.angles),
                                 sill=cm.sill, nugget=cm.nugget, radiuses=tuple(int(r) for r in ikd["radiuses"]),
                                 max_neighbours=ikd["max_neighbours"], marginal_prob=marginal_probs[i])
    return arr


@accepts_tuple("prop", 0)
def indicator_kriging(prop, grid, data, marginal_probs):
    """Indicator kriging over categories (geo.py:759-800).  `data` holds one
    dict per category: cov_model, radiuses, max_neighbours.  Two categories
    go to median_ik with marginal_probs (p0, 1 - p0), as in the reference."""
    GridValidator.validate_grid_dimensions(grid.x, grid.y, grid.z)
    ParameterValidator.validate_indicator_count(len(data))
    if len(marginal_probs) != len(data):
        raise ValueError(f"indicator_kriging: marginal_probs length ({len(marginal_probs)}) must match data "
                         f"length ({len(data)})")
    for i, p in enumerate(marginal_probs):
        ParameterValidator.validate_probability(p, f"marginal_probs[{i}]")
    for i, ikd in enumerate(data):
        ParameterValidator.validate_radius(ikd["radiuses"], f"data[{i}].radiuses")
        ParameterValidator.validate_max_neighbors(ikd["max_neighbours"])
        cm = ikd["cov_model"]
        ParameterValidator.validate_covariance_parameters(cm.sill, cm.nugget, cm.ranges, cm.angles)
    for i in range(len(data)):
        data[i]["marginal_prob"] = marginal_probs[i]
    if len(data) == 2:
        return median_ik(prop, grid, (data[0]["marginal_prob"], 1 - data[0]["marginal_prob"]), data[0]["radiuses"],
                         data[0]["max_neighbours"], data[0]["cov_model"])
    out_prop = _clone_prop(prop)
    sh = _shape((grid.x, grid.y, grid.z))
    inp = _HPGL_IND_MASKED_ARRAY(data=prop.data.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                 mask=prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh,
                                 indicator_count=prop.indicator_count)
    outp = _HPGL_IND_MASKED_ARRAY(data=out_prop.data.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                  mask=out_prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh,
                                  indicator_count=out_prop.indicator_count)
    _hpgl_so.hpgl_indicator_kriging(C.byref(inp), C.byref(outp), _ik_params(data, marginal_probs), len(data))
    _check_native("indicator_kriging")
    return out_prop


def simple_kriging_weights(center_point, n_x, n_y, n_z, ranges=(100000, 100000, 100000), sill=1,
                           cov_type=covariance.exponential, nugget=None, angles=None):
    """SK weights of scattered neighbours around a point (geo.py:897-935)."""
    if angles is None:
        angles = (0, 0, 0)
    if nugget is None:
        nugget = 0
    ParameterValidator.validate_covariance_parameters(sill, nugget, ranges, angles)
    if len(n_x) != len(n_y) or len(n_x) != len(n_z):
        raise RuntimeError("Invalid pointset. %s,%s,%s." % (len(n_x), len(n_y), len(n_z)))
    for name, arr in (("n_x", n_x), ("n_y", n_y), ("n_z", n_z)):
        a = numpy.asarray(arr, dtype="float32")
        if numpy.any(numpy.isnan(a)) or numpy.any(numpy.isinf(a)):
            raise ValueError(f"simple_kriging_weights: {name} contains NaN or infinite values")
    covp = _HPGL_COV_PARAMS(covariance_type=cov_type, ranges=tuple(ranges), angles=tuple(angles), sill=sill,
                            nugget=nugget)
    weights = numpy.array([0] * len(n_x), dtype="float32")
    rc = _hpgl_so.hpgl_simple_kriging_weights((C.c_float * 3)(*center_point), numpy.array(n_x, dtype="float32"),
                                              numpy.array(n_y, dtype="float32"), numpy.array(n_z, dtype="float32"),
                                              len(n_x), C.byref(covp), weights)
    if rc != 0:
        raise RuntimeError("simple_kriging_weights: " + last_error())
    return weights


_handlers = {}


def set_output_handler(handler, param):
    """geo.py:947-956."""
    if handler is None:
        _hpgl_so.hpgl_set_output_handler(C.cast(None, hpgl_output_handler), None)
        _handlers.pop("output", None)
    else:
        h = hpgl_output_handler(handler)
        _handlers["output"] = h
        _hpgl_so.hpgl_set_output_handler(h, param)


def set_progress_handler(handler, param):
    """geo.py:958-966."""
    if handler is None:
        _hpgl_so.hpgl_set_progress_handler(C.cast(None, hpgl_progress_handler), None)
        _handlers.pop("progress", None)
    else:
        h = hpgl_progress_handler(handler)
        _handlers["progress"] = h
        _hpgl_so.hpgl_set_progress_handler(h, param)


def set_thread_num(num):
    _hpgl_so.hpgl_set_thread_num(int(num))


def get_thread_num():
    return _hpgl_so.hpgl_get_thread_num()


# ------------------------------------------------------------ property files
def _native_error(prefix):
    msg = last_error()
    return RuntimeError(prefix + msg)


def _load_prop_cont_slow(filename, undefined_value):
    """geo.py:300-322: every line not starting with "--", every
    whitespace-separated part that float() accepts."""
    safe_path = PathValidator.validate_filepath(filename)
    values, mask = [], []
    with open(safe_path, "r", encoding="utf-8") as f:
        for line in f:
            if line.strip().startswith("--"):
                continue
            for part in line.split():
                try:
                    val = float(part.strip())
                except (ValueError, TypeError):
                    continue
                values.append(val)
                mask.append(0 if val == undefined_value else 1)
    return ContProperty(numpy.array(values, dtype="float32"), numpy.array(mask, dtype="uint8"))


def _load_prop_ind_slow(filename, undefined_value, ind_values):
    """geo.py:324-352."""
    dict_map = {v: i for i, v in enumerate(ind_values)}
    safe_path = PathValidator.validate_filepath(filename)
    values, mask = [], []
    with open(safe_path, "r", encoding="utf-8") as f:
        for line in f:
            if line.strip().startswith("--"):
                continue
            for part in line.split():
                try:
                    val = int(part.strip())
                    if val == undefined_value:
                        values.append(255)
                        mask.append(0)
                    else:
                        values.append(dict_map[val])
                        mask.append(1)
                except (ValueError, TypeError, KeyError):
                    pass
    return IndProperty(numpy.array(values, dtype="uint8", order="F"), numpy.array(mask, dtype="uint8", order="F"),
                       len(ind_values))


def _total(size):
    if isinstance(size, tuple) and len(size) == 3:
        GridValidator.validate_grid_dimensions(size[0], size[1], size[2])
        return int(size[0]) * int(size[1]) * int(size[2])
    return int(size)


def read_inc_file_float(filename, undefined_value, size):
    """geo.py:500-521: `size` values of an INC file (parsed on the GPU)."""
    safe_path = PathValidator.validate_filepath(filename, must_exist=True)
    n = _total(size)
    data = numpy.zeros(n, dtype="float32", order="F")
    mask = numpy.zeros(n, dtype="uint8", order="F")
    rc = _hpgl_so.hpgl_read_inc_file_float(safe_path.encode("utf-8"), undefined_value, n, data, mask)
    if rc != 0:
        raise _native_error("read_inc_file_float failed: ")
    return ContProperty(data, mask)


def read_inc_file_byte(filename, undefined_value, size, indicator_values):
    """geo.py:523-544: integer values as bytes, remapped to the index of
    their value in indicator_values."""
    safe_path = PathValidator.validate_filepath(filename, must_exist=True)
    n = _total(size)
    data = numpy.zeros(n, dtype="uint8", order="F")
    mask = numpy.zeros(n, dtype="uint8", order="F")
    vals = numpy.array(indicator_values, dtype="uint8")
    rc = _hpgl_so.hpgl_read_inc_file_byte(safe_path.encode("utf-8"), undefined_value, n, data, mask, vals, len(vals))
    if rc != 0:
        raise _native_error("read_inc_file_byte failed: ")
    return IndProperty(data, mask, len(indicator_values))


def load_cont_property(filename, undefined_value, size=None):
    """geo.py:490-498."""
    safe_path = PathValidator.validate_filepath(filename, must_exist=True)
    if size is None:
        print("[WARNING]. load_cont_property: Size is not specified. Using slow and ineficcient method.")
        return _load_prop_cont_slow(safe_path, undefined_value)
    return read_inc_file_float(safe_path, undefined_value, size)


def load_ind_property(filename, undefined_value, indicator_values, size=None):
    """geo.py:546-551."""
    if size is None:
        print("[WARNING]. load_ind_property: Size is not specified. Using slow and ineficcient method.")
        return _load_prop_ind_slow(filename, undefined_value, indicator_values)
    return read_inc_file_byte(filename, undefined_value, size, indicator_values)


def _file_shape(prop):
    if prop.data.ndim == 3:
        return _shape(prop.data.shape)
    return _shape((prop.data.size, 1, 1))


@accepts_tuple("prop", 0)
def write_property(prop, filename, prop_name, undefined_value, indicator_values=None):
    """geo.py:418-461: INC file (values printed on the GPU)."""
    safe_path = PathValidator.validate_filepath(filename)
    if indicator_values is None:
        indicator_values = []
    sh = _file_shape(prop)
    if isinstance(prop, ContProperty):
        marr = _HPGL_CONT_MASKED_ARRAY(data=prop.data.ctypes.data_as(C.POINTER(C.c_float)),
                                       mask=prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh)
        rc = _hpgl_so.hpgl_write_inc_file_float(safe_path.encode("utf-8"), C.byref(marr), undefined_value,
                                                prop_name.encode("utf-8"))
    else:
        vals = numpy.array(indicator_values, dtype="uint8")
        marr = _HPGL_IND_MASKED_ARRAY(data=prop.data.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                      mask=prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh,
                                      indicator_count=prop.indicator_count)
        rc = _hpgl_so.hpgl_write_inc_file_byte(safe_path.encode("utf-8"), C.byref(marr), undefined_value,
                                               prop_name.encode("utf-8"), vals.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                               len(indicator_values))
    if rc != 0:
        raise _native_error("write_property failed: ")


def get_gslib_property(prop_dict, prop_name, undefined_value):
    """geo.py:937-945: (values, informed) for one column of a dict of GSLIB
    columns; a cell is informed where its value differs from
    undefined_value."""
    prop = prop_dict[prop_name]
    informed = (numpy.asarray(prop) != undefined_value).astype(numpy.uint8)
    return (prop, informed)


@accepts_tuple("prop", 0)
def write_gslib_property(prop, filename, prop_name, undefined_value, indicator_values=None):
    """geo.py:463-488: GSLIB file (values printed on the GPU)."""
    safe_path = PathValidator.validate_filepath(filename)
    if indicator_values is None:
        indicator_values = []
    sh = _shape((prop.data.size, 1, 1)) if prop.data.ndim != 3 else _shape(prop.data.shape)
    if isinstance(prop, ContProperty):
        marr = _HPGL_CONT_MASKED_ARRAY(data=prop.data.ctypes.data_as(C.POINTER(C.c_float)),
                                       mask=prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh)
        rc = _hpgl_so.hpgl_write_gslib_cont_property(C.byref(marr), safe_path.encode("utf-8"),
                                                     prop_name.encode("utf-8"), float(undefined_value))
    else:
        vals = (C.c_ubyte * max(1, len(indicator_values)))(*indicator_values)
        marr = _HPGL_IND_MASKED_ARRAY(data=prop.data.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                      mask=prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh,
                                      indicator_count=prop.indicator_count)
        rc = _hpgl_so.hpgl_write_gslib_byte_property(C.byref(marr), safe_path.encode("utf-8"),
                                                     prop_name.encode("utf-8"), float(undefined_value), vals,
                                                     len(indicator_values))
    if rc != 0:
        raise _native_error("write_gslib_property failed: ")
