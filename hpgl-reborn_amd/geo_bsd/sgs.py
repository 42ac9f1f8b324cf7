"""Sequential Gaussian simulation (the reference's sgs.py:52-160) on the
MI355X path: hpgl_sgs_simulation in libhpgl.so."""
import ctypes as C

import numpy

from .cdf import CdfData
from .geo import ContProperty, CovarianceModel, _check_native, _clone_prop, _empty_clone, _shape, accepts_tuple
from .hpgl_wrap import (_HPGL_CONT_MASKED_ARRAY, _HPGL_FLOAT_ARRAY, _HPGL_KRIGING_KIND, _HPGL_SGS_PARAMS,
                        _HPGL_UBYTE_ARRAY, _hpgl_so, hpgl_non_parametric_cdf_t)
from .validation import GridValidator, ParameterValidator


def normed_cov_model(cov_model):
    """sgs.py:43-50: the model scaled to unit sill."""
    coef = cov_model.sill
    return CovarianceModel(cov_model.type, cov_model.ranges, cov_model.angles, cov_model.sill / coef,
                           cov_model.nugget / coef)


@accepts_tuple("prop", 0)
def sgs_simulation(prop, grid, cdf_data, radiuses, max_neighbours, cov_model, seed, kriging_type="sk", mean=None,
                   use_harddata=True, use_regions=False, region_size=None, mask=None, force_single_thread=False,
                   force_parallel=False, min_neighbours=0, **params):
    """Performs sequential Gaussian simulation (sgs.py:52-160).

    cdf_data: None (no transform) or CdfData of the normal-score transform.
    radiuses, max_neighbours, cov_model: the kriging search and model (the
    model is scaled to unit sill).  seed: path and draws.  kriging_type:
    "sk" (mean 0 in normal-score space) or "ok".  mean: a scalar is accepted
    and unused, as in the reference; an array of one mean per cell selects
    simple kriging with a locally varying mean (hpgl_sgs_lvm_simulation;
    the means go through the normal-score transform with the data).
    use_harddata: condition on the property's informed
    cells.  mask: 1 where cells are simulated (None: every cell).
    use_regions / force_* / min_neighbours: accepted for compatibility; the
    reference's one-path algorithm ignores them as well.
    Returns a new ContProperty.
    """
    GridValidator.validate_grid_dimensions(grid.x, grid.y, grid.z)
    valid_radiuses = tuple(int(r) for r in ParameterValidator.validate_radius(radiuses, "radiuses"))
    ParameterValidator.validate_max_neighbors(max_neighbours)
    ParameterValidator.validate_covariance_parameters(cov_model.sill, cov_model.nugget, cov_model.ranges,
                                                      cov_model.angles)
    ParameterValidator.validate_seed(seed)
    ParameterValidator.validate_min_neighbors(min_neighbours, max_neighbours)
    if kriging_type not in ("sk", "ok"):
        # the reference's lookup of the kind (sgs.py:122) raises KeyError
        raise KeyError(kriging_type)
    prop.fix_shape(grid)
    cov_model = normed_cov_model(cov_model)
    out_prop = _clone_prop(prop) if use_harddata else _empty_clone(prop)
    sgsp = _HPGL_SGS_PARAMS(covariance_type=cov_model.type, ranges=tuple(cov_model.ranges),
                            angles=tuple(cov_model.angles), sill=cov_model.sill, nugget=cov_model.nugget,
                            radiuses=valid_radiuses, max_neighbours=max_neighbours,
                            kriging_kind={"sk": _HPGL_KRIGING_KIND.simple, "ok": _HPGL_KRIGING_KIND.ordinary}[
                                kriging_type], seed=int(seed), min_neighbours=min_neighbours)
    sh = _shape((grid.x, grid.y, grid.z))
    cdf = None
    keep = []
    if cdf_data is not None:
        assert isinstance(cdf_data, CdfData)
        cv = numpy.ascontiguousarray(cdf_data.values, numpy.float32)
        cp = numpy.ascontiguousarray(cdf_data.probs, numpy.float32)
        keep += [cv, cp]
        cdf = C.byref(hpgl_non_parametric_cdf_t(values=cv.ctypes.data_as(C.POINTER(C.c_float)),
                                                probs=cp.ctypes.data_as(C.POINTER(C.c_float)), size=cv.size))
    sim = None
    if mask is not None:
        m = numpy.require(mask, "uint8", "F")
        if m.size != grid.x * grid.y * grid.z:
            raise ValueError("sgs_simulation: the mask does not match the grid")
        keep.append(m)
        sim = C.byref(_HPGL_UBYTE_ARRAY(data=m.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh))
    data = _HPGL_CONT_MASKED_ARRAY(data=out_prop.data.ctypes.data_as(C.POINTER(C.c_float)),
                                   mask=out_prop.mask.ctypes.data_as(C.POINTER(C.c_ubyte)), shape=sh)
    if mean is None or numpy.isscalar(mean):
        mean_arg = C.byref(C.c_double(mean)) if mean is not None else None
        _hpgl_so.hpgl_sgs_simulation(C.byref(data), C.byref(sgsp), cdf, mean_arg, sim)
    else:
        mf = numpy.require(mean, "float32", "F")
        if mf.size != grid.x * grid.y * grid.z:
            raise ValueError("sgs_simulation: the mean array does not match the grid")
        keep.append(mf)
        means = _HPGL_FLOAT_ARRAY(data=mf.ctypes.data_as(C.POINTER(C.c_float)), shape=sh)
        _hpgl_so.hpgl_sgs_lvm_simulation(C.byref(data), C.byref(sgsp), cdf, C.byref(means), sim)
    _check_native("sgs_simulation")
    return out_prop


def sgs_schedule(shape, radiuses, informed, seed, simulation_mask=None):
    """The host schedule of sgs_simulation for a grid of `shape` (nx, ny, nz):
    the cells to simulate along the reference's random path, grouped into
    levels that the GPU runs in parallel.  informed and simulation_mask are
    flat uint8 arrays of nx*ny*nz cells (x fastest).  Returns (cells,
    deviates, levels) in level order, path order within a level; deviates
    are the standard normal quantiles of the cells' draws.  Needs no GPU."""
    nx, ny, nz = (int(v) for v in shape)
    n = nx * ny * nz
    inf = numpy.ascontiguousarray(informed, dtype=numpy.uint8).reshape(-1)
    if inf.size != n:
        raise ValueError("informed must hold nx*ny*nz values")
    sm = None
    if simulation_mask is not None:
        sm = numpy.ascontiguousarray(simulation_mask, dtype=numpy.uint8).reshape(-1)
        if sm.size != n:
            raise ValueError("simulation_mask must hold nx*ny*nz values")
    dims = (C.c_int * 3)(nx, ny, nz)
    rad = (C.c_int * 3)(*(int(r) for r in radiuses))
    args = (dims, rad, inf.ctypes.data, None if sm is None else sm.ctypes.data, int(seed))
    m = _hpgl_so.hpglx_sgs_schedule(*args, None, None, None, 0)
    if m < 0:
        raise ValueError("sgs_schedule: " + _hpgl_so.hpglx_last_error().decode())
    cells = numpy.empty(m, numpy.int32)
    dev = numpy.empty(m, numpy.float64)
    lev = numpy.empty(m, numpy.int32)
    if m:
        _hpgl_so.hpglx_sgs_schedule(*args, cells.ctypes.data, dev.ctypes.data, lev.ctypes.data, m)
    return cells, dev, lev
