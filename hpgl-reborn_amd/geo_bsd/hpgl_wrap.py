"""ctypes binding of libhpgl.so (the MI355X kriging library).

Mirrors the reference binding (src/geo_bsd/hpgl_wrap.py): the same C struct
layouts (hpgl_wrap.py:44-135), and the library is loaded from this package
directory only (hpgl_wrap.py:170-273).  The library is the HIP build; there is no CPU
fallback, and importing this module fails loudly when it is missing.
"""
import ctypes as C
import os

import numpy
import numpy.ctypeslib as NC

hpgl_output_handler = C.CFUNCTYPE(C.c_int, C.c_char_p, C.py_object)
hpgl_progress_handler = C.CFUNCTYPE(C.c_int, C.c_char_p, C.c_int, C.py_object)


class _HPGL_SHAPE(C.Structure):
    _fields_ = [("data", C.c_int * 3), ("strides", C.c_int * 3)]


class _HPGL_CONT_MASKED_ARRAY(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("mask", C.POINTER(C.c_ubyte)), ("shape", _HPGL_SHAPE)]


class _HPGL_IND_MASKED_ARRAY(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_ubyte)), ("mask", C.POINTER(C.c_ubyte)), ("shape", _HPGL_SHAPE),
                ("indicator_count", C.c_int)]


class _HPGL_OK_PARAMS(C.Structure):
    _fields_ = [("covariance_type", C.c_int), ("ranges", C.c_double * 3), ("angles", C.c_double * 3),
                ("sill", C.c_double), ("nugget", C.c_double), ("radiuses", C.c_int * 3),
                ("max_neighbours", C.c_int)]


class _HPGL_SK_PARAMS(C.Structure):
    _fields_ = [("covariance_type", C.c_int), ("ranges", C.c_double * 3), ("angles", C.c_double * 3),
                ("sill", C.c_double), ("nugget", C.c_double), ("radiuses", C.c_int * 3),
                ("max_neighbours", C.c_int), ("automatic_mean", C.c_ubyte), ("mean", C.c_double)]


class _HPGL_MEDIAN_IK_PARAMS(C.Structure):
    _fields_ = [("covariance_type", C.c_int), ("ranges", C.c_double * 3), ("angles", C.c_double * 3),
                ("sill", C.c_double), ("nugget", C.c_double), ("radiuses", C.c_int * 3),
                ("max_neighbours", C.c_int), ("marginal_probs", C.c_double * 2)]


class _HPGL_IK_PARAMS(C.Structure):
    _fields_ = [("covariance_type", C.c_int), ("ranges", C.c_double * 3), ("angles", C.c_double * 3),
                ("sill", C.c_double), ("nugget", C.c_double), ("radiuses", C.c_int * 3),
                ("max_neighbours", C.c_int), ("marginal_prob", C.c_double)]


class __hpgl_cov_params_t(C.Structure):
    _fields_ = [("covariance_type", C.c_int), ("ranges", C.c_double * 3), ("angles", C.c_double * 3),
                ("sill", C.c_double), ("nugget", C.c_double)]


_HPGL_COV_PARAMS = __hpgl_cov_params_t


class _HPGL_FLOAT_ARRAY(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("shape", _HPGL_SHAPE)]


class _HPGL_UBYTE_ARRAY(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_ubyte)), ("shape", _HPGL_SHAPE)]


class _HPGL_KRIGING_KIND:
    ordinary = 0
    simple = 1


class _HPGL_SGS_PARAMS(C.Structure):
    _fields_ = [("covariance_type", C.c_int), ("ranges", C.c_double * 3), ("angles", C.c_double * 3),
                ("sill", C.c_double), ("nugget", C.c_double), ("radiuses", C.c_int * 3),
                ("max_neighbours", C.c_int), ("kriging_kind", C.c_int), ("seed", C.c_long),
                ("min_neighbours", C.c_int)]


class hpgl_non_parametric_cdf_t(C.Structure):
    _fields_ = [("values", C.POINTER(C.c_float)), ("probs", C.POINTER(C.c_float)), ("size", C.c_longlong)]


class _HPGLX_GRID_REQUEST(C.Structure):
    """hpglx_grid_request_t (include/hpglx.h)."""
    _fields_ = [
        ("kind", C.c_int), ("precision", C.c_int), ("dims", C.c_int * 3), ("radiuses", C.c_int * 3),
        ("max_neighbours", C.c_int), ("cov", _HPGL_COV_PARAMS),
        ("values", C.c_void_p), ("mask", C.c_void_p), ("ind", C.c_void_p), ("zfield", C.c_void_p),
        ("thresholds", C.c_void_p), ("nind", C.c_int), ("marginal", C.POINTER(C.c_double)),
        ("auto_mean", C.c_int), ("mean", C.c_double),
        ("out_val", C.c_void_p), ("out_mask", C.c_void_p), ("out_class", C.c_void_p), ("out_prob", C.c_void_p),
        ("out_var", C.c_void_p), ("out_est64", C.c_void_p), ("out_status", C.c_void_p),
        ("order_relations", C.c_int), ("out_count", C.c_void_p), ("out_nboff", C.c_void_p),
        ("pivot_tol", C.c_double), ("refine_steps", C.c_int), ("lvm_mean", C.c_void_p)]


LIB_NAME = "libhpgl.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)


def _load():
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it with `python hpgl-reborn_amd/build.py` "
                      "(there is no CPU fallback)")
    # PyTorch-ROCm ships its own HIP runtime under the same soname
    # (libamdhip64.so.7).  Whichever copy is loaded first serves the whole
    # process, so when torch is installed it is imported first: torch then runs
    # on the runtime it was built with, and this library shares it (one device
    # context, interoperable streams and pointers).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return C.CDLL(LIB_PATH)


_hpgl_so = _load()

_hpgl_so.hpgl_get_last_exception_message.restype = C.c_char_p
_hpgl_so.hpgl_get_last_exception_message.argtypes = []
_hpgl_so.hpgl_set_output_handler.restype = None
_hpgl_so.hpgl_set_output_handler.argtypes = [hpgl_output_handler, C.py_object]
_hpgl_so.hpgl_set_progress_handler.restype = None
_hpgl_so.hpgl_set_progress_handler.argtypes = [hpgl_progress_handler, C.py_object]
_hpgl_so.hpgl_set_thread_num.restype = None
_hpgl_so.hpgl_set_thread_num.argtypes = [C.c_int]
_hpgl_so.hpgl_get_thread_num.restype = C.c_int
_hpgl_so.hpgl_get_thread_num.argtypes = []

_hpgl_so.hpgl_ordinary_kriging.restype = None
_hpgl_so.hpgl_ordinary_kriging.argtypes = [C.POINTER(_HPGL_CONT_MASKED_ARRAY), C.POINTER(_HPGL_OK_PARAMS),
                                           C.POINTER(_HPGL_CONT_MASKED_ARRAY)]
_hpgl_so.hpgl_simple_kriging.restype = None
_hpgl_so.hpgl_simple_kriging.argtypes = [
    NC.ndpointer(dtype=numpy.float32, flags=["F", "W", "A"]), NC.ndpointer(dtype=numpy.ubyte, flags=["F", "W", "A"]),
    C.POINTER(_HPGL_SHAPE), C.POINTER(_HPGL_SK_PARAMS),
    NC.ndpointer(dtype=numpy.float32, flags=["F", "W", "A"]), NC.ndpointer(dtype=numpy.ubyte, flags=["F", "W", "A"]),
    C.POINTER(_HPGL_SHAPE)]
_hpgl_so.hpgl_simple_kriging_weights.restype = C.c_int
_hpgl_so.hpgl_simple_kriging_weights.argtypes = [
    (C.c_float * 3), NC.ndpointer(dtype=numpy.float32), NC.ndpointer(dtype=numpy.float32),
    NC.ndpointer(dtype=numpy.float32), C.c_int, C.POINTER(__hpgl_cov_params_t), NC.ndpointer(dtype=numpy.float32)]
_hpgl_so.hpgl_median_ik.restype = None
_hpgl_so.hpgl_median_ik.argtypes = [C.POINTER(_HPGL_IND_MASKED_ARRAY), C.POINTER(_HPGL_MEDIAN_IK_PARAMS),
                                    C.POINTER(_HPGL_IND_MASKED_ARRAY)]

_hpgl_so.hpglx_krige.restype = C.c_int
_hpgl_so.hpglx_krige.argtypes = [C.POINTER(_HPGLX_GRID_REQUEST), C.c_void_p]
_hpgl_so.hpgl_lvm_kriging.restype = None
_hpgl_so.hpgl_lvm_kriging.argtypes = [
    NC.ndpointer(dtype=numpy.float32, flags=["F", "W", "A"]), NC.ndpointer(dtype=numpy.ubyte, flags=["F", "W", "A"]),
    C.POINTER(_HPGL_SHAPE), NC.ndpointer(dtype=numpy.float32, flags=["F", "W", "A"]), C.POINTER(_HPGL_SHAPE),
    C.POINTER(_HPGL_OK_PARAMS),
    NC.ndpointer(dtype=numpy.float32, flags=["F", "W", "A"]), NC.ndpointer(dtype=numpy.ubyte, flags=["F", "W", "A"]),
    C.POINTER(_HPGL_SHAPE)]
_hpgl_so.hpgl_read_inc_file_float.restype = C.c_int
_hpgl_so.hpgl_read_inc_file_float.argtypes = [C.c_char_p, C.c_float, C.c_int,
                                              NC.ndpointer(dtype=numpy.float32, flags=["F", "W", "A"]),
                                              NC.ndpointer(dtype=numpy.ubyte, flags=["F", "W", "A"])]
_hpgl_so.hpgl_read_inc_file_byte.restype = C.c_int
_hpgl_so.hpgl_read_inc_file_byte.argtypes = [C.c_char_p, C.c_int, C.c_int,
                                             NC.ndpointer(dtype=numpy.ubyte, flags=["F", "W", "A"]),
                                             NC.ndpointer(dtype=numpy.ubyte, flags=["F", "W", "A"]),
                                             NC.ndpointer(dtype=numpy.ubyte), C.c_int]
_hpgl_so.hpgl_write_inc_file_float.restype = C.c_int
_hpgl_so.hpgl_write_inc_file_float.argtypes = [C.c_char_p, C.POINTER(_HPGL_CONT_MASKED_ARRAY), C.c_float, C.c_char_p]
_hpgl_so.hpgl_write_inc_file_byte.restype = C.c_int
_hpgl_so.hpgl_write_inc_file_byte.argtypes = [C.c_char_p, C.POINTER(_HPGL_IND_MASKED_ARRAY), C.c_int, C.c_char_p,
                                              C.POINTER(C.c_ubyte), C.c_int]
_hpgl_so.hpgl_write_gslib_cont_property.restype = C.c_int
_hpgl_so.hpgl_write_gslib_cont_property.argtypes = [C.POINTER(_HPGL_CONT_MASKED_ARRAY), C.c_char_p, C.c_char_p,
                                                    C.c_double]
_hpgl_so.hpgl_write_gslib_byte_property.restype = C.c_int
_hpgl_so.hpgl_write_gslib_byte_property.argtypes = [C.POINTER(_HPGL_IND_MASKED_ARRAY), C.c_char_p, C.c_char_p,
                                                    C.c_double, C.POINTER(C.c_ubyte), C.c_int]
_hpgl_so.hpglx_read_inc_float.restype = C.c_int
_hpgl_so.hpglx_read_inc_float.argtypes = [C.c_char_p, C.c_float, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
_hpgl_so.hpglx_read_inc_byte.restype = C.c_int
_hpgl_so.hpglx_read_inc_byte.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                         C.c_void_p]
_hpgl_so.hpglx_write_inc_float.restype = C.c_int
_hpgl_so.hpglx_write_inc_float.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_char_p,
                                           C.c_int, C.c_void_p]
_hpgl_so.hpglx_write_inc_byte.restype = C.c_int
_hpgl_so.hpglx_write_inc_byte.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_char_p,
                                          C.c_void_p, C.c_int, C.c_int, C.c_void_p]
_hpgl_so.hpgl_indicator_kriging.restype = None
_hpgl_so.hpgl_indicator_kriging.argtypes = [C.POINTER(_HPGL_IND_MASKED_ARRAY), C.POINTER(_HPGL_IND_MASKED_ARRAY),
                                            C.POINTER(_HPGL_IK_PARAMS), C.c_int]
_hpgl_so.hpglx_indicator_kriging.restype = C.c_int
_hpgl_so.hpglx_indicator_kriging.argtypes = [C.POINTER(_HPGL_IK_PARAMS), C.c_int, C.c_int * 3, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
_hpgl_so.hpgl_sgs_simulation.restype = None
_hpgl_so.hpgl_sgs_simulation.argtypes = [C.POINTER(_HPGL_CONT_MASKED_ARRAY), C.POINTER(_HPGL_SGS_PARAMS),
                                         C.POINTER(hpgl_non_parametric_cdf_t), C.POINTER(C.c_double),
                                         C.POINTER(_HPGL_UBYTE_ARRAY)]
_hpgl_so.hpgl_sgs_lvm_simulation.restype = None
_hpgl_so.hpgl_sgs_lvm_simulation.argtypes = [C.POINTER(_HPGL_CONT_MASKED_ARRAY), C.POINTER(_HPGL_SGS_PARAMS),
                                             C.POINTER(hpgl_non_parametric_cdf_t), C.POINTER(_HPGL_FLOAT_ARRAY),
                                             C.POINTER(_HPGL_UBYTE_ARRAY)]
_hpgl_so.hpglx_sgs_simulation.restype = C.c_int
_hpgl_so.hpglx_sgs_simulation.argtypes = [C.c_void_p, C.c_void_p, C.c_int * 3, C.POINTER(_HPGL_SGS_PARAMS),
                                          C.POINTER(hpgl_non_parametric_cdf_t), C.c_void_p, C.c_void_p,
                                          C.c_void_p]
_hpgl_so.hpglx_sgs_realizations.restype = C.c_int
_hpgl_so.hpglx_sgs_realizations.argtypes = [C.c_void_p, C.c_void_p, C.c_int * 3, C.POINTER(_HPGL_SGS_PARAMS),
                                            C.POINTER(C.c_long), C.c_int, C.POINTER(hpgl_non_parametric_cdf_t),
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
_hpgl_so.hpglx_sgs_schedule.restype = C.c_int
_hpgl_so.hpglx_sgs_schedule.argtypes = [C.c_int * 3, C.c_int * 3, C.c_void_p, C.c_void_p, C.c_long, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_int]
_hpgl_so.hpglx_template.restype = C.c_int
_hpgl_so.hpglx_template.argtypes = [C.POINTER(__hpgl_cov_params_t), C.c_int * 3, C.c_void_p, C.c_int]
_hpgl_so.hpglx_set_profiling.restype = None
_hpgl_so.hpglx_set_profiling.argtypes = [C.c_int]
_hpgl_so.hpglx_last_timing.restype = C.c_int
_hpgl_so.hpglx_last_timing.argtypes = [C.POINTER(C.c_float)]
_hpgl_so.hpglx_profile_summary.restype = C.c_int
_hpgl_so.hpglx_profile_summary.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_int)]
_hpgl_so.hpglx_last_call_stats.restype = C.c_int
_hpgl_so.hpglx_last_call_stats.argtypes = [C.POINTER(C.c_uint), C.POINTER(C.c_int)]
_hpgl_so.hpglx_set_default_precision.restype = None
_hpgl_so.hpglx_set_default_precision.argtypes = [C.c_int]
_hpgl_so.hpglx_host_introsort_perm.restype = None
_hpgl_so.hpglx_host_introsort_perm.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
_hpgl_so.hpglx_host_introsort_first.restype = None
_hpgl_so.hpglx_host_introsort_first.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
_hpgl_so.hpglx_knn.restype = C.c_int
_hpgl_so.hpglx_knn.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
_hpgl_so.hpglx_loo_state_size.restype = C.c_size_t
_hpgl_so.hpglx_loo_state_size.argtypes = [C.c_int, C.c_int, C.c_int]
_hpgl_so.hpglx_loo_forward.restype = C.c_int
_hpgl_so.hpglx_loo_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                       C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
_hpgl_so.hpglx_loo_backward.restype = C.c_int
_hpgl_so.hpglx_loo_backward.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                        C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
_hpgl_so.hpglx_last_error.restype = C.c_char_p
_hpgl_so.hpglx_last_error.argtypes = []
_hpgl_so.hpglx_version.restype = C.c_char_p
_hpgl_so.hpglx_version.argtypes = []
_hpgl_so.hpglx_unpack_offset.restype = None
_hpgl_so.hpglx_unpack_offset.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]


def last_error():
    msg = _hpgl_so.hpgl_get_last_exception_message()
    return msg.decode("utf-8", errors="replace") if msg else ""
