"""The C ABI library without a GPU: loading, exported symbols, struct
layouts, host-side plan (search template) and parameter validation."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in ("hpgl.h", "hpglx.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(hpglx?_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from geo_bsd import hpgl_wrap
    lib = C.CDLL(hpgl_wrap.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_reference_abi():
    """Sizes of the reference's POD structs on x86-64 (api.h:18-153)."""
    from geo_bsd import hpgl_wrap as W
    assert C.sizeof(W._HPGL_SHAPE) == 24
    assert C.sizeof(W._HPGL_CONT_MASKED_ARRAY) == 40
    assert C.sizeof(W._HPGL_IND_MASKED_ARRAY) == 48
    assert C.sizeof(W._HPGL_OK_PARAMS) == 88 and W._HPGL_OK_PARAMS.max_neighbours.offset == 84
    assert C.sizeof(W._HPGL_SK_PARAMS) == 104 and W._HPGL_SK_PARAMS.mean.offset == 96
    assert W._HPGL_SK_PARAMS.automatic_mean.offset == 88
    assert C.sizeof(W._HPGL_MEDIAN_IK_PARAMS) == 104 and W._HPGL_MEDIAN_IK_PARAMS.marginal_probs.offset == 88
    assert C.sizeof(W._HPGL_COV_PARAMS) == 72


def test_request_struct_matches_header(tmp_path):
    """hpglx_grid_request_t as compiled by the C compiler vs the ctypes mirror."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    from geo_bsd import hpgl_wrap as W
    src = tmp_path / "layout.c"
    fields = [f for f, _ in W._HPGLX_GRID_REQUEST._fields_]
    body = "\n".join(f'printf("{f} %zu\\n", offsetof(hpglx_grid_request_t, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "hpglx.h"\nint main(void){\n'
                   f'printf("size %zu\\n", sizeof(hpglx_grid_request_t));\n{body}\nreturn 0;}}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    assert int(out["size"]) == C.sizeof(W._HPGLX_GRID_REQUEST)
    for f in fields:
        assert int(out[f]) == getattr(W._HPGLX_GRID_REQUEST, f).offset, f


@pytest.mark.parametrize("wl_name", ["sk", "ok", "mik", "wells"])
def test_library_template_equals_oracle(orc, wl_name):
    from geo_bsd import CovarianceModel, device as D, synthetic as S
    wl = {"sk": S.sk_c1, "ok": S.ok_c2, "mik": S.mik_c3, "wells": S.ok_wells_c4}[wl_name](0.25)
    c = wl.cov
    t_lib = D.search_template(CovarianceModel(**c), wl.radiuses)
    t_orc, _ = orc.template(orc.make_cov(c["type"], c["ranges"], c["angles"], c["sill"], c["nugget"]),
                            wl.radiuses)
    assert np.array_equal(t_lib, t_orc)


def test_library_template_rejects_zero_range():
    from geo_bsd import CovarianceModel, device as D
    with pytest.raises(RuntimeError):
        D.search_template(CovarianceModel(type=0, ranges=(0, 1, 1)), (2, 2, 2))


def test_validation_matches_reference_rules():
    import geo_bsd
    from geo_bsd.validation import CriticalValidationError, ValidationWarning, ParameterValidator
    with pytest.raises(CriticalValidationError):
        geo_bsd.CovarianceModel(type=0, ranges=(1, 1, 1), sill=1.0, nugget=2.0)
    with pytest.raises(CriticalValidationError):
        geo_bsd.SugarboxGrid(0, 5, 5)
    with pytest.raises(CriticalValidationError):
        ParameterValidator.validate_radius((1, 2), "radiuses")
    with pytest.raises(CriticalValidationError):
        ParameterValidator.validate_max_neighbors(0)
    with pytest.raises(ValidationWarning):
        ParameterValidator.validate_max_neighbors(5000)
    # zero ranges pass validation in the reference (validation.py:468-472)
    geo_bsd.CovarianceModel(type=0, ranges=(0, 0, 0))
    assert ParameterValidator.validate_radius((3, 4, 5)) == (3, 4, 5)


def test_ind_property_checks_indicator_range():
    import geo_bsd
    with pytest.raises(RuntimeError):
        geo_bsd.IndProperty(np.array([0, 1, 2], np.uint8), np.array([1, 1, 1], np.uint8), 2)


def test_thread_num_roundtrip():
    import geo_bsd
    geo_bsd.set_thread_num(3)
    assert geo_bsd.get_thread_num() == 3


def test_empty_clone_keeps_kind_and_indicator_count():
    """geo.py:360-370 (imported from geo_bsd.geo by the reference's tests)."""
    import numpy as np
    from geo_bsd.geo import ContProperty, IndProperty, _empty_clone
    ind = IndProperty(np.array([0, 2, 1, 2], np.uint8), np.ones(4, np.uint8), 5)
    c = _empty_clone(ind)
    assert isinstance(c, IndProperty) and c.indicator_count == 5
    assert not c.data.any() and not c.mask.any()
    cont = ContProperty(np.array([1.5, -2.0], np.float32), np.array([1, 0], np.uint8))
    c = _empty_clone(cont)
    assert isinstance(c, ContProperty) and c.data.dtype == np.float32
    assert not c.data.any() and not c.mask.any() and cont.data[0] == 1.5


def test_reference_array_helpers():
    """geo.py:191-214, 382-390 helpers the reference's tests import."""
    import numpy as np
    from geo_bsd.geo import (SugarboxGrid, _create_hpgl_float_array, _create_hpgl_ubyte_array, _require_cont_data,
                             _requite_ind_data)
    a = _require_cont_data(np.arange(24, dtype=np.float64).reshape(2, 3, 4))
    assert a.dtype == np.float32 and a.flags["F"]
    b = _requite_ind_data(np.ones((2, 3, 4), np.int64))
    assert b.dtype == np.uint8 and b.flags["F"]
    assert _require_cont_data(None) is None and _requite_ind_data(None) is None
    fa = _create_hpgl_float_array(a, None)
    assert list(fa.shape.data) == [2, 3, 4] and fa.data[5] == a.reshape(-1, order="F")[5]
    ua = _create_hpgl_ubyte_array(b, SugarboxGrid(2, 3, 4))
    assert list(ua.shape.data) == [2, 3, 4] and ua.data[0] == 1
