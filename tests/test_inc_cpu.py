"""Pins the oracle's INC/GSLIB reader and writer: against the reference's
own read_inc_file.cpp / property_writer.cpp (oracle/_ref), and the decimal
to float rounding against exact rational arithmetic."""
import os
from fractions import Fraction

import numpy as np
import pytest

import inc_cases

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "inc", "NEW_TEST_PROP_01.INC")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _count_values(text):
    """Values of a well-formed file: tokens after the name line, outside
    comment lines, before '/'."""
    lines = text.decode().split("\n")
    i = 0
    while not lines[i][:1].isalpha():
        i += 1
    n = 0
    for line in lines[i + 1:]:
        for tok in line.split():
            if tok.startswith("--"):
                break
            if tok.startswith("/"):
                return n
            n += 1
    return n


def test_reader_fixture_matches_reference(ref):
    text = open(FIXTURE, "rb").read()
    n = _count_values(text)
    assert n > 1000
    d, m = ref.read_inc(text, n)
    rd, rm = ref.ref_read_inc(FIXTURE, n)
    assert np.array_equal(_bits(d), _bits(rd)) and np.array_equal(m, rm)
    db, mb = ref.read_inc(text, n, as_int=True)
    rdb, rmb = ref.ref_read_inc(FIXTURE, n, as_int=True)
    assert np.array_equal(db, rdb) and np.array_equal(mb, rmb)


def test_reader_mixed_tokens_match_reference(ref, tmp_path):
    text, n = inc_cases.mixed_file(n_values=6000, n_midpoints=800)
    path = tmp_path / "mixed.inc"
    path.write_bytes(text)
    d, m = ref.read_inc(text, n)
    rd, rm = ref.ref_read_inc(path, n)
    assert np.array_equal(_bits(d), _bits(rd)) and np.array_equal(m, rm)


def test_reader_bytes_match_reference(ref, tmp_path):
    text, n = inc_cases.byte_file(n_values=3000)
    path = tmp_path / "bytes.inc"
    path.write_bytes(text)
    d, m = ref.read_inc(text, n, as_int=True)
    rd, rm = ref.ref_read_inc(path, n, as_int=True)
    assert np.array_equal(d, rd) and np.array_equal(m, rm)


@pytest.mark.parametrize("body,size,ok", [
    (b"1 2 3 /\n", 3, True),
    (b"1 2 /\n", 3, False),                  # '/' before size values
    (b"1 abc 3 /\n", 3, False),              # sscanf fails
    (b"1 2 -- 9 9 9\n3 /\n", 3, True),       # mid-line comment
    (b"1 2 3 4 5", 8, True),                 # end of file: the last token repeats
    (b"1 2 3 x", 3, True),                   # tokens past size are not read
])
def test_reader_edge_cases_match_reference(ref, tmp_path, body, size, ok):
    text = b"-- header\nNAME\n" + body
    path = tmp_path / "edge.inc"
    path.write_bytes(text)
    if ok:
        d, m = ref.read_inc(text, size)
        rd, rm = ref.ref_read_inc(path, size)
        assert np.array_equal(_bits(d), _bits(rd)) and np.array_equal(m, rm)
    else:
        with pytest.raises(ValueError):
            ref.read_inc(text, size)
        with pytest.raises(ValueError):
            ref.ref_read_inc(path, size)


@pytest.mark.parametrize("tok", inc_cases.BAD_FLOAT_TOKENS)
def test_reader_rejects_what_sscanf_rejects(ref, tmp_path, tok):
    text = b"NAME\n1 " + tok.encode() + b" 3 /\n"
    path = tmp_path / "bad.inc"
    path.write_bytes(text)
    with pytest.raises(ValueError, match="Error parsing"):
        ref.read_inc(text, 3)
    with pytest.raises(ValueError):
        ref.ref_read_inc(path, 3)


def test_reader_eof_repeats_last_value(orc):
    d, m = orc.read_inc(b"P\n1 2 3.5\n", 6)
    assert list(d) == [1, 2, 3.5, 3.5, 3.5, 3.5]


def test_reader_messages(orc):
    with pytest.raises(ValueError, match="Unexpected end of data"):
        orc.read_inc(b"P\n1 /\n", 2)
    with pytest.raises(ValueError, match="Error parsing 'zz' string"):
        orc.read_inc(b"P\n1 zz 3\n", 3)
    with pytest.raises(ValueError, match="Property name not found"):
        orc.read_inc(b"-- only comments\n12 13\n", 2)


def _round_to_f32(q):
    """Exact round-to-nearest-even of a positive rational to float32 (by
    search over the float32 ordering)."""
    if q == 0:
        return np.float32(0)
    lo, hi = 0, 0x7F800000  # bit patterns; value(hi) = inf
    while hi - lo > 1:
        mid = (lo + hi) // 2
        v = np.array([mid], np.uint32).view(np.float32)[0]
        if Fraction(float(v)) <= q:
            lo = mid
        else:
            hi = mid
    a = Fraction(float(np.array([lo], np.uint32).view(np.float32)[0]))
    if hi == 0x7F800000:
        b = Fraction(2) ** 128
    else:
        b = Fraction(float(np.array([hi], np.uint32).view(np.float32)[0]))
    if q - a < b - q or (q - a == b - q and lo % 2 == 0):
        return np.array([lo], np.uint32).view(np.float32)[0]
    return np.array([hi], np.uint32).view(np.float32)[0]


def test_midpoint_tokens_round_exactly(orc):
    """Decimal midpoints between adjacent floats (and nudged by one unit in
    their 60th digit) against exact rational rounding."""
    rng = np.random.default_rng(3)
    toks = inc_cases.midpoint_tokens(rng, 300) + ["1.000000059604644775390625", "1.0000000596046447753906251",
                                                   "3.4028235677973366e38", "7.0064923216240854e-46",
                                                   "7.0064923216240862e-46", "1.4012984643248171e-45"]
    text = ("P\n" + "\n".join(toks) + "\n/\n").encode()
    d, _ = orc.read_inc(text, len(toks))
    for t, v in zip(toks, d):
        from decimal import Decimal
        exp = _round_to_f32(Fraction(Decimal(t)))
        assert _bits([v])[0] == _bits([exp])[0], t


@pytest.mark.parametrize("gslib", [False, True])
def test_writer_float_matches_reference(ref, tmp_path, gslib):
    rng = np.random.default_rng(4)
    n = 3000
    f = (rng.standard_normal(n) * 10.0 ** rng.integers(-40, 38, n)).astype(np.float32)
    f[:12] = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.4e-45, -1.4e-45, 3.4028235e38, 9999999.5, 0.99999995,
              1e-5, 123456789.0]
    m = (rng.random(n) > 0.2).astype(np.uint8)
    path = tmp_path / "w.txt"
    ref.ref_write(path, "PROP", m, floats=f, undefined=-99.0, gslib=gslib)
    body = ref.format_values(m, floats=f, undefined=-99.0)
    head = b"HPGL saved GSLIB file\n1\nPROP\n" if gslib else b"PROP\n"
    tail = b"" if gslib else b"/\n"
    assert path.read_bytes() == head + body + tail


@pytest.mark.parametrize("gslib", [False, True])
def test_writer_bytes_match_reference(ref, tmp_path, gslib):
    rng = np.random.default_rng(5)
    n = 2000
    b = rng.integers(0, 4, n).astype(np.uint8)
    m = (rng.random(n) > 0.3).astype(np.uint8)
    remap = np.array([0, 7, 130, 255], np.uint8)
    path = tmp_path / "wb.txt"
    ref.ref_write(path, "FACIES", m, bytes_=b, undefined=-99, remap=remap, indicator_count=4, gslib=gslib)
    body = ref.format_values(m, bytes_=b, undefined=-99, remap=remap)
    head = b"HPGL saved GSLIB file\n1\nFACIES\n" if gslib else b"FACIES\n"
    tail = b"" if gslib else b"/\n"
    assert path.read_bytes() == head + body + tail


def test_get_gslib_property_marks_undefined():
    """geo.py:937-945: informed where the value differs from the undefined
    marker; the column itself is returned as given."""
    import numpy as np
    from geo_bsd.geo import get_gslib_property
    col = np.array([1.5, -99.0, 0.0, -99.0, 7.0], np.float32)
    vals, informed = get_gslib_property({"poro": col, "perm": col * 2}, "poro", -99.0)
    assert vals is col
    assert informed.dtype == np.uint8 and informed.tolist() == [1, 0, 1, 0, 1]
