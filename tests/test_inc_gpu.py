"""The GPU INC/GSLIB reader and writer against the oracle (which the CPU
tests pin to the reference's own reader and writer): parsed values bit for
bit, masks, error messages, and written files byte for byte."""
import os

import numpy as np
import pytest

import inc_cases

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "inc", "NEW_TEST_PROP_01.INC")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _read(path, n, as_int=False, values=None):
    import geo_bsd
    print(f"[gpu] read {os.path.basename(str(path))} n={n} int={as_int}", flush=True)
    if as_int:
        p = geo_bsd.read_inc_file_byte(str(path), -99, n, values or [])
    else:
        p = geo_bsd.read_inc_file_float(str(path), -99, n)
    return p.data.reshape(-1, order="F"), p.mask.reshape(-1, order="F")


def test_read_reference_fixture(orc):
    text = open(FIXTURE, "rb").read()
    n = 1200
    d, m = _read(FIXTURE, n)
    od, om = orc.read_inc(text, n)
    assert np.array_equal(_bits(d), _bits(od)) and np.array_equal(m, om)


@pytest.mark.parametrize("n_values", [20000, 300000])
def test_read_mixed_tokens_bitwise(orc, tmp_path, n_values):
    text, n = inc_cases.mixed_file(n_values=n_values, n_midpoints=n_values // 10)
    path = tmp_path / "mixed.inc"
    path.write_bytes(text)
    d, m = _read(path, n)
    od, om = orc.read_inc(text, n)
    bad = np.nonzero(_bits(d) != _bits(od))[0]
    assert bad.size == 0, [(int(i), float(d[i]), float(od[i])) for i in bad[:5]]
    assert np.array_equal(m, om)


def test_read_long_lines_and_comments_across_chunks(orc, tmp_path):
    """Comment state and tokens carried across 32-byte lanes and 8 KB blocks:
    a 20 KB comment line full of number-like text, a 60 KB data line, tokens
    straddling every boundary."""
    rng = np.random.default_rng(8)
    toks = inc_cases.value_tokens(rng, 12000)
    lines = ["P", "-- " + " ".join(toks[:3000]), " ".join(toks[3000:9000]),
             "1 2 -- " + " ".join(toks[:2000]), " \t ".join(toks[9000:]), "/"]
    text = ("\n".join(lines) + "\n").encode()
    n = 6000 + 2 + 3000
    path = tmp_path / "long.inc"
    path.write_bytes(text)
    d, m = _read(path, n)
    od, om = orc.read_inc(text, n)
    assert np.array_equal(_bits(d), _bits(od)) and np.array_equal(m, om)


def test_read_bytes_with_remap(orc, tmp_path):
    """Integer tokens incl. out-of-range and overflowing ones, through the
    device entry point (IndProperty would reject values outside the list)."""
    import torch
    from geo_bsd import device as D
    text, n = inc_cases.byte_file(n_values=20000)
    path = tmp_path / "bytes.inc"
    path.write_bytes(text)
    values = [0, 1, 2, 5, 7]
    print("[gpu] read bytes (device)", flush=True)
    d, m = D.read_inc_byte(str(path), -99, n, values)
    torch.cuda.synchronize()
    od, om = orc.read_inc(text, n, as_int=True)
    od = orc.remap_bytes(od, om, values)
    assert np.array_equal(d.cpu().numpy(), od) and np.array_equal(m.cpu().numpy(), om)


def test_geo_read_bytes_clean_file(orc, tmp_path):
    rng = np.random.default_rng(13)
    vals = [3, 8, 11]
    toks = [str(int(v)) for v in rng.choice(vals + [-99], size=4000)]
    text = ("FACIES\n" + " ".join(toks) + "\n/\n").encode()
    path = tmp_path / "facies.inc"
    path.write_bytes(text)
    d, m = _read(path, len(toks), as_int=True, values=vals)
    od, om = orc.read_inc(text, len(toks), as_int=True)
    assert np.array_equal(d, orc.remap_bytes(od, om, vals)) and np.array_equal(m, om)


@pytest.mark.parametrize("tok", inc_cases.BAD_FLOAT_TOKENS)
def test_read_rejects_bad_tokens(tmp_path, tok):
    path = tmp_path / "bad.inc"
    path.write_bytes(b"NAME\n1 " + tok.encode() + b" 3 /\n")
    with pytest.raises(RuntimeError, match="Error parsing '" + __import__("re").escape(tok) + "' string"):
        _read(path, 3)


def test_read_errors_and_end_of_file(tmp_path):
    path = tmp_path / "e.inc"
    path.write_bytes(b"-- c\nNAME\n1 2 /\n")
    with pytest.raises(RuntimeError, match="Unexpected end of data"):
        _read(path, 3)
    _read(path, 2)
    path.write_bytes(b"-- c\n12\n")
    with pytest.raises(RuntimeError, match="Property name not found"):
        _read(path, 1)
    path.write_bytes(b"NAME\n1 2 3.5")
    d, m = _read(path, 6)
    assert list(d) == [1, 2, 3.5, 3.5, 3.5, 3.5] and m.all()
    path.write_bytes(b"NAME\n1 2 -- x\n")
    with pytest.raises(RuntimeError, match="end of file"):
        _read(path, 4)
    with pytest.raises(Exception):
        _read(tmp_path / "missing.inc", 4)


def _special_floats(rng, n):
    f = (rng.standard_normal(n) * 10.0 ** rng.integers(-45, 39, n)).astype(np.float32)
    f[:16] = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.4e-45, -1.4e-45, 3.4028235e38, 9999999.5, 0.99999995, 1e-5,
              123456789.0, 1e-16, 1e-17, 1e22, 1e29]
    bits = rng.integers(0, 0xFFFFFFFF, n // 4, dtype=np.uint64).astype(np.uint32).view(np.float32)
    f[16:16 + bits.size] = bits
    return f


@pytest.mark.parametrize("gslib", [False, True])
def test_write_float_bytes_identical(orc, tmp_path, gslib):
    import geo_bsd
    rng = np.random.default_rng(9)
    n = 50000
    f = _special_floats(rng, n)
    m = (rng.random(n) > 0.1).astype(np.uint8)
    path = tmp_path / "w.txt"
    print(f"[gpu] write float gslib={gslib}", flush=True)
    writer = geo_bsd.write_gslib_property if gslib else geo_bsd.write_property
    writer(geo_bsd.ContProperty(f, m), str(path), "PROP", -99.0)
    head = b"HPGL saved GSLIB file\n1\nPROP\n" if gslib else b"PROP\n"
    tail = b"" if gslib else b"/\n"
    exp = head + orc.format_values(m, floats=f, undefined=-99.0) + tail
    got = path.read_bytes()
    if got != exp:
        gl, el = got.split(b"\n"), exp.split(b"\n")
        diff = [(i, gl[i], el[i]) for i in range(min(len(gl), len(el))) if gl[i] != el[i]][:5]
        raise AssertionError(diff)


@pytest.mark.parametrize("gslib", [False, True])
def test_write_bytes_identical(orc, tmp_path, gslib):
    import geo_bsd
    rng = np.random.default_rng(10)
    n = 20000
    b = rng.integers(0, 4, n).astype(np.uint8)
    m = (rng.random(n) > 0.3).astype(np.uint8)
    values = [0, 7, 130, 255]
    path = tmp_path / "wb.txt"
    print(f"[gpu] write bytes gslib={gslib}", flush=True)
    prop = geo_bsd.IndProperty(b, m, 4)
    writer = geo_bsd.write_gslib_property if gslib else geo_bsd.write_property
    writer(prop, str(path), "FACIES", -99, values)
    head = b"HPGL saved GSLIB file\n1\nFACIES\n" if gslib else b"FACIES\n"
    tail = b"" if gslib else b"/\n"
    exp = head + orc.format_values(m, bytes_=b, undefined=-99, remap=np.array(values, np.uint8)) + tail
    assert path.read_bytes() == exp


def test_device_round_trip(tmp_path):
    """Device tensors -> file -> device tensors: the 7 printed digits read
    back as the float nearest to them."""
    import torch
    from geo_bsd import device as D
    rng = np.random.default_rng(11)
    n = 100000
    f = (rng.standard_normal(n) * 3).astype(np.float32)
    m = (rng.random(n) > 0.2).astype(np.uint8)
    dev = torch.device("cuda:0")
    path = tmp_path / "rt.inc"
    D.write_inc(str(path), torch.from_numpy(f).to(dev), torch.from_numpy(m).to(dev), -99.0, "RT")
    d, mm = D.read_inc_float(str(path), -99.0, n)
    torch.cuda.synchronize()
    d = d.cpu().numpy()
    exp = np.array(["%E" % v for v in np.where(m == 1, f, np.float32(-99.0))], dtype=np.float32)
    assert np.array_equal(d, exp) and np.array_equal(mm.cpu().numpy(), m)


def test_slow_loader_agrees_on_a_clean_file(tmp_path):
    import geo_bsd
    rng = np.random.default_rng(12)
    toks = ["%E" % float(np.float32(v)) for v in rng.standard_normal(3000)] + ["-99"] * 10
    path = tmp_path / "clean.inc"
    path.write_bytes(("-- c\nP\n" + "\n".join(toks) + "\n/\n").encode())
    fast = geo_bsd.load_cont_property(str(path), -99, len(toks))
    slow = geo_bsd.load_cont_property(str(path), -99)
    assert np.array_equal(fast.data.reshape(-1), slow.data.reshape(-1))
    assert np.array_equal(fast.mask.reshape(-1), slow.mask.reshape(-1))
