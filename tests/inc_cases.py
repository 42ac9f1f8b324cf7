"""Synthetic INC property files for the reader and writer tests.

Token mix: the reference writer's "%E" form, short and long decimals,
integers, signs, exponents across the float range, subnormals, exact decimal
midpoints between adjacent floats (and those midpoints nudged by one unit in
their 60th digit), hex floats, inf/nan spellings, tokens with trailing text
that sscanf ignores, "--" comments at line starts and mid-line, a header with
a line longer than 1023 characters, and a closing '/'.
"""
from decimal import Decimal, getcontext

import numpy as np

getcontext().prec = 200


def _exact(f32):
    """Exact decimal expansion of a float32 value (as a Decimal)."""
    return Decimal(float(f32))


def midpoint_tokens(rng, count):
    out = []
    bits = rng.integers(1, 0x7F7FFFFF, size=count, dtype=np.int64).astype(np.uint32)
    for b in bits:
        lo = np.array([b], np.uint32).view(np.float32)[0]
        hi = np.array([b + 1], np.uint32).view(np.float32)[0]
        if not np.isfinite(hi):
            continue
        mid = (_exact(lo) + _exact(hi)) / 2
        s = format(mid, "f") if abs(mid.adjusted()) < 30 else format(mid, "E")
        nudge = Decimal(1).scaleb(mid.adjusted() - 60)
        choice = rng.integers(0, 3)
        if choice == 1:
            s = format(mid + nudge, "E")
        elif choice == 2:
            s = format(mid - nudge, "E")
        out.append(s)
    return out


def value_tokens(rng, count):
    toks = []
    for _ in range(count):
        k = rng.integers(0, 14)
        x = np.float32(rng.standard_normal() * 10.0 ** rng.integers(-8, 9))
        if k == 0:
            toks.append("%E" % float(x))
        elif k == 1:
            toks.append("%.9g" % float(x))
        elif k == 2:
            toks.append(str(int(rng.integers(-200, 200))))
        elif k == 3:
            toks.append("%.17e" % (float(x) * 1.0000001))
        elif k == 4:
            e = int(rng.integers(-46, 39))
            toks.append("%de%d" % (int(rng.integers(1, 99999)), e - 4))
        elif k == 5:
            toks.append("".join(str(int(d)) for d in rng.integers(0, 10, size=int(rng.integers(20, 60)))) +
                        "e-%d" % int(rng.integers(0, 80)))
        elif k == 6:
            toks.append(rng.choice(["1.4e-45", "7e-46", "7.006e-46", "1e-45", "3.4028235e38", "3.4028236e38",
                                    "3.40282357e38", "3.4028236e+38", "1e39", "1e-50", "0.0", "-0", "+0.000",
                                    "1.17549435e-38", "1.1754942e-38", "9.99999999999999999999e-1"]))
        elif k == 7:
            toks.append(rng.choice(["0x1.8p3", "-0x1p-149", "0x1.fffffep127", "0X10", "0x.8", "0x1p-150",
                                    "0x1.000001p0", "0x1.0000018p0", "0x.", "0x1p", "0xffffffffffffffffffp0"]))
        elif k == 8:
            toks.append(rng.choice(["inf", "-INF", "Infinity", "nan", "-NaN", "+inf", "infx", "nan(12)", "NaN(",
                                    "infinityy"]))
        elif k == 9:
            toks.append(rng.choice(["1.5abc", "2e", "3e+", "-.5", "+.5e1", "12,5", "1.2.3", "4e2x", "00012", ".25",
                                    "0.e5", "5.", "1e99999", "1e-99999"]))
        elif k == 10:
            toks.append("%.6f" % float(x))
        elif k == 11:
            toks.append("%E" % -99.0)
        elif k == 12:
            toks.append("-99")
        else:
            toks.append("%.1f" % float(x))
    return toks


def make_text(rng, tokens, name="PROP", slash=True, comments=True, header_long_line=True):
    """Lays tokens out on lines of 1..8 with comments between; returns bytes."""
    lines = ["-- synthetic property for the INC reader tests", "--"]
    if header_long_line:
        lines.append("-- " + "x" * 1500)
    lines += ["", name + "  ", ""]
    i = 0
    while i < len(tokens):
        k = int(rng.integers(1, 9))
        row = tokens[i:i + k]
        i += k
        sep = [" ", "\t", "   ", " \t "][int(rng.integers(0, 4))]
        line = sep.join(row)
        if comments and rng.random() < 0.08:
            line += " -- trailing comment 1 2 3"
        lines.append(line)
        if comments and rng.random() < 0.03:
            lines.append("-- a comment line 4 5 6")
    if slash:
        lines.append("/")
    return ("\n".join(lines) + "\n").encode()


def mixed_file(seed=11, n_values=20000, n_midpoints=2000, **kw):
    rng = np.random.default_rng(seed)
    toks = value_tokens(rng, n_values - n_midpoints) + midpoint_tokens(rng, n_midpoints)
    order = rng.permutation(len(toks))
    toks = [toks[j] for j in order]
    return make_text(rng, toks, **kw), len(toks)


def byte_file(seed=12, n_values=5000, values=(0, 1, 2, 5, 7), undefined=-99):
    rng = np.random.default_rng(seed)
    toks = []
    for _ in range(n_values):
        r = rng.random()
        if r < 0.2:
            toks.append(str(undefined))
        elif r < 0.95:
            toks.append(str(int(rng.choice(values))))
        else:
            toks.append(rng.choice(["300", "-1", "+7", "0007", "2147483648", "99999999999999999999", "5x", "-129"]))
    return make_text(rng, toks), len(toks)


# tokens sscanf("%f") rejects: each makes a file fail to read
BAD_FLOAT_TOKENS = ["0x", "0xg", ".", "-.", "in", "infinit", "na", "-", "+", "e5", ".e5", "+-1", "abc"]
