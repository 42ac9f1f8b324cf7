"""Writes tests/golden/sk_problem_7_2.json from the reference's stored output.

Source (reference repository, read as text):
  solved_problems_book/7.2. Conditioning by Kriging/transect.txt        (inputs)
  solved_problems_book/7.2. Conditioning by Kriging/problem_7_2.py:28-75 (setup)
  solved_problems_book/7.2. Conditioning by Kriging/Result/problem_7.2.txt (outputs)

Two simple-kriging runs share a 20x1x1 grid, an exponential model with ranges
(5,1,1) and sill 1, radiuses (5,1,1), max_neighbours 2 and the automatic mean:
  * "SK result": hard data = the transect values;
  * "Simple Kriging result": hard data = the "SGS result" values at the same
    cells.
Cell of a sample at x: int((x - x0) / dx) with x0 = min(x) - 5 and
dx = 1 + (max(x) - min(x)) / 20 (grid_3d.Grid.get_ijk).
Printed numpy float32 values carry 8 decimals; the expected arrays are those
decimals.
"""
import json
import os
import re
import sys

import numpy as np

REF = "/root/reference/solved_problems_book/7.2. Conditioning by Kriging"


def parse_arrays(text):
    out = {}
    for m in re.finditer(r"([A-Za-z .]+):\s*\[([^\]]*)\]", text):
        out[m.group(1).strip()] = [float(v) for v in m.group(2).split()]
    return out


def main():
    lines = open(os.path.join(REF, "transect.txt")).read().split("\n")
    ncols = int(lines[1].split()[0])
    rows = [l.split() for l in lines[2 + ncols:] if l.strip()]
    xs = [float(r[0]) for r in rows]
    vals = [float(r[1]) for r in rows]
    i_max = 20
    x0 = min(xs) - 5
    dx = 1 + (max(xs) - min(xs)) / i_max
    cells = [int((x - x0) / dx) for x in xs]
    res = parse_arrays(open(os.path.join(REF, "Result", "problem_7.2.txt")).read())
    sgs = res["SGS result"]
    runs = [
        dict(name="SK result", data={c: v for c, v in zip(cells, vals)}, expected=res["SK result"]),
        dict(name="Simple Kriging result", data={c: sgs[c] for c in cells}, expected=res["Simple Kriging result"]),
    ]
    doc = dict(
        source="solved_problems_book/7.2. Conditioning by Kriging (Result/problem_7.2.txt)",
        grid=[i_max, 1, 1], cov=dict(type=1, ranges=[5.0, 1.0, 1.0], angles=[0.0, 0.0, 0.0], sill=1.0, nugget=0.0),
        radiuses=[5, 1, 1], max_neighbours=2, undefined=-99.0, printed_decimals=8,
        runs=[dict(name=r["name"], cells=sorted(r["data"]), values=[r["data"][c] for c in sorted(r["data"])],
                   expected=r["expected"]) for r in runs])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sk_problem_7_2.json")
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    sys.exit(main())
