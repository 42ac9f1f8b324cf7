"""Non-parametric CDF tables for the normal-score transform of sequential
Gaussian simulation (the reference's cdf.py)."""
import numpy


class CdfData:
    """cdf.py:3-6: ascending values and their cumulative probabilities."""

    def __init__(self, values, probs):
        self.values = numpy.require(values, "float32")
        self.probs = numpy.require(probs, "float32")


def calc_cdf(prop):
    """cdf.py:8-37: the distinct informed values in ascending order, the last
    one dropped (when there are two or more), each with the running sum of
    the value counts over the informed count, accumulated in double."""
    data = numpy.asarray(prop.data)
    mask = numpy.asarray(prop.mask)
    vals = data[mask != 0]
    if vals.size == 0:
        raise ValueError("calc_cdf: no informed values (all cells are masked)")
    values, counts = numpy.unique(vals, return_counts=True)
    size = 1 if values.size == 1 else values.size - 1
    values = values[:size]
    probs = numpy.add.accumulate(counts[:size] / float(vals.size))
    return CdfData(values=values, probs=probs)
