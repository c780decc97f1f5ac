"""The benchmark's NMI (src/evaluate.cpp:63-98: MI / sqrt(H1 H2)) on small partitions, including the one-block
case where the reference's value is 0/0 and the benchmark reports null."""
import math

import numpy as np

import bench


def test_nmi_matches_the_closed_form():
    a = np.array([0, 0, 1, 1, 2, 2])
    b = np.array([0, 0, 1, 1, 1, 1])
    # H(a) = ln 3, H(b) = -(1/3 ln 1/3 + 2/3 ln 2/3), MI = H(b) since b is a coarsening of a
    hb = -(1 / 3 * math.log(1 / 3) + 2 / 3 * math.log(2 / 3))
    assert abs(bench.nmi(a, b) - hb / math.sqrt(math.log(3) * hb)) < 1e-12
    assert abs(bench.nmi(a, a) - 1.0) < 1e-12
    assert abs(bench.nmi(a, np.array([5, 5, 7, 7, 9, 9])) - 1.0) < 1e-12  # labels do not matter


def test_nmi_of_one_block_is_null():
    assert bench.nmi(np.zeros(10), np.arange(10) % 3) is None
    assert bench.nmi(np.arange(10) % 3, np.zeros(10)) is None
    assert bench.mean_nmi([None, 0.5, 1.0]) == 0.75
    assert bench.mean_nmi([None]) is None
