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


def test_roofline_reads_the_committed_per_class_traffic():
    """The bench's roofline pairs each degree class's launch with the committed rocprofv3 traffic of the same
    kernel, and picks the class with the largest device time as the headline figure."""
    traffic = bench.committed_traffic(bench.TRAFFIC_FILE)
    assert sorted(traffic) == [0, 1, 2, 3, 4, 5]
    from sbp_amd.autotune import tuned_shape
    shape = tuned_shape()  # the launch shape the committed traffic was measured with
    names = bench.class_kernels(shape, True, B=5000)
    assert [traffic[b][0] for b in range(6)] == names
    tot = {"ft_class_launches": [2, 2, 2, 2, 2, 0], "ft_class_ms": [6.0, 8.0, 4.0, 2.0, 3.0, 0.0],
           "ft_class_bytes": [6e9, 8e10, 4e9, 2e9, 3e9, 0]}
    r = bench.roofline(tot, shape, traffic_file=bench.TRAFFIC_FILE, B=5000)
    assert r["kernel"] == names[1] and r["degree_class"] == "8 < k <= 32"
    assert abs(r["achieved"] - 4e10 / 4e-3 / 1e9) < 1e-6 and abs(r["frac"] - r["achieved"] / 8000.0) < 1e-12
    assert r["traffic"] == traffic[1][1] and abs(r["hbm_achieved"] - traffic[1][1] / 4e-3 / 1e9) < 1e-6
    assert set(r["classes"]) == {"0", "1", "2", "3", "4"}
    # another shape names other kernels: the committed traffic no longer applies to them
    r2 = bench.roofline(tot, ([16, 16, 32, 64, 128], [4, 4, 4, 4]), traffic_file=bench.TRAFFIC_FILE, B=5000)
    assert r2["traffic"] is None and r2["hbm_frac"] is None
