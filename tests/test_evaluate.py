"""Partition evaluation (evaluate::hungarian, calculate_f1_score, calculate_nmi; src/evaluate.cpp:9-210) in
the library's host code, against the reference's known answers (test/evaluate_test.cpp: the F1 scores of two
contingency tables, and the Rand index, recall and precision of the first), scikit-learn's NMI, SciPy's
maximum-weight assignment and permutation invariance.  No device is used."""
import json
import math
import os

import numpy as np
import pytest
import scipy.optimize
from sklearn.metrics import normalized_mutual_info_score

from sbp_amd import native

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "evaluate_tables.json")))


def float_eq(a, b, ulps=4):
    """gtest EXPECT_FLOAT_EQ: within 4 single-precision ULPs."""
    return abs(int(np.float32(a).view(np.int32)) - int(np.float32(b).view(np.int32))) <= ulps


def test_f1_known_answers():
    e = native.evaluate_table(GOLD["n"], GOLD["matrix"])
    assert float_eq(e.f1_score, GOLD["f1"])
    for k in ("rand_index", "recall", "precision"):  # given to 6 digits in the reference's fixture
        assert abs(getattr(e, k) - GOLD[k]) < 1e-6, k
    e2 = native.evaluate_table(GOLD["n2"], GOLD["matrix2"])
    assert float_eq(e2.f1_score, GOLD["f1_matrix2"])


def test_scores_do_not_depend_on_order():
    t = np.array(GOLD["matrix2"])
    base = native.evaluate_table(GOLD["n2"], t)
    rng = np.random.default_rng(0)
    p = native.evaluate_table(GOLD["n2"], t[rng.permutation(t.shape[0])][:, rng.permutation(t.shape[1])])
    for k in ("f1_score", "rand_index", "recall", "precision"):
        assert getattr(p, k) == pytest.approx(getattr(base, k), rel=1e-12)
    assert p.nmi == pytest.approx(base.nmi, rel=1e-9)


@pytest.mark.parametrize("seed", range(5))
def test_evaluate_against_libraries(seed):
    rng = np.random.default_rng(seed)
    N = 3000
    truth = rng.integers(0, 12, N)
    found = np.where(rng.random(N) < 0.7, truth * 2 % 17, rng.integers(0, 17, N))  # 17 blocks, mostly aligned
    e, table = native.evaluate(truth, found)
    assert (e.found_blocks, e.true_blocks) == table.shape == (17, 12)
    assert table.sum() == N
    # NMI: the entropies use logf, so agreement is to float precision
    assert e.nmi == pytest.approx(normalized_mutual_info_score(truth, found, average_method="geometric"), abs=1e-6)
    # accuracy = the maximum-weight matching of communities, whatever the tie-breaking
    raw = np.zeros((17, 12), np.int64)
    np.add.at(raw, (found, truth), 1)
    r, c = scipy.optimize.linear_sum_assignment(raw, maximize=True)
    assert e.accuracy * N == pytest.approx(raw[r, c].sum())
    assert np.trace(table) == raw[r, c].sum()
    # the same scores from the matched table
    t = native.evaluate_table(N, table)
    assert t.f1_score == e.f1_score and t.nmi == e.nmi and t.accuracy == e.accuracy


def test_fewer_found_than_true_and_unknown_labels():
    rng = np.random.default_rng(9)
    N = 2000
    truth = rng.integers(0, 10, N)
    found = truth // 3  # 4 blocks
    truth_u = truth.copy()
    truth_u[:100] = -1  # unknown: left out of the table
    e, table = native.evaluate(truth_u, found)
    assert table.shape == (4, 10) and table.sum() == N - 100
    raw = np.zeros((4, 10), np.int64)
    np.add.at(raw, (found[100:], truth[100:]), 1)
    r, c = scipy.optimize.linear_sum_assignment(raw, maximize=True)
    assert np.trace(table) == raw[r, c].sum()
    # a perfect partition scores 1
    p, _ = native.evaluate(truth, truth)
    assert p.f1_score == pytest.approx(1.0) and p.nmi == pytest.approx(1.0, abs=1e-6) and p.accuracy == 1.0
    with pytest.raises(native.SbpError):
        native.evaluate(np.full(10, -1), np.zeros(10))
