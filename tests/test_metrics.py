"""Ranking-quality metrics: rank NDCG, edit distance and error counts against the reference's own
evaluation_utils.hpp:112-267 (compiled into oracle/_ref), score NDCG (quantify_util.py:313-329) against
the oracle, which hand-computed lists pin."""
import math

import numpy as np
import pytest

import oracle_lib as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference gold library not built")


def lists(rng, n):
    g = rng.permutation(4 * n + 10)[:n].astype(np.int32)
    t = g.copy()
    for _ in range(int(rng.integers(0, 6))):
        i, j = rng.integers(0, n, 2)
        t[i], t[j] = t[j], t[i]
    for _ in range(int(rng.integers(0, 4))):
        t[rng.integers(0, n)] = rng.integers(0, 5 * n + 10)  # misses and repeated rows
    return g, t


def test_rank_metrics_hand_computed(acc):
    assert acc.ndcg_rank([4, 7, 9], [4, 7, 9]) == 1.0
    # golden [1, 2], test [2, 1]: rel(1) = 1, rel(2) = 2; dcg = 1/log2(2) + 2/log2(3), idcg = 2/log2(2) + 1/log2(3)
    assert acc.ndcg_rank([1, 2], [2, 1]) == pytest.approx((1 + 2 / math.log2(3)) / (2 + 1 / math.log2(3)), rel=1e-12)
    assert acc.ndcg_rank([1, 2], [3, 4]) == 0.0
    assert acc.edit_distance([1, 2, 3], [1, 2, 3]) == 0
    assert acc.edit_distance([1, 2, 3], [2, 3]) == 1
    assert acc.edit_distance([1, 2, 3], [3, 2, 1]) == 2
    assert acc.edit_distance([], [5, 6]) == 2
    assert acc.bounded_count_errors([1, 2, 3, 4], [1, 5, 3, 6], (2, 4, 8)) == [1, 2]
    assert len(acc.bounded_ndcg(list(range(30)), list(range(30)))) == 2  # 50 > 30 stops the list


@needs_ref
def test_rank_metrics_match_reference(acc):
    rng = np.random.default_rng(1)
    for _ in range(400):
        n = int(rng.integers(1, 80))
        g, t = lists(rng, n)
        assert acc.ndcg_rank(g, t) == O.ref().ref_ndcg(g, t, n)
        m = int(rng.integers(0, n + 1))
        assert acc.edit_distance(g, t[:m]) == O.ref().ref_edit_distance(g, n, np.ascontiguousarray(t[:m]), m)
        assert acc.bounded_count_errors(g, t, (n,))[0] == O.ref().ref_count_errors(g, t, n)


def test_score_ndcg_oracle_hand_computed():
    sw, val = [1, 2, 3], [3.0, 2.0, 1.0]
    idcg = 3 / 1 + 2 / math.log2(3) + 1 / 2
    assert O.ndcg_score(sw, val, [1, 2, 3]) == pytest.approx(1.0, rel=1e-15)
    assert O.ndcg_score(sw, val, [2, 1, 3]) == pytest.approx((2 / 1 + 3 / math.log2(3) + 1 / 2) / idcg, rel=1e-15)
    assert O.ndcg_score(sw, val, [9, 1, 2]) == pytest.approx((3 / math.log2(3) + 2 / 2) / idcg, rel=1e-15)
    # a row listed twice in sw: the later score counts (a dict built in order)
    assert O.ndcg_score([5, 5], [1.0, 2.0], [5]) == pytest.approx(2 / (1 + 2 / math.log2(3)), rel=1e-15)


def test_score_ndcg_matches_oracle(acc):
    rng = np.random.default_rng(2)
    for _ in range(300):
        n = int(rng.integers(1, 100))
        g, t = lists(rng, n)
        val = np.sort(rng.standard_normal(n))[::-1].copy()
        h = t[: int(rng.integers(0, n + 1))]
        ndcg, dcg, idcg = acc.ndcg_score(g, val, h)
        assert ndcg == O.ndcg_score(g, val, h)
        assert idcg == pytest.approx(sum(v / math.log2(i + 2) for i, v in enumerate(val)), rel=1e-12)
        assert ndcg == pytest.approx(dcg / idcg, rel=1e-15)
