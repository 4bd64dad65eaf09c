"""ReSparse (resparse_csr, csr_util.py:152-187; resparse_spm, quantify_util.py:375-403): the oracle's
keep mask against hand-computed rows and a sort-based restatement, then the library against the oracle."""
import numpy as np
import pytest

import oracle_lib as O


def small():
    # r0: pruned (min 0.5 <= mean); r1: every |v| above either mean, kept whole; r2: empty;
    # r3: ties; r4: kept whole only under the all-cells mean (1.045) and pruned under the nonzero mean (1.9)
    rows = [[(0, 1.0), (1, -3.0), (2, 2.0), (3, 0.5)], [(0, 5.0), (3, 6.0)], [], [(0, 0.1), (1, 0.1), (2, 0.1)],
            [(0, 1.5), (1, 1.6)]]
    ptr = np.cumsum([0] + [len(r) for r in rows]).astype(np.int64)
    idx = np.array([c for r in rows for c, _ in r], np.int32)
    val = np.array([v for r in rows for _, v in r], np.float32)
    return ptr, idx, val, 4


def kept(ptr, idx, val, mask):
    return [[(int(idx[j]), round(float(val[j]), 6)) for j in range(ptr[r], ptr[r + 1]) if mask[j]]
            for r in range(len(ptr) - 1)]


def test_oracle_hand_computed():
    ptr, idx, val, cols = small()
    m = O.resparse_mask(ptr, val, cols, 0.5)  # mean over nonzeros = 20.9 / 11 = 1.9
    assert kept(ptr, idx, val, m) == [[(1, -3.0), (2, 2.0)], [(0, 5.0), (3, 6.0)], [], [(0, 0.1), (1, 0.1)],
                                      [(1, 1.6)]]
    m = O.resparse_mask(ptr, val, cols, 0.5, mean_over_cells=True)  # mean over 20 cells = 1.045
    assert kept(ptr, idx, val, m)[4] == [(0, 1.5), (1, 1.6)]
    m = O.resparse_mask(ptr, val, cols, 0.7)  # ceil(4 * 0.7) = 3, ceil(3 * 0.7) = 3, ceil(2 * 0.7) = 2
    k = kept(ptr, idx, val, m)
    assert k[0] == [(0, 1.0), (1, -3.0), (2, 2.0)] and len(k[3]) == 3 and len(k[4]) == 2
    assert O.resparse_mask(ptr, val, cols, 1.0).all()
    assert kept(ptr, idx, val, O.resparse_mask(ptr, val, cols, 0.0)) == [[], [(0, 5.0), (3, 6.0)], [], [], []]


def sort_based(ptr, val, cols, keep, cells):
    """The rule restated with a stable sort (independent of the oracle's rank counting)."""
    a = np.abs(val.astype(np.float64))
    avg = a.sum() / (len(ptr[:-1]) * cols if cells else max(len(val), 1))
    mask = np.zeros(len(val), bool)
    for r in range(len(ptr) - 1):
        b, e = ptr[r], ptr[r + 1]
        if e == b:
            continue
        if a[b:e].min() > avg:
            mask[b:e] = True
        else:
            order = np.argsort(-a[b:e], kind="stable")
            mask[b + order[:int(np.ceil((e - b) * keep))]] = True
    return mask


@pytest.mark.parametrize("keep", [0.0, 0.25, 0.5, 0.777, 1.0])
@pytest.mark.parametrize("cells", [False, True])
def test_oracle_against_sort_restatement(keep, cells):
    rng = np.random.default_rng(int(keep * 1000) + cells)
    lens = rng.integers(0, 12, 300)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    val = rng.integers(-6, 7, ptr[-1]).astype(np.float32) / 4  # many ties
    val[val == 0] = 0.25
    np.testing.assert_array_equal(O.resparse_mask(ptr, val, 40, keep, cells), sort_based(ptr, val, 40, keep, cells))


@pytest.mark.parametrize("keep", [0.0, 0.3, 0.5, 0.9, 1.0])
@pytest.mark.parametrize("cells", [False, True])
def test_library_matches_oracle(acc, keep, cells):
    csr = acc.CSRMatrix.generate(3000, 120, deg=(0, 40), seed=int(keep * 100) + 7 * cells)
    ptr, idx, val = csr.arrays()
    r = csr.resparse(keep, mean_over_cells=cells)
    p2, i2, v2 = r.arrays()
    m = O.resparse_mask(ptr, val, 120, keep, cells)
    assert r.rows == csr.rows and r.cols == csr.cols and r.nnz == int(m.sum())
    cs = np.concatenate([[0], np.cumsum(m)])
    np.testing.assert_array_equal(np.diff(p2), cs[ptr[1:]] - cs[ptr[:-1]])
    np.testing.assert_array_equal(i2, idx[m])
    np.testing.assert_array_equal(v2, val[m])


def test_library_rejects_bad_ratio(acc):
    csr = acc.CSRMatrix.generate(10, 8, deg=(1, 3), seed=1)
    with pytest.raises(acc.AccelesError, match="keep"):
        csr.resparse(1.5)


# --- the keep-ratio search against the reference's own search logs ---------------------------------

def search_logs():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "resparse_search_logs.json")) as f:
        return json.load(f)


def logged_convergence(e):
    """An epoch's convergence: the value the log printed, else the metric of its printed precisions."""
    from acceles_amd.resparse import convergence
    if e["convergence"] is not None:
        return e["convergence"]
    return convergence([[p / 100 for p in e["precision_pct"]]])


def test_convergence_metric_matches_the_logged_values():
    """convergence() of each epoch's printed mean precisions is the convergence the reference printed
    (quantify_util.py:433-436; self_generated_data/csr_util.py:260): exactly up to rounding where the
    precisions are printed unrounded, within the 0.005 % print rounding (x sum of weights) otherwise."""
    from acceles_amd.resparse import WEIGHTS, convergence
    n = 0
    for log in search_logs():
        for e in log["epochs"]:
            if e["convergence"] is None or len(e["precision_pct"]) != 7:
                continue
            got = convergence([[p / 100 for p in e["precision_pct"]]])
            tol = 1e-12 if not e["rounded"] else 0.5e-4 * sum(WEIGHTS) + 1e-12
            assert abs(got - e["convergence"]) <= tol, (log["log"], e["epoch"], got, e["convergence"])
            n += 1
    assert n >= 150
    # a dropped weight or a transposed one is far outside that tolerance
    e = search_logs()[6]["epochs"][1]
    p = [x / 100 for x in e["precision_pct"]]
    assert abs(convergence([p[::-1]]) - e["convergence"]) > 0.1


def test_bisection_replays_every_logged_search():
    """bisect_keep fed each epoch's logged convergence tries the logged ratios in the logged order and
    settles where the reference did ("Convergence reached ...%" prints high * 100)."""
    from acceles_amd.resparse import bisect_keep
    for log in search_logs():
        by_keep = {round(e["top_percent"] / 100, 3): logged_convergence(e) for e in log["epochs"]}
        tried = []

        def measure(keep):
            tried.append(keep)
            return (by_keep[keep],)

        n = len(log["epochs"])
        res = bisect_keep(measure, log["accuracy_threshold"], max_epochs=n)
        assert [round(k * 100, 1) for k in tried] == [e["top_percent"] for e in log["epochs"]], log["log"]
        assert res["threshold"] == log["epochs"][0]["convergence"] * log["accuracy_threshold"]
        if log["converged_percent"] is None:  # a log that ends before the search stopped
            assert not res["converged"]
        else:
            assert res["converged"] and res["keep"] * 100 == log["converged_percent"], log["log"]


def test_glove_search_settles_at_the_published_ratio():
    """The GloVe 6B search at threshold 0.999, CK=512, SPLIT=32 keeps 52.9 % (precision@100 96.61 %)."""
    from acceles_amd.resparse import bisect_keep
    log = next(l for l in search_logs() if "glove.6B.50d.txt_search_resparse_0.999_" in l["log"])
    by_keep = {round(e["top_percent"] / 100, 3): logged_convergence(e) for e in log["epochs"]}
    res = bisect_keep(lambda k: (by_keep[k],), 0.999)
    assert res["keep"] == 0.529 and len(res["history"]) == 11
    last = log["epochs"][-1]
    assert last["top_percent"] == 52.9 and last["precision_pct"][-1] == 96.61
