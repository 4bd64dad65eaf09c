"""Pins the oracle (oracle/oracle.c) before anything is checked against it.

Each check is independent of the oracle's own code: the reference's gold
outputs (tests/golden/*.npz, written by tests/golden/make_golden.py from the
reference's headers), brute force in numpy (int64 / float64 / lexsort),
known-answer tests and closed forms.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def rank_order(scores):
    """Brute-force ranking: score descending, row ascending."""
    rows = np.arange(len(scores))
    return np.lexsort((rows, -np.asarray(scores, np.float64)))


def random_csr(rng, rows, cols, lo, hi, values="absnormal"):
    deg = rng.integers(lo, hi + 1, rows)
    ptr = np.zeros(rows + 1, np.int64)
    ptr[1:] = np.cumsum(deg)
    idx = np.concatenate([np.sort(rng.choice(cols, d, replace=False)) for d in deg]).astype(np.int32) \
        if rows else np.zeros(0, np.int32)
    if values == "absnormal":
        val = np.abs(rng.standard_normal(len(idx))).astype(np.float32)
    else:
        val = rng.standard_normal(len(idx)).astype(np.float32)
    return ptr, idx, val


def dense(ptr, idx, val, cols, dtype=np.float64):
    n = len(ptr) - 1
    A = np.zeros((n, cols), dtype)
    for r in range(n):
        A[r, idx[ptr[r]:ptr[r + 1]]] = val[ptr[r]:ptr[r + 1]]
    return A


# --------------------------------------------------------------------------
# coo2csr (utils.hpp:494-547)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("sort", [0, 1])
def test_coo2csr_matches_reference_gold(sort):
    g = golden("coo2csr.npz")
    ptr, idx, val = O.coo2csr(g["row"], g["col"], g["val"], int(g["nrows"]), int(g["ncols"]), sort)
    np.testing.assert_array_equal(ptr, g[f"ptr_s{sort}"].astype(np.int64))
    np.testing.assert_array_equal(idx, g[f"idx_s{sort}"].astype(np.int32))
    np.testing.assert_array_equal(val, g[f"val_s{sort}"])


def test_coo2csr_invariants_and_range():
    rng = np.random.default_rng(1)
    row = rng.integers(0, 50, 400).astype(np.int32)
    col = rng.integers(0, 30, 400).astype(np.int32)
    val = rng.standard_normal(400).astype(np.float32)
    ptr, idx, v = O.coo2csr(row, col, val, 50, 30, 0)
    assert ptr[0] == 0 and ptr[-1] == 400 and np.all(np.diff(ptr) >= 0)
    np.testing.assert_array_equal(np.diff(ptr), np.bincount(row, minlength=50))
    for r in range(50):  # stable: input order within a row
        np.testing.assert_array_equal(v[ptr[r]:ptr[r + 1]], val[row == r])
    with pytest.raises(ValueError):
        O.coo2csr(np.array([0, 50], np.int32), np.array([0, 0], np.int32), np.ones(2, np.float32), 50, 30)


# --------------------------------------------------------------------------
# quantization (quantify_util.py:8-49)
# --------------------------------------------------------------------------

def test_quantize_known_answers():
    # B=2 symmetric: M=1; 0.5 is a tie -> half to even gives 0 (half away would give 1)
    q, s = O.quantize(np.array([0.5, 1.0, -0.5, -1.0, 0.25, -0.75], np.float32), 2, True)
    assert s == 1.0
    np.testing.assert_array_equal(q, [0, 1, 0, -1, 0, -1])
    # B=8 unsigned: M=255, max maps to 255, zero to 0
    q, s = O.quantize(np.array([0.0, 2.0, 1.0, 0.004], np.float32), 8, False)
    assert s == 2.0
    np.testing.assert_array_equal(q, [0, 255, 128, 1])  # 127.5 -> 128 (even); 0.51 -> 1
    # B=8 symmetric: M=127, clip keeps -128 reachable only through rounding of -1 * 127
    q, s = O.quantize(np.array([-4.0, 4.0, 2.0, -2.0], np.float32), 8, True)
    np.testing.assert_array_equal(q, [-127, 127, 64, -64])  # 63.5 -> 64 (even)
    q, _ = O.quantize(np.zeros(5, np.float32), 8, True)
    np.testing.assert_array_equal(q, 0)


@pytest.mark.parametrize("bits,sym", [(4, False), (8, False), (8, True), (6, True)])
def test_quantize_equals_reference_numpy_expression(bits, sym):
    """The reference's own numpy expression (quantify_util.py:21-27) on float32 data."""
    rng = np.random.default_rng(bits * 10 + sym)
    v = rng.standard_normal(20000).astype(np.float32)
    if not sym:
        v = np.abs(v)
    m = np.float32(2 ** bits - 1)
    M = int(m / 2) if sym else int(m)
    s = np.float32(max(abs(v.max()), abs(v.min())))
    expect = np.clip(np.int32(np.round(M * (v / s))), -M - 1, M)
    q, so = O.quantize(v, bits, sym)
    assert so == s
    np.testing.assert_array_equal(q, expect)


# --------------------------------------------------------------------------
# SpMV
# --------------------------------------------------------------------------

def test_spmv_i32_equals_dense_bruteforce():
    rng = np.random.default_rng(2)
    ptr, idx, _ = random_csr(rng, 300, 97, 0, 40)
    qv = rng.integers(-128, 256, len(idx)).astype(np.int32)
    qx = rng.integers(-128, 128, 97).astype(np.int32)
    y = O.spmv_i32(ptr, idx, qv, qx)
    A = dense(ptr, idx, qv.astype(np.int64), 97, np.int64)
    np.testing.assert_array_equal(y, A @ qx.astype(np.int64))


def test_spmv_i32_overflow_is_reported():
    ptr = np.array([0, 3], np.int64)
    with pytest.raises(OverflowError):
        O.spmv_i32(ptr, np.array([0, 1, 2], np.int32), np.full(3, 2 ** 30, np.int32), np.ones(3, np.int32))


@pytest.mark.parametrize("G", [1, 2, 4, 8, 16, 32])
def test_spmv_f32_grouped_within_fp32_bound_of_fp64(G):
    rng = np.random.default_rng(3 + G)
    ptr, idx, val = random_csr(rng, 400, 128, 0, 100, values="normal")
    x = rng.standard_normal(128).astype(np.float32)
    y = O.spmv_f32(ptr, idx, val, x, G)
    A = dense(ptr, idx, val, 128)
    exact = A @ x.astype(np.float64)
    bound = np.abs(A) @ np.abs(x.astype(np.float64)) * (100 * np.finfo(np.float32).eps)
    assert np.all(np.abs(y - exact) <= bound + 1e-30)


def test_spmv_f32_g1_is_sequential_fmaf_and_close_to_reference_gold():
    g = golden("spmv_gold.npz")
    ptr, idx, val, x = g["ptr"].astype(np.int64), g["idx"].astype(np.int32), g["val"], g["x"]
    y = O.spmv_f32(ptr, idx, val, x, 1)
    A = dense(ptr, idx, val, len(x))
    bound = np.abs(A) @ np.abs(x.astype(np.float64)) * (64 * np.finfo(np.float32).eps)
    assert np.all(np.abs(y.astype(np.float64) - g["y"]) <= bound + 1e-30)
    # a single-entry row is exact in every order
    one = np.where(np.diff(ptr) == 1)[0]
    for r in one:
        assert y[r] == g["y"][r]


def test_spmv_f32_butterfly_is_independent_of_partition_order():
    """G lanes of one row: permuting entries within a lane's stride class keeps
    the sum of each lane's partial exact only for G == nnz; here we check the
    closed case nnz == G, where the result is the fixed add tree of products."""
    x = np.array([1.0, 2.0 ** -24, -1.0, 2.0 ** -24], np.float32)
    ptr = np.array([0, 4], np.int64)
    idx = np.arange(4, dtype=np.int32)
    val = np.ones(4, np.float32)
    # G=4: tree ((p0+p2)+(p1+p3)) = (0 + 2^-23) = 2^-23; sequential ((1+2^-24)-1)+2^-24 = 2^-24
    assert O.spmv_f32(ptr, idx, val, x, 4)[0] == np.float32(2.0 ** -23)
    assert O.spmv_f32(ptr, idx, val, x, 1)[0] == np.float32(2.0 ** -24)


# --------------------------------------------------------------------------
# Top-k (cpu_mkl_baseline.cpp:128-149 with the declared tie rule)
# --------------------------------------------------------------------------

def test_topk_equals_lexsort_bruteforce_with_ties():
    rng = np.random.default_rng(4)
    y = rng.integers(-5, 6, 3000).astype(np.float64)  # heavy ties
    for k in (1, 7, 100, 2999, 3000, 3005):
        idx, val = O.topk(y, k)
        order = rank_order(y)[:k]
        m = min(k, len(y))
        np.testing.assert_array_equal(idx[:m], order)
        np.testing.assert_array_equal(val[:m], y[order])
        assert np.all(idx[m:] == -1) and np.all(np.isneginf(val[m:]))


def test_topk_matches_reference_float_gold():
    g = golden("gold_topk.npz")
    row, col, fv, fx, k = g["row"], g["col"], g["fval"], g["fx"], int(g["k"])
    nrows = int(row.max()) + 1
    ptr, idx, val = O.coo2csr(row.astype(np.int32), col.astype(np.int32), fv, nrows, len(fx))
    y = O.spmv_f32(ptr, idx, val, fx, 1)
    ti, tv = O.topk(y, k)
    # the gold accumulates with separate multiply and add: compare the set and
    # the values within FP32 rounding; scores are distinct (tie-free input)
    assert set(ti.tolist()) == set(g["f_idx"].tolist())
    np.testing.assert_allclose(np.sort(tv)[::-1], g["f_val"], rtol=1e-5)


def test_integer_gold_topk_2_scores_and_rows():
    g = golden("gold_topk.npz")
    row, col = g["row"].astype(np.int32), g["col"].astype(np.int32)
    nrows, ncols, ck = int(row.max()) + 1, len(g["qx"]), int(g["ck"])
    ptr, idx, qv = O.coo2csr(row, col, g["qval"].astype(np.float32), nrows, ncols)
    y = O.spmv_i32(ptr, idx, qv.astype(np.int32), g["qx"])
    ti, tv = O.topk(y, ck)
    np.testing.assert_array_equal(np.sort(tv)[::-1].astype(np.int32), g["i_val"])
    if len(np.unique(y[ti])) == ck and tv[-1] > np.max(np.delete(y, ti)):  # tie-free cut
        assert set(ti.tolist()) == set(g["i_idx"].tolist())
    # tied variant: the reference gold keeps a later row at the cut (">=",
    # gold_algorithms.hpp:281); the declared rule keeps the lower row.  The
    # multiset of kept scores is tie-rule independent and must agree.
    ptr, idx, qv = O.coo2csr(row, col, g["qval_tied"].astype(np.float32), nrows, ncols)
    y = O.spmv_i32(ptr, idx, qv.astype(np.int32), g["qx_tied"])
    ti, tv = O.topk(y, ck)
    np.testing.assert_array_equal(np.sort(tv)[::-1].astype(np.int32), g["t_val"])
    cut = tv[-1]
    np.testing.assert_array_equal(sorted(r for r in ti if y[r] > cut),
                                  sorted(r for r in g["t_idx"] if y[r] > cut))


def test_reference_sort_tuples_order_documented():
    g = golden("sort_tuples.npz")
    # the reference's final order puts the HIGHER index first on ties
    np.testing.assert_array_equal(g["out_idx"], [7, 1, 9, 5, 2, 3])


# --------------------------------------------------------------------------
# partitioned select (quantify_util.py:173-196)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("nrows,split,ck", [(1000, 32, 512), (1003, 32, 512), (777, 7, 70), (64, 64, 64),
                                            (500, 3, 10)])
def test_split_select_equals_per_partition_bruteforce(nrows, split, ck):
    rng = np.random.default_rng(nrows + split)
    y = rng.integers(-50, 50, nrows).astype(np.int32)
    rows, scores = O.split_select(y, split, ck)
    size, m = nrows // split, ck // split
    expect = []
    for p in range(split):
        b = p * size
        e = nrows if p == split - 1 else b + size
        order = rank_order(y[b:e])[:m] + b
        expect += order.tolist()
    np.testing.assert_array_equal(rows, expect)
    np.testing.assert_array_equal(scores, y[rows])


@pytest.mark.parametrize("nrows,split,ck", [(1000, 8, 64), (777, 5, 40), (300, 3, 300)])
def test_split_select_order_equals_bruteforce_over_the_shuffled_partitions(nrows, split, ck):
    """Random-CSR select: partition p holds rows order[p*size ...]; within a partition the rows rank by
    score descending, then original row ascending (ties are frequent: 100 score values)."""
    rng = np.random.default_rng(nrows * split)
    y = rng.integers(-50, 50, nrows).astype(np.int32)
    order = rng.permutation(nrows).astype(np.int32)
    rows, scores = O.split_select_order(y, order, split, ck)
    size, m = nrows // split, ck // split
    expect = []
    for p in range(split):
        part = order[p * size:] if p == split - 1 else order[p * size:(p + 1) * size]
        expect += part[np.lexsort((part, -y[part]))][:m].tolist()
    np.testing.assert_array_equal(rows, expect)
    np.testing.assert_array_equal(scores, y[rows])
    # the identity order is the plain select
    r0, s0 = O.split_select_order(y, np.arange(nrows, dtype=np.int32), split, ck)
    r1, s1 = O.split_select(y, split, ck)
    np.testing.assert_array_equal(r0, r1)
    np.testing.assert_array_equal(s0, s1)


def test_split_select_rejects_bad_arguments():
    with pytest.raises(ValueError):
        O.split_select(np.zeros(10, np.int32), 32, 512)


def test_split_select_statistics_match_closed_form():
    """Coverage of the true top-K by a partitioned select (the question of
    topk_errors.py:29-41), for i.i.d. scores: the top-K positions are a uniform
    K-subset, and all are kept iff no partition holds more than pk of them:
        P = [z^K] (sum_{i<=pk} C(n/b, i) z^i)^b / C(n, K).
    (topk_errors.py's closed_form_approx subtracts single-partition overflow
    terms only and overestimates this; the exact generating function is used.)"""
    from math import comb
    n, b, pk = 2048, 16, 4
    rng = np.random.default_rng(5)
    trials = 400
    hits = {8: 0, 32: 0}
    for _ in range(trials):
        y = rng.permutation(n).astype(np.int32)
        rows, _ = O.split_select(y, b, b * pk)
        s = set(rows.tolist())
        order = rank_order(y)
        for K in hits:
            hits[K] += all(t in s for t in order[:K])

    def exact(K):
        poly = [comb(n // b, i) for i in range(pk + 1)]
        acc = [1]
        for _ in range(b):
            nxt = [0] * (len(acc) + pk)
            for a, ca in enumerate(acc):
                for i, ci in enumerate(poly):
                    nxt[a + i] += ca * ci
            acc = nxt
        return acc[K] / comb(n, K) if K < len(acc) else 0.0

    for K, h in hits.items():
        p = exact(K)
        sigma = np.sqrt(max(p * (1 - p), 1e-4) / trials)
        assert abs(h / trials - p) <= 4 * sigma + 1e-9, (K, h / trials, p)


# --------------------------------------------------------------------------
# rescoring and the whole two-stage query
# --------------------------------------------------------------------------

def test_two_stage_with_all_rows_as_candidates_equals_fp32_topk():
    rng = np.random.default_rng(6)
    ptr, idx, val = random_csr(rng, 256, 64, 1, 20)
    x = rng.standard_normal(64).astype(np.float32)
    qv, _ = O.quantize(val, 8, False)
    cand, oi, ov = O.two_stage(ptr, idx, val, qv, x, 64, 8, split=1, ck=256, k=50)
    assert sorted(cand.tolist()) == list(range(256))
    y = O.spmv_f32(ptr, idx, val, x, 8)
    order = rank_order(y)[:50]
    np.testing.assert_array_equal(oi, order)
    np.testing.assert_array_equal(ov, y[order])


def test_two_stage_components_compose():
    rng = np.random.default_rng(7)
    ptr, idx, val = random_csr(rng, 3000, 100, 5, 30)
    x = rng.standard_normal(100).astype(np.float32)
    qv, _ = O.quantize(val, 4, False)
    cand, oi, ov = O.two_stage(ptr, idx, val, qv, x, 100, 4, split=16, ck=128, k=20)
    qx, _ = O.quantize(x, 8, True)
    y = O.spmv_i32(ptr, idx, qv, qx)
    rows, _ = O.split_select(y, 16, 128)
    np.testing.assert_array_equal(cand, rows)
    ri, rv = O.rescore_topk(ptr, idx, val, x, 4, rows, 20)
    np.testing.assert_array_equal(oi, ri)
    np.testing.assert_array_equal(ov, rv)
    yf = O.spmv_f32(ptr, idx, val, x, 4)
    np.testing.assert_array_equal(ov, yf[oi])  # rescored values are the FP32 scores


def test_precision_known_answers():
    exact = np.arange(100, dtype=np.int32)
    cand = np.arange(0, 100, 2, dtype=np.int32)  # even rows only
    p = O.precision(exact, cand)
    np.testing.assert_allclose(p, [1.0, 0.5, 0.5, 0.5, 0.5, 38 / 75, 0.5])
    p = O.precision(exact, np.arange(200, dtype=np.int32))
    np.testing.assert_array_equal(p, 1.0)
    p = O.precision(exact, np.array([], np.int32))
    np.testing.assert_array_equal(p, 0.0)


# --------------------------------------------------------------------------
# Ultra-CSR packets and the x-decoder (fpga_utils.hpp, test_x_decoder.cpp)
# --------------------------------------------------------------------------

def test_xdecode_known_answer():
    # src/fpga/src/x_decoder/test_x_decoder.cpp:6: flags at bits 9, 19, 29
    x = O.xdecode(0x20080200, 30, 0)
    np.testing.assert_array_equal(x, [0] * 9 + [1] * 10 + [2] * 10 + [3])
    np.testing.assert_array_equal(O.xdecode(0x20080200, 30, 5), np.array([0] * 9 + [1] * 10 + [2] * 10 + [3]) + 5)


def test_ultracsr_geometry_and_bit_positions():
    assert O.oracle().orc_ultracsr_entries(10, 6) == 30  # types.hpp:161
    # one entry: row 7 (differs from last_row 0 -> flag 1), col 0x2AB, value -3 (6-bit 0b111101)
    pk = O.ultracsr_pack(np.array([7], np.int32), np.array([0x2AB], np.int32), np.array([-3], np.int32), 0, 10, 6)
    bits = np.unpackbits(pk[0].view(np.uint8), bitorder="little")
    assert bits[0] == 1 and bits[1:30].sum() == 0
    col = sum(int(bits[30 + i]) << i for i in range(10))
    val = sum(int(bits[330 + i]) << i for i in range(6))
    assert col == 0x2AB and val == 0b111101
    # nothing else is set: flag + column bits + value bits
    assert bits.sum() == 1 + bin(0x2AB).count("1") + bin(0b111101).count("1")


def test_ultracsr_roundtrip():
    rng = np.random.default_rng(8)
    deg = rng.integers(1, 50, 200)
    row = np.repeat(np.arange(200, dtype=np.int32) + 11, deg)
    col = rng.integers(0, 1024, len(row)).astype(np.int32)
    qv = rng.integers(-32, 32, len(row)).astype(np.int32)
    pk = O.ultracsr_pack(row, col, qv, 10, 10, 6)
    xr, c, v = O.ultracsr_unpack(pk, len(row), 10, 6, 0)
    np.testing.assert_array_equal(c, col)
    np.testing.assert_array_equal(v, qv)
    np.testing.assert_array_equal(xr - xr[0], row - row[0])  # row counter advances once per row


@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_all_core_legs_equal_one_thread(acc, nthreads):
    """The all-core CPU baselines give the one-thread ports' results: the
    FP32 Top-k (heap per thread, then a heap of the threads' lists; ties at
    the cut resolved to the lower row in both) and the two-stage path."""
    csr = acc.CSRMatrix.generate(30000, 200, deg=(5, 30), seed=nthreads)
    ptr, idx, val = csr.arrays()
    X = acc.generate_queries(3, 200, seed=nthreads)
    for x in X:
        for k in (1, 100, 700):
            np.testing.assert_array_equal(np.stack(O.top_k_spmv_f32_par(ptr, idx, val, x, k, nthreads)),
                                          np.stack(O.top_k_spmv_f32(ptr, idx, val, x, k)))
    # ties everywhere: every score equal, the lowest rows win
    ones = np.ones_like(val)
    xo = np.ones(200, np.float32)
    oi, _ = O.top_k_spmv_f32_par(np.arange(0, 30001, dtype=np.int64), np.zeros(30000, np.int32),
                                 ones[:30000], xo, 50, nthreads)
    np.testing.assert_array_equal(oi, np.arange(50))
    q, _ = O.quantize(val, 8, False)
    for x in X:
        a = O.two_stage_par(ptr, idx, val, q, x, 200, 4, 16, 256, 100, nthreads)
        b = O.two_stage(ptr, idx, val, q, x, 200, 4, 16, 256, 100)
        for u, v in zip(a, b):
            np.testing.assert_array_equal(u, v)
