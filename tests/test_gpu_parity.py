"""GPU parity: the HIP path against the oracle, through the C ABI.

Bars: INT32 stage-1 scores, candidate sets and Top-K row sets bit-exact;
FP32 scores bit-exact to the oracle's declared summation order (and within
1e-5 relative of FP64); every query in a run checked.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

KS = [1, 8, 16, 32, 50, 75, 100]


def dev(acc, a):
    return acc.DeviceBuffer.from_numpy(np.ascontiguousarray(a))


def empty(acc, n, dtype):
    return acc.DeviceBuffer((n,), dtype)


def make(acc, rows, cols, deg, bits=8, split=32, signed=False, deg_kind=0, seed=1, l2=True):
    vk = acc.VAL_NORMAL if signed else acc.VAL_ABSNORMAL
    csr = acc.CSRMatrix.generate(rows, cols, deg=deg, deg_kind=deg_kind, val_kind=vk, l2norm=l2, seed=seed)
    pk = csr.pack(bits=bits, split=split, is_signed=signed)
    return csr, pk, acc.GpuMatrix(csr, pk)


def oracle_qv(csr, bits, signed):
    ptr, idx, val = csr.arrays()
    q, _ = O.quantize(val, bits, signed)
    return ptr, idx, val, q


def els(out):
    a = out.numpy().view(np.int32).reshape(-1, 2)
    return a[:, 0].view(np.float32), a[:, 1]


# --------------------------------------------------------------------------
# FP32 CSR path
# --------------------------------------------------------------------------

@pytest.mark.parametrize("rows,cols,deg", [(20000, 500, (50, 90)), (5000, 1024, (0, 40)), (3000, 64, (60, 64)),
                                           (777, 33, (1, 3))])
def test_fp32_spmv_bitexact(acc, gpu, rows, cols, deg):
    csr, _, g = make(acc, rows, cols, deg, seed=rows)
    ptr, idx, val = csr.arrays()
    x = acc.generate_queries(1, cols, seed=rows)[0]
    y = empty(acc, rows, np.float32)
    g.spmv_fp32(dev(acc, x), y)
    acc.synchronize()
    yo = O.spmv_f32(ptr, idx, val, x, g.info.group)
    np.testing.assert_array_equal(y.numpy(), yo)
    # FP32 vs FP64: within the dot-product error bound n*eps*sum|a_ij x_j|, and
    # within 1e-5 relative wherever the score is not dominated by cancellation
    y64 = O.spmv_f64(ptr, idx, val, x.astype(np.float64))
    mag = O.spmv_f64(ptr, idx, np.abs(val), np.abs(x).astype(np.float64))
    assert np.all(np.abs(yo - y64) <= 128 * np.finfo(np.float32).eps * mag + 1e-30)
    ok = np.abs(y64) > 0.1 * mag
    np.testing.assert_allclose(yo[ok], y64[ok], rtol=1e-5)


@pytest.mark.parametrize("k", [1, 8, 100, 960])
def test_fp32_topk_bitexact(acc, gpu, k):
    csr, _, g = make(acc, 100000, 500, (50, 90), seed=7)  # configs[0] shape
    ptr, idx, val = csr.arrays()
    X = acc.generate_queries(3, 500, l2norm=True, seed=k)
    for x in X:  # consecutive queries reuse the device state
        out = empty(acc, k * 2, np.int32)
        g.topk_fp32(dev(acc, x), k, out)
        acc.synchronize()
        v, i = els(out)
        yo = O.spmv_f32(ptr, idx, val, x, g.info.group)
        oi, ov = O.topk(yo, k)
        np.testing.assert_array_equal(i, oi)
        np.testing.assert_array_equal(v, ov.astype(np.float32))


def test_fp32_topk_more_than_rows(acc, gpu):
    csr, _, g = make(acc, 50, 20, (1, 5), seed=3)
    x = acc.generate_queries(1, 20, seed=3)[0]
    out = g.topk_fp32_host(x, 64)
    assert np.all(out["index"][50:] == -1) and np.all(np.isneginf(out["value"][50:]))
    ptr, idx, val = csr.arrays()
    oi, _ = O.topk(O.spmv_f32(ptr, idx, val, x, g.info.group), 64)
    np.testing.assert_array_equal(out["index"], oi)


@pytest.mark.parametrize("n,k", [(300000, 100), (300000, 300), (2500000, 256), (5000, 7), (0, 5)])
def test_dense_top_k_heap(acc, gpu, n, k):
    # radix passes for k <= 256 (2.5M: a second pass over the first pass's keys), sorted lists above
    rng = np.random.default_rng(n + k)
    y = rng.integers(-1000, 1000, n).astype(np.float32)  # many ties: lower row must win
    y[:: 7] = -0.0
    out = empty(acc, 2 * k, np.int32)
    acc.top_k_heap(dev(acc, y) if n else empty(acc, 1, np.float32), n, k, out)
    acc.synchronize()
    v, i = els(out)
    oi, ov = O.topk(y, k)
    m = min(n, k)
    np.testing.assert_array_equal(i[:m], oi[:m])
    np.testing.assert_array_equal(v[:m], ov[:m])
    assert (i[m:] == -1).all() and np.isneginf(v[m:]).all()


# --------------------------------------------------------------------------
# low-bit path: stage-1 INT32 scores, candidates, two-stage output
# --------------------------------------------------------------------------

CASES = [
    # rows, cols, deg, bits, split, signed, deg_kind
    (40000, 500, (40, 100), 8, 32, False, 0),   # configs[1] shape, reduced rows
    (40000, 500, (40, 100), 4, 32, False, 0),
    (30000, 1024, (0.5, 40.0), 8, 64, False, 1),  # power-law rows up to 1024
    (20000, 900, (30, 90), 8, 16, True, 0),     # signed values
    (5000, 7, (0, 7), 6, 5, True, 0),           # tiny columns, empty rows, 6-bit
    (1000, 4096, (0, 300), 8, 3, False, 0),     # wide rows
    (64, 100, (1, 10), 8, 64, False, 0),        # one row per partition
    (600000, 64, (2, 8), 8, 1, False, 0),       # one partition of 9375 slices (maxima re-read per digit)
    (200000, 4, (1, 4), 2, 1, False, 0),        # 2-bit, 4 columns: thousands of slices tie at the top
    (60000, 4, (1, 1), 2, 1, False, 0),         # one entry per row: thousands of tied rows in listed slices
]


@pytest.mark.parametrize("rows,cols,deg,bits,split,signed,kind", CASES)
def test_lowbit_scores_bitexact(acc, gpu, rows, cols, deg, bits, split, signed, kind):
    csr, pk, g = make(acc, rows, cols, deg, bits, split, signed, kind, seed=rows + bits)
    ptr, idx, val, q = oracle_qv(csr, bits, signed)
    for qs in range(2):
        x = acc.generate_queries(1, cols, seed=qs)[0]
        y = empty(acc, rows, np.int32)
        g.scores_lowbit(dev(acc, x), y)
        acc.synchronize()
        qx, _ = O.quantize(x, 8, True)
        np.testing.assert_array_equal(y.numpy(), O.spmv_i32(ptr, idx, q, qx))


@pytest.mark.parametrize("rows,cols,deg,bits,split,signed,kind", CASES)
def test_lowbit_candidates_bitexact(acc, gpu, rows, cols, deg, bits, split, signed, kind):
    csr, pk, g = make(acc, rows, cols, deg, bits, split, signed, kind, seed=rows + bits)
    ptr, idx, val, q = oracle_qv(csr, bits, signed)
    for ck in sorted({split, min(64, 512 // split) * split, 16 * split if 16 * split <= 4096 else split}):
        x = acc.generate_queries(1, cols, seed=ck)[0]
        rr = empty(acc, ck, np.int32)
        ss = empty(acc, ck, np.int32)
        n = g.candidates_lowbit(dev(acc, x), ck, rr, ss)
        qx, _ = O.quantize(x, 8, True)
        y = O.spmv_i32(ptr, idx, q, qx)
        orow, osc = O.split_select(y, split, ck)
        assert n == len(orow)
        np.testing.assert_array_equal(rr.numpy()[:n], orow)
        np.testing.assert_array_equal(ss.numpy()[:n], osc)


@pytest.mark.parametrize("rows,cols,deg,bits,split,signed,kind", CASES[:4])
def test_lowbit_two_stage_bitexact(acc, gpu, rows, cols, deg, bits, split, signed, kind):
    csr, pk, g = make(acc, rows, cols, deg, bits, split, signed, kind, seed=rows + bits)
    ptr, idx, val, q = oracle_qv(csr, bits, signed)
    ck = min(64, 512 // split) * split
    X = acc.generate_queries(4, cols, seed=99)
    for qi, x in enumerate(X):
        k = KS[(qi * 3) % len(KS)] if qi else 100
        out = empty(acc, 2 * k, np.int32)
        g.topk_lowbit(dev(acc, x), k, ck, out)
        acc.synchronize()
        v, i = els(out)
        _, oi, ov = O.two_stage(ptr, idx, val, q, x, cols, g.info.group, split, ck, k)
        np.testing.assert_array_equal(i, oi)
        np.testing.assert_array_equal(v, ov)


@pytest.mark.parametrize("k", [200, 257, 960])
def test_lowbit_two_stage_large_k(acc, gpu, k):
    # K above 256 ranks by merging sorted lists, K <= 256 by radix selection
    rows, cols, deg, bits, split, signed, kind = CASES[0]
    csr, pk, g = make(acc, rows, cols, deg, bits, split, signed, kind, seed=rows + bits)
    ptr, idx, val, q = oracle_qv(csr, bits, signed)
    x = acc.generate_queries(1, cols, seed=k)[0]
    out = empty(acc, 2 * k, np.int32)
    g.topk_lowbit(dev(acc, x), k, 1024, out)
    acc.synchronize()
    v, i = els(out)
    _, oi, ov = O.two_stage(ptr, idx, val, q, x, cols, g.info.group, split, 1024, k)
    np.testing.assert_array_equal(i, oi)
    np.testing.assert_array_equal(v, ov)


def test_lowbit_stream_and_host_form_agree(acc, gpu):
    csr, pk, g = make(acc, 50000, 500, (40, 100), 8, 32, seed=5)
    x = acc.generate_queries(1, 500, seed=5)[0]
    xd = dev(acc, x)
    out = empty(acc, 200, np.int32)
    st = acc.Stream()
    g.topk_lowbit(xd, 100, 512, out, st)
    st.synchronize()
    v, i = els(out)
    h = g.topk_lowbit_host(x, 100, 512)
    np.testing.assert_array_equal(h["index"], i)
    np.testing.assert_array_equal(h["value"], v)


BATCH_MODES = pytest.mark.parametrize("mode", [1, 2], ids=["packets", "dense"])


def check_batch(acc, g, csr, bits, signed, X, ks, ck, split, out_stride=100):
    """Batched Top-K through the C ABI against the oracle, every query."""
    ptr, idx, val, q = oracle_qv(csr, bits, signed)
    nq, cols = X.shape
    out = empty(acc, nq * out_stride * 2, np.int32)
    g.topk_lowbit_batch(dev(acc, X), ks, ck, out, out_stride)
    acc.synchronize()
    a = out.numpy().reshape(nq, out_stride, 2)
    for qi in range(nq):
        _, oi, ov = O.two_stage(ptr, idx, val, q, X[qi], cols, g.info.group, split, ck, int(ks[qi]))
        np.testing.assert_array_equal(a[qi, :ks[qi], 1], oi, err_msg=f"query {qi}")
        np.testing.assert_array_equal(a[qi, :ks[qi], 0].view(np.float32), ov, err_msg=f"query {qi}")


@BATCH_MODES
def test_lowbit_batch_uneven_k(acc, gpu, mode):
    csr, pk, g = make(acc, 30000, 900, (30, 90), 8, 32, seed=8)
    g.set_batch_mode(mode)
    ptr, idx, val, q = oracle_qv(csr, 8, False)
    nq = 12
    X = acc.generate_queries(nq, 900, seed=8)
    ks = np.array([KS[i % len(KS)] for i in range(nq)], np.int32)
    out = empty(acc, nq * 100 * 2, np.int32)
    g.topk_lowbit_batch(dev(acc, X), ks, 512, out, 100)
    acc.synchronize()
    a = out.numpy().reshape(nq, 100, 2)
    for qi in range(nq):
        _, oi, ov = O.two_stage(ptr, idx, val, q, X[qi], 900, g.info.group, 32, 512, int(ks[qi]))
        np.testing.assert_array_equal(a[qi, :ks[qi], 1], oi)
        np.testing.assert_array_equal(a[qi, :ks[qi], 0].view(np.float32), ov)


@BATCH_MODES
@pytest.mark.parametrize("rows,cols,deg,bits,split,signed,nq", [
    (20000, 500, (40, 100), 8, 32, False, 40),   # unsigned 8-bit (offset bytes), 64-query blocks
    (20000, 900, (30, 90), 8, 16, True, 37),     # signed 8-bit
    (20000, 500, (40, 100), 4, 32, False, 9),    # 4-bit, one 32-query block
    (30000, 1024, (0.5, 40.0), 8, 64, False, 33),  # power-law rows, 1024 columns (32 column blocks)
    (20000, 900, (30, 90), 8, 32, False, 256),   # 900 columns, 256 queries: two 128-query blocks
    (20000, 500, (40, 100), 8, 32, False, 256),  # 256 queries: one block, two tiles per MFMA wavefront
    (20000, 500, (40, 100), 8, 32, False, 160),  # five tiles: three two-tile wavefronts, a padding tile
    (20000, 200, (10, 30), 8, 8, False, 200),    # 8 column blocks, seven tiles: a ragged tile, a padding tile
    (12000, 300, (5, 30), 4, 8, False, 300),     # 300 queries: two blocks of three two-tile wavefronts, a ragged tile, a padding tile
    (40000, 4096, (0, 300), 4, 16, False, 24),   # 4096 columns: query columns past what the kernels hold in registers
    (50000, 200, (10, 30), 8, 2, False, 20),     # partitions of 391 slices (threshold by extraction)
    (150000, 200, (10, 30), 8, 2, False, 20),    # partitions of 1172 slices (threshold by collectors)
    (20000, 512, (40, 100), 6, 16, True, 40),    # signed 6-bit, 512 columns: image entries of 8 + 10 bits
    (20000, 256, (10, 60), 8, 8, False, 64),     # 256 columns: image entries one column bit wider than the packets'
    (5000, 3, (1, 3), 8, 4, False, 20),          # three columns: the narrowest image entries (8 + 2 bits)
    (20000, 700, (30, 90), 8, 16, False, 50),    # 22 column blocks: the 24-block chain
    (20000, 700, (30, 90), 8, 16, False, 300),   # 24 blocks, 300 queries: two blocks of five one-tile wavefronts, two per SIMD
    (20000, 1024, (0, 40), 8, 32, False, 160),   # 32 blocks, five tiles in one block of one-tile wavefronts, two per SIMD
])
def test_lowbit_batch_matches_oracle(acc, gpu, mode, rows, cols, deg, bits, split, signed, nq):
    kind = 1 if isinstance(deg[0], float) else 0
    csr, pk, g = make(acc, rows, cols, deg, bits, split, signed, kind, seed=rows + nq)
    if mode == 2 and cols > 1024:  # the matrix cores hold B registers for up to 1024 columns
        with pytest.raises(acc.AccelesError, match="more than 1024 columns"):
            g.set_batch_mode(mode)
        mode = 0  # AUTO: the packet batch
    g.set_batch_mode(mode)
    X = acc.generate_queries(nq, cols, seed=nq)
    ck = min(64, 512 // split) * split
    ks = np.array([KS[i % len(KS)] for i in range(nq)], np.int32)
    check_batch(acc, g, csr, bits, signed, X, ks, ck, split)


@BATCH_MODES
def test_lowbit_batch_heavy_ties(acc, gpu, mode):
    """2-bit values and a 0/1 query: few distinct scores, so slice lists
    overflow at the threshold and the dense path rescans slices exactly."""
    csr, pk, g = make(acc, 6000, 64, (20, 40), 2, 8, seed=31)
    g.set_batch_mode(mode)
    X = (acc.generate_queries(20, 64, seed=31) > 0).astype(np.float32)
    ks = np.array([[1, 8, 16, 32, 50, 64][i % 6] for i in range(20)], np.int32)
    check_batch(acc, g, csr, 2, False, X, ks, 64, 8)  # M = 8 of 12 slices per partition


@BATCH_MODES
def test_lowbit_batch_all_rows_tie(acc, gpu, mode):
    """Every row holds the same entries, so every score ties: every slice list
    reaches the threshold (all slices rescanned, far more candidates than a
    wavefront's register path holds), and the results are the lowest rows in
    order."""
    rows, cols = 5000, 64
    ptr = np.arange(rows + 1, dtype=np.int64) * 8
    idx = np.tile(np.arange(0, 64, 8, dtype=np.int32), rows)
    val = np.full(rows * 8, 0.5, np.float32)
    csr = acc.CSRMatrix.from_arrays(rows, cols, ptr, idx, val)
    g = acc.GpuMatrix(csr, csr.pack(bits=8, split=8))
    g.set_batch_mode(mode)
    X = np.ones((12, cols), np.float32)
    ks = np.array([1, 8, 16, 32, 50, 100] * 2, np.int32)
    check_batch(acc, g, csr, 8, False, X, ks, 256, 8)  # M = 32 > the 10 slices of a partition
    out = empty(acc, 12 * 100 * 2, np.int32)
    g.topk_lowbit_batch(dev(acc, X), ks, 256, out, 100)
    acc.synchronize()
    a = out.numpy().reshape(12, 100, 2)
    # each partition (625 rows) keeps its 32 lowest rows; on equal scores the
    # lower row ranks first
    expect = np.sort(np.concatenate([p * 625 + np.arange(32) for p in range(8)]))
    for qi, k in enumerate(ks):
        np.testing.assert_array_equal(a[qi, :k, 1], expect[:k])


@BATCH_MODES
def test_lowbit_batch_tiny_partitions(acc, gpu, mode):
    """Partitions of one or two slices (at most M slices: every slice is
    scanned), empty rows, K larger than some partitions hold."""
    csr, pk, g = make(acc, 333, 40, (0, 6), 8, 16, seed=5)
    g.set_batch_mode(mode)
    X = acc.generate_queries(18, 40, seed=5)
    ks = np.array([1, 8, 16, 32, 50, 75, 100, 100, 3] * 2, np.int32)
    check_batch(acc, g, csr, 8, False, X, ks, 1024, 16)


def test_dense_layout_refuses_duplicate_columns(acc, gpu):
    """A row naming a column twice is one byte per column only while the two
    values' sum fits: here it does not (255 + 255), so the dense mode is
    refused with a message, AUTO keeps the packet path, and the results still
    match the oracle (which sums both entries)."""
    rng = np.random.default_rng(9)
    rows, cols = 3000, 128
    ptr = np.arange(0, rows * 12 + 1, 12, dtype=np.int64)
    idx = rng.integers(0, cols, rows * 12).astype(np.int32)
    idx[5 * 12 + 1] = idx[5 * 12]  # row 5 repeats a column
    val = np.abs(rng.standard_normal(rows * 12)).astype(np.float32)
    val[5 * 12:5 * 12 + 2] = 3.0  # both copies keep a non-zero quantized value
    csr = acc.CSRMatrix.from_arrays(rows, cols, ptr, idx, val)
    g = acc.GpuMatrix(csr, csr.pack(bits=8, split=8))
    assert not g.info.dense_ok
    with pytest.raises(acc.AccelesError, match="sum past one byte"):
        g.set_batch_mode(acc.BATCH_DENSE)
    g.set_batch_mode(acc.BATCH_AUTO)
    X = acc.generate_queries(24, cols, seed=9)
    ks = np.array([KS[i % len(KS)] for i in range(24)], np.int32)
    check_batch(acc, g, csr, 8, False, X, ks, 256, 8)


@pytest.mark.parametrize("bits,signed", [(4, False), (8, False), (6, True)])
def test_dense_layout_sums_repeated_columns(acc, gpu, bits, signed):
    """Columns sampled with replacement (as the reference generators do): a
    row's values in one column are summed into its tile byte when the sums
    fit, and the dense results equal the oracle's (which sums the entries)."""
    rng = np.random.default_rng(bits)
    rows, cols, d = 4000, 96, 16
    ptr = np.arange(0, rows * d + 1, d, dtype=np.int64)
    idx = np.sort(rng.integers(0, cols, (rows, d)), axis=1).astype(np.int32)
    idx[0] = np.arange(d) * (cols // d)  # row 0: distinct columns, it holds the largest value
    idx = idx.ravel()
    val = rng.standard_normal(rows * d).astype(np.float32)
    if not signed:
        val = np.abs(val)
    val[0] = 8.0 if not signed else -8.0  # the scale; the others stay near half of it or below
    csr = acc.CSRMatrix.from_arrays(rows, cols, ptr, idx, val)
    pk = csr.pack(bits=bits, split=8, is_signed=signed)
    assert pk.info.nnz_kept > 0
    g = acc.GpuMatrix(csr, pk)
    assert g.info.dense_ok, "repeated columns whose sums fit a byte keep the dense layout"
    g.set_batch_mode(acc.BATCH_DENSE)
    X = acc.generate_queries(40, cols, seed=bits)
    ks = np.array([KS[i % len(KS)] for i in range(40)], np.int32)
    check_batch(acc, g, csr, bits, signed, X, ks, 256, 8)


def test_loaded_files_give_the_same_results(acc, gpu, tmp_path):
    csr, pk, g = make(acc, 30000, 500, (40, 100), 8, 32, seed=12)
    csr.save(str(tmp_path / "m.csr"))
    pk.save(str(tmp_path / "m.pk"))
    g2 = acc.GpuMatrix(acc.CSRMatrix.load(str(tmp_path / "m.csr")), acc.PackedMatrix.load(str(tmp_path / "m.pk")))
    X = acc.generate_queries(3, 500, seed=12)
    for x in X:
        a, b = empty(acc, 200, np.int32), empty(acc, 200, np.int32)
        g.topk_lowbit(dev(acc, x), 100, 512, a)
        g2.topk_lowbit(dev(acc, x), 100, 512, b)
        acc.synchronize()
        np.testing.assert_array_equal(a.numpy(), b.numpy())


def test_fp64_exact_scores(acc, gpu):
    csr, _, g = make(acc, 20000, 500, (40, 100), seed=4)
    ptr, idx, val = csr.arrays()
    x = acc.generate_queries(1, 500, seed=4)[0]
    y = empty(acc, 20000, np.float64)
    g.spmv_fp64(dev(acc, x), y)
    acc.synchronize()
    np.testing.assert_allclose(y.numpy(), O.spmv_f64(ptr, idx, val, x.astype(np.float64)), rtol=1e-12,
                               atol=1e-15)


# --------------------------------------------------------------------------
# full-size checks (400,000 x 500)
# --------------------------------------------------------------------------

def test_embedding_scale_parity_and_recall(acc, gpu):
    """400,000 x 500, row nnz U[40,100], |N(0,1)| L2-normalised, 8-bit,
    SPLIT=32, CK=512: INT32 scores bit-exact at full size, two-stage output
    bit-exact, and recall@K vs FP64 exact Top-K."""
    csr, pk, g = make(acc, 400000, 500, (40, 100), 8, 32, seed=0xACCE1E5 + 1)
    ptr, idx, val, q = oracle_qv(csr, 8, False)
    X = acc.generate_queries(6, 500, seed=2)
    full = []
    for n, x in enumerate(X):
        xd = dev(acc, x)
        if n < 2:
            y = empty(acc, 400000, np.int32)
            g.scores_lowbit(xd, y)
            acc.synchronize()
            qx, _ = O.quantize(x, 8, True)
            np.testing.assert_array_equal(y.numpy(), O.spmv_i32(ptr, idx, q, qx))
        out = empty(acc, 200, np.int32)
        g.topk_lowbit(xd, 100, 512, out)
        y64 = empty(acc, 400000, np.float64)
        g.spmv_fp64(xd, y64)
        acc.synchronize()
        v, i = els(out)
        if n < 2:
            _, oi, ov = O.two_stage(ptr, idx, val, q, x, 500, g.info.group, 32, 512, 100)
            np.testing.assert_array_equal(i, oi)
            np.testing.assert_array_equal(v, ov)
        yy = y64.numpy()
        exact = np.lexsort((np.arange(len(yy)), -yy))[:100]
        full.append([len(set(i[:K]) & set(exact[:K])) / K for K in KS])
    rec = np.mean(full, axis=0)
    assert rec[-1] >= 0.99, rec


# --------------------------------------------------------------------------
# plans (HIP graphs): every replay equals the eager call on the same queries
# --------------------------------------------------------------------------

@pytest.mark.parametrize("kind,nq,mode", [("lowbit", 1, 0), ("fp32", 1, 0), ("batch", 40, 1), ("batch", 40, 2),
                                          ("batch", 130, 2)])
def test_plan_replays_match_eager(acc, gpu, kind, nq, mode):
    csr, pk, g = make(acc, 20000, 500, (40, 100), 8, 32, seed=77)
    if mode:
        g.set_batch_mode(mode)
    K, CK = 50, 512
    x = empty(acc, nq * 500, np.float32)
    out = empty(acc, nq * K * 2, np.int32)
    ref = empty(acc, nq * K * 2, np.int32)
    other = empty(acc, 7 * K * 2, np.int32)
    ks = np.array([[1, 8, 16, 32, 50][i % 5] for i in range(nq)], np.int32) if kind == "batch" else np.array([K], np.int32)
    X0 = acc.generate_queries(nq, 500, seed=1000)
    acc.lib().acc_memcpy_h2d(x.ptr, X0.ctypes.data, X0.nbytes)
    pkind = {"lowbit": acc.PLAN_LOWBIT, "fp32": acc.PLAN_FP32, "batch": acc.PLAN_LOWBIT_BATCH}[kind]
    plan = acc.Plan(g, pkind, x, ks, CK, out, K)
    Xo = dev(acc, acc.generate_queries(7, 500, seed=5))
    for rep in range(4):
        X = acc.generate_queries(nq, 500, seed=2000 + rep)
        acc.lib().acc_memcpy_h2d(x.ptr, X.ctypes.data, X.nbytes)
        plan.launch()
        acc.synchronize()
        # eager work on the same handle between replays (another batch shape)
        g.topk_lowbit_batch(Xo, np.full(7, 9, np.int32), CK, other, K)
        acc.synchronize()
        if kind == "lowbit":
            g.topk_lowbit(x, K, CK, ref)
        elif kind == "fp32":
            g.topk_fp32(x, K, ref)
        else:
            g.topk_lowbit_batch(x, ks, CK, ref, K)
        acc.synchronize()
        a, b = out.numpy().reshape(nq, K, 2), ref.numpy().reshape(nq, K, 2)
        for qi in range(nq):
            np.testing.assert_array_equal(a[qi, :ks[qi]], b[qi, :ks[qi]], err_msg=f"replay {rep} query {qi}")
    plan.close()


@pytest.mark.parametrize("kind", ["lowbit", "fp32"])
def test_plans_in_flight_on_two_streams(acc, gpu, kind):
    # plans own their query workspaces: two plans of one handle run concurrently on two streams
    csr, pk, g = make(acc, 60000, 500, (40, 100), 8, 32, seed=78)
    K, CK = 100, 512
    pkind = {"lowbit": acc.PLAN_LOWBIT, "fp32": acc.PLAN_FP32}[kind]
    X = acc.generate_queries(2, 500, seed=4242)
    xs = [dev(acc, X[j]) for j in range(2)]
    outs = [empty(acc, K * 2, np.int32) for _ in range(2)]
    plans = [acc.Plan(g, pkind, xs[j], [K], CK, outs[j], K) for j in range(2)]
    streams = [acc.Stream(0), acc.Stream(0)]
    for rep in range(6):
        for j in range(2):
            plans[j].launch(streams[j])
    for s in streams:
        s.synchronize()
    ref = empty(acc, K * 2, np.int32)
    for j in range(2):
        if kind == "lowbit":
            g.topk_lowbit(xs[j], K, CK, ref)
        else:
            g.topk_fp32(xs[j], K, ref)
        acc.synchronize()
        np.testing.assert_array_equal(outs[j].numpy(), ref.numpy(), err_msg=f"plan {j}")
    for p in plans:
        p.close()


def test_autotune_keeps_results(acc, gpu):
    """The tuned layout choice per batch size changes speed, not results."""
    csr, pk, g = make(acc, 20000, 500, (40, 100), 8, 32, seed=91)
    pts = g.autotune(512, 50, 64)
    assert [p.nq for p in pts] == [1, 2, 4, 8, 16, 32, 64]
    for p in pts:
        assert p.layout in (acc.BATCH_SINGLE, acc.BATCH_SPARSE, acc.BATCH_DENSE)
        assert p.us_single > 0 and p.us_sparse > 0 and p.us_dense > 0
    for nq in (1, 3, 20, 64):
        X = acc.generate_queries(nq, 500, seed=nq + 90)
        ks = np.array([[1, 8, 16, 32, 50][i % 5] for i in range(nq)], np.int32)
        check_batch(acc, g, csr, 8, False, X, ks, 512, 32, out_stride=50)
