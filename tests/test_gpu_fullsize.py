"""GPU parity at the full sizes of the benchmark workloads, through the C ABI.

- 20,000,000 x 1024 single queries (uniform and power-law rows), SPLIT=64,
  CK=1024, K=100: every INT32 score, the candidates and the Top-100 bit-exact
  against the oracle; the FP32 CSR Top-100 bit-exact in the declared G-lane
  order, and equal to the sequential (one lane per row) order of the
  reference's CPU path wherever rounding cannot move a row across the cut.
- 2,000,000 x 900, 256 queries of mixed K per HIP-graph launch, 8-bit and
  4-bit: every 16th query bit-exact, all 256 against the FP64 exact Top-K.
- The serving loop over three resident matrices (1M x 500, 5M x 900,
  10M x 1024): probe queries alone, in a mix of other requests, and against
  the oracle.
- 400,000 x 500 at 4 bits; at 8 bits, batches of 100 and 256 queries on the
  matrix cores (one and two query tiles per MFMA wavefront).
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

KS = [1, 8, 16, 32, 50, 75, 100]
SEED = 0xACCE1E5
EPS32 = float(np.finfo(np.float32).eps)


def dev(acc, a):
    return acc.DeviceBuffer.from_numpy(np.ascontiguousarray(a))


def empty(acc, n, dtype):
    return acc.DeviceBuffer((n,), dtype)


def els(out):
    a = out.numpy().view(np.int32).reshape(-1, 2)
    return a[:, 0].view(np.float32), a[:, 1]


def exact_topk(acc, g, xd, rows, k):
    """FP64 exact Top-k rows (score descending, row ascending) and the FP64 scores."""
    y64 = empty(acc, rows, np.float64)
    g.spmv_fp64(xd, y64)
    acc.synchronize()
    yy = y64.numpy()
    top = np.argpartition(-yy, k)[:k + 1]
    return top[np.lexsort((top, -yy[top]))][:k], yy


def check_sequential_order(ptr, idx, val, x, rows_gpu, k):
    """The GPU's FP32 Top-k rows (declared G-lane order) against the Top-k of
    the sequential order (one fmaf chain per row: the reference's spmv_gold /
    MKL single-thread order, cpu_mkl_baseline.cpp:128-149,261).  The two row
    sets may differ only in rows whose FP64 score lies within both orders'
    rounding bound of the cut; returns how many rows differ."""
    y1 = O.spmv_f32(ptr, idx, val, x, 1)
    seq, _ = O.topk(y1, k)
    diff = np.setxor1d(seq, rows_gpu)
    if len(diff) == 0:
        return 0
    y64 = O.spmv_f64(ptr, idx, val, x.astype(np.float64))
    mag = O.spmv_f64(ptr, idx, np.abs(val), np.abs(x).astype(np.float64))
    n = np.diff(ptr)
    order = np.argsort(-y64, kind="stable")
    cut = 0.5 * (y64[order[k - 1]] + y64[order[k]])
    bound = 2.0 * n * EPS32 * mag + 1e-30
    assert np.all(np.abs(y64[diff] - cut) <= 2 * bound[diff]), "rows differ beyond the rounding bound"
    return len(diff)


# --------------------------------------------------------------------------
# 20,000,000 x 1024 single queries
# --------------------------------------------------------------------------

@pytest.mark.parametrize("dist", ["uniform", "gamma"])
def test_large_single_query_full_size(acc, gpu, dist):
    rows, cols, split, ck, K = 20_000_000, 1024, 64, 1024, 100
    print(f"\n[fullsize] {rows}x{cols} {dist}", flush=True)
    if dist == "uniform":
        csr = acc.CSRMatrix.generate(rows, cols, deg=(10, 30), l2norm=False, seed=SEED + 2)
    else:
        csr = acc.CSRMatrix.generate(rows, cols, deg=(0.5, 40.0), deg_kind=acc.DEG_GAMMA, l2norm=False,
                                     seed=SEED + 2)
    pk = csr.pack(bits=8, split=split)
    g = acc.GpuMatrix(csr, pk)
    G = g.info.group
    ptr, idx, val = csr.arrays()
    q, _ = O.quantize(val, 8, False)
    X = acc.generate_queries(2, cols, seed=SEED + 3)
    for n, x in enumerate(X):
        print(f"[fullsize] query {n}", flush=True)
        xd = dev(acc, x)
        qx, _ = O.quantize(x, 8, True)
        yo = O.spmv_i32(ptr, idx, q, qx)
        y = empty(acc, rows, np.int32)
        g.scores_lowbit(xd, y)
        acc.synchronize()
        np.testing.assert_array_equal(y.numpy(), yo, err_msg="INT32 scores")
        del y
        rr, ss = empty(acc, ck, np.int32), empty(acc, ck, np.int32)
        nc = g.candidates_lowbit(xd, ck, rr, ss)
        orow, osc = O.split_select(yo, split, ck)
        assert nc == len(orow) == ck
        np.testing.assert_array_equal(rr.numpy(), orow, err_msg="candidate rows")
        np.testing.assert_array_equal(ss.numpy(), osc, err_msg="candidate scores")
        out = empty(acc, 2 * K, np.int32)
        g.topk_lowbit(xd, K, ck, out)
        acc.synchronize()
        v, i = els(out)
        _, oi, ov = O.two_stage(ptr, idx, val, q, x, cols, G, split, ck, K)
        np.testing.assert_array_equal(i, oi, err_msg="two-stage rows")
        np.testing.assert_array_equal(v, ov, err_msg="two-stage FP32 scores")
        # the plain FP32 CSR Top-100 on the same GPU
        out32 = empty(acc, 2 * K, np.int32)
        g.topk_fp32(xd, K, out32)
        acc.synchronize()
        v32, i32 = els(out32)
        yf = O.spmv_f32(ptr, idx, val, x, G)
        fi, fv = O.topk(yf, K)
        np.testing.assert_array_equal(i32, fi, err_msg="FP32 Top-K rows")
        np.testing.assert_array_equal(v32, fv.astype(np.float32), err_msg="FP32 Top-K scores")
        moved = check_sequential_order(ptr, idx, val, x, i32, K)
        print(f"[fullsize] rows differing from the sequential order: {moved}", flush=True)
        # recall of the two-stage output against the FP64 exact Top-100
        exact, _ = exact_topk(acc, g, xd, rows, K)
        assert len(set(i) & set(exact)) / K >= 0.99


# --------------------------------------------------------------------------
# 2,000,000 x 900, 256 queries of mixed K per graph launch
# --------------------------------------------------------------------------

@pytest.mark.parametrize("bits", [8, 4])
def test_batched_graph_full_size(acc, gpu, bits):
    rows, cols, B, ck, split = 2_000_000, 900, 256, 512, 32
    print(f"\n[fullsize] {rows}x{cols} batched, {bits}-bit", flush=True)
    csr = acc.CSRMatrix.generate(rows, cols, deg=(30, 90), seed=SEED + 4)
    pk = csr.pack(bits=bits, split=split)
    g = acc.GpuMatrix(csr, pk)
    g.autotune(ck, 100, B)
    G = g.info.group
    rng = np.random.default_rng(SEED + 5)
    ks = rng.choice(KS, B).astype(np.int32)
    ptr, idx, val = csr.arrays()
    q, _ = O.quantize(val, bits, False)
    xbuf = empty(acc, B * cols, np.float32)
    out = empty(acc, B * 100 * 2, np.int32)
    X0 = acc.generate_queries(B, cols, seed=SEED + 6)
    xbuf.upload(X0)
    plan = acc.Plan(g, acc.PLAN_LOWBIT_BATCH, xbuf, ks, ck, out, 100)
    for rep in range(2):
        X = acc.generate_queries(B, cols, seed=SEED + 60 + rep)
        xbuf.upload(X)
        print(f"[fullsize] replay {rep}", flush=True)
        plan.launch()
        acc.synchronize()
        a = out.numpy().reshape(B, 100, 2)
        for qi in range(0, B, 16):
            _, oi, ov = O.two_stage(ptr, idx, val, q, X[qi], cols, G, split, ck, int(ks[qi]))
            np.testing.assert_array_equal(a[qi, :ks[qi], 1], oi, err_msg=f"replay {rep} query {qi}")
            np.testing.assert_array_equal(a[qi, :ks[qi], 0].view(np.float32), ov, err_msg=f"replay {rep} query {qi}")
    plan.close()
    # every query of the last replay against the FP64 exact Top-K
    xd = dev(acc, X)
    prec = []
    for qi in range(B):
        k = int(ks[qi])
        exact, _ = exact_topk(acc, g, xd.at(qi * cols), rows, k)
        prec.append(len(set(a[qi, :k, 1]) & set(exact)) / k)
    prec = np.array(prec)
    print(f"[fullsize] {bits}-bit mean precision@K {prec.mean():.4f}, all found {np.mean(prec == 1.0):.4f}", flush=True)
    assert prec.mean() >= (0.99 if bits == 8 else 0.9)


# --------------------------------------------------------------------------
# serving: three resident matrices
# --------------------------------------------------------------------------

def test_serving_full_size(acc, gpu):
    specs = [(1_000_000, 500, 60), (5_000_000, 900, 40), (10_000_000, 1024, 20)]
    print("\n[fullsize] serving setup", flush=True)
    srv = acc.Server(max_batch=256, slots=3, ck=512)
    mats = []
    for i, (r, c, d) in enumerate(specs):
        csr = acc.CSRMatrix.generate(r, c, deg=(d // 2, d + d // 2), l2norm=False, seed=SEED + 10 + i)
        pk = csr.pack(bits=8, split=32)
        mid = srv.add_matrix(csr, pk)
        gm = acc.GpuMatrix(csr, None)  # the G the server's copy uses (a function of the mean row length)
        mats.append((mid, csr, gm.info.group))
        del gm, pk
    rng = np.random.default_rng(SEED + 30)
    probes = []
    for j in range(9):
        mi = j % 3
        probes.append((mi, acc.generate_queries(1, specs[mi][1], seed=SEED + 40 + j)[0], KS[(2 * j) % 7]))
    print("[fullsize] probes alone", flush=True)
    alone = [srv.wait(srv.submit(mats[mi][0], x, k), k) for mi, x, k in probes]
    print("[fullsize] probes in the mix", flush=True)
    order = rng.permutation(600 + len(probes))
    tickets = []
    for t in order:
        if t < len(probes):
            mi, x, k = probes[t]
            tickets.append((t, srv.submit(mats[mi][0], x, k), k))
        else:
            mi = int(rng.integers(0, 3))
            k = int(rng.integers(1, 101))
            tickets.append((-1, srv.submit(mats[mi][0], rng.standard_normal(specs[mi][1]).astype(np.float32), k), k))
    mixed = {}
    for t, tk, k in tickets:
        res = srv.wait(tk, k)
        if t >= 0:
            mixed[t] = res
    st = srv.stats()
    assert st["batches"] < st["completed"], "requests were batched"
    print("[fullsize] oracle", flush=True)
    for j, (mi, x, k) in enumerate(probes):
        np.testing.assert_array_equal(mixed[j], alone[j], err_msg=f"probe {j}: mix vs alone")
        _, csr, G = mats[mi]
        ptr, idx, val = csr.arrays()
        q, _ = O.quantize(val, 8, False)
        _, oi, ov = O.two_stage(ptr, idx, val, q, x, csr.cols, G, 32, 512, k)
        np.testing.assert_array_equal(alone[j]["index"], oi, err_msg=f"probe {j}: rows")
        np.testing.assert_array_equal(alone[j]["value"], ov, err_msg=f"probe {j}: FP32 scores")
    srv.close()


# --------------------------------------------------------------------------
# 400,000 x 500 at 4 bits
# --------------------------------------------------------------------------

def test_embedding_scale_4bit(acc, gpu):
    rows, cols, split, ck, K = 400_000, 500, 32, 512, 100
    csr = acc.CSRMatrix.generate(rows, cols, deg=(40, 100), seed=SEED + 1)
    pk = csr.pack(bits=4, split=split)
    g = acc.GpuMatrix(csr, pk)
    ptr, idx, val = csr.arrays()
    q, _ = O.quantize(val, 4, False)
    X = acc.generate_queries(4, cols, seed=SEED + 2)
    prec = []
    for n, x in enumerate(X):
        xd = dev(acc, x)
        qx, _ = O.quantize(x, 8, True)
        if n < 2:
            y = empty(acc, rows, np.int32)
            g.scores_lowbit(xd, y)
            acc.synchronize()
            np.testing.assert_array_equal(y.numpy(), O.spmv_i32(ptr, idx, q, qx))
        out = empty(acc, 2 * K, np.int32)
        g.topk_lowbit(xd, K, ck, out)
        acc.synchronize()
        v, i = els(out)
        _, oi, ov = O.two_stage(ptr, idx, val, q, x, cols, g.info.group, split, ck, K)
        np.testing.assert_array_equal(i, oi)
        np.testing.assert_array_equal(v, ov)
        exact, _ = exact_topk(acc, g, xd, rows, K)
        prec.append(len(set(i) & set(exact)) / K)
    # the same queries through the batched path
    Xd = dev(acc, X)
    outb = empty(acc, len(X) * K * 2, np.int32)
    g.topk_lowbit_batch(Xd, np.full(len(X), K, np.int32), ck, outb, K)
    acc.synchronize()
    a = outb.numpy().reshape(len(X), K, 2)
    for n, x in enumerate(X):
        _, oi, _ = O.two_stage(ptr, idx, val, q, x, cols, g.info.group, split, ck, K)
        np.testing.assert_array_equal(a[n, :, 1], oi)
    print(f"\n[fullsize] 4-bit precision@100 {np.mean(prec):.4f}", flush=True)


@pytest.mark.parametrize("B", [100, 256])
def test_embedding_scale_batch(acc, gpu, B):
    """400,000 x 500, 8-bit, a batch on the matrix cores in one launch (100
    queries: four 32-query tiles, one per MFMA wavefront; 256: eight, two
    per wavefront): the first and last query of every tile against the
    oracle."""
    rows, cols, split, ck, K = 400_000, 500, 32, 512, 100
    print(f"\n[fullsize] 400000x500 batch of {B}", flush=True)
    csr = acc.CSRMatrix.generate(rows, cols, deg=(40, 100), seed=SEED + 1)
    pk = csr.pack(bits=8, split=split)
    g = acc.GpuMatrix(csr, pk)
    g.set_batch_mode(acc.BATCH_DENSE)
    ptr, idx, val = csr.arrays()
    q, _ = O.quantize(val, 8, False)
    X = acc.generate_queries(B, cols, seed=SEED + 7)
    out = empty(acc, B * K * 2, np.int32)
    g.topk_lowbit_batch(dev(acc, X), np.full(B, K, np.int32), ck, out, K)
    acc.synchronize()
    a = out.numpy().reshape(B, K, 2)
    for qi in sorted({q for t in range(0, B, 32) for q in (t, min(t + 31, B - 1))}):
        _, oi, ov = O.two_stage(ptr, idx, val, q, X[qi], cols, g.info.group, split, ck, K)
        np.testing.assert_array_equal(a[qi, :, 1], oi, err_msg=f"query {qi}")
        np.testing.assert_array_equal(a[qi, :, 0].view(np.float32), ov, err_msg=f"query {qi}")


# --------------------------------------------------------------------------
# the src/cpu drop-in: top_k_spmv (cpu_mkl_baseline.cpp:249-275)
# --------------------------------------------------------------------------

def test_top_k_spmv_drop_in(acc, gpu, capfd):
    rows, cols, k, num_test = 100_000, 500, 100, 5
    csr = acc.CSRMatrix.generate(rows, cols, deg=(50, 90), seed=SEED)
    g = acc.GpuMatrix(csr, None)
    res = g.top_k_spmv(k=k, num_test=num_test)
    out = capfd.readouterr().out
    for name in ("compute time:", "sort time:", "exec time:"):
        assert name in out
    # the last query is U(-1, 1) from the seed 0x5EED + num_test - 1
    x = acc.generate_queries(1, cols, val_kind=acc.VAL_UNIFORM, l2norm=False, seed=0x5EED + num_test - 1)[0]
    ptr, idx, val = csr.arrays()
    oi, ov = O.topk(O.spmv_f32(ptr, idx, val, x, g.info.group), k)
    np.testing.assert_array_equal(res["index"], oi)
    np.testing.assert_array_equal(res["value"], ov.astype(np.float32))
    bad = np.empty(k, acc.ELEMENT_DTYPE)
    with pytest.raises(acc.AccelesError, match="row/col"):  # the shape is checked against the handle
        acc._check(acc.lib().acc_top_k_spmv(g._h, rows + 1, cols, k, 1, bad, 1))
