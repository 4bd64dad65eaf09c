"""ReSparse keep-ratio search on the GPU (datasets/search_best_resparse_ratio.py:63-116).

The reference pins its search only by result logs of datasets that are not available here (GloVe, its
generated 5,000,000 x 1024 matrices), so these checks are the search's own invariants: the bisection
replays from its history, the returned ratio meets the convergence threshold, pruning is monotone in
nnz, and keep = 1 leaves the candidates unchanged."""
import numpy as np
import pytest

from acceles_amd.resparse import KS, search_keep_ratio

pytestmark = pytest.mark.gpu


def test_search_keep_ratio_invariants(acc, gpu):
    csr = acc.CSRMatrix.generate(20000, 200, deg=(20, 60), seed=31)
    res = search_keep_ratio(csr, bits=8, split=8, ck=256, nq=24, accuracy_threshold=0.99, seed=5)
    h = res["history"]
    assert h[0][0] == 1.0 and abs(res["threshold"] - 0.99 * h[0][1]) < 1e-12
    # replay the bisection
    low, high, keep, tried = 0.0, 1.0, 0.5, {1.0}
    for k, c, _ in h[1:]:
        assert k == keep
        tried.add(k)
        if c >= res["threshold"]:
            high = k
        else:
            low = k
        keep = round((low + high) / 2, 3)
    assert keep in tried and res["keep"] == high
    conv = {k: c for k, c, _ in h}
    assert res["keep"] == 1.0 or conv[res["keep"]] >= res["threshold"]
    nnz = sorted((k, n) for k, _, n in h)
    assert all(a[1] <= b[1] for a, b in zip(nnz, nnz[1:])), "fewer entries kept at a smaller ratio"
    assert len(KS) == 7


def test_keep_one_leaves_candidates_unchanged(acc, gpu):
    csr = acc.CSRMatrix.generate(5000, 100, deg=(10, 30), seed=3)
    x = acc.DeviceBuffer.from_numpy(acc.generate_queries(1, 100, seed=1)[0])
    outs = []
    for m in (csr, csr.resparse(1.0)):
        g = acc.GpuMatrix(None, m.pack(bits=8, split=4))
        rows = acc.DeviceBuffer((64,), np.int32)
        sc = acc.DeviceBuffer((64,), np.int32)
        n = g.candidates_lowbit(x, 64, rows, sc)
        acc.synchronize()
        outs.append((rows.numpy()[:n], sc.numpy()[:n]))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
