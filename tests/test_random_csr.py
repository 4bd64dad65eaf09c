"""Random-CSR (README.md:5): rows shuffled before the SPLIT partitions
(quantify_util.py:491 shuffles the rows; csr_util.py:101-114 is the CSR form),
so that the Top-K rows of a query spread over the partitions as the coverage
model of topk_errors.py:29-41 assumes.  Packing takes a seed; results keep
naming the original rows."""
import numpy as np
import pytest
from scipy.stats import hypergeom

import oracle_lib as O


def packed_partitions(acc, pk, split, tmp_path):
    """Rows of each partition, read from the packet file (the slot -> row map and the partition slice
    ranges; the layout is this library's own, see test_files.packet_offsets)."""
    from test_files import packet_offsets
    path = str(tmp_path / "p.pk")
    pk.save(path)
    data = open(path, "rb").read()
    o = packet_offsets(pk, split)
    psb = np.frombuffer(data, np.int64, split + 1, o["part_slice_begin"] + 8)
    perm = np.frombuffer(data, np.int32, pk.info.nslices * 64, o["perm"] + 8)
    return [np.sort(perm[psb[p] * 64:psb[p + 1] * 64][perm[psb[p] * 64:psb[p + 1] * 64] >= 0]) for p in range(split)]


@pytest.mark.parametrize("seed", [0, 7, 0xACCE1E5])
def test_partitions_follow_the_seeded_order(acc, tmp_path, seed):
    rows, split = 3000, 7
    csr = acc.CSRMatrix.generate(rows, 90, deg=(1, 20), seed=11)
    pk = csr.pack(bits=8, split=split, shuffle_seed=seed)
    assert pk.info.shuffle_seed == seed
    order = np.arange(rows) if seed == 0 else csr.permute_rows(seed)[1]
    size = rows // split
    got = packed_partitions(acc, pk, split, tmp_path)
    for p in range(split):
        want = order[p * size:(p + 1) * size] if p < split - 1 else order[p * size:]
        np.testing.assert_array_equal(got[p], np.sort(want))
    # the entries of every row are unchanged by the shuffle
    for a, b in zip(pk.unpack(), csr.pack(bits=8, split=split).unpack()):
        np.testing.assert_array_equal(a, b)


def test_shuffled_packets_round_trip_through_a_file(acc, tmp_path):
    csr = acc.CSRMatrix.generate(2000, 64, deg=(1, 12), seed=5)
    pk = csr.pack(bits=4, split=5, shuffle_seed=99)
    p = str(tmp_path / "s.pk")
    pk.save(p)
    back = acc.PackedMatrix.load(p)
    assert back.info.shuffle_seed == 99
    assert [r.tolist() for r in packed_partitions(acc, back, 5, tmp_path)] == \
           [r.tolist() for r in packed_partitions(acc, pk, 5, tmp_path)]


def clustered(acc, n, cols, k, seed):
    """n rows; rows 0 .. k-1 lie close to the query direction (the true Top-k, clustered at the front,
    hence all in partition 0 without a shuffle), the rest are random |N(0,1)| rows; L2-normalised."""
    rng = np.random.default_rng(seed)
    x = np.abs(rng.standard_normal(cols)).astype(np.float32)
    x /= np.linalg.norm(x)
    dense = np.abs(rng.standard_normal((n, cols))).astype(np.float32)
    dense[rng.random((n, cols)) < 0.75] = 0
    dense[:k] = x + 0.02 * np.abs(rng.standard_normal((k, cols))).astype(np.float32)
    dense /= np.maximum(np.linalg.norm(dense, axis=1, keepdims=True), 1e-12)
    r, c = np.nonzero(dense)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=n))]).astype(np.int64)
    csr = acc.CSRMatrix.from_arrays(n, cols, ptr, c.astype(np.int32), dense[r, c])
    return csr, x


def coverage(n, split, k, m):
    """Expected fraction of k rows, placed at random over split equal partitions of n rows, that the
    partitions' top-m keep: split * E[min(X, m)] / k with X ~ Hypergeom(n, k, n / split)."""
    X = hypergeom(n, k, n // split)
    xs = np.arange(0, k + 1)
    return split * float(np.sum(np.minimum(xs, m) * X.pmf(xs))) / k


def test_shuffle_restores_the_coverage_of_clustered_top_rows(acc):
    """Without a shuffle the clustered Top-k share one partition and only its top m = ck/split survive
    (recall m/k); with seeded shuffles the recall approaches the random-placement coverage."""
    n, cols, k, split, ck = 20000, 64, 100, 32, 128
    m = ck // split
    csr, x = clustered(acc, n, cols, k, seed=3)
    ptr, idx, val = csr.arrays()
    qv, _ = O.quantize(val, 8, False)
    truth = set(range(k))
    _, oi, _ = O.two_stage_order(ptr, idx, val, qv, x, cols, 1, split, ck, k, np.arange(n, dtype=np.int32))
    plain = len(truth & set(oi.tolist())) / k
    assert plain == pytest.approx(m / k, abs=0.011)
    rec = []
    for seed in range(1, 41):
        _, order = csr.permute_rows(seed)
        _, oi, _ = O.two_stage_order(ptr, idx, val, qv, x, cols, 1, split, ck, k, order)
        rec.append(len(truth & set(oi.tolist())) / k)
    expect = coverage(n, split, k, m)
    assert 0.6 < expect < 0.95  # a regime where the partitions' capacity matters
    # 40 shuffles: the mean recall has a standard error of about 0.01
    assert np.mean(rec) == pytest.approx(expect, abs=0.04)
    assert np.mean(rec) > plain + 0.5


@pytest.mark.gpu
@pytest.mark.parametrize("bits", [8, 4])
def test_gpu_random_csr_matches_oracle(acc, gpu, bits):
    """Single-query and batched low-bit Top-K on shuffled packets: the oracle over the same partitions
    (orc_two_stage_order), rows named by their original ids, ties to the lower original id."""
    rows, cols, split, ck, seed = 60000, 300, 16, 256, 0x5EED
    csr = acc.CSRMatrix.generate(rows, cols, deg=(5, 40), seed=21)
    pk = csr.pack(bits=bits, split=split, shuffle_seed=seed)
    _, order = csr.permute_rows(seed)
    g = acc.GpuMatrix(csr, pk)
    X = acc.generate_queries(40, cols, seed=22)
    ks = np.array([1, 8, 16, 32, 50, 75, 100, 100] * 5, np.int32)
    ptr, idx, val = csr.arrays()
    qv, _ = O.quantize(val, bits, False)
    refs = [O.two_stage_order(ptr, idx, val, qv, X[i], cols, g.info.group, split, ck, int(ks[i]), order)
            for i in range(len(X))]
    # one query at a time
    for i in range(8):
        res = g.topk_lowbit_host(X[i], int(ks[i]), ck)
        np.testing.assert_array_equal(res["index"], refs[i][1])
        np.testing.assert_array_equal(res["value"], refs[i][2])
    # batched, both scoring layouts
    st = acc.Stream(0)
    Xd = acc.DeviceBuffer.from_numpy(X)
    for mode in (acc.BATCH_SPARSE, acc.BATCH_DENSE):
        g.set_batch_mode(mode)
        out = acc.DeviceBuffer((len(X), 100, 2), np.int32)
        g.topk_lowbit_batch(Xd, ks, ck, out, 100, st)
        st.synchronize()
        got = out.numpy()
        for i in range(len(X)):
            k = int(ks[i])
            np.testing.assert_array_equal(got[i, :k, 1], refs[i][1])
            np.testing.assert_array_equal(got[i, :k, 0].view(np.float32), refs[i][2])
