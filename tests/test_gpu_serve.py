"""Serving loop (acc_server_*) on the GPU: results alone equal results in the
mix, against the oracle, memory and queue admission, the load generator."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def matrices(acc):
    specs = [  # rows, cols, degree, kind, bits, split
        (20000, 500, (40, 100), acc.DEG_UNIFORM, 8, 32),
        (30000, 900, (30, 90), acc.DEG_UNIFORM, 8, 16),
        (40000, 1024, (0.5, 40.0), acc.DEG_GAMMA, 4, 32),
    ]
    out = []
    for i, (r, c, d, kind, bits, split) in enumerate(specs):
        csr = acc.CSRMatrix.generate(r, c, deg=d, deg_kind=kind, seed=500 + i)
        out.append((csr, csr.pack(bits=bits, split=split), bits, split))
    return out


def test_server_alone_equals_mix_and_oracle(acc, gpu):
    mats = matrices(acc)
    srv = acc.Server(max_batch=64, slots=3, ck=512)
    ids = [srv.add_matrix(csr, pk) for csr, pk, _, _ in mats]
    rng = np.random.default_rng(3)
    probes = []
    for j in range(12):
        mi = j % 3
        x = acc.generate_queries(1, mats[mi][0].cols, seed=900 + j)[0]
        probes.append((mi, x, [1, 8, 16, 32, 50, 75, 100][j % 7]))
    alone = []
    for mi, x, k in probes:
        alone.append(srv.wait(srv.submit(ids[mi], x, k), k))
    # the same probes among 400 other requests of all matrices and K
    tickets, order = [], rng.permutation(400 + len(probes))
    for t in order:
        if t < len(probes):
            mi, x, k = probes[t]
            tickets.append(("p", t, srv.submit(ids[mi], x, k), k))
        else:
            mi = int(rng.integers(0, 3))
            x = rng.standard_normal(mats[mi][0].cols).astype(np.float32)
            k = int(rng.integers(1, 101))
            tickets.append(("o", t, srv.submit(ids[mi], x, k), k))
    mixed = {}
    for kind, t, tk, k in tickets:
        res = srv.wait(tk, k)
        if kind == "p":
            mixed[t] = res
    st = srv.stats()
    assert st["batches"] < st["completed"], "requests were batched"
    for j, (mi, x, k) in enumerate(probes):
        np.testing.assert_array_equal(mixed[j], alone[j], err_msg=f"probe {j}")
        csr, pk, bits, split = mats[mi]
        ptr, idx, val = csr.arrays()
        q, _ = O.quantize(val, bits, False)
        gm = acc.GpuMatrix(csr, pk)  # G of the server's copy of this matrix
        _, oi, ov = O.two_stage(ptr, idx, val, q, x, csr.cols, gm.info.group, split, 512, k)
        np.testing.assert_array_equal(alone[j]["index"], oi, err_msg=f"probe {j}")
        np.testing.assert_array_equal(alone[j]["value"], ov, err_msg=f"probe {j}")
    srv.close()


def test_server_admission(acc, gpu):
    csr = acc.CSRMatrix.generate(20000, 500, deg=(40, 100), seed=7)
    pk = csr.pack(bits=8, split=32)
    # slot buffers alone: 2 x 64 x (8192*4 + 960*8) bytes
    small = acc.Server(memory_budget=12 << 20, max_batch=64, slots=2)
    with pytest.raises(acc.AccelesError, match="budget"):
        small.add_matrix(csr, pk)
    small.close()
    srv = acc.Server(max_batch=8, slots=1, max_pending=4)
    mid = srv.add_matrix(csr, pk)
    x = acc.generate_queries(1, 500, seed=1)[0]
    refused = 0
    tickets = []
    for _ in range(64):
        try:
            tickets.append(srv.submit(mid, x, 10))
        except acc.AccelesError as e:
            assert "queue full" in str(e)
            refused += 1
    for t in tickets:
        srv.wait(t, 10)
    assert refused > 0 and srv.stats()["refused"] == refused
    srv.close()


def test_server_poisson_load(acc, gpu):
    mats = matrices(acc)
    srv = acc.Server(max_batch=128, slots=3)
    for csr, pk, _, _ in mats:
        srv.add_matrix(csr, pk)
    st = srv.bench_poisson(5000, 0.5, kmax=100, seed=11)
    assert st["completed"] == st["submitted"] > 1000
    assert 0 < st["p50_us"] <= st["p99_us"] <= st["max_us"]
    srv.close()


def test_concurrent_add_matrix_admission(acc, gpu):
    """Two threads add a matrix each to a server whose budget holds one: the
    reservation is taken under the lock, so exactly one is admitted and the
    other is refused with ACC_ERR_BUDGET."""
    import threading
    # 2048 columns at 8 bits: scores may pass 2^25, so the dense layout is
    # never used and a matrix's footprint does not depend on the room left
    csrs = [acc.CSRMatrix.generate(20000, 2048, deg=(40, 100), seed=700 + i) for i in range(2)]
    pks = [c.pack(bits=8, split=32) for c in csrs]
    probe = acc.Server(max_batch=16, slots=1, ck=512)
    base = probe.stats()["device_bytes"]
    probe.add_matrix(csrs[0], pks[0])
    one = probe.stats()["device_bytes"] - base
    probe.close()
    srv = acc.Server(memory_budget=base + int(1.5 * one), max_batch=16, slots=1, ck=512)
    results = [None, None]

    def add(i):
        try:
            results[i] = srv.add_matrix(csrs[i], pks[i])
        except acc.AccelesError as e:
            results[i] = e

    th = [threading.Thread(target=add, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    ok = [r for r in results if isinstance(r, int)]
    bad = [r for r in results if isinstance(r, acc.AccelesError)]
    # never both: the second reservation or its resize meets the first's bytes
    assert len(ok) <= 1 and len(ok) + len(bad) == 2, results
    assert all(e.code == -8 for e in bad)  # ACC_ERR_BUDGET
    assert srv.stats()["device_bytes"] <= base + int(1.5 * one)
    if not ok:  # both refused while each held its estimate: room for one remains
        ok = [srv.add_matrix(csrs[0], pks[0])]
    with pytest.raises(acc.AccelesError, match="budget"):
        srv.add_matrix(csrs[1], pks[1])
    x = acc.generate_queries(1, 2048, seed=1)[0]
    assert len(srv.wait(srv.submit(ok[0], x, 10), 10)) == 10
    srv.close()


def test_one_matrix_several_batches_in_flight(acc, gpu):
    """Every slot has its own workspace: a burst against one matrix runs on several slots at once, and
    each result equals the one the request gets alone and the oracle's."""
    csr = acc.CSRMatrix.generate(200000, 500, deg=(40, 100), seed=17)
    pk = csr.pack(bits=8, split=32)
    srv = acc.Server(max_batch=16, slots=4, ck=512)
    mid = srv.add_matrix(csr, pk)
    X = acc.generate_queries(160, 500, seed=18)
    ks = [[1, 8, 16, 32, 50, 75, 100][i % 7] for i in range(len(X))]
    alone = [srv.wait(srv.submit(mid, X[i], ks[i]), ks[i]) for i in range(0, len(X), 20)]
    srv.reset_stats()
    tickets = [srv.submit(mid, X[i], ks[i]) for i in range(len(X))]
    burst = [srv.wait(t, k) for t, k in zip(tickets, ks)]
    st = srv.stats()
    assert st["peak_inflight"] > 1, st
    assert st["completed"] == len(X)
    for j, i in enumerate(range(0, len(X), 20)):
        np.testing.assert_array_equal(burst[i], alone[j])
    ptr, idx, val = csr.arrays()
    q, _ = O.quantize(val, 8, False)
    gm = acc.GpuMatrix(csr, pk)
    for i in range(0, len(X), 40):
        _, oi, ov = O.two_stage(ptr, idx, val, q, X[i], 500, gm.info.group, 32, 512, ks[i])
        np.testing.assert_array_equal(burst[i]["index"], oi)
        np.testing.assert_array_equal(burst[i]["value"], ov)
    srv.close()


def test_requests_admitted_by_memory(acc, gpu):
    """max_request_bytes bounds the host memory of waiting and unread requests: submissions past it are
    refused with ACC_ERR_BUDGET, and reading results frees room again."""
    csr = acc.CSRMatrix.generate(20000, 500, deg=(40, 100), seed=19)
    pk = csr.pack(bits=8, split=32)
    budget = 16 * (500 * 4 + 100 * 8 + 512)
    srv = acc.Server(max_batch=8, slots=1, max_request_bytes=budget)
    mid = srv.add_matrix(csr, pk)
    x = acc.generate_queries(1, 500, seed=20)[0]
    held, refused = [], 0
    for _ in range(40):
        try:
            held.append(srv.submit(mid, x, 100))
        except acc.AccelesError as e:
            assert e.code == -8 and "request memory" in str(e)
            refused += 1
    # each request holds at least its query and its K results
    assert 1 <= len(held) <= budget // (500 * 4 + 100 * 8) and refused == 40 - len(held) > 0
    assert 0 < srv.stats()["request_bytes"] <= budget
    first = srv.wait(held[0], 100)
    for t in held[1:]:
        np.testing.assert_array_equal(srv.wait(t, 100), first)
    assert srv.stats()["request_bytes"] == 0
    srv.wait(srv.submit(mid, x, 100), 100)  # room again
    srv.close()
