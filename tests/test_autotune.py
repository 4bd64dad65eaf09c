"""Autotuner over SPLIT, CK and the queries per launch (acceles_amd/autotune.py)."""
import numpy as np
import pytest

import oracle_lib as O
from acceles_amd.autotune import TunePoint, choose, valid_cks, tune_lowbit


def test_valid_cks_follow_the_path_limits():
    # M = ck / split in [1, 64], ck <= 4096, k <= split * M
    assert valid_cks(32, 100, (64, 128, 256, 512, 1024, 2048, 4096)) == [128, 256, 512, 1024, 2048]
    assert valid_cks(64, 100, (64, 128, 256, 4096, 8192)) == [128, 256, 4096]
    assert valid_cks(16, 100, (96, 100, 128)) == [128]  # 96 and 100 keep 6 per partition: 96 < k
    assert valid_cks(8, 1, (4, 8, 16, 512, 1024)) == [8, 16, 512]


def test_choose_takes_the_fastest_point_above_the_floor():
    P = lambda s, c, b, r, q: TunePoint(s, c, b, r, q, 1e6 * b / q if q else 0.0, r >= 0.99)
    pts = [P(16, 256, 64, 0.985, 0.0), P(16, 512, 64, 0.995, 900.0), P(32, 512, 128, 0.999, 1200.0),
           P(32, 1024, 128, 1.0, 1200.0), P(64, 256, 256, 0.97, 0.0)]
    best = choose(pts, 0.99)
    assert (best.split, best.ck, best.batch) == (32, 1024, 128)  # equal speed: the larger CK
    assert choose(pts, 0.9995).ck == 1024
    assert choose(pts, 1.01) is None
    assert choose([], 0.5) is None


@pytest.mark.gpu
def test_tuned_configuration_meets_the_floor_and_matches_the_oracle(acc, gpu):
    rows, cols, bits, k = 100000, 300, 8, 100
    csr = acc.CSRMatrix.generate(rows, cols, deg=(20, 60), seed=31)
    best, pts = tune_lowbit(csr, bits=bits, k=k, splits=(8, 32, 64), cks=(128, 512, 1024), batches=(32, 128),
                            nq=32, floor=0.99, reps=2)
    assert best is not None and best.recall >= 0.99
    assert len(pts) == sum(len(valid_cks(s, k, (128, 512, 1024))) for s in (8, 32, 64)) * 2
    assert all(p.qps > 0 for p in pts if p.feasible) and all(p.qps == 0 for p in pts if not p.feasible)
    # results at the chosen configuration are the oracle's
    pk = csr.pack(bits=bits, split=best.split)
    g = acc.GpuMatrix(csr, pk)
    ptr, idx, val = csr.arrays()
    qv, _ = O.quantize(val, bits, False)
    X = acc.generate_queries(4, cols, seed=32)
    for x in X:
        _, oi, ov = O.two_stage(ptr, idx, val, qv, x, cols, g.info.group, best.split, best.ck, k)
        res = g.topk_lowbit_host(x, k, best.ck)
        np.testing.assert_array_equal(res["index"], oi)
        np.testing.assert_array_equal(res["value"], ov)
