"""Autotuner of the batched low-bit path over SPLIT, CK and the queries per launch, under a recall floor.

The reference fixes the partition count and the candidate count at build time (src/common/types.hpp:39
SPLIT, :54 the partitions of the FPGA kernel, :81 CK) or takes them as script arguments
(datasets/search_best_resparse_ratio.py:28-32).  Here they are searched per matrix:

  * SPLIT changes the packets (rows are dealt to SPLIT partitions when packing), so every SPLIT is packed
    and uploaded once;
  * CK (with M = CK / SPLIT candidates kept per partition, 1 <= M <= 64, K <= CK <= 4096) is a query
    argument;
  * the queries per launch (B) set how many queries share one pass over the packets; the scoring layout
    for each B is the handle's own choice (acc_gpu_autotune: one query per pass, vector-ALU packets or
    matrix cores), so the layout is tuned too.

Each (SPLIT, CK) is first scored for accuracy: recall@K (the share of the FP64 exact Top-K that the
two-stage path returns, K = k) over sample queries; those below the floor are dropped.  The survivors
are timed at every B, and the fastest (queries/s) wins; ties go to the larger CK (more recall margin).
"""
import time
from dataclasses import dataclass, asdict

import numpy as np


@dataclass
class TunePoint:
    split: int
    ck: int
    batch: int
    recall: float          # mean recall@k over the sample queries
    qps: float             # queries/s at this batch size (0 when not timed: below the floor)
    us_per_batch: float
    feasible: bool

    def as_dict(self):
        return asdict(self)


def valid_cks(split, k, cks):
    """The CK values the low-bit path accepts for this SPLIT and k."""
    out = []
    for ck in cks:
        m = ck // split
        if ck >= split and 1 <= m <= 64 and split * m <= 4096 and k <= split * m:
            out.append(ck)
    return out


def choose(points, floor):
    """The fastest point at or above the recall floor (ties: larger CK, then smaller SPLIT); None if none."""
    ok = [p for p in points if p.feasible and p.recall >= floor]
    if not ok:
        return None
    return max(ok, key=lambda p: (p.qps, p.ck, -p.split))


def _exact(g, acc, xs, rows, k, stream):
    y = acc.DeviceBuffer((rows,), np.float64)
    out = []
    for x in xs:
        g.spmv_fp64(x, y, stream)
        stream.synchronize()
        yy = y.numpy()
        top = np.argpartition(-yy, k)[:k + 1] if rows > k else np.arange(rows)
        out.append(top[np.lexsort((top, -yy[top]))][:k])
    return out


def tune_lowbit(csr, bits=8, k=100, splits=(16, 32, 64), cks=(256, 512, 1024), batches=(32, 64, 128, 256),
                nq=64, floor=0.99, is_signed=False, shuffle_seed=0, reps=5, seed=0x7E57, device=0, log=None):
    """Searches SPLIT x CK x queries-per-launch on csr.  Returns (best TunePoint or None, [TunePoint]).

    Recall is measured on nq seeded N(0,1) queries against the FP64 exact Top-k; throughput on B copies
    of those queries per launch, reps launches after one warm-up, with the layout acc_gpu_autotune picks
    for that B."""
    import acceles_amd as acc

    say = log or (lambda *a: None)
    cols, rows = csr.cols, csr.rows
    X = acc.generate_queries(nq, cols, seed=seed)
    bmax = max(batches)
    Xb = np.tile(X, ((bmax + nq - 1) // nq, 1))[:bmax]
    Xd = acc.DeviceBuffer.from_numpy(X, device=device)
    Xbd = acc.DeviceBuffer.from_numpy(Xb, device=device)
    stream = acc.Stream(device)
    ks = np.full(bmax, k, np.int32)
    exact = None
    points = []
    for split in splits:
        if split > rows:
            continue
        options = valid_cks(split, k, cks)
        if not options:
            continue
        pk = csr.pack(bits=bits, split=split, is_signed=is_signed, shuffle_seed=shuffle_seed)
        g = acc.GpuMatrix(csr, pk, device=device)
        if exact is None:
            exact = _exact(g, acc, [Xd.at(i * cols) for i in range(nq)], rows, k, stream)
        g.autotune(max(options), k, bmax)  # the layout per batch size
        out = acc.DeviceBuffer((bmax, k, 2), np.int32, device=device)
        for ck in options:
            g.topk_lowbit_batch(Xd, ks[:nq], ck, out, k, stream)
            stream.synchronize()
            got = out.numpy()[:nq, :, 1]
            recall = float(np.mean([len(set(got[i].tolist()) & set(exact[i].tolist())) / k for i in range(nq)]))
            feasible = recall >= floor
            for b in batches:
                us, qps = 0.0, 0.0
                if feasible:
                    g.topk_lowbit_batch(Xbd, ks[:b], ck, out, k, stream)
                    stream.synchronize()
                    t0 = time.perf_counter()
                    for _ in range(reps):
                        g.topk_lowbit_batch(Xbd, ks[:b], ck, out, k, stream)
                    stream.synchronize()
                    us = (time.perf_counter() - t0) / reps * 1e6
                    qps = b / (us * 1e-6)
                points.append(TunePoint(split, ck, b, round(recall, 5), round(qps, 1), round(us, 2), feasible))
            say(f"[tune] split={split} ck={ck}: recall@{k} {recall:.4f}" +
                ("" if feasible else f" < {floor}") +
                ("".join(f", B={p.batch} {p.qps:.0f} q/s" for p in points[-len(batches):]) if feasible else ""))
        del g, pk
    return choose(points, floor), points
