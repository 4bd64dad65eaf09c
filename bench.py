#!/usr/bin/env python3
"""Benchmark: low-bit two-stage Top-K SpMV vs the plain FP32 CSR Top-K on one MI355X.

Workload (default): 400,000 x 500 matrix, row nnz uniform in [40, 100],
values |N(0,1)| L2-normalised per row, 8-bit unsigned packets, SPLIT=32,
CK=512, K=100; queries dense N(0,1) of length 500 (synthetic, seeded).
One step = one batch of 100 queries (the reference's num_test) through the
batched low-bit two-stage path (one matrix pass per block of queries).
value = queries/s over the timed steps, inputs resident in HBM.

Also reported: the same queries one at a time (single-query path), the plain
FP32 CSR Top-K on the same GPU, per-launch times from HIP events on the
query stream, the rooflines of the batched scoring kernel and of the
single-query stage-1 kernel, precision/recall@K vs an FP64 exact Top-K, and
the reference CPU path (FP32 SpMV + heap Top-K, one thread) on a bounded
sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
INT8_PEAK_TOPS = 5000.0     # MI355X dense int8 matrix peak (no sparsity)
KS = [1, 8, 16, 32, 50, 75, 100]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=400000)
    ap.add_argument("--cols", type=int, default=500)
    ap.add_argument("--deg", type=int, nargs=2, default=[40, 100])
    ap.add_argument("--bits", type=int, default=8)
    ap.add_argument("--split", type=int, default=32)
    ap.add_argument("--ck", type=int, default=512)
    ap.add_argument("--k", type=int, default=100)
    ap.add_argument("--batch", type=int, default=100)
    ap.add_argument("--batches-per-step", type=int, default=10,
                    help="embedding: batches of --batch queries per timed step")
    ap.add_argument("--cpu-seconds", type=float, default=16.0, help="bound on the CPU baseline sample (four legs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="embedding: skip the 20,000,000x1024 slice-pass roofline")
    ap.add_argument("--inflight", type=int, default=0, choices=[0, 1, 2, 3, 4],
                    help="batches (embedding, batched) or queries (large-single) in flight on as many streams; "
                         "0: 3 batches (embedding), 2 batches (batched), 2 queries (large-single)")
    ap.add_argument("--batch-mode", choices=["auto", "packets", "dense"], default="auto",
                    help="scoring layout of the batched path")
    ap.add_argument("--workload", choices=["embedding", "large-single", "batched", "serve", "copy", "resparse",
                                           "bitwidth"],
                    default="embedding",
                    help="embedding: 400,000x500 batched (default line); large-single: 20,000,000x1024 single "
                         "queries; batched: 2,000,000x900, 256 queries of mixed K per graph launch; serve: three "
                         "resident matrices under Poisson arrivals; copy: 1 GiB device-to-device copies, the "
                         "calibration run for the HBM traffic counters; resparse: the ReSparse keep-ratio search on "
                         "the embedding matrix, then single queries on the pruned packets; bitwidth: "
                         "precision/recall@K against bytes/nnz for 4-, 6- and 8-bit packets")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="serve: seconds per offered rate")
    ap.add_argument("--deg-dist", choices=["uniform", "gamma"], default="uniform",
                    help="large-single: row lengths U[10,30] or Gamma(0.5, 40) (mean 20)")
    return ap.parse_args()


def cpu_threads():
    """(cores visible to this process, threads the CPU legs use): the affinity mask, capped by
    OMP_NUM_THREADS when that is set (the box's CPU share)."""
    aff = len(os.sched_getaffinity(0))
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return aff, (min(aff, env) if env > 0 else aff)


def cpu_baseline(csr, X, k, seconds, bits, split, ck, group):
    """The reference CPU path on this box's host cores, via the oracle's C restatement
    (oracle/oracle.c), each leg over as many queries as fit in seconds/4:
      fp32_1t    FP32 CSR SpMV then a min-heap Top-k (cpu_mkl_baseline.cpp:249-275), one thread;
      fp32_all   the same on every core: rows split across threads, per-thread heaps merged (the
                 reference's multicore mode, cpu_mkl_baseline.cpp:195-245, with its race removed);
      lowbit_1t / lowbit_all
                 the low-bit two-stage path (quantify_util.py:173-196, 354-367): INT32 scores of the
                 quantized matrix, Top-(ck/split) per partition, FP32 rescoring, Top-k.
    The line's value is fp32_all (the CPU path on all cores)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    ptr, idx, val = csr.arrays()
    nproc, nt = cpu_threads()
    qv, _ = O.quantize(val, bits, False)
    cols = X.shape[1]
    legs = {}

    def run(name, fn):
        n, t0 = 0, time.perf_counter()
        while n < 4 or time.perf_counter() - t0 < seconds / 4:
            fn(X[n % len(X)])
            n += 1
        legs[name] = {"queries_per_s": round(n / (time.perf_counter() - t0), 3), "queries": n}

    run("fp32_1t", lambda x: O.top_k_spmv_f32(ptr, idx, val, x, k))
    # the reference times the SpMV and the heap apart (cpu_mkl_baseline.cpp:260-273): the SpMV alone in the
    # same sequential row order, the heap as the rest of the query
    m = min(8, legs["fp32_1t"]["queries"])
    t0 = time.perf_counter()
    for i in range(m):
        O.spmv_f32(ptr, idx, val, X[i % len(X)], 1)
    spmv_ms = (time.perf_counter() - t0) / m * 1e3
    legs["fp32_1t"]["spmv_ms"] = round(spmv_ms, 3)
    legs["fp32_1t"]["topk_ms"] = round(max(0.0, 1e3 / legs["fp32_1t"]["queries_per_s"] - spmv_ms), 3)
    run("fp32_all", lambda x: O.top_k_spmv_f32_par(ptr, idx, val, x, k, nt))
    run("lowbit_1t", lambda x: O.two_stage_par(ptr, idx, val, qv, x, cols, group, split, ck, k, 1))
    run("lowbit_all", lambda x: O.two_stage_par(ptr, idx, val, qv, x, cols, group, split, ck, k, nt))
    return {"value": legs["fp32_all"]["queries_per_s"], "unit": "queries/s", "cores": nt, "nproc": nproc,
            "kind": "port",
            "sample": f"queries of the same workload, about {seconds / 4:.0f} s per leg, through the reference CPU "
                      f"path semantics restated in oracle/oracle.c; value = FP32 CSR SpMV + min-heap Top-{k} "
                      f"(cpu_mkl_baseline.cpp:249-275) on {nt} threads",
            "legs": legs}


def timed(stream, fn, steps, warmup):
    return timed_multi([stream], fn, steps, warmup)


def timed_multi(streams, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    for s in streams:
        s.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    for s in streams:
        s.synchronize()
    return time.perf_counter() - t0


def exact_topk(g, acc, x_dev, rows, K, stream):
    """FP64 exact Top-K rows (score desc, row asc) of one device query."""
    y64 = acc.DeviceBuffer((rows,), np.float64)
    g.spmv_fp64(x_dev, y64, stream)
    stream.synchronize()
    yy = y64.numpy()
    top = np.argpartition(-yy, K)[:K + 1]
    return top[np.lexsort((top, -yy[top]))][:K]


def precision_at(got, exact):
    return [len(set(got[:k]) & set(exact[:k])) / k for k in KS]


# the sources each profiled kernel is built from: a counter file counts only while they are unchanged
KERNEL_SOURCES = {
    "pk_scores": ["acceles_amd/csrc/k_pkmfma.hpp", "acceles_amd/csrc/k_pkmfma_a.hip", "acceles_amd/csrc/k_pkmfma_b.hip",
                  "acceles_amd/csrc/k_pkmfma_c.hip", "acceles_amd/csrc/k_pkmfma_d.hip", "acceles_amd/csrc/k_pkmfma_e.hip",
                  "acceles_amd/csrc/k_pkmfma_f.hip", "acceles_amd/csrc/k_pkmfma_g.hip",
                  "acceles_amd/csrc/k_pkmfma_h.hip",
                  "acceles_amd/csrc/kernels_common.hpp", "acceles_amd/csrc/device_common.hpp"],
    "lowbit_slicemax": ["acceles_amd/csrc/k_lowbit.hip", "acceles_amd/csrc/kernels_common.hpp",
                        "acceles_amd/csrc/device_common.hpp"],
}


def sources_digest(kernel):
    """sha256 over the kernel's source files (in KERNEL_SOURCES order)."""
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES[kernel]:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def pmc_traffic(kernel, rows, cols, bits):
    """HBM bytes per launch of `kernel` on this workload, from rocprofv3 FETCH_SIZE and WRITE_SIZE passes
    scaled by the same counters' reading for a copy kernel of known size; None when not profiled, or when
    the file was measured on other kernel sources (its sources_sha256 differs from this tree's)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{rows}x{cols}_{bits}bit_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if kernel in KERNEL_SOURCES and d.get("sources_sha256") != sources_digest(kernel):
        log(f"[bench] {os.path.basename(path)}: measured on other sources of {kernel}; traffic not reported")
        return None
    return d.get("hbm_bytes_per_launch")


def line_common(a, metric, value, unit, ms, dtype, workload, extra_cfg):
    return {"metric": metric, "value": round(value, 2), "unit": unit, "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic",
            "config": dict({"workload": workload, "parallelism": "single GPU"}, **extra_cfg)}


def copy_peak(acc, reps=8):
    """Read + write GB/s of 1 GiB device-to-device copies: a measured streaming figure beside the 8 TB/s
    published peak."""
    n = 1 << 28
    src = acc.DeviceBuffer((n,), np.float32)
    dst = acc.DeviceBuffer((n,), np.float32)
    dst.copy_from(src)
    acc.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_from(src)
    acc.synchronize()
    dt = time.perf_counter() - t0
    del src, dst
    return 2 * 4 * n * reps / dt / 1e9


def c3_slicemax(acc, stream, reps=20):
    """The single-query slice pass at HBM scale: 20,000,000 x 1024, row nnz U[10,30], |N(0,1)|, 8-bit
    packets, SPLIT=64 (0.93 GB of packets, beyond the 256 MB MALL).  lowbit_slicemax is timed as a burst
    of reps back-to-back launches between two HIP events on the query stream; achieved = (packets + the
    FP32 query) / average launch.  traffic: rocprofv3 counter bytes per launch from profiles/, if there."""
    rows, cols, K, CK = 20_000_000, 1024, 100, 1024
    t0 = time.perf_counter()
    csr = acc.CSRMatrix.generate(rows, cols, deg=(10, 30), l2norm=False, seed=0xACCE1E5 + 2)
    pk = csr.pack(bits=8, split=64)
    g = acc.GpuMatrix(csr, pk, device=0)
    del csr
    x = acc.DeviceBuffer.from_numpy(acc.generate_queries(1, cols, seed=0xACCE1E5 + 3))
    out = acc.DeviceBuffer((K, 2), np.int32)
    g.topk_lowbit(x, K, CK, out, stream)
    stream.synchronize()
    ms = g.time_lowbit_stage(x, K, CK, 0, reps, stream)
    nbytes = pk.info.stream_bytes + cols * 4
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"[bench] 20000000x1024 slice pass: {ms * 1e3:.1f}us for {nbytes / 1e9:.3f} GB = {gbs:.0f} GB/s "
        f"(setup {time.perf_counter() - t0:.1f}s)")
    traffic = pmc_traffic("lowbit_slicemax", rows, cols, 8)
    copy_gbs = copy_peak(acc)
    log(f"[bench] measured copy peak: {copy_gbs:.0f} GB/s (1 GiB device-to-device, read + write)")
    res = {"bound": "hbm", "kernel": "lowbit_slicemax", "workload": f"{rows}x{cols}, row nnz U[10,30], 8-bit, "
           "SPLIT=64, one query", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "measured_copy_gbs": round(copy_gbs, 1),
           "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
           "traffic_ratio": None if traffic is None else round(traffic / nbytes, 4),
           "algorithmic_bytes_per_launch": int(nbytes), "avg_launch_us": round(ms * 1e3, 2),
           "bytes_per_nnz_packed": round(pk.info.bytes_per_nnz, 4)}
    del g, pk
    return res


def bench_large_single(a, acc, stream):
    """20,000,000 x 1024, average degree 20, 8-bit, SPLIT=64, CK=1024, K=100: one query per step."""
    rows, cols, K, CK = 20_000_000, 1024, 100, 1024
    t0 = time.perf_counter()
    if a.deg_dist == "uniform":
        csr = acc.CSRMatrix.generate(rows, cols, deg=(10, 30), l2norm=False, seed=0xACCE1E5 + 2)
    else:
        csr = acc.CSRMatrix.generate(rows, cols, deg=(0.5, 40.0), deg_kind=acc.DEG_GAMMA, l2norm=False,
                                     seed=0xACCE1E5 + 2)
    pk = csr.pack(bits=8, split=64)
    g = acc.GpuMatrix(csr, pk, device=0)
    log(f"[bench] setup {time.perf_counter() - t0:.1f}s: nnz={csr.nnz} packed={pk.info.stream_bytes / 1e9:.3f}GB "
        f"({pk.info.bytes_per_nnz:.3f} B/nnz) fp32={g.info.csr_bytes / 1e9:.3f}GB")
    nq = max(a.steps, 8)
    X = acc.generate_queries(nq, cols, seed=0xACCE1E5 + 3)
    Xd = acc.DeviceBuffer.from_numpy(X)
    out = acc.DeviceBuffer((nq, K, 2), np.int32)
    xs = [Xd.at(i * cols) for i in range(nq)]
    os_ = [out.at(i * K * 2) for i in range(nq)]
    i = [0]

    def step():
        g.topk_lowbit(xs[i[0] % nq], K, CK, os_[i[0] % nq], stream)
        i[0] += 1
    dt_seq = timed(stream, step, a.steps, a.warmup)
    # queries in flight: HIP-graph plans with their own workspaces on as many streams, so one query's
    # selection kernels (a few dozen workgroups) overlap the next query's streaming pass
    NF = a.inflight or 2
    streams = [stream] + [acc.Stream(0) for _ in range(1, NF)]
    plans = [acc.Plan(g, acc.PLAN_LOWBIT, xs[j], [K], CK, os_[j], K) for j in range(NF)]

    def step2():
        j = i[0] % NF
        plans[j].launch(streams[j])
        i[0] += 1
    dt = timed_multi(streams, step2, a.steps, a.warmup)
    for p_ in plans:
        p_.close()
    qps = a.steps / dt
    g.set_profiling(True)
    fp = []
    for j in range(4):
        g.topk_fp32(xs[j % nq], K, os_[0], stream)
        fp.append(g.stage_times())
    g.set_profiling(False)
    # per-dispatch events carry the dispatch overhead; the low-bit stages are also timed as bursts of back-to-back
    # launches (events around the burst), which is how they run inside back-to-back queries
    lb_ms = np.array([g.time_lowbit_stage(xs[0], K, CK, st_, 20, stream) for st_ in range(3)])
    fp_ms = np.mean(fp, axis=0)
    s1_bytes = pk.info.stream_bytes + cols * 4
    s1_gbs = s1_bytes / (lb_ms[0] * 1e-3) / 1e9
    fqps = 1e3 / float(np.sum(fp_ms))
    for j in range(nq):
        g.topk_lowbit(xs[j], K, CK, os_[j], stream)
    stream.synchronize()
    res = out.numpy()
    prec = [precision_at(res[j, :, 1], exact_topk(g, acc, xs[j], rows, K, stream)) for j in range(min(nq, 8))]
    precision = dict(zip(map(str, KS), np.round(np.mean(prec, axis=0), 5).tolist()))
    log(f"[bench] single query: {qps:.0f} q/s with {NF} in flight ({a.steps / dt_seq:.0f} q/s one at a time); "
        f"slicemax {lb_ms[0] * 1e3:.1f}us ({s1_gbs:.0f} GB/s), select "
        f"{lb_ms[1] * 1e3:.1f}us, final {lb_ms[2] * 1e3:.1f}us; fp32 CSR Top-K {fqps:.0f} q/s; precision {precision}")
    line = line_common(a, "single-query low-bit Top-100 SpMV queries/s", qps, "queries/s", dt * 1e3 / a.steps, "u8",
                       f"single-query low-bit Top-100: {rows}x{cols}, row nnz {a.deg_dist} mean 20, |N(0,1)|, 8-bit "
                       f"packets, SPLIT=64, CK=1024; one query per step, {NF} in flight (graph plans on as many streams)",
                       {"rows": rows, "cols": cols, "nnz": csr.nnz})
    line.update({
        "roofline": {"bound": "hbm", "kernel": "lowbit_slicemax", "achieved": round(s1_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(s1_gbs / HBM_PEAK_GBS, 4), "traffic": pmc_traffic("lowbit_slicemax", rows, cols, 8),
                     "algorithmic_bytes_per_launch": int(s1_bytes), "avg_launch_us": round(float(lb_ms[0]) * 1e3, 2)},
        "cpu_baseline": None,
        "one_at_a_time_qps": round(a.steps / dt_seq, 2),
        "lowbit_launch_us": [round(float(v) * 1e3, 2) for v in lb_ms],
        "fp32_launch_us": [round(float(v) * 1e3, 2) for v in fp_ms], "fp32_csr_gpu_qps": round(fqps, 2),
        "bytes_per_nnz_packed": round(pk.info.bytes_per_nnz, 4), "precision_at_k": precision,
    })
    print(json.dumps(line), flush=True)


def bench_batched(a, acc, stream):
    """2,000,000 x 900, row nnz U[30,90], L2-normalised; 256 queries per step with K drawn from
    {1,8,16,32,50,75,100}; CK=512, SPLIT from --split (default 32); autotuned layout; one HIP-graph launch per step."""
    rows, cols, B, CK = 2_000_000, 900, 256, 512
    t0 = time.perf_counter()
    csr = acc.CSRMatrix.generate(rows, cols, deg=(30, 90), seed=0xACCE1E5 + 4)
    # partition count (SPLIT) and scoring layout tuned under recall@100 >= 0.99 (acceles_amd/autotune.py)
    from acceles_amd.autotune import tune_lowbit
    best, tpts = tune_lowbit(csr, bits=a.bits, k=100, splits=(16, 32, 64), cks=(CK,), batches=(B,), nq=64,
                             floor=0.99, reps=3, log=log)
    split = best.split if best is not None else a.split
    a.split = split
    pk = csr.pack(bits=a.bits, split=split)
    g = acc.GpuMatrix(csr, pk, device=0)
    log(f"[bench] setup {time.perf_counter() - t0:.1f}s: nnz={csr.nnz} packed={pk.info.stream_bytes / 1e6:.0f}MB, "
        f"tuned SPLIT={split}")
    pts = g.autotune(CK, 100, B)
    log("[bench] autotune: " + ", ".join(f"nq={p.nq}:{['auto', 'packets', 'dense', 'single'][p.layout]}"
                                         f"({p.us_single:.0f}/{p.us_sparse:.0f}/{p.us_dense:.0f}us)" for p in pts))
    rng = np.random.default_rng(0xACCE1E5 + 5)
    ks = rng.choice(KS, B).astype(np.int32)
    X = acc.generate_queries(B, cols, seed=0xACCE1E5 + 6)
    Xd = acc.DeviceBuffer.from_numpy(X)
    out = acc.DeviceBuffer((B, 100, 2), np.int32)
    plan = acc.Plan(g, acc.PLAN_LOWBIT_BATCH, Xd, ks, CK, out, 100)
    dt1 = timed(stream, lambda: plan.launch(stream), a.steps, a.warmup)
    res = out.numpy()
    # batches in flight: one plan (own workspace) per stream; two beat three and four at this size
    # (a 256-query tile pass keeps the chip busy; a third launch only contends for it)
    NF = a.inflight or 2
    streams = [stream] + [acc.Stream(0) for _ in range(1, NF)]
    xs2 = [Xd] + [acc.DeviceBuffer.from_numpy(acc.generate_queries(B, cols, seed=0xACCE1E5 + 60 + j))
                  for j in range(1, NF)]
    outs2 = [out] + [acc.DeviceBuffer((B, 100, 2), np.int32) for _ in range(1, NF)]
    plans = [plan] + [acc.Plan(g, acc.PLAN_LOWBIT_BATCH, xs2[j], ks, CK, outs2[j], 100) for j in range(1, NF)]
    it = [0]

    def pipe_step():
        j = it[0] % NF
        plans[j].launch(streams[j])
        it[0] += 1
    dt = timed_multi(streams, pipe_step, a.steps, a.warmup)
    qps = a.steps * B / dt
    same = bool(np.array_equal(out.numpy(), res))
    prec = []
    for j in range(0, B, 32):
        k = int(ks[j])
        ex = exact_topk(g, acc, Xd.at(j * cols), rows, 100, stream)
        prec.append(len(set(res[j, :k, 1]) & set(ex[:k])) / k)
    log(f"[bench] batched: {qps:.0f} q/s with {NF} launches in flight ({dt * 1e3 / a.steps:.3f} ms per {B}-query "
        f"graph launch; {a.steps * B / dt1:.0f} q/s one at a time); mean precision@K {np.mean(prec):.4f}")
    line = line_common(a, "batched low-bit Top-K queries/s (mixed K)", qps, "queries/s", dt * 1e3 / a.steps,
                       "u8" if a.bits == 8 else f"u{a.bits}",
                       f"low-bit two-stage, batched: {rows}x{cols}, row nnz U[30,90], L2-normalised, {a.bits}-bit, "
                       f"SPLIT={a.split}, CK=512, {B} queries per HIP-graph launch, K from {{1,8,16,32,50,75,100}}, "
                       f"{NF} launches in flight (a plan per stream)",
                       {"rows": rows, "cols": cols, "nnz": csr.nnz, "split": a.split, "queries_per_step": B})
    line.update({"roofline": None, "cpu_baseline": None, "mean_precision_at_k": round(float(np.mean(prec)), 5),
                 "tuned": [p.as_dict() for p in tpts],
                 "one_at_a_time_qps": round(a.steps * B / dt1, 2), "in_flight_equals_one_at_a_time": same,
                 "autotune": [{"nq": p.nq, "layout": int(p.layout), "us_single": round(p.us_single, 1),
                               "us_packets": round(p.us_sparse, 1), "us_dense": round(p.us_dense, 1)} for p in pts]})
    for p_ in plans:
        p_.close()
    print(json.dumps(line), flush=True)


def bench_serve(a, acc, stream):
    """Three resident matrices (1,000,000x500 deg 60, 5,000,000x900 deg 40, 10,000,000x1024 deg 20;
    |N(0,1)|, 8-bit), Poisson arrivals at 2,000 / 10,000 / 50,000 q/s, random matrix, K uniform in [1,100]."""
    specs = [(1_000_000, 500, 60), (5_000_000, 900, 40), (10_000_000, 1024, 20)]
    t0 = time.perf_counter()
    srv = acc.Server(max_batch=256, slots=3, ck=512, autotune=True)
    # probe queries per matrix with their FP64 exact Top-100 (scored on the GPU before the matrix is handed over)
    NP = 8
    probes = []
    for i, (r, c, d) in enumerate(specs):
        csr = acc.CSRMatrix.generate(r, c, deg=(d // 2, d + d // 2), l2norm=False, seed=0xACCE1E5 + 10 + i)
        X = acc.generate_queries(NP, c, seed=0xACCE1E5 + 20 + i)
        Xd = acc.DeviceBuffer.from_numpy(X)
        g_exact = acc.GpuMatrix(csr, None)
        probes.append((X, [exact_topk(g_exact, acc, Xd.at(j * c), r, 100, stream) for j in range(NP)]))
        del g_exact, Xd
        srv.add_matrix(csr, csr.pack(bits=8, split=32))
        del csr
    log(f"[bench] serve setup {time.perf_counter() - t0:.1f}s, {srv.stats()['device_bytes'] / 1e9:.2f} GB resident")
    runs = {}
    for rate in (2000, 10000, 50000):
        st = srv.bench_poisson(rate, a.serve_seconds, kmax=100, seed=rate)
        runs[str(rate)] = {k: (round(v, 2) if isinstance(v, float) else v) for k, v in st.items()}
        log(f"[bench] serve {rate} q/s offered: {st['qps']:.0f} q/s done, p50 {st['p50_us']:.0f}us, "
            f"p99 {st['p99_us']:.0f}us, mean batch {st['mean_batch']:.1f}")
    # precision@K of the served results against the FP64 exact Top-K; recall@K = share of probes with all K found
    prec = []
    for mid, (X, exact) in enumerate(probes):
        tickets = [srv.submit(mid, X[j], 100) for j in range(len(X))]
        for j, t in enumerate(tickets):
            prec.append(precision_at(srv.wait(t, 100)["index"], exact[j]))
    prec = np.array(prec)
    precision = dict(zip(map(str, KS), np.round(prec.mean(axis=0), 5).tolist()))
    recall = dict(zip(map(str, KS), np.round((prec == 1.0).mean(axis=0), 5).tolist()))
    log(f"[bench] serve precision@K {precision}; recall@K (all K found) {recall} over {len(prec)} probes")
    srv.close()
    top = runs["50000"]
    line = line_common(a, "serving queries/s with p50/p99 latency (continuous batching)", top["qps"], "queries/s",
                       1e3 / max(top["qps"], 1e-9), "u8",
                       "serving: 1,000,000x500 deg 60 + 5,000,000x900 deg 40 + 10,000,000x1024 deg 20, 8-bit, CK=512, "
                       "Poisson arrivals, random matrix, K uniform in [1,100]", {"offered_qps": [2000, 10000, 50000]})
    line.update({"roofline": None, "cpu_baseline": None, "runs": runs, "precision_at_k": precision,
                 "recall_all_found_at_k": recall})
    print(json.dumps(line), flush=True)


def topk_rows(y, k):
    """Top-k rows of a score vector: score descending, row ascending."""
    top = np.argpartition(-y, k)[:k + 1] if len(y) > k else np.arange(len(y))
    top = top[np.lexsort((top, -y[top]))]
    return top[:k]


def bench_bitwidth(a, acc, stream):
    """Precision/recall@K against bytes/nnz per value width (4-bit unsigned, 6-bit signed -- the
    reference's B of types.hpp:148 -- and 8-bit unsigned) on the embedding matrix (400,000x500,
    SPLIT=32, CK=512) and the 20,000,000x1024 matrix (SPLIT=64, CK=1024).  Per width: the stage-1
    approximate Top-K from the INT32 scores alone (the quantity of plot_quantize_with_4to6bit.py:14-16)
    and the two-stage result after FP32 rescoring, both against the FP64 exact Top-K."""
    configs = [("400000x500", 400000, 500, (40, 100), True, 32, 512, 64),
               ("20000000x1024", 20_000_000, 1024, (10, 30), False, 64, 1024, 16)]
    widths = [(4, False), (6, True), (8, False)]
    K = 100
    table = []
    for name, rows, cols, deg, l2, split, ck, nq in configs:
        t0 = time.perf_counter()
        csr = acc.CSRMatrix.generate(rows, cols, deg=deg, l2norm=l2, seed=0xACCE1E5 + (1 if l2 else 2))
        X = acc.generate_queries(nq, cols, seed=0xACCE1E5 + 9)
        Xd = acc.DeviceBuffer.from_numpy(X)
        xs = [Xd.at(i * cols) for i in range(nq)]
        exact = None
        for bits, signed in widths:
            pk = csr.pack(bits=bits, split=split, is_signed=signed)
            g = acc.GpuMatrix(csr, pk, device=0)
            if exact is None:
                exact = [exact_topk(g, acc, xs[i], rows, K, stream) for i in range(nq)]
            y = acc.DeviceBuffer((rows,), np.int32)
            out = acc.DeviceBuffer((K, 2), np.int32)
            p1, p2 = [], []
            for i in range(nq):
                g.scores_lowbit(xs[i], y, stream)
                stream.synchronize()
                p1.append(precision_at(topk_rows(y.numpy().astype(np.int64), K), exact[i]))
                g.topk_lowbit(xs[i], K, ck, out, stream)
                stream.synchronize()
                p2.append(precision_at(out.numpy()[:, 1], exact[i]))
            ms = g.time_lowbit_stage(xs[0], K, ck, 0, 20, stream)
            row = {"matrix": name, "bits": bits, "signed": signed, "bytes_per_nnz": round(pk.info.bytes_per_nnz, 4),
                   "packed_bytes": int(pk.info.stream_bytes), "slicemax_us": round(ms * 1e3, 2),
                   "stage1_precision_at_k": dict(zip(map(str, KS), np.round(np.mean(p1, axis=0), 4).tolist())),
                   "two_stage_precision_at_k": dict(zip(map(str, KS), np.round(np.mean(p2, axis=0), 4).tolist())),
                   "two_stage_all_found_at_k": dict(zip(map(str, KS), np.round(np.mean(np.array(p2) == 1.0, axis=0),
                                                                               4).tolist())),
                   "queries": nq}
            table.append(row)
            log(f"[bench] {name} {bits}-bit{' signed' if signed else ''}: {row['bytes_per_nnz']} B/nnz, slicemax "
                f"{row['slicemax_us']}us; stage-1 P@K {row['stage1_precision_at_k']}; two-stage P@K "
                f"{row['two_stage_precision_at_k']}")
            del g, pk, y
        log(f"[bench] {name}: {time.perf_counter() - t0:.1f}s")
        del csr
    c2_4 = next(r for r in table if r["matrix"] == "400000x500" and r["bits"] == 4)
    line = line_common(a, "two-stage precision@100 at 4-bit (embedding matrix)", c2_4["two_stage_precision_at_k"]["100"],
                       "precision", 0.0, "u4",
                       "precision/recall@K vs bytes/nnz for 4-, 6- (signed) and 8-bit packets on 400000x500 "
                       "(SPLIT=32, CK=512) and 20000000x1024 (SPLIT=64, CK=1024); K=1..100 vs FP64 exact",
                       {})
    line.update({"roofline": None, "cpu_baseline": None, "table": table})
    print(json.dumps(line), flush=True)


def bench_copy(a, acc, stream):
    """1 GiB device-to-device copies: rocprofv3 FETCH_SIZE / WRITE_SIZE of this copy kernel against its known
    size (1 GiB read, 1 GiB written per copy) give the counters' scale on this GPU."""
    n = 1 << 28
    src = acc.DeviceBuffer((n,), np.float32)
    dst = acc.DeviceBuffer((n,), np.float32)
    dst.copy_from(src)
    acc.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        dst.copy_from(src)
    acc.synchronize()
    dt = time.perf_counter() - t0
    gbs = 2 * 4 * n * a.steps / dt / 1e9
    log(f"[bench] copy: {a.steps} x 1 GiB, {gbs:.0f} GB/s read+write")
    print(json.dumps(line_common(a, "device copy GB/s (read+write)", gbs, "GB/s", dt * 1e3 / a.steps, "u8",
                                 "1 GiB device-to-device copy (traffic-counter calibration)", {"bytes": 4 * n})),
          flush=True)


def bench_resparse(a, acc, stream):
    """The embedding workload's matrix: ReSparse keep ratio searched at 0.999 of the unpruned convergence metric
    (100 queries, its SPLIT and CK), then single queries on the pruned packets, rescored on the unpruned FP32 rows."""
    from acceles_amd.resparse import search_keep_ratio
    K, CK = a.k, a.ck
    csr = acc.CSRMatrix.generate(a.rows, a.cols, deg=tuple(a.deg), seed=0xACCE1E5 + 1)
    t0 = time.perf_counter()
    res = search_keep_ratio(csr, bits=a.bits, split=a.split, ck=CK, nq=100, accuracy_threshold=0.999, log=log)
    t_search = time.perf_counter() - t0
    log(f"[bench] resparse: keep {res['keep']:.3f} after {len(res['history'])} epochs ({t_search:.1f}s)")
    nq = 64
    X = acc.generate_queries(nq, a.cols, seed=0xACCE1E5 + 9)
    Xd = acc.DeviceBuffer.from_numpy(X)
    xs = [Xd.at(i * a.cols) for i in range(nq)]
    out = acc.DeviceBuffer((nq, K, 2), np.int32)
    os_ = [out.at(i * K * 2) for i in range(nq)]
    g_exact = acc.GpuMatrix(csr, None)
    exact = [exact_topk(g_exact, acc, xs[i], a.rows, K, stream) for i in range(nq)]
    del g_exact
    runs = {}
    for name, m in (("unpruned", csr), ("pruned", csr.resparse(res["keep"]))):
        pk = m.pack(bits=a.bits, split=a.split)
        g = acc.GpuMatrix(csr, pk)  # pruned packets for stage 1, unpruned FP32 rows for the rescoring
        i = [0]

        def step():
            g.topk_lowbit(xs[i[0] % nq], K, CK, os_[i[0] % nq], stream)
            i[0] += 1
        dt = timed(stream, step, a.steps, a.warmup)
        for j in range(nq):
            g.topk_lowbit(xs[j], K, CK, os_[j], stream)
        stream.synchronize()
        res_ = out.numpy()
        prec = np.mean([precision_at(res_[j, :, 1], exact[j]) for j in range(nq)], axis=0)
        s1 = g.time_lowbit_stage(xs[0], K, CK, 0, 40, stream)
        runs[name] = {"nnz": m.nnz, "stream_bytes": pk.info.stream_bytes, "qps": round(a.steps / dt, 2),
                      "slicemax_us": round(s1 * 1e3, 2),
                      "precision_at_k": dict(zip(map(str, KS), np.round(prec, 5).tolist()))}
        log(f"[bench] {name}: nnz {m.nnz}, packets {pk.info.stream_bytes / 1e6:.1f} MB, {a.steps / dt:.0f} q/s, "
            f"slicemax {s1 * 1e3:.1f}us, precision@K {runs[name]['precision_at_k']}")
        del g
    line = line_common(a, "single-query low-bit Top-100 queries/s after ReSparse", runs["pruned"]["qps"],
                       "queries/s", 1e3 / runs["pruned"]["qps"], "u8",
                       f"ReSparse keep ratio searched at 0.999 convergence (100 queries), then single queries: "
                       f"{a.rows}x{a.cols}, row nnz U[{a.deg[0]},{a.deg[1]}], {a.bits}-bit, SPLIT={a.split}, CK={CK}",
                       {"rows": a.rows, "cols": a.cols})
    line.update({"roofline": None, "cpu_baseline": None, "keep": res["keep"], "search_seconds": round(t_search, 2),
                 "history": [(round(k, 3), round(c, 5), n) for k, c, n in res["history"]], "runs": runs})
    print(json.dumps(line), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the row-sharded multi-GPU "
              "design is described in DESIGN.md and is not measured here.", file=sys.stderr)
        sys.exit(2)
    import acceles_amd as acc
    if a.workload != "embedding":
        stream = acc.Stream(0)
        {"large-single": bench_large_single, "batched": bench_batched, "serve": bench_serve,
         "copy": bench_copy, "resparse": bench_resparse, "bitwidth": bench_bitwidth}[a.workload](a, acc, stream)
        return

    assert acc.device_count() > 0, "no GPU visible"
    stream = acc.Stream(0)

    t0 = time.perf_counter()
    csr = acc.CSRMatrix.generate(a.rows, a.cols, deg=tuple(a.deg), seed=0xACCE1E5 + 1)
    pk = csr.pack(bits=a.bits, split=a.split)
    g = acc.GpuMatrix(csr, pk, device=0)
    g.set_batch_mode({"auto": acc.BATCH_AUTO, "packets": acc.BATCH_SPARSE, "dense": acc.BATCH_DENSE}[a.batch_mode])
    X = acc.generate_queries(a.batch, a.cols, seed=0xACCE1E5 + 2)
    Xd = acc.DeviceBuffer.from_numpy(X)
    log(f"[bench] setup {time.perf_counter() - t0:.1f}s: nnz={csr.nnz} packed={pk.info.stream_bytes / 1e6:.1f}MB "
        f"({pk.info.bytes_per_nnz:.3f} B/nnz) fp32={g.info.csr_bytes / 1e6:.1f}MB group={g.info.group} "
        f"items={g.info.work_items}")

    B, K, CK = a.batch, a.k, a.ck
    ks = np.full(B, K, np.int32)
    out = acc.DeviceBuffer((B, K, 2), np.int32)
    out1 = acc.DeviceBuffer((B, K, 2), np.int32)
    xs = [Xd.at(i * a.cols) for i in range(B)]
    outs1 = [out1.at(i * K * 2) for i in range(B)]

    BPS = max(1, a.batches_per_step)

    def batch_step():
        g.topk_lowbit_batch(Xd, ks, CK, out, K, stream)

    def batch_steps():
        for _ in range(BPS):
            batch_step()

    def single_step():
        for i in range(B):
            g.topk_lowbit(xs[i], K, CK, outs1[i], stream)

    def fp32_step():
        for i in range(B):
            g.topk_fp32(xs[i], K, outs1[i], stream)

    # ---- timed region: batched low-bit two-stage, one batch at a time (eager calls) ----
    dt1 = timed(stream, batch_steps, a.steps, a.warmup)
    qps1 = a.steps * BPS * B / dt1
    res = out.numpy()

    # ---- timed region: batches in flight (HIP-graph plans with their own workspaces on as many
    # streams): one batch's selection, rescoring and Top-K kernels overlap the next batch's tile pass ----
    NF = a.inflight or 3
    Xs = [Xd] + [acc.DeviceBuffer.from_numpy(acc.generate_queries(B, a.cols, seed=0xACCE1E5 + 7 + j))
                 for j in range(1, NF)]
    outs2 = [acc.DeviceBuffer((B, K, 2), np.int32) for _ in range(NF)]
    streams = [stream] + [acc.Stream(0) for _ in range(1, NF)]
    plans = [acc.Plan(g, acc.PLAN_LOWBIT_BATCH, Xs[j], ks, CK, outs2[j], K) for j in range(NF)]
    it = [0]

    def pipe_step():
        for _ in range(BPS):
            j = it[0] % NF
            plans[j].launch(streams[j])
            it[0] += 1
    dt = timed_multi(streams, pipe_step, a.steps, a.warmup)
    qps = a.steps * BPS * B / dt
    ms_step = dt * 1e3 / a.steps
    pipelined_same = bool(np.array_equal(outs2[0].numpy(), res))
    for p_ in plans:
        p_.close()
    log(f"[bench] low-bit two-stage, batched: {qps:.0f} queries/s with {NF} batches in flight ({ms_step:.3f} ms per "
        f"{BPS * B}-query step); one batch at a time {qps1:.0f} q/s; in-flight results equal eager: {pipelined_same}")

    # ---- the other batched layout, for comparison ----
    other_qps = None
    if g.info.dense_ok:
        mode = g.info.batch_mode
        g.set_batch_mode(acc.BATCH_SPARSE if mode != acc.BATCH_SPARSE else acc.BATCH_DENSE)
        other_qps = a.steps * BPS * B / timed(stream, batch_steps, a.steps, 1)
        g.set_batch_mode(mode)
        log(f"[bench] batched, other layout: {other_qps:.0f} queries/s")

    # ---- the same queries one at a time, and the plain FP32 CSR Top-K ----
    s_steps = max(2, a.steps // 4)
    single_qps = s_steps * B / timed(stream, single_step, s_steps, 1)
    res1 = out1.numpy()
    fqps = s_steps * B / timed(stream, fp32_step, s_steps, 1)
    log(f"[bench] low-bit single-query {single_qps:.0f} q/s; fp32 CSR Top-K {fqps:.0f} q/s; "
        f"batched/fp32 {qps1 / fqps:.2f}x (both one at a time)")
    same = bool(np.array_equal(res, res1))

    # ---- per-launch times from HIP events on the query stream ----
    g.set_profiling(True)
    bt, fp = [], []
    for i in range(10):
        batch_step()
        bt.append(g.stage_times())
    for i in range(40):
        g.topk_fp32(xs[i % B], K, outs1[0], stream)
        fp.append(g.stage_times())
    g.set_profiling(False)
    bt_ms, fp_ms = np.mean(bt, axis=0), np.mean(fp, axis=0)
    lb_ms = np.array([g.time_lowbit_stage(xs[0], K, CK, st_, 40, stream) for st_ in range(3)])
    kept = pk.info.nnz_kept
    nslots = pk.info.nslices * 64
    dense = len(bt_ms) == 4
    if dense:
        # matrix cores over the packets: prep, scores, select + FP32 rescoring, final
        names = ["prep", "scores", "select_rescore", "final"]
        t_sc = bt_ms[1]
        cols_pad = (a.cols + 31) // 32 * 32
        # algorithmic bytes: the packed stream and the queries (the image packets
        # pk_scores reads instead are a wider encoding of the same entries)
        bs_bytes = pk.info.stream_bytes + B * a.cols * 4
        bs_ops = 2.0 * nslots * cols_pad * ((B + 31) // 32 * 32)
        sc_kernel = "pk_scores"
    else:
        names = ["scores", "select", "final"]
        t_sc = bt_ms[0]
        bs_bytes = pk.info.stream_bytes + B * a.cols * 4 + B * nslots * 4
        bs_ops = 2.0 * kept * B
        sc_kernel = "batch_scores"
    bs_gbs = bs_bytes / (t_sc * 1e-3) / 1e9
    bs_tops = bs_ops / (t_sc * 1e-3) / 1e12
    s1_bytes = pk.info.stream_bytes + a.cols * 4
    s1_gbs = s1_bytes / (lb_ms[0] * 1e-3) / 1e9
    fp32_spmv_gbs = g.info.csr_bytes / (fp_ms[0] * 1e-3) / 1e9
    log("[bench] batch: " + ", ".join(f"{n} {v * 1e3:.1f}us" for n, v in zip(names, bt_ms)) +
        f"; {sc_kernel} {bs_gbs:.0f} GB/s, {bs_tops:.1f} TOP/s int8")
    log(f"[bench] single: slicemax {lb_ms[0] * 1e3:.1f}us ({s1_gbs:.0f} GB/s), select {lb_ms[1] * 1e3:.1f}us, "
        f"final {lb_ms[2] * 1e3:.1f}us; fp32 spmv {fp_ms[0] * 1e3:.1f}us ({fp32_spmv_gbs:.0f} GB/s), "
        f"select {fp_ms[1] * 1e3:.1f}us + {fp_ms[2] * 1e3:.1f}us")

    # ---- accuracy vs FP64 exact Top-K ----
    y64 = acc.DeviceBuffer((a.rows,), np.float64)
    prec, allfound, ndcgs = [], [], []
    for i in range(B):
        g.spmv_fp64(xs[i], y64, stream)
        stream.synchronize()
        yy = y64.numpy()
        top = np.argpartition(-yy, K)[:K + 1]
        exact = top[np.lexsort((top, -yy[top]))][:K]
        got = res[i, :, 1]
        p = [len(set(got[:k]) & set(exact[:k])) / k for k in KS]
        prec.append(p)
        allfound.append([float(x == 1.0) for x in p])
        # score NDCG (quantify_util.py:313-329): relevance = exact FP64 score
        ndcgs.append([acc.ndcg_score(exact[:k], yy[exact[:k]], got[:k])[0] for k in KS])
    precision = dict(zip(map(str, KS), np.round(np.mean(prec, axis=0), 5).tolist()))
    recall_all = dict(zip(map(str, KS), np.round(np.mean(allfound, axis=0), 5).tolist()))
    ndcg_at_k = dict(zip(map(str, KS), np.round(np.mean(ndcgs, axis=0), 6).tolist()))
    log(f"[bench] precision@K {precision}; NDCG@K {ndcg_at_k}")

    # ---- HBM traffic from rocprof counters, when profiled for this workload ----
    traffic = pmc_traffic(sc_kernel, a.rows, a.cols, a.bits)

    c3 = None if a.no_c3 else c3_slicemax(acc, stream)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(csr, X, K, a.cpu_seconds, a.bits, a.split, CK, g.info.group)
        log(f"[bench] cpu baseline ({cpu['cores']} of {cpu['nproc']} cores): " +
            ", ".join(f"{n} {v['queries_per_s']:.1f} q/s" for n, v in cpu["legs"].items()))

    line = {
        "metric": "queries/s + HBM GB/s, low-bit Top-100 SpMV vs FP32 CSR at recall@100>=0.99",
        "value": round(qps, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"low-bit two-stage Top-{K}, batched: {a.rows}x{a.cols}, row nnz U[{a.deg[0]},{a.deg[1]}], "
                        f"|N(0,1)| L2-normalised, {a.bits}-bit packets, SPLIT={a.split}, CK={CK}, "
                        f"{BPS} batches of {B} N(0,1) queries per step, {NF} batches in flight (graph plans on "
                        f"as many streams)",
            "rows": a.rows, "cols": a.cols, "nnz": csr.nnz, "bits": a.bits, "split": a.split, "ck": CK,
            "k": K, "queries_per_step": BPS * B, "batch": B, "parallelism": "single GPU",
        },
        "roofline": ({
            "bound": "hbm",
            "kernel": sc_kernel,
            "achieved": round(bs_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(bs_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_ratio": None if traffic is None else round(traffic / bs_bytes, 4),
            "algorithmic_bytes_per_launch": int(bs_bytes),
            "avg_launch_us": round(float(t_sc) * 1e3, 2),
            "mfma_int8_tops": round(bs_tops, 2),
            "mfma_frac": round(bs_tops / INT8_PEAK_TOPS, 4),
        } if dense else {
            "bound": "mfma",
            "kernel": sc_kernel,
            "achieved": round(bs_tops, 2),
            "peak": INT8_PEAK_TOPS,
            "unit": "TFLOP/s",
            "frac": round(bs_tops / INT8_PEAK_TOPS, 5),
            "traffic": traffic,
            "algorithmic_ops_per_launch": int(bs_ops),
            "avg_launch_us": round(float(t_sc) * 1e3, 2),
            "note": "int8 MACs on VALU dot4 (one useful byte lane); peak is the dense int8 matrix-core figure",
        }),
        "roofline_hbm_slice_pass": c3,
        "roofline_single_query_slicemax": {
            "bound": "hbm", "kernel": "lowbit_slicemax", "achieved": round(s1_gbs, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(s1_gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": s1_bytes,
            "avg_launch_us": round(float(lb_ms[0]) * 1e3, 2),
            "note": f"{s1_bytes / 1e6:.0f} MB of packets: resident in the 256 MB MALL between launches",
        },
        "cpu_baseline": cpu,
        "one_batch_at_a_time_qps": round(qps1, 2),
        "in_flight_equals_eager": pipelined_same,
        "single_query_qps": round(single_qps, 2),
        "fp32_csr_gpu_qps": round(fqps, 2),
        "lowbit_batched_vs_fp32_speedup": round(qps1 / fqps, 3),  # both one at a time
        "batched_equals_single": same,
        "batch_launch_us": dict(zip(names, [round(float(v) * 1e3, 2) for v in bt_ms])),
        "batch_scoring": "matrix cores, slices expanded from the image packets in LDS" if dense else "sliced packets (dot4)",
        "image_packet_bytes": g.info.dense_bytes,
        "batched_other_layout_qps": None if other_qps is None else round(other_qps, 2),
        "lowbit_launch_us": [round(float(v) * 1e3, 2) for v in lb_ms],
        "fp32_launch_us": [round(float(v) * 1e3, 2) for v in fp_ms],
        "fp32_spmv_gbs": round(fp32_spmv_gbs, 1),
        "bytes_per_nnz_packed": round(pk.info.bytes_per_nnz, 4),
        "bytes_per_nnz_fp32": round(g.info.csr_bytes / csr.nnz, 4),
        "precision_at_k": precision,
        "recall_all_found_at_k": recall_all,
        "ndcg_at_k": ndcg_at_k,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
