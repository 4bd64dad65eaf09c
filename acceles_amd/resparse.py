"""ReSparse keep-ratio search (datasets/search_best_resparse_ratio.py:63-116 of the reference).

The reference bisects the share of each row's entries that ReSparse keeps: epoch 0 measures the
convergence metric of the unpruned matrix (quantify_util.py:413-437: the precision@K of the stage-1
candidates against the exact Top-K, weighted 1, 1/8, 1/16, 1/32, 1/50, 1/75, 1/100 over
K = 1, 8, 16, 32, 50, 75, 100); later epochs prune with `keep`, quantize, pack and measure again, moving
`high` down when the metric stays at or above accuracy_threshold x the unpruned metric and `low` up
otherwise, with keep = round((low + high) / 2, 3) until a keep value repeats; the result is `high`.

Here every epoch runs on the GPU: the exact Top-K is the FP64 SpMV of the unpruned matrix (computed
once, the queries are fixed by a seed rather than redrawn per epoch), the candidates are the low-bit
stage-1 candidates (acc_gpu_candidates_lowbit) of the pruned, packed matrix.
"""
import numpy as np

KS = (1, 8, 16, 32, 50, 75, 100)
WEIGHTS = (1.0, 1 / 8, 1 / 16, 1 / 32, 1 / 50, 1 / 75, 1 / 100)


def _exact_topk(y, k):
    top = np.argpartition(-y, k - 1)[:k] if len(y) > k else np.arange(len(y))
    return top[np.lexsort((top, -y[top]))][:k]


def convergence(precisions):
    """Weighted sum of the mean precision@K over queries (quantify_util.py:433-436)."""
    avg = np.mean(np.asarray(precisions, np.float64), axis=0)
    return float(np.dot(avg, WEIGHTS))


def search_keep_ratio(csr, bits=8, split=32, ck=512, nq=100, accuracy_threshold=0.999, mean_over_cells=False,
                      seed=0xACCE1E5, device=0, max_epochs=100, log=None):
    """Returns {"keep": ratio, "history": [(keep, convergence, nnz)], "threshold": ...}."""
    import acceles_amd as acc

    K = max(KS)
    X = acc.generate_queries(nq, csr.cols, seed=seed)
    Xd = acc.DeviceBuffer.from_numpy(X, device=device)
    exact = []
    g0 = acc.GpuMatrix(csr, None, device=device)
    y = acc.DeviceBuffer((csr.rows,), np.float64, device)
    for i in range(nq):
        g0.spmv_fp64(Xd.at(i * csr.cols), y)
        acc.synchronize()
        exact.append(_exact_topk(y.numpy(), K))
    del g0, y

    rows = acc.DeviceBuffer((ck,), np.int32, device)
    scores = acc.DeviceBuffer((ck,), np.int32, device)

    def measure(keep):
        m = csr if keep >= 1.0 else csr.resparse(keep, mean_over_cells)
        g = acc.GpuMatrix(None, m.pack(bits=bits, split=split), device=device)
        prec = []
        for i in range(nq):
            n = g.candidates_lowbit(Xd.at(i * csr.cols), ck, rows, scores)
            acc.synchronize()
            cand = set(rows.numpy()[:n].tolist())
            prec.append([len(set(exact[i][:k].tolist()) & cand) / k for k in KS])
        return convergence(prec), m.nnz

    def logged(keep):
        c, nnz = measure(keep)
        if log:
            log(f"resparse keep {keep:.3f}: convergence {c:.5f}, nnz {nnz}")
        return c, nnz

    return bisect_keep(logged, accuracy_threshold, max_epochs)


def bisect_keep(measure, accuracy_threshold, max_epochs=100):
    """The keep-ratio bisection of search_best_resparse_ratio.py:55-100 (both the embedding and the
    generated-matrix scripts run it): measure(keep) returns (convergence, ...); epoch 0 measures keep = 1
    and sets the threshold to accuracy_threshold times its convergence; epoch e >= 1 measures keep
    (0.5 first), moves `high` to keep when the convergence reaches the threshold and `low` otherwise,
    and the next keep is round((low + high) / 2, 3); the search stops when that value was measured
    before, or after max_epochs epochs.  Returns {"keep": high, "threshold", "history": [(keep, *measure
    (keep))], "converged": whether it stopped on a repeated value}."""
    first = tuple(measure(1.0))
    history = [(1.0,) + first]
    threshold = first[0] * accuracy_threshold
    low, high, keep = 0.0, 1.0, 0.5
    tried = {1.0}
    converged = False
    for _ in range(1, max_epochs):
        tried.add(keep)
        r = tuple(measure(keep))
        history.append((keep,) + r)
        if r[0] >= threshold:
            high = keep
        else:
            low = keep
        keep = round((low + high) / 2, 3)
        if keep in tried:
            converged = True
            break
    return {"keep": high, "threshold": threshold, "history": history, "converged": converged}
