"""Writes the golden fixtures of tests/golden/ from the reference's own gold code.

Run in a container that holds the reference tree, after
`make -C oracle ref` has built oracle/_ref/libref_gold.so from it:

    python tests/golden/make_golden.py

Every fixture is data only (inputs and the reference's outputs) in .npz files
(no pickles).  Inputs are small and seeded; the functions called are the
reference's readMtx, coo2csr, spmv_gold, spmv_coo_gold_top_k,
spmv_coo_gold_top_k_2 and sort_tuples (see oracle/ref_gold.cpp).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


def u32(a):
    return np.ascontiguousarray(a, np.uint32)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref_gold.so is missing: run `make -C oracle ref` first")
    R = O.ref()
    rng = np.random.default_rng(0xACCE1E5)

    # 1) readMtx on the committed sample file (comments, 1-based, unsorted)
    mtx = os.path.join(HERE, "sample.mtx")
    dims = np.zeros(3, np.int32)
    n = R.ref_readMtx(mtx.encode(), 0, 0, dims, np.zeros(1, np.int32), np.zeros(1, np.int32),
                      np.zeros(1, np.float32), 0)
    out = {}
    for sort in (0, 1):
        row = np.zeros(n, np.int32)
        col = np.zeros(n, np.int32)
        val = np.zeros(n, np.float32)
        R.ref_readMtx(mtx.encode(), 0, sort, dims, row, col, val, n)
        out[f"row_s{sort}"], out[f"col_s{sort}"], out[f"val_s{sort}"] = row, col, val
    out["dims"] = dims.copy()
    np.savez(os.path.join(HERE, "readmtx_sample.npz"), **out)

    # 2) coo2csr: unsorted COO with distinct (row, col) pairs and empty rows
    nrows, ncols, nnz = 40, 23, 300
    flat = rng.choice(nrows * ncols, size=nnz, replace=False)
    row, col = (flat // ncols).astype(np.int32), (flat % ncols).astype(np.int32)
    row[row % 7 == 3] = 0  # leave some rows empty (their entries move to row 0)
    keep = np.unique(row.astype(np.int64) * ncols + col, return_index=True)[1]
    keep.sort()
    row, col = row[keep], col[keep]
    val = rng.standard_normal(len(row)).astype(np.float32)
    res = {"row": row, "col": col, "val": val, "nrows": np.int32(nrows), "ncols": np.int32(ncols)}
    for sort in (0, 1):
        ptr = np.zeros(nrows + 1, np.uint32)
        idx = np.zeros(len(row), np.uint32)
        vo = np.zeros(len(row), np.float32)
        R.ref_coo2csr(nrows, ncols, len(row), u32(row), u32(col), val, sort, ptr, idx, vo)
        res[f"ptr_s{sort}"], res[f"idx_s{sort}"], res[f"val_s{sort}"] = ptr, idx, vo
    np.savez(os.path.join(HERE, "coo2csr.npz"), **res)

    # 3) spmv_gold (FP32, sequential multiply-add) on a CSR with varied rows
    ptr, idx, v = res["ptr_s1"], res["idx_s1"], res["val_s1"]
    x = rng.uniform(-1, 1, ncols).astype(np.float32)
    y = np.zeros(nrows, np.float32)
    R.ref_spmv_gold_f32(nrows, ptr.copy(), idx.copy(), v.copy(), x.copy(), y)
    np.savez(os.path.join(HERE, "spmv_gold.npz"), ptr=ptr, idx=idx, val=v, x=x, y=y)

    # 4) streaming Top-k golds on row-sorted COO without empty rows
    nrows, ncols = 500, 64
    rows, cols = [], []
    for r in range(nrows):
        d = int(rng.integers(3, 12))
        cs = np.sort(rng.choice(ncols, d, replace=False))
        rows += [r] * d
        cols += list(cs)
    rows = np.array(rows, np.uint32)
    cols = np.array(cols, np.uint32)
    fv = np.abs(rng.standard_normal(len(rows))).astype(np.float32)
    fx = rng.uniform(0.01, 1.0, ncols).astype(np.float32)  # positive: every score > 0
    k = 16
    ri = np.zeros(k, np.uint32)
    rv = np.zeros(k, np.float32)
    R.ref_spmv_coo_gold_top_k_f32(len(rows), rows, cols, fv, fx.copy(), k, ri, rv)
    R.ref_sort_tuples_f32(k, ri, rv)
    # integer gold, tie-free (distinct scores) and tied variants
    qv = rng.integers(1, 256, len(rows)).astype(np.int32)
    qx = rng.integers(1, 128, ncols).astype(np.int32)
    ck = 32
    ii = np.zeros(ck, np.uint32)
    iv = np.zeros(ck, np.int32)
    R.ref_spmv_coo_gold_top_k_2_i32(len(rows), rows, cols, qv, qx.copy(), ck, ii, iv)
    R.ref_sort_tuples_i32(ck, ii, iv)
    qv_t = rng.integers(1, 4, len(rows)).astype(np.int32)  # few distinct values: many ties
    qx_t = np.ones(ncols, np.int32)
    it = np.zeros(ck, np.uint32)
    vt = np.zeros(ck, np.int32)
    R.ref_spmv_coo_gold_top_k_2_i32(len(rows), rows, cols, qv_t, qx_t.copy(), ck, it, vt)
    R.ref_sort_tuples_i32(ck, it, vt)
    np.savez(os.path.join(HERE, "gold_topk.npz"), row=rows, col=cols, fval=fv, fx=fx, k=np.int32(k), f_idx=ri,
             f_val=rv, qval=qv, qx=qx, ck=np.int32(ck), i_idx=ii, i_val=iv, qval_tied=qv_t, qx_tied=qx_t,
             t_idx=it, t_val=vt)

    # 5) sort_tuples order on ties (value desc, higher index first)
    si = np.array([5, 1, 9, 3, 7, 2], np.uint32)
    sv = np.array([4, 9, 4, 1, 9, 4], np.int32)
    R.ref_sort_tuples_i32(len(si), si, sv)
    np.savez(os.path.join(HERE, "sort_tuples.npz"), in_idx=np.array([5, 1, 9, 3, 7, 2], np.uint32),
             in_val=np.array([4, 9, 4, 1, 9, 4], np.int32), out_idx=si, out_val=sv)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
