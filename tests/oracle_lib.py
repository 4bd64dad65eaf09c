"""ctypes bindings of the checker libraries (test infrastructure only).

- oracle/liboracle.so: the C restatement of the path (oracle/oracle.c);
- oracle/_ref/libref_gold.so: the reference's own header-only gold code,
  built in this container from the reference tree (absent elsewhere, so
  tests that need it skip).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libref_gold.so")

_I32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_U32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_I64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_F32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_F64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_U64 = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")

_orc = None
_ref = None


def oracle() -> C.CDLL:
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", ORACLE_DIR, "all"], check=True, capture_output=True)
        L = C.CDLL(ORACLE_SO)
        L.orc_coo2csr.argtypes = [C.c_int32, C.c_int32, C.c_int64, _I32, _I32, _F32, C.c_int, _I64, _I32, _F32]
        L.orc_quantize.argtypes = [C.c_int64, _F32, C.c_int, C.c_int, C.POINTER(C.c_float), _I32]
        L.orc_quantize.restype = None
        L.orc_spmv_f32_grouped.argtypes = [C.c_int32, _I64, _I32, _F32, _F32, C.c_int, _F32]
        L.orc_spmv_f32_grouped.restype = None
        L.orc_spmv_f64.argtypes = [C.c_int32, _I64, _I32, _F32, _F64, _F64]
        L.orc_spmv_f64.restype = None
        L.orc_spmv_i32.argtypes = [C.c_int32, _I64, _I32, _I32, _I32, _I32]
        L.orc_topk.argtypes = [C.c_int64, _F64, C.c_int, _I32, _F64]
        L.orc_topk.restype = None
        L.orc_split_select.argtypes = [C.c_int32, _I32, C.c_int, C.c_int, _I32, _I32]
        L.orc_split_select_order.argtypes = [C.c_int32, _I32, _I32, C.c_int, C.c_int, _I32, _I32]
        L.orc_two_stage_order.argtypes = [C.c_int32, C.c_int32, _I64, _I32, _F32, _I32, _F32, C.c_int, C.c_int,
                                          C.c_int, C.c_int, _I32, _I32, _I32, _F32]
        L.orc_rescore_topk.argtypes = [_I64, _I32, _F32, _F32, C.c_int, C.c_int, _I32, C.c_int, _I32, _F32]
        L.orc_rescore_topk.restype = None
        L.orc_two_stage.argtypes = [C.c_int32, C.c_int32, _I64, _I32, _F32, _I32, _F32, C.c_int, C.c_int, C.c_int,
                                    C.c_int, _I32, _I32, _F32]
        L.orc_precision.argtypes = [_I32, C.c_int, _I32, C.c_int, _I32, _F64]
        L.orc_precision.restype = None
        L.orc_xdecode.argtypes = [C.c_uint32, C.c_int, C.c_int32, _I32]
        L.orc_xdecode.restype = None
        L.orc_ultracsr_entries.argtypes = [C.c_int, C.c_int]
        L.orc_ultracsr_pack.argtypes = [C.c_int64, _I32, _I32, _I32, C.c_int32, C.c_int, C.c_int, _U64, C.c_int64]
        L.orc_ultracsr_pack.restype = C.c_int64
        L.orc_ultracsr_unpack.argtypes = [_U64, C.c_int64, C.c_int, C.c_int, C.c_int32, _I32, _I32, _I32]
        L.orc_ultracsr_unpack.restype = None
        L.orc_top_k_spmv_f32.argtypes = [C.c_int32, _I64, _I32, _F32, _F32, C.c_int, _F32, _I32, _F32]
        L.orc_top_k_spmv_f32.restype = None
        L.orc_resparse_mask.argtypes = [C.c_int32, C.c_int32, _I64, _F32, C.c_double, C.c_int, _U8]
        L.orc_resparse_mask.restype = None
        L.orc_ndcg_score.argtypes = [_I32, _F64, C.c_int, _I32, C.c_int]
        L.orc_ndcg_score.restype = C.c_double
        L.orc_top_k_spmv_f32_par.argtypes = [C.c_int32, _I64, _I32, _F32, _F32, C.c_int, C.c_int, _F32, _I32, _F32]
        L.orc_top_k_spmv_f32_par.restype = None
        L.orc_two_stage_par.argtypes = [C.c_int32, C.c_int32, _I64, _I32, _F32, _I32, _F32, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, _I32, _I32, _F32]
        _orc = L
    return _orc


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        L.ref_readMtx.argtypes = [C.c_char_p, C.c_int, C.c_int, _I32, _I32, _I32, _F32, C.c_longlong]
        L.ref_readMtx.restype = C.c_longlong
        L.ref_coo2csr.argtypes = [C.c_uint, C.c_uint, C.c_uint, _U32, _U32, _F32, C.c_int, _U32, _U32, _F32]
        L.ref_coo2csr.restype = None
        L.ref_spmv_gold_f32.argtypes = [C.c_uint, _U32, _U32, _F32, _F32, _F32]
        L.ref_spmv_gold_f32.restype = None
        L.ref_spmv_coo_gold_top_k_f32.argtypes = [C.c_uint, _U32, _U32, _F32, _F32, C.c_int, _U32, _F32]
        L.ref_spmv_coo_gold_top_k_f32.restype = None
        L.ref_spmv_coo_gold_top_k_2_i32.argtypes = [C.c_uint, _U32, _U32, _I32, _I32, C.c_int, _U32, _I32]
        L.ref_spmv_coo_gold_top_k_2_i32.restype = None
        L.ref_sort_tuples_i32.argtypes = [C.c_uint, _U32, _I32]
        L.ref_sort_tuples_i32.restype = None
        L.ref_sort_tuples_f32.argtypes = [C.c_uint, _U32, _F32]
        L.ref_sort_tuples_f32.restype = None
        L.ref_ndcg.argtypes = [_I32, _I32, C.c_int]
        L.ref_ndcg.restype = C.c_double
        L.ref_edit_distance.argtypes = [_I32, C.c_int, _I32, C.c_int]
        L.ref_edit_distance.restype = C.c_uint
        L.ref_count_errors.argtypes = [_I32, _I32, C.c_int]
        L.ref_count_errors.restype = C.c_int
        _ref = L
    return _ref


# ---------------------------------------------------------------------------
# numpy-friendly wrappers of the oracle
# ---------------------------------------------------------------------------

def coo2csr(row, col, val, nrows, ncols, sort_tuples=False):
    row, col = np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float32)
    ptr = np.empty(nrows + 1, np.int64)
    idx = np.empty(len(row), np.int32)
    vo = np.empty(len(row), np.float32)
    rc = oracle().orc_coo2csr(nrows, ncols, len(row), row, col, val, int(sort_tuples), ptr, idx, vo)
    if rc != 0:
        raise ValueError("index out of range")
    return ptr, idx, vo


def quantize(v, bits, symmetric):
    v = np.ascontiguousarray(v, np.float32).ravel()
    q = np.empty(v.size, np.int32)
    s = C.c_float()
    oracle().orc_quantize(v.size, v, bits, int(symmetric), C.byref(s), q)
    return q, s.value


def spmv_f32(ptr, idx, val, x, G):
    n = len(ptr) - 1
    y = np.empty(n, np.float32)
    oracle().orc_spmv_f32_grouped(n, np.ascontiguousarray(ptr, np.int64), np.ascontiguousarray(idx, np.int32),
                                  np.ascontiguousarray(val, np.float32), np.ascontiguousarray(x, np.float32), G, y)
    return y


def spmv_f64(ptr, idx, val, x):
    n = len(ptr) - 1
    y = np.empty(n, np.float64)
    oracle().orc_spmv_f64(n, np.ascontiguousarray(ptr, np.int64), np.ascontiguousarray(idx, np.int32),
                          np.ascontiguousarray(val, np.float32), np.ascontiguousarray(x, np.float64), y)
    return y


def spmv_i32(ptr, idx, qv, qx):
    n = len(ptr) - 1
    y = np.empty(n, np.int32)
    rc = oracle().orc_spmv_i32(n, np.ascontiguousarray(ptr, np.int64), np.ascontiguousarray(idx, np.int32),
                               np.ascontiguousarray(qv, np.int32), np.ascontiguousarray(qx, np.int32), y)
    if rc != 0:
        raise OverflowError("int32 score overflow")
    return y


def topk(y, k):
    y = np.ascontiguousarray(y, np.float64)
    idx = np.empty(k, np.int32)
    val = np.empty(k, np.float64)
    oracle().orc_topk(len(y), y, k, idx, val)
    return idx, val


def split_select(y, split, ck):
    y = np.ascontiguousarray(y, np.int32)
    rows = np.empty(ck, np.int32)
    sc = np.empty(ck, np.int32)
    n = oracle().orc_split_select(len(y), y, split, ck, rows, sc)
    if n < 0:
        raise ValueError("bad split arguments")
    return rows[:n], sc[:n]


def split_select_order(y, order, split, ck):
    """split_select with the rows dealt to partitions in `order` (Random-CSR)."""
    y = np.ascontiguousarray(y, np.int32)
    order = np.ascontiguousarray(order, np.int32)
    rows = np.empty(ck, np.int32)
    sc = np.empty(ck, np.int32)
    n = oracle().orc_split_select_order(len(y), y, order, split, ck, rows, sc)
    if n < 0:
        raise ValueError("bad split/ck")
    return rows[:n], sc[:n]


def rescore_topk(ptr, idx, val, x, G, cand, k):
    oi = np.empty(k, np.int32)
    ov = np.empty(k, np.float32)
    oracle().orc_rescore_topk(np.ascontiguousarray(ptr, np.int64), np.ascontiguousarray(idx, np.int32),
                              np.ascontiguousarray(val, np.float32), np.ascontiguousarray(x, np.float32), G,
                              len(cand), np.ascontiguousarray(cand, np.int32), k, oi, ov)
    return oi, ov


def two_stage(ptr, idx, val, qv, x, cols, G, split, ck, k):
    n = len(ptr) - 1
    cand = np.empty(ck, np.int32)
    oi = np.empty(k, np.int32)
    ov = np.empty(k, np.float32)
    m = oracle().orc_two_stage(n, cols, np.ascontiguousarray(ptr, np.int64), np.ascontiguousarray(idx, np.int32),
                               np.ascontiguousarray(val, np.float32), np.ascontiguousarray(qv, np.int32),
                               np.ascontiguousarray(x, np.float32), G, split, ck, k, cand, oi, ov)
    if m < 0:
        raise ValueError("two-stage failed")
    return cand[:m], oi, ov


KS = np.array([1, 8, 16, 32, 50, 75, 100], np.int32)


def precision(exact, cand, ks=KS):
    out = np.empty(len(ks), np.float64)
    oracle().orc_precision(np.ascontiguousarray(exact, np.int32), len(cand), np.ascontiguousarray(cand, np.int32),
                           len(ks), np.ascontiguousarray(ks, np.int32), out)
    return out


def xdecode(flags, n=30, carry=0):
    x = np.empty(n, np.int32)
    oracle().orc_xdecode(flags, n, carry, x)
    return x


def ultracsr_pack(row, col, qval, last_row, cb, vb):
    n = oracle().orc_ultracsr_entries(cb, vb)
    cap = max(1, (len(row) + n - 1) // n)
    out = np.zeros(cap * 8, np.uint64)
    np_ = oracle().orc_ultracsr_pack(len(row), np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32),
                                     np.ascontiguousarray(qval, np.int32), last_row, cb, vb, out, cap)
    return out[: np_ * 8].reshape(np_, 8)


def ultracsr_unpack(packets, nnz, cb, vb, carry0=0):
    xr = np.empty(nnz, np.int32)
    c = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.int32)
    oracle().orc_ultracsr_unpack(np.ascontiguousarray(packets, np.uint64).ravel(), nnz, cb, vb, carry0, xr, c, v)
    return xr, c, v


def top_k_spmv_f32(ptr, idx, val, x, k):
    n = len(ptr) - 1
    y = np.empty(n, np.float32)
    oi = np.empty(k, np.int32)
    ov = np.empty(k, np.float32)
    oracle().orc_top_k_spmv_f32(n, ptr, idx, val, x, k, y, oi, ov)
    return oi, ov


def top_k_spmv_f32_par(ptr, idx, val, x, k, nthreads):
    """top_k_spmv on nthreads cores (the same results as top_k_spmv_f32)."""
    n = len(ptr) - 1
    y = np.empty(n, np.float32)
    oi = np.empty(k, np.int32)
    ov = np.empty(k, np.float32)
    oracle().orc_top_k_spmv_f32_par(n, ptr, idx, val, x, k, nthreads, y, oi, ov)
    return oi, ov


def two_stage_order(ptr, idx, val, qv, x, cols, G, split, ck, k, order):
    """two_stage over the partitions of a row order (Random-CSR)."""
    n = len(ptr) - 1
    cand = np.empty(ck, np.int32)
    oi = np.empty(k, np.int32)
    ov = np.empty(k, np.float32)
    m = oracle().orc_two_stage_order(n, cols, np.ascontiguousarray(ptr, np.int64), np.ascontiguousarray(idx, np.int32),
                                     np.ascontiguousarray(val, np.float32), np.ascontiguousarray(qv, np.int32),
                                     np.ascontiguousarray(x, np.float32), G, split, ck, k,
                                     np.ascontiguousarray(order, np.int32), cand, oi, ov)
    if m < 0:
        raise ValueError("two-stage failed")
    return cand[:m], oi, ov


def two_stage_par(ptr, idx, val, qv, x, cols, G, split, ck, k, nthreads):
    """two_stage on nthreads cores (the same results)."""
    n = len(ptr) - 1
    cand = np.empty(ck, np.int32)
    oi = np.empty(k, np.int32)
    ov = np.empty(k, np.float32)
    m = oracle().orc_two_stage_par(n, cols, np.ascontiguousarray(ptr, np.int64), np.ascontiguousarray(idx, np.int32),
                                   np.ascontiguousarray(val, np.float32), np.ascontiguousarray(qv, np.int32),
                                   np.ascontiguousarray(x, np.float32), G, split, ck, k, nthreads, cand, oi, ov)
    if m < 0:
        raise ValueError("two-stage failed")
    return cand[:m], oi, ov


def resparse_mask(ptr, val, cols, keep, mean_over_cells=False):
    ptr = np.ascontiguousarray(ptr, np.int64)
    mask = np.empty(int(ptr[-1]), np.uint8)
    oracle().orc_resparse_mask(len(ptr) - 1, cols, ptr, np.ascontiguousarray(val, np.float32), float(keep),
                               int(mean_over_cells), mask)
    return mask.astype(bool)


def ndcg_score(sw_idx, sw_val, hw_idx):
    sw_idx = np.ascontiguousarray(sw_idx, np.int32)
    hw_idx = np.ascontiguousarray(hw_idx, np.int32)
    return oracle().orc_ndcg_score(sw_idx, np.ascontiguousarray(sw_val, np.float64), len(sw_idx), hw_idx, len(hw_idx))
