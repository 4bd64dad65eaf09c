"""acceles_amd -- MI355X-native Top-K SpMV (AccelES), Python binding.

A ctypes layer over ``libacceles.so`` (the C ABI in ``include/acceles.h``).
Names follow the reference: ``read_coo_matrix`` / ``coo_to_csr`` /
``top_k_spmv`` (src/cpu/cpu_mkl_baseline.cpp:43,79,249), ``readMtx`` /
``coo2csr`` (src/common/utils/utils.hpp:449,494), ``quantify_matrix`` /
``quantify_vector`` (src/resources/python/quantify/quantify_util.py:8,31),
``Element`` (cpu_mkl_baseline.cpp:16-19).

The library must be built first (``make -C acceles_amd`` or
``__graft_entry__.build()``); importing without it raises ImportError, and GPU
calls without a visible GPU raise ``AccelesError`` -- there is no CPU
fallback for any query.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libacceles.so")

ACC_OK = 0
ERRORS = {
    -1: "ACC_ERR_ARG", -2: "ACC_ERR_IO", -3: "ACC_ERR_FORMAT", -4: "ACC_ERR_RANGE",
    -5: "ACC_ERR_NOMEM", -6: "ACC_ERR_DEVICE", -7: "ACC_ERR_NODEVICE", -8: "ACC_ERR_BUDGET",
}
DEG_UNIFORM, DEG_GAMMA = 0, 1
VAL_ABSNORMAL, VAL_NORMAL, VAL_UNIFORM = 0, 1, 2
BATCH_AUTO, BATCH_SPARSE, BATCH_DENSE, BATCH_SINGLE = 0, 1, 2, 3
PLAN_LOWBIT, PLAN_LOWBIT_BATCH, PLAN_FP32 = 1, 2, 3


class AccelesError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Element(C.Structure):
    """One ranked result (cpu_mkl_baseline.cpp:16-19)."""
    _fields_ = [("value", C.c_float), ("index", C.c_int32)]


ELEMENT_DTYPE = np.dtype([("value", np.float32), ("index", np.int32)])


class PackOpts(C.Structure):
    _fields_ = [("bits", C.c_int), ("is_signed", C.c_int), ("split", C.c_int), ("col_bits", C.c_int),
                ("shuffle_seed", C.c_uint64)]


class PackedInfo(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("split", C.c_int32), ("bits", C.c_int32),
                ("col_bits", C.c_int32), ("entry_bits", C.c_int32), ("is_signed", C.c_int32),
                ("nnz", C.c_int64), ("nnz_kept", C.c_int64), ("nslices", C.c_int64),
                ("stream_bytes", C.c_int64), ("perm_bytes", C.c_int64), ("bytes_per_nnz", C.c_double),
                ("scale", C.c_float), ("shuffle_seed", C.c_uint64)]


class ServerOpts(C.Structure):
    _fields_ = [("device", C.c_int), ("memory_budget", C.c_int64), ("max_batch", C.c_int), ("slots", C.c_int),
                ("ck", C.c_int), ("max_pending", C.c_int), ("autotune", C.c_int), ("max_request_bytes", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("refused", C.c_int64), ("batches", C.c_int64),
                ("mean_batch", C.c_double), ("p50_us", C.c_double), ("p99_us", C.c_double), ("max_us", C.c_double),
                ("qps", C.c_double), ("device_bytes", C.c_int64), ("request_bytes", C.c_int64),
                ("peak_inflight", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TunePoint(C.Structure):
    _fields_ = [("nq", C.c_int32), ("layout", C.c_int32), ("us_single", C.c_float), ("us_sparse", C.c_float),
                ("us_dense", C.c_float)]


class GpuInfo(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("split", C.c_int32), ("bits", C.c_int32),
                ("has_csr", C.c_int32), ("has_packed", C.c_int32), ("group", C.c_int32),
                ("work_items", C.c_int32), ("nnz", C.c_int64), ("csr_bytes", C.c_int64),
                ("stream_bytes", C.c_int64), ("device_bytes", C.c_int64),
                ("batch_mode", C.c_int32), ("dense_ok", C.c_int32), ("dense_bytes", C.c_int64)]


_P = C.c_void_p
_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_U64P = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_EP = np.ctypeslib.ndpointer(ELEMENT_DTYPE, flags="C_CONTIGUOUS")

# (name, restype, argtypes) of every exported symbol declared in include/acceles.h
SIGNATURES = [
    ("acc_last_error", C.c_char_p, []),
    ("acc_version", C.c_char_p, []),
    ("acc_read_coo_matrix", C.c_int, [C.c_char_p, C.c_int, C.POINTER(_P)]),
    ("acc_read_mtx", C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(_P)]),
    ("acc_write_mtx", C.c_int, [C.c_char_p, _P]),
    ("acc_coo_create", C.c_int, [C.c_int32, C.c_int32, C.c_int64, _I32P, _I32P, _F32P, C.POINTER(_P)]),
    ("acc_coo_info", C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    ("acc_coo_export", C.c_int, [_P, _I32P, _I32P, _F32P]),
    ("acc_coo_destroy", None, [_P]),
    ("acc_coo2csr", C.c_int, [C.c_int32, C.c_int32, C.c_int64, _I32P, _I32P, _F32P, C.c_int, _I64P, _I32P, _F32P]),
    ("acc_coo_to_csr", C.c_int, [_P, C.POINTER(_P)]),
    ("acc_csr_create", C.c_int, [C.c_int32, C.c_int32, C.c_int64, _I64P, _I32P, _F32P, C.POINTER(_P)]),
    ("acc_csr_info", C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    ("acc_csr_export", C.c_int, [_P, _I64P, _I32P, _F32P]),
    ("acc_csr_destroy", None, [_P]),
    ("acc_quantize", C.c_int, [C.c_int64, _F32P, C.c_int, C.c_int, C.POINTER(C.c_float), _I32P]),
    ("acc_permute_rows", C.c_int, [_P, C.c_uint64, C.POINTER(_P), _I32P]),
    ("acc_resparse_csr", C.c_int, [_P, C.c_double, C.c_int, C.POINTER(_P)]),
    ("acc_ndcg_rank", C.c_int, [_I32P, _I32P, C.c_int, C.POINTER(C.c_double)]),
    ("acc_ndcg_score", C.c_int, [_I32P, C.POINTER(C.c_double), C.c_int, _I32P, C.c_int, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("acc_edit_distance", C.c_int, [_I32P, C.c_int, _I32P, C.c_int, C.POINTER(C.c_uint32)]),
    ("acc_count_errors", C.c_int, [_I32P, _I32P, C.c_int, C.POINTER(C.c_int32)]),
    ("acc_ultracsr_entries", C.c_int, [C.c_int, C.c_int]),
    ("acc_pack_ultracsr", C.c_int64, [C.c_int64, _I32P, _I32P, _I32P, C.c_int32, C.c_int, C.c_int, _U64P, C.c_int64]),
    ("acc_unpack_ultracsr", C.c_int, [_U64P, C.c_int64, C.c_int, C.c_int, C.c_int32, _I32P, _I32P, _I32P]),
    ("acc_xdecode", C.c_int, [C.c_uint32, C.c_int, C.c_int32, _I32P]),
    ("acc_pack_lowbit", C.c_int, [_P, C.POINTER(PackOpts), C.POINTER(_P)]),
    ("acc_packed_info", C.c_int, [_P, C.POINTER(PackedInfo)]),
    ("acc_unpack_lowbit", C.c_int, [_P, _I64P, _P, _P]),
    ("acc_packed_destroy", None, [_P]),
    ("acc_csr_save", C.c_int, [_P, C.c_char_p]),
    ("acc_csr_load", C.c_int, [C.c_char_p, C.POINTER(_P)]),
    ("acc_packed_save", C.c_int, [_P, C.c_char_p]),
    ("acc_packed_load", C.c_int, [C.c_char_p, C.POINTER(_P)]),
    ("acc_generate_csr", C.c_int, [C.c_int32, C.c_int32, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                   C.c_uint64, C.POINTER(_P)]),
    ("acc_generate_queries", C.c_int, [C.c_int32, C.c_int32, C.c_int, C.c_int, C.c_uint64, _F32P]),
    ("acc_gpu_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("acc_dev_malloc", C.c_int, [C.c_int, C.c_size_t, C.POINTER(_P)]),
    ("acc_dev_free", C.c_int, [_P]),
    ("acc_memcpy_h2d", C.c_int, [_P, _P, C.c_size_t]),
    ("acc_memcpy_d2h", C.c_int, [_P, _P, C.c_size_t]),
    ("acc_memcpy_d2d", C.c_int, [_P, _P, C.c_size_t]),
    ("acc_stream_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("acc_stream_destroy", C.c_int, [_P]),
    ("acc_stream_synchronize", C.c_int, [_P]),
    ("acc_event_create", C.c_int, [C.POINTER(_P)]),
    ("acc_event_record", C.c_int, [_P, _P]),
    ("acc_event_elapsed_ms", C.c_int, [_P, _P, C.POINTER(C.c_float)]),
    ("acc_event_destroy", C.c_int, [_P]),
    ("acc_gpu_upload", C.c_int, [_P, _P, C.c_int, C.POINTER(_P)]),
    ("acc_gpu_matrix_info", C.c_int, [_P, C.POINTER(GpuInfo)]),
    ("acc_gpu_matrix_destroy", None, [_P]),
    ("acc_gpu_set_profiling", C.c_int, [_P, C.c_int]),
    ("acc_gpu_set_batch_mode", C.c_int, [_P, C.c_int, C.c_int]),
    ("acc_gpu_autotune", C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, C.c_int, C.POINTER(C.c_int)]),
    ("acc_server_create", C.c_int, [_P, C.POINTER(_P)]),
    ("acc_server_add_matrix", C.c_int, [_P, _P, _P, C.POINTER(C.c_int)]),
    ("acc_server_submit", C.c_int, [_P, C.c_int, _P, C.c_int, C.POINTER(C.c_uint64)]),
    ("acc_server_wait", C.c_int, [_P, C.c_uint64, _P, C.c_int]),
    ("acc_server_stats", C.c_int, [_P, _P]),
    ("acc_server_reset_stats", C.c_int, [_P]),
    ("acc_server_bench_poisson", C.c_int, [_P, C.c_double, C.c_double, C.c_int, C.c_uint64, _P]),
    ("acc_server_destroy", None, [_P]),
    ("acc_gpu_plan_create", C.c_int, [_P, C.c_int, _P, C.c_int, _P, C.c_int, _P, C.c_int, C.POINTER(_P)]),
    ("acc_gpu_plan_launch", C.c_int, [_P, _P]),
    ("acc_gpu_plan_destroy", None, [_P]),
    ("acc_gpu_stage_times", C.c_int, [_P, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]),
    ("acc_gpu_time_lowbit_stage", C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P, C.POINTER(C.c_float)]),
    ("acc_gpu_topk_fp32", C.c_int, [_P, _P, C.c_int, _P, _P]),
    ("acc_gpu_topk_lowbit", C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P]),
    ("acc_gpu_spmv_fp64", C.c_int, [_P, _P, _P, _P]),
    ("acc_gpu_topk_lowbit_batch", C.c_int, [_P, _P, C.c_int, _I32P, C.c_int, _P, C.c_int, _P]),
    ("acc_gpu_spmv_fp32", C.c_int, [_P, _P, _P, _P]),
    ("acc_gpu_scores_lowbit", C.c_int, [_P, _P, _P, _P]),
    ("acc_gpu_candidates_lowbit", C.c_int, [_P, _P, C.c_int, _P, _P, C.POINTER(C.c_int), _P]),
    ("acc_gpu_top_k_heap", C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    ("acc_gpu_top_k_heap_noalloc", C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    ("acc_topk_fp32_host", C.c_int, [_P, _F32P, C.c_int, _EP]),
    ("acc_topk_lowbit_host", C.c_int, [_P, _F32P, C.c_int, C.c_int, _EP]),
    ("acc_top_k_spmv", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _EP, C.c_int]),
]

_lib = None


def lib() -> C.CDLL:
    """Loads libacceles.so once; raises ImportError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int):
    if rc < 0:
        raise AccelesError(rc, lib().acc_last_error().decode())
    return rc


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


# ---------------------------------------------------------------------------
# COO / CSR
# ---------------------------------------------------------------------------

class COOMatrix:
    """COOMatrix / coo_t (cpu_mkl_baseline.cpp:33-40, coo_matrix.hpp:9-24)."""

    def __init__(self, handle):
        self._h = handle if isinstance(handle, _P) else _P(handle)
        r, c, n = C.c_int32(), C.c_int32(), C.c_int64()
        _check(lib().acc_coo_info(self._h, C.byref(r), C.byref(c), C.byref(n)))
        self.rows, self.cols, self.size = r.value, c.value, n.value

    @classmethod
    def from_arrays(cls, rows, cols, row, col, val):
        row, col, val = _i32(row), _i32(col), _f32(val)
        h = _P()
        _check(lib().acc_coo_create(rows, cols, len(row), row, col, val, C.byref(h)))
        return cls(h)

    def arrays(self):
        row = np.empty(self.size, np.int32)
        col = np.empty(self.size, np.int32)
        val = np.empty(self.size, np.float32)
        _check(lib().acc_coo_export(self._h, row, col, val))
        return row, col, val

    def write_mtx(self, path):
        _check(lib().acc_write_mtx(path.encode(), self._h))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.acc_coo_destroy(self._h)
            self._h = _P()


class CSRMatrix:
    """Host FP32 CSR handle (the MKL handle of cpu_mkl_baseline.cpp:79-109)."""

    def __init__(self, handle):
        self._h = handle if isinstance(handle, _P) else _P(handle)
        r, c, n = C.c_int32(), C.c_int32(), C.c_int64()
        _check(lib().acc_csr_info(self._h, C.byref(r), C.byref(c), C.byref(n)))
        self.rows, self.cols, self.nnz = r.value, c.value, n.value

    @classmethod
    def from_arrays(cls, rows, cols, ptr, idx, val):
        ptr, idx, val = _i64(ptr), _i32(idx), _f32(val)
        h = _P()
        _check(lib().acc_csr_create(rows, cols, len(idx), ptr, idx, val, C.byref(h)))
        return cls(h)

    @classmethod
    def generate(cls, rows, cols, deg=(40, 100), deg_kind=DEG_UNIFORM, val_kind=VAL_ABSNORMAL, l2norm=True,
                 seed=0xACCE1E5):
        h = _P()
        _check(lib().acc_generate_csr(rows, cols, deg_kind, float(deg[0]), float(deg[1]), val_kind, int(l2norm),
                                      seed, C.byref(h)))
        return cls(h)

    def arrays(self):
        ptr = np.empty(self.rows + 1, np.int64)
        idx = np.empty(self.nnz, np.int32)
        val = np.empty(self.nnz, np.float32)
        _check(lib().acc_csr_export(self._h, ptr, idx, val))
        return ptr, idx, val

    def permute_rows(self, seed):
        """Random-CSR: seeded row shuffle; returns (csr, perm), out row i = in row perm[i]."""
        h = _P()
        perm = np.empty(self.rows, np.int32)
        _check(lib().acc_permute_rows(self._h, seed, C.byref(h), perm))
        return CSRMatrix(h), perm

    def save(self, path):
        """Binary CSR file (acc_csr_save)."""
        _check(lib().acc_csr_save(self._h, os.fsencode(path)))

    @classmethod
    def load(cls, path):
        """CSR matrix from a binary file (acc_csr_load)."""
        h = _P()
        _check(lib().acc_csr_load(os.fsencode(path), C.byref(h)))
        return cls(h)

    def resparse(self, keep, mean_over_cells=False):
        """ReSparse (resparse_csr, csr_util.py:152-187; resparse_spm with mean_over_cells): rows whose every |v|
        exceeds the mean |v| stay whole, the others keep their ceil(n*keep) largest |v|."""
        h = _P()
        _check(lib().acc_resparse_csr(self._h, float(keep), int(mean_over_cells), C.byref(h)))
        return CSRMatrix(h)

    def pack(self, bits=8, split=32, is_signed=False, col_bits=0, shuffle_seed=0):
        """Low-bit sliced packets; shuffle_seed != 0 deals the rows to partitions in a seeded shuffle
        (Random-CSR), results still naming the original rows."""
        return PackedMatrix(self, bits=bits, split=split, is_signed=is_signed, col_bits=col_bits,
                            shuffle_seed=shuffle_seed)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.acc_csr_destroy(self._h)
            self._h = _P()


# ---- ranking-quality metrics (evaluation_utils.hpp:112-267, quantify_util.py:313-329) ----

def ndcg_rank(golden, test):
    """normalized_discounted_cumulative_gain (evaluation_utils.hpp:112-147) of two row lists of one length."""
    g, t = _i32(golden), _i32(test)
    if len(g) != len(t):
        raise AccelesError(ACC_ERR_ARG, "ndcg_rank: lists of different lengths")
    out = C.c_double()
    _check(lib().acc_ndcg_rank(g, t, len(g), C.byref(out)))
    return out.value


def bounded_ndcg(golden, test, bounds=(10, 20, 50)):
    """bounded_ndcg (evaluation_utils.hpp:152-183): ndcg_rank of the first b rows for each bound b <= len."""
    out = []
    for b in bounds:
        if b > len(golden):
            break
        out.append(ndcg_rank(golden[:b], test[:b]))
    return out


def ndcg_score(sw_idx, sw_val, hw_idx):
    """ndcg (quantify_util.py:313-329): returns (ndcg, dcg, idcg)."""
    si, hi = _i32(sw_idx), _i32(hw_idx)
    sv = np.ascontiguousarray(sw_val, np.float64)
    n, d, i = C.c_double(), C.c_double(), C.c_double()
    _check(lib().acc_ndcg_score(si, sv.ctypes.data_as(C.POINTER(C.c_double)), len(si), hi, len(hi), C.byref(n),
                                C.byref(d), C.byref(i)))
    return n.value, d.value, i.value


def edit_distance(a, b):
    """edit_distance (evaluation_utils.hpp:197-209)."""
    x, y = _i32(a), _i32(b)
    out = C.c_uint32()
    _check(lib().acc_edit_distance(x, len(x), y, len(y), C.byref(out)))
    return out.value


def bounded_edit_distance(golden, test, bounds=(10, 20, 50)):
    """bounded_edit_distance (evaluation_utils.hpp:211-237)."""
    return [edit_distance(golden[:b], test[:b]) for b in bounds if b <= len(golden)]


def bounded_count_errors(golden, test, bounds=(10, 20, 50)):
    """bounded_count_errors (evaluation_utils.hpp:239-263)."""
    out = []
    for b in bounds:
        if b > len(golden):
            break
        e = C.c_int32()
        _check(lib().acc_count_errors(_i32(golden[:b]), _i32(test[:b]), b, C.byref(e)))
        out.append(e.value)
    return out


def read_coo_matrix(path: str, zero_index: bool = True) -> COOMatrix:
    """read_coo_matrix (cpu_mkl_baseline.cpp:43-77)."""
    h = _P()
    _check(lib().acc_read_coo_matrix(path.encode(), int(zero_index), C.byref(h)))
    return COOMatrix(h)


def readMtx(path: str, zero_indexed_file: bool = False, sort_tuples: bool = True) -> COOMatrix:
    """readMtx (utils.hpp:449-492)."""
    h = _P()
    _check(lib().acc_read_mtx(path.encode(), int(zero_indexed_file), int(sort_tuples), C.byref(h)))
    return COOMatrix(h)


def coo_to_csr(coo: COOMatrix) -> CSRMatrix:
    """coo_to_csr (cpu_mkl_baseline.cpp:79-109)."""
    h = _P()
    _check(lib().acc_coo_to_csr(coo._h, C.byref(h)))
    return CSRMatrix(h)


def coo2csr(row, col, val, nrows, ncols, sort_tuples=True):
    """coo2csr (utils.hpp:494-547): returns (ptr, idx, val)."""
    row, col, val = _i32(row), _i32(col), _f32(val)
    ptr = np.empty(nrows + 1, np.int64)
    idx = np.empty(len(row), np.int32)
    vo = np.empty(len(row), np.float32)
    _check(lib().acc_coo2csr(nrows, ncols, len(row), row, col, val, int(sort_tuples), ptr, idx, vo))
    return ptr, idx, vo


def quantize(v, bits, symmetric):
    """quantify_matrix / quantify_vector (quantify_util.py:8-49): returns (q, scale)."""
    v = _f32(v).ravel()
    q = np.empty(v.size, np.int32)
    s = C.c_float()
    _check(lib().acc_quantize(v.size, v, bits, int(symmetric), C.byref(s), q))
    return q, s.value


def quantify_matrix(matrix, B, symmetric):
    return quantize(matrix, B, symmetric)[0].reshape(np.shape(matrix))


def quantify_vector(vector, B, symmetric):
    return quantize(vector, B, symmetric)[0].reshape(np.shape(vector))


def generate_queries(nq, cols, val_kind=VAL_NORMAL, l2norm=False, seed=0xACCE1E5):
    out = np.empty((nq, cols), np.float32)
    _check(lib().acc_generate_queries(nq, cols, val_kind, int(l2norm), seed, out))
    return out


# ---------------------------------------------------------------------------
# Ultra-CSR reference packets
# ---------------------------------------------------------------------------

def ultracsr_entries(col_bits=10, val_bits=6):
    return _check(lib().acc_ultracsr_entries(col_bits, val_bits))


def pack_ultracsr(row, col, qval, last_row=0, col_bits=10, val_bits=6):
    row, col, qval = _i32(row), _i32(col), _i32(qval)
    n = ultracsr_entries(col_bits, val_bits)
    cap = max(1, (len(row) + n - 1) // n)
    out = np.zeros(cap * 8, np.uint64)
    np_ = _check(lib().acc_pack_ultracsr(len(row), row, col, qval, last_row, col_bits, val_bits, out, cap))
    return out[: np_ * 8].reshape(np_, 8)


def unpack_ultracsr(packets, nnz, col_bits=10, val_bits=6, carry0=0):
    packets = np.ascontiguousarray(packets, dtype=np.uint64).ravel()
    xrow = np.empty(nnz, np.int32)
    col = np.empty(nnz, np.int32)
    qv = np.empty(nnz, np.int32)
    _check(lib().acc_unpack_ultracsr(packets, nnz, col_bits, val_bits, carry0, xrow, col, qv))
    return xrow, col, qv


def xdecode(flags, n=30, carry=0):
    x = np.empty(n, np.int32)
    _check(lib().acc_xdecode(flags, n, carry, x))
    return x


# ---------------------------------------------------------------------------
# low-bit sliced packets
# ---------------------------------------------------------------------------

class PackedMatrix:
    """Quantized, sliced low-bit packets of a CSR matrix (acc_pack_lowbit)."""

    def __init__(self, csr: CSRMatrix, bits=8, split=32, is_signed=False, col_bits=0, shuffle_seed=0):
        o = PackOpts(bits, int(is_signed), split, col_bits, int(shuffle_seed))
        h = _P()
        _check(lib().acc_pack_lowbit(csr._h, C.byref(o), C.byref(h)))
        self._h = h
        self.info = PackedInfo()
        _check(lib().acc_packed_info(self._h, C.byref(self.info)))

    def save(self, path):
        """Binary packet file (acc_packed_save)."""
        _check(lib().acc_packed_save(self._h, os.fsencode(path)))

    @classmethod
    def load(cls, path):
        """Packets from a binary file (acc_packed_load)."""
        h = _P()
        _check(lib().acc_packed_load(os.fsencode(path), C.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        self.info = PackedInfo()
        _check(lib().acc_packed_info(self._h, C.byref(self.info)))
        return self

    def unpack(self):
        """(ptr, idx, qval) of the kept (non-zero quantized) entries, CSR order."""
        ptr = np.empty(self.info.rows + 1, np.int64)
        _check(lib().acc_unpack_lowbit(self._h, ptr, None, None))
        idx = np.empty(int(ptr[-1]), np.int32)
        qv = np.empty(int(ptr[-1]), np.int32)
        _check(lib().acc_unpack_lowbit(self._h, ptr, idx.ctypes.data_as(_P), qv.ctypes.data_as(_P)))
        return ptr, idx, qv

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.acc_packed_destroy(self._h)
            self._h = _P()


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

def device_count() -> int:
    n = C.c_int()
    _check(lib().acc_gpu_device_count(C.byref(n)))
    return n.value


class DeviceBuffer:
    """Device memory owned by this object (acc_dev_malloc); typed like a numpy array."""

    def __init__(self, shape, dtype, device=0):
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.device = device
        p = _P()
        _check(lib().acc_dev_malloc(device, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a, device=0):
        a = np.ascontiguousarray(a)
        b = cls(a.shape, a.dtype, device)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        _check(lib().acc_memcpy_h2d(self.ptr, a.ctypes.data_as(_P), a.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().acc_memcpy_d2h(out.ctypes.data_as(_P), self.ptr, self.nbytes))
        return out

    def copy_from(self, other):
        """Device-to-device copy of another buffer of the same size."""
        if other.nbytes != self.nbytes:
            raise AccelesError(ACC_ERR_ARG, "copy_from: sizes differ")
        _check(lib().acc_memcpy_d2d(self.ptr, other.ptr, self.nbytes))

    def at(self, offset_elems):
        """Raw device address offset_elems elements into the buffer."""
        return self.ptr.value + int(offset_elems) * self.dtype.itemsize

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value and _lib is not None:
            _lib.acc_dev_free(self.ptr)
            self.ptr = None


class Stream:
    def __init__(self, device=0):
        s = _P()
        _check(lib().acc_stream_create(device, C.byref(s)))
        self.handle = s

    def synchronize(self):
        _check(lib().acc_stream_synchronize(self.handle))

    def __del__(self):
        if getattr(self, "handle", None) is not None and self.handle.value and _lib is not None:
            _lib.acc_stream_destroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = _P()
        _check(lib().acc_event_create(C.byref(e)))
        self.handle = e

    def record(self, stream=None):
        _check(lib().acc_event_record(self.handle, _stream(stream)))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _check(lib().acc_event_elapsed_ms(self.handle, end.handle, C.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "handle", None) is not None and self.handle.value and _lib is not None:
            _lib.acc_event_destroy(self.handle)
            self.handle = None


def synchronize(stream=None):
    _check(lib().acc_stream_synchronize(_stream(stream)))


def _ptr(t):
    """Device address of a DeviceBuffer, a raw int, or an object with data_ptr()."""
    if isinstance(t, DeviceBuffer):
        return t.ptr
    if isinstance(t, int):
        return _P(t)
    return _P(t.data_ptr())


def _stream(stream):
    if stream is None:
        return _P()
    if isinstance(stream, Stream):
        return stream.handle
    return _P(stream if isinstance(stream, int) else stream.cuda_stream)


class Plan:
    """A query launch sequence captured into a HIP graph (acc_gpu_plan_create).

    Reads the device buffer x and writes out on every launch; refill x between
    launches.  The GpuMatrix must outlive the plan."""

    def __init__(self, matrix, kind, x, ks, ck, out, out_stride):
        ks = _i32(np.atleast_1d(ks))
        h = _P()
        _check(lib().acc_gpu_plan_create(matrix._h, int(kind), _ptr(x), len(ks), ks.ctypes.data, int(ck), _ptr(out),
                                         int(out_stride), C.byref(h)))
        self._h, self._m = h, matrix

    def launch(self, stream=None):
        _check(lib().acc_gpu_plan_launch(self._h, _stream(stream)))

    def close(self):
        if self._h:
            lib().acc_gpu_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop over resident matrices with continuous batching (acc_server_*)."""

    def __init__(self, device=0, memory_budget=0, max_batch=256, slots=4, ck=512, max_pending=0, autotune=False,
                 max_request_bytes=0):
        o = ServerOpts(device, memory_budget, max_batch, slots, ck, max_pending, int(autotune), int(max_request_bytes))
        h = _P()
        _check(lib().acc_server_create(C.byref(o), C.byref(h)))
        self._h = h
        self.cols = []

    def add_matrix(self, csr, packed):
        mid = C.c_int()
        _check(lib().acc_server_add_matrix(self._h, csr._h, packed._h, C.byref(mid)))
        self.cols.append(csr.cols)
        return mid.value

    def submit(self, matrix_id, x, k):
        x = _f32(x)
        t = C.c_uint64()
        _check(lib().acc_server_submit(self._h, int(matrix_id), x.ctypes.data, int(k), C.byref(t)))
        return t.value

    def wait(self, ticket, k, timeout_ms=-1):
        out = np.zeros(k, ELEMENT_DTYPE)
        _check(lib().acc_server_wait(self._h, C.c_uint64(ticket), out.ctypes.data, int(timeout_ms)))
        return out

    def stats(self):
        st = ServerStats()
        _check(lib().acc_server_stats(self._h, C.byref(st)))
        return st.as_dict()

    def reset_stats(self):
        _check(lib().acc_server_reset_stats(self._h))

    def bench_poisson(self, rate_qps, seconds, kmax=100, seed=1):
        st = ServerStats()
        _check(lib().acc_server_bench_poisson(self._h, float(rate_qps), float(seconds), int(kmax), int(seed),
                                              C.byref(st)))
        return st.as_dict()

    def close(self):
        if self._h:
            lib().acc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuMatrix:
    """Device-resident matrix (acc_gpu_upload): FP32 CSR and/or low-bit packets."""

    def __init__(self, csr: CSRMatrix | None, packed: PackedMatrix | None = None, device: int = 0):
        h = _P()
        _check(lib().acc_gpu_upload(csr._h if csr is not None else None,
                                    packed._h if packed is not None else None, device, C.byref(h)))
        self._h = h
        info = self.info
        self.rows, self.cols = info.rows, info.cols

    @property
    def info(self):
        """acc_gpu_info_t, read afresh (the batch mode changes after upload)."""
        info = GpuInfo()
        _check(lib().acc_gpu_matrix_info(self._h, C.byref(info)))
        return info

    def set_batch_mode(self, mode, min_dense_batch=0):
        """BATCH_AUTO, BATCH_SPARSE or BATCH_DENSE scoring for topk_lowbit_batch."""
        _check(lib().acc_gpu_set_batch_mode(self._h, int(mode), int(min_dense_batch)))

    def autotune(self, ck, k, max_batch):
        """acc_gpu_autotune: per batch size, the fastest layout (list of TunePoint)."""
        pts = (TunePoint * 16)()
        n = C.c_int()
        _check(lib().acc_gpu_autotune(self._h, int(ck), int(k), int(max_batch), pts, 16, C.byref(n)))
        return [pts[i] for i in range(n.value)]

    def set_profiling(self, enable=True):
        _check(lib().acc_gpu_set_profiling(self._h, int(enable)))

    def stage_times(self):
        """Launch durations (ms) of the last query while profiling is on."""
        ms = (C.c_float * 8)()
        n = C.c_int()
        _check(lib().acc_gpu_stage_times(self._h, ms, 8, C.byref(n)))
        return [ms[i] for i in range(n.value)]

    def time_lowbit_stage(self, x, k, ck, stage, reps=20, stream=None):
        """Average ms of one low-bit stage (0 slice maxima, 1 select, 2 final) launched reps times back to back."""
        ms = C.c_float()
        _check(lib().acc_gpu_time_lowbit_stage(self._h, _ptr(x), k, ck, stage, reps, _stream(stream), C.byref(ms)))
        return ms.value

    # host-pointer forms (numpy in, numpy out)
    def topk_fp32_host(self, x, k):
        out = np.empty(k, ELEMENT_DTYPE)
        _check(lib().acc_topk_fp32_host(self._h, _f32(x), k, out))
        return out

    def topk_lowbit_host(self, x, k, ck):
        out = np.empty(k, ELEMENT_DTYPE)
        _check(lib().acc_topk_lowbit_host(self._h, _f32(x), k, ck, out))
        return out

    # device-pointer forms (torch tensors or raw ints)
    def topk_fp32(self, x, k, out, stream=None):
        _check(lib().acc_gpu_topk_fp32(self._h, _ptr(x), k, _ptr(out), _stream(stream)))

    def topk_lowbit(self, x, k, ck, out, stream=None):
        _check(lib().acc_gpu_topk_lowbit(self._h, _ptr(x), k, ck, _ptr(out), _stream(stream)))

    def spmv_fp64(self, x, y, stream=None):
        _check(lib().acc_gpu_spmv_fp64(self._h, _ptr(x), _ptr(y), _stream(stream)))

    def topk_lowbit_batch(self, X, ks, ck, out, out_stride, stream=None):
        ks = _i32(ks)
        _check(lib().acc_gpu_topk_lowbit_batch(self._h, _ptr(X), len(ks), ks, ck, _ptr(out), out_stride,
                                               _stream(stream)))

    def spmv_fp32(self, x, y, stream=None):
        _check(lib().acc_gpu_spmv_fp32(self._h, _ptr(x), _ptr(y), _stream(stream)))

    def scores_lowbit(self, x, y, stream=None):
        _check(lib().acc_gpu_scores_lowbit(self._h, _ptr(x), _ptr(y), _stream(stream)))

    def candidates_lowbit(self, x, ck, rows, scores, stream=None):
        n = C.c_int()
        _check(lib().acc_gpu_candidates_lowbit(self._h, _ptr(x), ck, _ptr(rows), _ptr(scores), C.byref(n),
                                               _stream(stream)))
        return n.value

    def top_k_spmv(self, k=100, num_test=100):
        """top_k_spmv drop-in (cpu_mkl_baseline.cpp:249-275)."""
        out = np.empty(k, ELEMENT_DTYPE)
        _check(lib().acc_top_k_spmv(self._h, self.rows, self.cols, k, 1, out, num_test))
        return out

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.acc_gpu_matrix_destroy(self._h)
            self._h = _P()


def top_k_heap(y, n, k, out, stream=None):
    """top_k_heap (cpu_mkl_baseline.cpp:128-149) on a device score vector."""
    _check(lib().acc_gpu_top_k_heap(_ptr(y), n, k, _ptr(out), _stream(stream)))


def top_k_heap_noalloc(y, n, k, out, stream=None):
    """top_k_heap with its selection scratch in a static device array: allocates nothing."""
    _check(lib().acc_gpu_top_k_heap_noalloc(_ptr(y), n, k, _ptr(out), _stream(stream)))
