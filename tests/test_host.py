"""CPU tests of the product library's host side (no GPU needed): the C ABI
loads and exports every declared symbol; loaders, converter, quantizer,
packers and generators agree with the oracle and the reference's gold."""
import os
import re

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_library_exports_every_declared_symbol(acc):
    src = open(os.path.join(ROOT, "include", "acceles.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    declared = set(re.findall(r"\b(acc_[a-z0-9_]+)\s*\(", src))
    assert len(declared) > 30
    lib = acc.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    bound = {n for n, _, _ in acc.SIGNATURES}
    assert declared == bound, declared ^ bound
    assert b"gfx950" in lib.acc_version()


def test_error_codes_and_messages(acc):
    with pytest.raises(acc.AccelesError) as e:
        acc.read_coo_matrix("/nonexistent/file.mtx")
    assert e.value.code == -2 and "Failed to open" in str(e.value)
    with pytest.raises(acc.AccelesError) as e:
        acc.coo2csr([0, 5], [0, 0], [1.0, 2.0], 5, 3)
    assert e.value.code == -4


# --------------------------------------------------------------------------
# loaders
# --------------------------------------------------------------------------

@pytest.mark.parametrize("sort", [0, 1])
def test_readMtx_matches_reference_gold(acc, sort):
    g = golden("readmtx_sample.npz")
    coo = acc.readMtx(os.path.join(GOLDEN, "sample.mtx"), zero_indexed_file=False, sort_tuples=bool(sort))
    assert (coo.rows, coo.cols, coo.size) == tuple(g["dims"])
    row, col, val = coo.arrays()
    np.testing.assert_array_equal(row, g[f"row_s{sort}"])
    np.testing.assert_array_equal(col, g[f"col_s{sort}"])
    np.testing.assert_array_equal(val, g[f"val_s{sort}"])


def test_read_coo_matrix_semantics(acc, tmp_path):
    # cpu_mkl_baseline.cpp:43-77: exactly two header lines, strtof values
    p = tmp_path / "m.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% c\n3 4 3\n1 1 0.1\n3 4 -2.5\n2 2 1e-2\n")
    coo = acc.read_coo_matrix(str(p), zero_index=False)
    row, col, val = coo.arrays()
    np.testing.assert_array_equal(row, [0, 2, 1])
    np.testing.assert_array_equal(col, [0, 3, 1])
    np.testing.assert_array_equal(val, np.array([0.1, -2.5, 1e-2], np.float32))
    with pytest.raises(acc.AccelesError):  # zero_index=True on a 1-based file: column 4 is out of range
        acc.read_coo_matrix(str(p), zero_index=True)
    bad = tmp_path / "short.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n%\n3 3 5\n1 1 1\n")
    with pytest.raises(acc.AccelesError) as e:
        acc.read_coo_matrix(str(bad), zero_index=False)
    assert e.value.code == -3


def test_mtx_write_read_roundtrip_large(acc, tmp_path):
    csr = acc.CSRMatrix.generate(5000, 300, deg=(1, 40), seed=3)
    ptr, idx, val = csr.arrays()
    row = np.repeat(np.arange(csr.rows, dtype=np.int32), np.diff(ptr))
    coo = acc.COOMatrix.from_arrays(csr.rows, csr.cols, row, idx, val)
    path = str(tmp_path / "big.mtx")
    coo.write_mtx(path)
    back = acc.read_coo_matrix(path, zero_index=False)
    r2, c2, v2 = back.arrays()
    np.testing.assert_array_equal(r2, row)
    np.testing.assert_array_equal(c2, idx)
    np.testing.assert_array_equal(v2, val)  # %.9g round-trips float32
    csr2 = acc.coo_to_csr(back)
    for a, b in zip(csr2.arrays(), (ptr, idx, val)):
        np.testing.assert_array_equal(a, b)


# --------------------------------------------------------------------------
# conversion
# --------------------------------------------------------------------------

@pytest.mark.parametrize("sort", [0, 1])
def test_coo2csr_matches_reference_gold_and_oracle(acc, sort):
    g = golden("coo2csr.npz")
    ptr, idx, val = acc.coo2csr(g["row"], g["col"], g["val"], int(g["nrows"]), int(g["ncols"]), bool(sort))
    np.testing.assert_array_equal(ptr, g[f"ptr_s{sort}"].astype(np.int64))
    np.testing.assert_array_equal(idx, g[f"idx_s{sort}"])
    np.testing.assert_array_equal(val, g[f"val_s{sort}"])
    rng = np.random.default_rng(sort)
    row = rng.integers(0, 1000, 20000).astype(np.int32)
    col = rng.integers(0, 77, 20000).astype(np.int32)
    v = rng.standard_normal(20000).astype(np.float32)
    for a, b in zip(acc.coo2csr(row, col, v, 1000, 77, bool(sort)), O.coo2csr(row, col, v, 1000, 77, sort)):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("bits,sym", [(4, False), (8, False), (8, True), (6, True), (2, True)])
def test_quantize_matches_oracle(acc, bits, sym):
    rng = np.random.default_rng(bits)
    v = rng.standard_normal(50000).astype(np.float32)
    if not sym:
        v = np.abs(v)
    q, s = acc.quantize(v, bits, sym)
    qo, so = O.quantize(v, bits, sym)
    assert s == so
    np.testing.assert_array_equal(q, qo)


def test_permute_rows(acc):
    csr = acc.CSRMatrix.generate(1000, 50, deg=(0, 10), seed=9)
    out, perm = csr.permute_rows(seed=42)
    assert sorted(perm.tolist()) == list(range(1000))
    p0, i0, v0 = csr.arrays()
    p1, i1, v1 = out.arrays()
    for r in range(1000):
        s = perm[r]
        np.testing.assert_array_equal(i1[p1[r]:p1[r + 1]], i0[p0[s]:p0[s + 1]])
        np.testing.assert_array_equal(v1[p1[r]:p1[r + 1]], v0[p0[s]:p0[s + 1]])
    _, perm2 = csr.permute_rows(seed=42)
    np.testing.assert_array_equal(perm, perm2)


# --------------------------------------------------------------------------
# generators (create_matrices.py:57-137 semantics)
# --------------------------------------------------------------------------

def test_generator_uniform_degree_distinct_columns_normalised(acc):
    csr = acc.CSRMatrix.generate(4000, 500, deg=(40, 100), seed=11)
    ptr, idx, val = csr.arrays()
    deg = np.diff(ptr)
    assert deg.min() >= 40 and deg.max() <= 100 and abs(deg.mean() - 70) < 1.0
    for r in range(0, 4000, 97):
        c = idx[ptr[r]:ptr[r + 1]]
        assert np.all(np.diff(c) > 0)  # sorted, distinct
        np.testing.assert_allclose(np.linalg.norm(val[ptr[r]:ptr[r + 1]].astype(np.float64)), 1.0, rtol=1e-5)
    assert np.all(val > 0)
    again = acc.CSRMatrix.generate(4000, 500, deg=(40, 100), seed=11)
    for a, b in zip(again.arrays(), (ptr, idx, val)):
        np.testing.assert_array_equal(a, b)


def test_generator_gamma_powerlaw(acc):
    csr = acc.CSRMatrix.generate(200000, 1024, deg=(0.5, 40.0), deg_kind=acc.DEG_GAMMA, l2norm=False, seed=12)
    deg = np.diff(csr.arrays()[0])
    assert deg.min() >= 1 and deg.max() <= 1024
    assert 17 < deg.mean() < 23  # Gamma(0.5, 40) has mean 20 (rounding and the floor shift it a little)
    assert np.median(deg) < 12  # heavy right tail


def test_queries_generator(acc):
    x = acc.generate_queries(8, 900, val_kind=acc.VAL_NORMAL, l2norm=True, seed=1)
    np.testing.assert_allclose(np.linalg.norm(x.astype(np.float64), axis=1), 1.0, rtol=1e-6)
    assert abs(x.mean()) < 0.01


# --------------------------------------------------------------------------
# low-bit sliced packets
# --------------------------------------------------------------------------

@pytest.mark.parametrize("bits,sym,cols", [(8, False, 500), (4, False, 1024), (8, True, 900), (6, True, 64),
                                           (2, False, 3)])
def test_pack_lowbit_roundtrip(acc, bits, sym, cols):
    csr = acc.CSRMatrix.generate(3001, cols, deg=(0, min(cols, 60)), val_kind=acc.VAL_NORMAL if sym else
                                 acc.VAL_ABSNORMAL, seed=bits + cols)
    pk = csr.pack(bits=bits, split=7, is_signed=sym)
    ptr, idx, val = csr.arrays()
    q, s = O.quantize(val, bits, sym)
    assert pk.info.scale == s
    keep = q != 0
    kept_ptr = np.zeros(csr.rows + 1, np.int64)
    kept_ptr[1:] = np.cumsum(np.add.reduceat(keep.astype(np.int64), ptr[:-1]) * (np.diff(ptr) > 0))
    uptr, uidx, uq = pk.unpack()
    np.testing.assert_array_equal(uptr, kept_ptr)
    np.testing.assert_array_equal(uidx, idx[keep])
    np.testing.assert_array_equal(uq, q[keep])
    assert pk.info.nnz_kept == keep.sum()
    E = pk.info.entry_bits
    # bytes accounting: stream words plus 8 bytes of metadata per slice
    assert pk.info.stream_bytes >= pk.info.nnz_kept * E / 8


def test_pack_lowbit_rejects_negative_unsigned(acc):
    csr = acc.CSRMatrix.generate(100, 20, deg=(1, 5), val_kind=acc.VAL_NORMAL, seed=1)
    with pytest.raises(acc.AccelesError):
        csr.pack(bits=8, split=4, is_signed=False)
    with pytest.raises(acc.AccelesError):
        csr.pack(bits=8, split=101, is_signed=True)


def test_pack_lowbit_bytes_per_nnz_at_embedding_scale(acc):
    csr = acc.CSRMatrix.generate(40000, 500, deg=(40, 100), seed=2)
    p8 = csr.pack(bits=8, split=32)
    p4 = csr.pack(bits=4, split=32)
    # 9-bit columns + 8 or 4 value bits, plus padding and slice metadata
    assert 17 / 8 <= p8.info.bytes_per_nnz < 2.35
    assert p4.info.bytes_per_nnz < p8.info.bytes_per_nnz


# --------------------------------------------------------------------------
# Ultra-CSR packets
# --------------------------------------------------------------------------

def test_xdecode_known_answer(acc):
    np.testing.assert_array_equal(acc.xdecode(0x20080200, 30, 0), [0] * 9 + [1] * 10 + [2] * 10 + [3])


@pytest.mark.parametrize("cb,vb", [(10, 6), (10, 8), (9, 4), (12, 8)])
def test_ultracsr_pack_matches_oracle(acc, cb, vb):
    rng = np.random.default_rng(cb * 100 + vb)
    deg = rng.integers(1, 40, 300)
    row = np.repeat(np.arange(300, dtype=np.int32) + 5, deg)
    col = rng.integers(0, 2 ** cb, len(row)).astype(np.int32)
    qv = rng.integers(-(2 ** (vb - 1)), 2 ** (vb - 1), len(row)).astype(np.int32)
    assert acc.ultracsr_entries(cb, vb) == O.oracle().orc_ultracsr_entries(cb, vb)
    a = acc.pack_ultracsr(row, col, qv, last_row=4, col_bits=cb, val_bits=vb)
    b = O.ultracsr_pack(row, col, qv, 4, cb, vb)
    np.testing.assert_array_equal(a, b)
    if acc.ultracsr_entries(cb, vb) <= 32:
        xr, c, v = acc.unpack_ultracsr(a, len(row), cb, vb, 0)
        xo, co, vo = O.ultracsr_unpack(b, len(row), cb, vb, 0)
        np.testing.assert_array_equal(xr, xo)
        np.testing.assert_array_equal(c, col)
        np.testing.assert_array_equal(v, qv)
