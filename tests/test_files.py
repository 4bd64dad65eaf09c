"""Binary CSR and packet files: round trips, and malformed files refused with ACC_ERR_FORMAT."""
import os

import numpy as np
import pytest


def test_csr_round_trip(acc, tmp_path):
    csr = acc.CSRMatrix.generate(3000, 300, deg=(0, 40), seed=4)
    p = str(tmp_path / "m.csr")
    csr.save(p)
    back = acc.CSRMatrix.load(p)
    assert (back.rows, back.cols, back.nnz) == (csr.rows, csr.cols, csr.nnz)
    for a, b in zip(csr.arrays(), back.arrays()):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("bits,signed", [(8, False), (4, False), (6, True)])
def test_packed_round_trip(acc, tmp_path, bits, signed):
    csr = acc.CSRMatrix.generate(5000, 700, deg=(1, 60), seed=bits, val_kind=acc.VAL_NORMAL if signed else 0)
    pk = csr.pack(bits=bits, split=7, is_signed=signed)
    p = str(tmp_path / "m.pk")
    pk.save(p)
    back = acc.PackedMatrix.load(p)
    for f, _ in pk.info._fields_:
        assert getattr(back.info, f) == getattr(pk.info, f), f
    for a, b in zip(pk.unpack(), back.unpack()):
        np.testing.assert_array_equal(a, b)


def corrupt(path, offset, data):
    with open(path, "r+b") as f:
        f.seek(offset)
        f.write(data)


def test_malformed_files_refused(acc, tmp_path):
    csr = acc.CSRMatrix.generate(200, 50, deg=(1, 10), seed=2)
    pk = csr.pack(bits=8, split=4)
    c, k = str(tmp_path / "m.csr"), str(tmp_path / "m.pk")
    csr.save(c)
    pk.save(k)
    with pytest.raises(acc.AccelesError, match="not a CSR file"):
        acc.CSRMatrix.load(k)  # wrong magic
    with pytest.raises(acc.AccelesError, match="not a packet file"):
        acc.PackedMatrix.load(c)
    for path, loader, msg in ((c, acc.CSRMatrix.load, "not a CSR file"), (k, acc.PackedMatrix.load, "not a packet")):
        size = os.path.getsize(path)
        t = str(tmp_path / "trunc")
        with open(path, "rb") as f, open(t, "wb") as g:
            g.write(f.read()[: size - 5])  # truncated
        with pytest.raises(acc.AccelesError, match=msg):
            loader(t)
    # a column index out of range: the last idx entry of the CSR file
    ptr, idx, val = csr.arrays()
    bad = str(tmp_path / "bad.csr")
    csr.save(bad)
    off = 8 + 4 + 4 + 8 + 8 * len(ptr) + 8 + 4 * (len(idx) - 1)
    corrupt(bad, off, np.int32(50).tobytes())
    with pytest.raises(acc.AccelesError, match="not a CSR file"):
        acc.CSRMatrix.load(bad)
    # a lane naming a row past the matrix: perm[0], after the 72-byte header and four vectors
    badk = str(tmp_path / "bad.pk")
    pk.save(badk)
    split, ns = 4, pk.info.nslices
    off = 72 + (8 + 4 * (split + 1)) + (8 + 8 * (split + 1)) + (8 + 4 * ns) + (8 + 4 * ns) + 8
    corrupt(badk, off, np.int32(200).tobytes())
    with pytest.raises(acc.AccelesError, match="not a packet file"):
        acc.PackedMatrix.load(badk)
    corrupt(badk, off, np.int32(-1).tobytes())  # a padding lane is fine
    acc.PackedMatrix.load(badk)


def test_missing_file(acc, tmp_path):
    with pytest.raises(acc.AccelesError, match="cannot open"):
        acc.CSRMatrix.load(str(tmp_path / "none.csr"))
