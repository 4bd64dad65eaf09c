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
    # a lane naming a row past the matrix: perm[0], after the 80-byte header and four vectors
    badk = str(tmp_path / "bad.pk")
    pk.save(badk)
    split, ns = 4, pk.info.nslices
    off = 80 + (8 + 4 * (split + 1)) + (8 + 8 * (split + 1)) + (8 + 4 * ns) + (8 + 4 * ns) + 8
    corrupt(badk, off, np.int32(200).tobytes())
    with pytest.raises(acc.AccelesError, match="not a packet file"):
        acc.PackedMatrix.load(badk)
    # a padding lane in the middle of a partition drops a row: refused
    corrupt(badk, off, np.int32(-1).tobytes())
    with pytest.raises(acc.AccelesError, match="not a packet file"):
        acc.PackedMatrix.load(badk)


def packet_offsets(pk, split):
    """Byte offsets of the vectors of a packet file (each: a uint64 length, then the data), after the
    80-byte header (magic, 7 int32 fields, scale, nnz, nnz_kept, dup_rows, merged max/min, shuffle seed)."""
    ns = pk.info.nslices
    o = {"part_row_begin": 80}
    o["part_slice_begin"] = o["part_row_begin"] + 8 + 4 * (split + 1)
    o["slice_off"] = o["part_slice_begin"] + 8 + 8 * (split + 1)
    o["slice_ndw"] = o["slice_off"] + 8 + 4 * ns
    o["perm"] = o["slice_ndw"] + 8 + 4 * ns
    o["words"] = o["perm"] + 8 + 4 * ns * 64
    return o


def test_malformed_packet_fields_refused(acc, tmp_path):
    """Every field the kernels or the host index with is checked on load."""
    csr = acc.CSRMatrix.generate(1000, 50, deg=(1, 10), seed=3)
    split = 4
    pk = csr.pack(bits=8, split=split)
    ns = pk.info.nslices
    o = packet_offsets(pk, split)
    good = str(tmp_path / "good.pk")
    pk.save(good)
    acc.PackedMatrix.load(good)
    cases = {
        # part_slice_begin: slice ranges must grow from 0 and give each partition ceil(rows_p/64) slices
        "slices_jump": (o["part_slice_begin"] + 8 + 8, np.int64(ns + 1000)),
        "slices_start": (o["part_slice_begin"] + 8, np.int64(1)),
        # part_row_begin: from 0, never decreasing, ending at rows
        "rows_start": (o["part_row_begin"] + 8, np.int32(5)),
        "rows_decrease": (o["part_row_begin"] + 8 + 4 * 2, np.int32(10)),
        # a vector length past what the file holds (no allocation is attempted)
        "huge_length": (o["words"], np.uint64(1 << 60)),
        "huge_slices": (o["slice_off"], np.uint64((1 << 63) - 1)),
        # a row named twice (lane 1 repeats lane 0's row)
        "row_twice": (o["perm"] + 8 + 4, None),
        # an entry whose column lies past cols (column field of lane 0's first entry = 63 >= 50)
        "column_range": (o["words"] + 8, None),
    }
    for name, (off, val) in cases.items():
        bad = str(tmp_path / f"{name}.pk")
        with open(good, "rb") as f:
            data = bytearray(f.read())
        if name == "row_twice":
            data[off:off + 4] = data[off - 4:off]
        elif name == "column_range":
            w = int(np.frombuffer(bytes(data[off:off + 4]), np.uint32)[0])
            data[off:off + 4] = np.uint32((w & ~0x3F) | 63).tobytes()  # col_bits = 6 for 50 columns
        else:
            data[off:off + val.nbytes] = val.tobytes()
        with open(bad, "wb") as f:
            f.write(bytes(data))
        with pytest.raises(acc.AccelesError, match="not a packet file"):
            acc.PackedMatrix.load(bad)


def test_loaded_packet_statistics_recomputed(acc, tmp_path):
    """The kept-entry count and the repeated-column statistics come from the
    decoded entries, not from the header."""
    csr = acc.CSRMatrix.generate(500, 40, deg=(1, 10), seed=4)
    pk = csr.pack(bits=8, split=2)
    path = str(tmp_path / "m.pk")
    pk.save(path)
    corrupt(path, 8 + 4 * 8 + 8, np.int64(12345).tobytes())  # nnz_kept in the header
    got = acc.PackedMatrix.load(path)
    assert got.info.nnz_kept == pk.info.nnz_kept


def test_missing_file(acc, tmp_path):
    with pytest.raises(acc.AccelesError, match="cannot open"):
        acc.CSRMatrix.load(str(tmp_path / "none.csr"))
