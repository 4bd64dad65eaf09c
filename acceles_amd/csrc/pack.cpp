// pack.cpp -- low-bit packing.
//
//   * Sliced packets (the GPU layout): quantize with one global FP32 scale
//     (quantify_util.py:8-28), drop entries that quantize to zero
//     (csr_util.py:13-22), split rows into SPLIT partitions
//     (quantify_util.py:184-189), order each partition's rows by entry count
//     and cut them into 64-row slices, one row per wavefront lane, each lane
//     a continuous bit stream of (column, value) entries with the dwords of
//     the 64 lanes interleaved so a wavefront's dword load is one 256-byte
//     contiguous segment.
//   * Ultra-CSR packets (the reference's FPGA format,
//     host_spmv_ucsr.cpp:388-425, fpga_utils.hpp:438-503), with the
//     x-decoder (fpga_utils.hpp:279-302) for reading them back.
#include "internal.hpp"

#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <exception>

using acc::fail;

// ---------------------------------------------------------------------------
// sliced packets
// ---------------------------------------------------------------------------

extern "C" int acc_pack_lowbit(const acc_csr *m, const acc_pack_opts *o, acc_packed **out) {
    if (!m || !o || !out) return fail(ACC_ERR_ARG, "null argument");
    if (o->bits < 2 || o->bits > 8) return fail(ACC_ERR_ARG, "value bits must be in [2, 8]");
    if (o->split < 1 || o->split > m->rows) return fail(ACC_ERR_ARG, "split must be in [1, rows]");
    int cb = o->col_bits > 0 ? o->col_bits : acc::bits_for(m->cols);
    if ((int64_t(1) << cb) < m->cols) return fail(ACC_ERR_ARG, "col_bits too small for the column count");
    int E = cb + o->bits;
    if (E < 10) { cb += 10 - E; E = 10; }  // narrowest decoder instantiated on the GPU
    if (E > 22) return fail(ACC_ERR_ARG, "column bits + value bits must be <= 22 (cols <= 16384 at 8 bits)");
    const int64_t nnz = m->nnz();

    acc_packed *pk = new acc_packed();
    pk->rows = m->rows;
    pk->cols = m->cols;
    pk->split = o->split;
    pk->bits = o->bits;
    pk->col_bits = cb;
    pk->entry_bits = E;
    pk->is_signed = o->is_signed ? 1 : 0;
    pk->nnz = nnz;

    if (!pk->is_signed) {
        int neg = 0;
#pragma omp parallel for reduction(| : neg)
        for (int64_t i = 0; i < nnz; ++i)
            if (m->val[(size_t)i] < 0.f) neg = 1;
        if (neg) { delete pk; return fail(ACC_ERR_ARG, "unsigned packing needs non-negative values (use is_signed=1)"); }
    }
    std::vector<int32_t> q((size_t)nnz);
    int rc = acc_quantize(nnz, m->val.data(), o->bits, pk->is_signed, &pk->scale, q.data());
    if (rc != ACC_OK) { delete pk; return rc; }

    const int32_t R = m->rows;
    std::vector<int32_t> kept((size_t)R);
    int64_t nkept = 0;
#pragma omp parallel for reduction(+ : nkept) schedule(static, 4096)
    for (int32_t r = 0; r < R; ++r) {
        int32_t c = 0;
        for (int64_t j = m->ptr[(size_t)r]; j < m->ptr[(size_t)r + 1]; ++j) c += q[(size_t)j] != 0;
        kept[(size_t)r] = c;
        nkept += c;
    }
    pk->nnz_kept = nkept;

    // Per (row, column) sums of the kept values: rows may name a column twice
    // (the reference generators sample columns with replacement,
    // create_matrices.py:52-54), and a dense layout holds a row's values in
    // one column as their sum.  Records the rows with repeats and the range
    // of the sums.
    int64_t dups = 0;
    int32_t mmax = 0, mmin = 0;
#pragma omp parallel reduction(+ : dups) reduction(max : mmax) reduction(min : mmin)
    {
        std::vector<uint32_t> seen((size_t)(m->cols + 31) / 32 + 1, 0u);
        std::vector<int32_t> sum((size_t)m->cols, 0);
#pragma omp for schedule(static, 4096)
        for (int32_t r = 0; r < R; ++r) {
            bool dup = false;
            const int64_t b = m->ptr[(size_t)r], e = m->ptr[(size_t)r + 1];
            for (int64_t j = b; j < e; ++j) {
                if (q[(size_t)j] == 0) continue;
                const uint32_t c = (uint32_t)m->idx[(size_t)j];
                uint32_t &w = seen[c >> 5];
                if (w & (1u << (c & 31))) dup = true;
                w |= 1u << (c & 31);
                sum[c] += q[(size_t)j];
            }
            for (int64_t j = b; j < e; ++j) {
                const uint32_t c = (uint32_t)m->idx[(size_t)j];
                if (seen[c >> 5] & (1u << (c & 31))) {
                    mmax = std::max(mmax, sum[c]);
                    mmin = std::min(mmin, sum[c]);
                    sum[c] = 0;
                    seen[c >> 5] &= ~(1u << (c & 31));
                }
            }
            dups += dup;
        }
    }
    pk->dup_rows = dups;
    pk->merged_max = mmax;
    pk->merged_min = mmin;

    // partitions (quantify_util.py:185-189): rows/split each, remainder to the last
    const int S = o->split;
    const int32_t psize = R / S;
    pk->part_row_begin.resize((size_t)S + 1);
    for (int p = 0; p < S; ++p) pk->part_row_begin[(size_t)p] = p * psize;
    pk->part_row_begin[(size_t)S] = R;
    pk->part_slice_begin.resize((size_t)S + 1);
    pk->part_slice_begin[0] = 0;
    for (int p = 0; p < S; ++p) {
        int32_t n = pk->part_row_begin[(size_t)p + 1] - pk->part_row_begin[(size_t)p];
        pk->part_slice_begin[(size_t)p + 1] = pk->part_slice_begin[(size_t)p] + (n + 63) / 64;
    }
    const int64_t NS = pk->part_slice_begin[(size_t)S];
    pk->perm.assign((size_t)NS * 64, -1);
    pk->slice_ndw.assign((size_t)NS, 0);
    pk->slice_off.assign((size_t)NS, 0);

    // Random-CSR (README.md:5; quantify_util.py:491 shuffles the rows before
    // the partitions): with a seed, partition p holds rows order[b .. e) of
    // the seeded shuffle instead of rows b .. e-1.  Rows keep their ids (the
    // slot -> row map names them), so results and ties are unchanged in kind.
    pk->shuffle_seed = o->shuffle_seed;
    std::vector<int32_t> order;
    if (o->shuffle_seed) order = acc::row_shuffle(R, o->shuffle_seed);
    const int32_t *ord = order.empty() ? nullptr : order.data();
    // order rows of each partition by kept count descending, then by their
    // place in the partition (counting sort: stable, O(rows))
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < S; ++p) {
        int32_t b = pk->part_row_begin[(size_t)p], e = pk->part_row_begin[(size_t)p + 1];
        auto row_at = [&](int32_t i) { return ord ? ord[i] : i; };
        int32_t maxc = 0;
        for (int32_t i = b; i < e; ++i) maxc = std::max(maxc, kept[(size_t)row_at(i)]);
        std::vector<int64_t> start((size_t)maxc + 2, 0);
        for (int32_t i = b; i < e; ++i) start[(size_t)(maxc - kept[(size_t)row_at(i)]) + 1]++;
        for (int32_t c = 0; c <= maxc; ++c) start[(size_t)c + 1] += start[(size_t)c];
        int32_t *dst = pk->perm.data() + pk->part_slice_begin[(size_t)p] * 64;
        for (int32_t i = b; i < e; ++i) {
            const int32_t r = row_at(i);
            dst[start[(size_t)(maxc - kept[(size_t)r])]++] = r;
        }
        for (int64_t s = pk->part_slice_begin[(size_t)p]; s < pk->part_slice_begin[(size_t)p + 1]; ++s) {
            int32_t r0 = pk->perm[(size_t)s * 64];  // longest row of the slice
            int64_t cnt = r0 >= 0 ? kept[(size_t)r0] : 0;
            pk->slice_ndw[(size_t)s] = (uint32_t)((cnt * E + 31) / 32);
        }
    }
    uint64_t off = 0;
    for (int64_t s = 0; s < NS; ++s) {
        if (off > 0xFFFFFFFFull) { delete pk; return fail(ACC_ERR_RANGE, "packed matrix exceeds 2^38 bytes"); }
        pk->slice_off[(size_t)s] = (uint32_t)off;
        off += pk->slice_ndw[(size_t)s];
    }
    pk->words.assign((size_t)off * 64, 0u);

    const uint32_t vmask = (1u << o->bits) - 1u;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t s = 0; s < NS; ++s) {
        uint32_t *base = pk->words.data() + (size_t)pk->slice_off[(size_t)s] * 64;
        for (int l = 0; l < 64; ++l) {
            int32_t r = pk->perm[(size_t)s * 64 + l];
            if (r < 0) continue;
            int64_t j = 0;
            for (int64_t t = m->ptr[(size_t)r]; t < m->ptr[(size_t)r + 1]; ++t) {
                int32_t qv = q[(size_t)t];
                if (qv == 0) continue;
                uint32_t x = (uint32_t)m->idx[(size_t)t] | (((uint32_t)qv & vmask) << cb);
                int64_t pos = j * E;
                int64_t w = pos >> 5;
                int sh = (int)(pos & 31);
                base[(size_t)w * 64 + l] |= x << sh;
                if (sh + E > 32) base[(size_t)(w + 1) * 64 + l] |= x >> (32 - sh);
                ++j;
            }
        }
    }
    *out = pk;
    return ACC_OK;
}

extern "C" int acc_packed_info(const acc_packed *pk, acc_packed_info_t *info) {
    if (!pk || !info) return fail(ACC_ERR_ARG, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->rows = pk->rows;
    info->cols = pk->cols;
    info->split = pk->split;
    info->bits = pk->bits;
    info->col_bits = pk->col_bits;
    info->entry_bits = pk->entry_bits;
    info->is_signed = pk->is_signed;
    info->nnz = pk->nnz;
    info->nnz_kept = pk->nnz_kept;
    info->nslices = pk->nslices();
    info->stream_bytes = (int64_t)pk->words.size() * 4 + pk->nslices() * 8;
    info->perm_bytes = (int64_t)pk->perm.size() * 4;
    info->bytes_per_nnz = pk->nnz > 0 ? (double)info->stream_bytes / (double)pk->nnz : 0.0;
    info->scale = pk->scale;
    info->shuffle_seed = pk->shuffle_seed;
    return ACC_OK;
}

namespace {

// The kept entries of lane l of slice s: (column, signed value), in stream order.
void lane_entries(const acc_packed *pk, int64_t s, int l, std::vector<std::pair<int32_t, int32_t>> &out) {
    const int E = pk->entry_bits, cb = pk->col_bits, vb = pk->bits;
    const uint32_t cmask = (1u << cb) - 1u, vmask = (1u << vb) - 1u;
    const uint32_t *base = pk->words.data() + (size_t)pk->slice_off[(size_t)s] * 64;
    const int64_t ndw = pk->slice_ndw[(size_t)s], nent = ndw * 32 / E;
    out.clear();
    for (int64_t j = 0; j < nent; ++j) {
        const int64_t pos = j * E, w = pos >> 5;
        const int sh = (int)(pos & 31);
        const uint64_t lo = base[(size_t)w * 64 + l];
        const uint64_t hi = w + 1 < ndw ? base[(size_t)(w + 1) * 64 + l] : 0;
        const uint32_t x = (uint32_t)(((hi << 32) | lo) >> sh);
        const uint32_t v = (x >> cb) & vmask;
        if (v == 0) break;
        int32_t sv = (int32_t)v;
        if (pk->is_signed && (v >> (vb - 1))) sv = (int32_t)(v | ~vmask);
        out.push_back({(int32_t)(x & cmask), sv});
    }
}

// The image entries of one lane: (image byte, column), repeated columns
// summed, in column order; false when a sum leaves the byte.
bool lane_image_entries(const acc_packed *pk, int64_t s, int l, int u8off, std::vector<std::pair<int32_t, int32_t>> &ent,
                        std::vector<std::pair<uint32_t, int32_t>> &out) {
    lane_entries(pk, s, l, ent);
    std::stable_sort(ent.begin(), ent.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    out.clear();
    for (size_t i = 0; i < ent.size();) {
        size_t j = i;
        int64_t sum = 0;
        while (j < ent.size() && ent[j].first == ent[i].first) sum += ent[j++].second;
        const int64_t b = u8off ? sum - 128 : sum;
        if (b < -128 || b > 127) return false;
        out.push_back({(uint32_t)(b & 0xFF), ent[i].first});
        i = j;
    }
    return true;
}

// The entries of decode wavefront w (rows 16w .. 16w+15) of slice s: byte |
// (offset << 8), offset = (row - 16w) * stride + column, rows in order.
bool wave_image_entries(const acc_packed *pk, int64_t s, int w, int u8off, uint32_t stride,
                        std::vector<std::pair<int32_t, int32_t>> &ent, std::vector<std::pair<uint32_t, int32_t>> &row,
                        std::vector<uint32_t> &out) {
    out.clear();
    for (int r = 0; r < 16; ++r) {
        const int l = 16 * w + r;
        if (pk->perm[(size_t)s * 64 + l] < 0) continue;
        if (!lane_image_entries(pk, s, l, u8off, ent, row)) return false;
        for (const auto &e : row) out.push_back(e.first | (((uint32_t)r * stride + (uint32_t)e.second) << 8));
    }
    return true;
}

}  // namespace

int acc::image_packets(const acc_packed *pk, int u8off, int ob, uint32_t stride, std::vector<uint32_t> &words,
                       std::vector<std::pair<uint32_t, uint32_t>> &slices) {
    const int64_t NS = pk->nslices();
    const int E = ob + 8;
    if (ob < 1 || E > 32 || (int64_t(16) * stride) > (int64_t(1) << ob) || (int64_t)stride <= pk->cols)
        return fail(ACC_ERR_ARG, "image packets: offset bits");
    slices.assign((size_t)NS, {0u, 0u});
    // per slice and wavefront: entries per lane, a multiple of 4 (<= 256)
    std::vector<uint32_t> n4((size_t)NS * 4, 0);
    int bad = 0;
#pragma omp parallel
    {
        std::vector<std::pair<int32_t, int32_t>> ent;
        std::vector<std::pair<uint32_t, int32_t>> row;
        std::vector<uint32_t> out;
#pragma omp for schedule(dynamic, 64) reduction(| : bad)
        for (int64_t s = 0; s < NS; ++s)
            for (int w = 0; w < 4; ++w) {
                if (!wave_image_entries(pk, s, w, u8off, stride, ent, row, out)) bad = 1;
                const uint64_t n = (out.size() + 63) / 64;
                n4[(size_t)s * 4 + w] = (uint32_t)((n + 3) / 4 * 4);
            }
    }
    if (bad) return fail(ACC_ERR_RANGE, "image packets: a row's values in one column sum past one byte");
    auto ndw_of = [E](uint32_t n) { return (n * (uint32_t)E + 31u) / 32u; };
    uint64_t off = 0;
    for (int64_t s = 0; s < NS; ++s) {
        uint32_t ndw = 0, packed = 0;
        for (int w = 0; w < 4; ++w) {
            const uint32_t n = n4[(size_t)s * 4 + w];
            if (n > 256) return fail(ACC_ERR_RANGE, "image packets: more than 256 entries per lane");
            ndw += ndw_of(n);
            packed |= (n / 4) << (8 * w);
        }
        if (off + ndw > 0xFFFFFFFFull) return fail(ACC_ERR_RANGE, "image packets exceed 2^40 bytes");
        slices[(size_t)s] = {(uint32_t)off, packed};
        off += ndw;
    }
    words.assign((size_t)off * 64, 0u);
#pragma omp parallel
    {
        std::vector<std::pair<int32_t, int32_t>> ent;
        std::vector<std::pair<uint32_t, int32_t>> row;
        std::vector<uint32_t> out;
#pragma omp for schedule(dynamic, 64)
        for (int64_t s = 0; s < NS; ++s) {
            uint32_t *base = words.data() + (size_t)slices[(size_t)s].first * 64;
            for (int w = 0; w < 4; ++w) {
                const uint32_t n = n4[(size_t)s * 4 + w];
                if (n == 0) continue;
                wave_image_entries(pk, s, w, u8off, stride, ent, row, out);
                // lane l takes entries l*m .. l*m + m-1 of the wavefront's
                // list, m = ceil(size / 64) (so one store instruction spreads
                // over the rows); its slots m .. n-1 repeat entries of the next
                // lanes, wrapping round to the list's start (the same byte to
                // the same place again, never twice in one instruction)
                const size_t m = (out.size() + 63) / 64;
                for (uint32_t j = 0; j < n; ++j)
                    for (int l = 0; l < 64; ++l) {
                        const size_t i = (size_t)l * m + j;
                        const uint32_t v = out[i % out.size()];
                        const uint32_t pos = j * (uint32_t)E, wi = pos >> 5, sh = pos & 31;
                        base[(size_t)wi * 64 + l] |= v << sh;
                        if (sh + E > 32) base[(size_t)(wi + 1) * 64 + l] |= v >> (32 - sh);
                    }
                base += (size_t)ndw_of(n) * 64;
            }
        }
    }
    return ACC_OK;
}

extern "C" int acc_unpack_lowbit(const acc_packed *pk, int64_t *ptr, int32_t *idx, int32_t *qval) {
    if (!pk || !ptr) return fail(ACC_ERR_ARG, "null argument");
    const int E = pk->entry_bits, cb = pk->col_bits, vb = pk->bits;
    const uint32_t cmask = (1u << cb) - 1u, vmask = (1u << vb) - 1u;
    // per row: its slot (slice, lane)
    std::vector<int64_t> slot_of((size_t)pk->rows, -1);
    for (int64_t i = 0; i < (int64_t)pk->perm.size(); ++i)
        if (pk->perm[(size_t)i] >= 0) slot_of[(size_t)pk->perm[(size_t)i]] = i;
    // decode every lane stream; entries end at the first zero value
    std::vector<std::vector<std::pair<int32_t, int32_t>>> rows((size_t)pk->rows);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t r = 0; r < pk->rows; ++r) {
        int64_t sl = slot_of[(size_t)r];
        if (sl < 0) continue;
        int64_t s = sl / 64;
        int l = (int)(sl % 64);
        const uint32_t *base = pk->words.data() + (size_t)pk->slice_off[(size_t)s] * 64;
        int64_t nent = (int64_t)pk->slice_ndw[(size_t)s] * 32 / E;
        for (int64_t j = 0; j < nent; ++j) {
            int64_t pos = j * E, w = pos >> 5;
            int sh = (int)(pos & 31);
            uint64_t lo = base[(size_t)w * 64 + l];
            uint64_t hi = (w + 1 < (int64_t)pk->slice_ndw[(size_t)s]) ? base[(size_t)(w + 1) * 64 + l] : 0;
            uint32_t x = (uint32_t)(((hi << 32) | lo) >> sh);
            uint32_t v = (x >> cb) & vmask;
            if (v == 0) break;
            int32_t sv = (int32_t)v;
            if (pk->is_signed && (v >> (vb - 1))) sv = (int32_t)(v | ~vmask);
            rows[(size_t)r].push_back({(int32_t)(x & cmask), sv});
        }
    }
    ptr[0] = 0;
    for (int32_t r = 0; r < pk->rows; ++r) ptr[(size_t)r + 1] = ptr[(size_t)r] + (int64_t)rows[(size_t)r].size();
    if (idx || qval) {
        for (int32_t r = 0; r < pk->rows; ++r) {
            int64_t d = ptr[(size_t)r];
            for (auto &e : rows[(size_t)r]) {
                if (idx) idx[d] = e.first;
                if (qval) qval[d] = e.second;
                ++d;
            }
        }
    }
    return ACC_OK;
}

extern "C" void acc_packed_destroy(acc_packed *pk) { delete pk; }

// ---------------------------------------------------------------------------
// Ultra-CSR reference packets
// ---------------------------------------------------------------------------

extern "C" int acc_ultracsr_entries(int col_bits, int val_bits) {
    if (col_bits < 1 || val_bits < 1 || 1 + col_bits + val_bits > 511)
        return fail(ACC_ERR_ARG, "bad Ultra-CSR field widths");
    return 511 / (1 + col_bits + val_bits);
}

namespace {
inline void set_field(uint64_t *w, int pos, int len, uint64_t v) {
    v &= (len >= 64) ? ~0ull : ((1ull << len) - 1);
    int i = pos >> 6, sh = pos & 63;
    w[i] |= v << sh;
    if (sh + len > 64) w[i + 1] |= v >> (64 - sh);
}
inline uint64_t get_field(const uint64_t *w, int pos, int len) {
    int i = pos >> 6, sh = pos & 63;
    uint64_t v = w[i] >> sh;
    if (sh + len > 64) v |= w[i + 1] << (64 - sh);
    return v & ((len >= 64) ? ~0ull : ((1ull << len) - 1));
}
}  // namespace

extern "C" int64_t acc_pack_ultracsr(int64_t nnz, const int32_t *row, const int32_t *col, const int32_t *qval,
                                     int32_t last_row, int col_bits, int val_bits, uint64_t *packets, int64_t cap) {
    int n = acc_ultracsr_entries(col_bits, val_bits);
    if (n < 0) return n;
    if (nnz < 0 || (nnz > 0 && (!row || !col || !qval || !packets))) return fail(ACC_ERR_ARG, "bad arguments");
    int64_t np = (nnz + n - 1) / n;
    if (np > cap) return fail(ACC_ERR_ARG, "packet buffer too small");
    std::memset(packets, 0, sizeof(uint64_t) * 8 * (size_t)np);
    int32_t prev = last_row;
    for (int64_t i = 0; i < nnz; ++i) {
        uint64_t *w = packets + 8 * (i / n);
        int j = (int)(i % n);
        if (col[i] < 0 || (col_bits < 31 && col[i] >= (1 << col_bits)))
            return fail(ACC_ERR_RANGE, "column does not fit the column field");
        set_field(w, j, 1, row[i] != prev ? 1u : 0u);
        prev = row[i];
        set_field(w, n + j * col_bits, col_bits, (uint32_t)col[i]);
        set_field(w, n + n * col_bits + j * val_bits, val_bits, (uint32_t)qval[i]);
    }
    return np;
}

extern "C" int acc_xdecode(uint32_t flags, int n, int32_t carry, int32_t *x) {
    if (!x || n < 1 || n > 32) return fail(ACC_ERR_ARG, "n must be in [1, 32]");
    for (int i = 0; i < n; ++i) {
        uint32_t mask = i >= 31 ? 0xFFFFFFFFu : ((2u << i) - 1u);
        x[i] = carry + __builtin_popcount(flags & mask);
    }
    return ACC_OK;
}

extern "C" int acc_unpack_ultracsr(const uint64_t *packets, int64_t nnz, int col_bits, int val_bits, int32_t carry0,
                                   int32_t *xrow, int32_t *col, int32_t *qval) {
    int n = acc_ultracsr_entries(col_bits, val_bits);
    if (n < 0) return n;
    if (n > 32) return fail(ACC_ERR_ARG, "more than 32 entries per packet: flags exceed the decoder width");
    int32_t carry = carry0, x[32];
    for (int64_t p = 0; p * n < nnz; ++p) {
        const uint64_t *w = packets + 8 * p;
        acc_xdecode((uint32_t)get_field(w, 0, n), n, carry, x);
        carry = x[n - 1];
        for (int j = 0; j < n && p * n + j < nnz; ++j) {
            int64_t i = p * n + j;
            if (xrow) xrow[i] = x[j];
            if (col) col[i] = (int32_t)get_field(w, n + j * col_bits, col_bits);
            uint32_t v = (uint32_t)get_field(w, n + n * col_bits + j * val_bits, val_bits);
            if (v >> (val_bits - 1)) v |= ~((1u << val_bits) - 1u);
            if (qval) qval[i] = (int32_t)v;
        }
    }
    return ACC_OK;
}

// ---------------------------------------------------------------------------
// Binary files of CSR matrices and packets (the load-time cache the reference
// lacks: its drivers re-read .mtx text and re-pack on every run)
// ---------------------------------------------------------------------------

namespace {

const char kCsrMagic[8] = {'A', 'C', 'C', 'C', 'S', 'R', '0', '1'};
const char kPackMagic[8] = {'A', 'C', 'C', 'P', 'K', 'T', '0', '2'};

struct File {
    FILE *f = nullptr;
    explicit File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

template <typename T>
bool put(FILE *f, const T &v) { return std::fwrite(&v, sizeof(T), 1, f) == 1; }
template <typename T>
bool get(FILE *f, T &v) { return std::fread(&v, sizeof(T), 1, f) == 1; }
template <typename T>
bool put_vec(FILE *f, const std::vector<T> &v) {
    const uint64_t n = v.size();
    return put(f, n) && (n == 0 || std::fwrite(v.data(), sizeof(T), n, f) == n);
}
// Bytes from the current position to the end of the file.
uint64_t bytes_left(FILE *f) {
    const long pos = std::ftell(f);
    if (pos < 0 || std::fseek(f, 0, SEEK_END) != 0) return 0;
    const long end = std::ftell(f);
    std::fseek(f, pos, SEEK_SET);
    return end > pos ? (uint64_t)(end - pos) : 0;
}

// A vector of at most max_n elements; a length field larger than what the
// file still holds is refused before anything is allocated.
template <typename T>
bool get_vec(FILE *f, std::vector<T> &v, uint64_t max_n) {
    uint64_t n = 0;
    if (!get(f, n) || n > max_n || n > bytes_left(f) / sizeof(T)) return false;
    v.resize(n);
    return n == 0 || std::fread(v.data(), sizeof(T), n, f) == n;
}

// What the kernels and the host index with, checked on a loaded packet file;
// the entry statistics are recomputed from the decoded entries rather than
// taken from the header.
bool valid_packed(acc_packed *pk) {
    const int S = pk->split;
    const int64_t NS = pk->nslices();
    if (pk->entry_bits != pk->col_bits + pk->bits || pk->entry_bits < 10 || pk->entry_bits > 22 || pk->bits < 2 ||
        pk->bits > 8 || pk->cols < 1 || pk->cols > (1 << pk->col_bits) || pk->rows < 1)
        return false;
    if (pk->part_row_begin.size() != (size_t)S + 1 || pk->part_slice_begin.size() != (size_t)S + 1 ||
        pk->slice_ndw.size() != (size_t)NS || pk->perm.size() != (size_t)NS * 64)
        return false;
    // partitions: row ranges from 0 to rows, slice ranges from 0 to NS, each
    // partition ceil(rows_p / 64) slices
    if (pk->part_row_begin[0] != 0 || pk->part_row_begin[(size_t)S] != pk->rows || pk->part_slice_begin[0] != 0 ||
        pk->part_slice_begin[(size_t)S] != NS)
        return false;
    for (int p = 0; p < S; ++p) {
        const int64_t nr = (int64_t)pk->part_row_begin[(size_t)p + 1] - pk->part_row_begin[(size_t)p];
        const int64_t ns = pk->part_slice_begin[(size_t)p + 1] - pk->part_slice_begin[(size_t)p];
        if (nr < 0 || ns != (nr + 63) / 64) return false;
    }
    for (int64_t s = 0; s < NS; ++s)
        if (((uint64_t)pk->slice_off[(size_t)s] + pk->slice_ndw[(size_t)s]) * 64 > pk->words.size()) return false;
    // lanes: padding (-1) only closes a partition's last slice, and the
    // partition's real lanes name each of its rows_p rows; over the whole
    // matrix every row appears exactly once
    std::vector<uint8_t> seen((size_t)pk->rows, 0);
    for (int p = 0; p < S; ++p) {
        const int64_t nr = (int64_t)pk->part_row_begin[(size_t)p + 1] - pk->part_row_begin[(size_t)p];
        const int64_t base = pk->part_slice_begin[(size_t)p] * 64;
        const int64_t end = pk->part_slice_begin[(size_t)p + 1] * 64;
        for (int64_t i = base; i < end; ++i) {
            const int32_t r = pk->perm[(size_t)i];
            if (i - base < nr) {
                if (r < 0 || r >= pk->rows || seen[(size_t)r]) return false;
                seen[(size_t)r] = 1;
            } else if (r != -1) {
                return false;
            }
        }
    }
    // entries: every non-zero value names a column below cols; recount the
    // kept entries, the rows naming a column twice and the range of a row's
    // per-column sums (the dense layout's eligibility depends on them)
    const int E = pk->entry_bits, cb = pk->col_bits, vb = pk->bits;
    const uint32_t cmask = (1u << cb) - 1u, vmask = (1u << vb) - 1u;
    bool ok = true;
    int64_t kept = 0, dups = 0;
    int32_t mmax = 0, mmin = 0;
#pragma omp parallel reduction(&& : ok) reduction(+ : kept, dups) reduction(max : mmax) reduction(min : mmin)
    {
        std::vector<int32_t> sum((size_t)pk->cols, 0);
        std::vector<uint32_t> mark((size_t)pk->cols / 32 + 1, 0u);
        std::vector<int32_t> used;
#pragma omp for schedule(dynamic, 64)
        for (int64_t s = 0; s < NS; ++s) {
            const uint32_t *w = pk->words.data() + (size_t)pk->slice_off[(size_t)s] * 64;
            const int64_t ndw = pk->slice_ndw[(size_t)s], nent = ndw * 32 / E;
            for (int l = 0; l < 64 && ok; ++l) {
                bool dup = false;
                used.clear();
                for (int64_t j = 0; j < nent; ++j) {
                    const int64_t pos = j * E, wi = pos >> 5;
                    const int sh = (int)(pos & 31);
                    const uint64_t lo = w[(size_t)wi * 64 + l];
                    const uint64_t hi = wi + 1 < ndw ? w[(size_t)(wi + 1) * 64 + l] : 0;
                    const uint32_t x = (uint32_t)(((hi << 32) | lo) >> sh);
                    const uint32_t v = (x >> cb) & vmask, c = x & cmask;
                    if (v == 0) continue;
                    if ((int64_t)c >= pk->cols) {
                        ok = false;
                        break;
                    }
                    int32_t sv = (int32_t)v;
                    if (pk->is_signed && (v >> (vb - 1))) sv = (int32_t)(v | ~vmask);
                    if (pk->perm[(size_t)s * 64 + l] < 0) continue;
                    ++kept;
                    uint32_t &mk = mark[c >> 5];
                    if (mk & (1u << (c & 31))) {
                        dup = true;
                    } else {
                        mk |= 1u << (c & 31);
                        used.push_back((int32_t)c);
                    }
                    sum[c] += sv;
                }
                for (int32_t c : used) {
                    mmax = std::max(mmax, sum[(size_t)c]);
                    mmin = std::min(mmin, sum[(size_t)c]);
                    sum[(size_t)c] = 0;
                    mark[(size_t)c >> 5] = 0u;
                }
                dups += dup;
            }
        }
    }
    if (!ok) return false;
    pk->nnz_kept = kept;
    pk->dup_rows = dups;
    pk->merged_max = mmax;
    pk->merged_min = mmin;
    return true;
}

}  // namespace

extern "C" int acc_csr_save(const acc_csr *m, const char *path) {
    if (!m || !path) return fail(ACC_ERR_ARG, "null argument");
    File w(path, "wb");
    if (!w.f) return fail(ACC_ERR_IO, std::string("cannot open ") + path + " for writing");
    bool ok = std::fwrite(kCsrMagic, 1, 8, w.f) == 8 && put(w.f, m->rows) && put(w.f, m->cols) &&
              put_vec(w.f, m->ptr) && put_vec(w.f, m->idx) && put_vec(w.f, m->val);
    return ok ? ACC_OK : fail(ACC_ERR_IO, std::string("write failed: ") + path);
}

extern "C" int acc_csr_load(const char *path, acc_csr **out) {
    if (!path || !out) return fail(ACC_ERR_ARG, "null argument");
    try {
        File r(path, "rb");
        if (!r.f) return fail(ACC_ERR_IO, std::string("cannot open ") + path);
        char magic[8];
        acc_csr *m = new acc_csr();
        bool ok = std::fread(magic, 1, 8, r.f) == 8 && std::memcmp(magic, kCsrMagic, 8) == 0 && get(r.f, m->rows) &&
                  get(r.f, m->cols) && m->rows >= 0 && m->cols >= 0 && get_vec(r.f, m->ptr, (uint64_t)m->rows + 1) &&
                  m->ptr.size() == (size_t)m->rows + 1 && get_vec(r.f, m->idx, UINT64_MAX / 8) &&
                  get_vec(r.f, m->val, m->idx.size());
        // the arrays must describe a CSR matrix: monotone ptr, in-range columns
        if (ok) ok = m->ptr[0] == 0 && m->ptr[(size_t)m->rows] == (int64_t)m->idx.size() && m->val.size() == m->idx.size();
        for (int32_t i = 0; ok && i < m->rows; ++i) ok = m->ptr[(size_t)i] <= m->ptr[(size_t)i + 1];
        for (size_t j = 0; ok && j < m->idx.size(); ++j) ok = m->idx[j] >= 0 && m->idx[j] < m->cols;
        if (!ok) {
            delete m;
            return fail(ACC_ERR_FORMAT, std::string("not a CSR file of this library: ") + path);
        }
        *out = m;
        return ACC_OK;
    } catch (const std::exception &e) {
        return fail(ACC_ERR_NOMEM, std::string("loading ") + path + ": " + e.what());
    }
}

extern "C" int acc_packed_save(const acc_packed *pk, const char *path) {
    if (!pk || !path) return fail(ACC_ERR_ARG, "null argument");
    File w(path, "wb");
    if (!w.f) return fail(ACC_ERR_IO, std::string("cannot open ") + path + " for writing");
    bool ok = std::fwrite(kPackMagic, 1, 8, w.f) == 8 && put(w.f, pk->rows) && put(w.f, pk->cols) &&
              put(w.f, pk->split) && put(w.f, pk->bits) && put(w.f, pk->col_bits) && put(w.f, pk->entry_bits) &&
              put(w.f, pk->is_signed) && put(w.f, pk->scale) && put(w.f, pk->nnz) && put(w.f, pk->nnz_kept) &&
              put(w.f, pk->dup_rows) && put(w.f, pk->merged_max) && put(w.f, pk->merged_min) &&
              put(w.f, pk->shuffle_seed) && put_vec(w.f, pk->part_row_begin) && put_vec(w.f, pk->part_slice_begin) && put_vec(w.f, pk->slice_off) &&
              put_vec(w.f, pk->slice_ndw) && put_vec(w.f, pk->perm) && put_vec(w.f, pk->words);
    return ok ? ACC_OK : fail(ACC_ERR_IO, std::string("write failed: ") + path);
}

extern "C" int acc_packed_load(const char *path, acc_packed **out) {
    if (!path || !out) return fail(ACC_ERR_ARG, "null argument");
    try {
        File r(path, "rb");
        if (!r.f) return fail(ACC_ERR_IO, std::string("cannot open ") + path);
        char magic[8];
        acc_packed *pk = new acc_packed();
        bool ok = std::fread(magic, 1, 8, r.f) == 8 && std::memcmp(magic, kPackMagic, 8) == 0 && get(r.f, pk->rows) &&
                  get(r.f, pk->cols) && get(r.f, pk->split) && get(r.f, pk->bits) && get(r.f, pk->col_bits) &&
                  get(r.f, pk->entry_bits) && get(r.f, pk->is_signed) && get(r.f, pk->scale) && get(r.f, pk->nnz) &&
                  get(r.f, pk->nnz_kept) && get(r.f, pk->dup_rows) && get(r.f, pk->merged_max) &&
                  get(r.f, pk->merged_min) && get(r.f, pk->shuffle_seed) && pk->split >= 1 && pk->rows >= pk->split && pk->col_bits >= 1 &&
                  pk->col_bits <= 20 && get_vec(r.f, pk->part_row_begin, (uint64_t)pk->split + 1) &&
                  get_vec(r.f, pk->part_slice_begin, (uint64_t)pk->split + 1) &&
                  get_vec(r.f, pk->slice_off, UINT64_MAX / 8) && get_vec(r.f, pk->slice_ndw, pk->slice_off.size()) &&
                  get_vec(r.f, pk->perm, pk->slice_off.size() * 64) && get_vec(r.f, pk->words, UINT64_MAX / 8);
        if (ok) ok = valid_packed(pk);
        if (!ok) {
            delete pk;
            return fail(ACC_ERR_FORMAT, std::string("not a packet file of this library: ") + path);
        }
        *out = pk;
        return ACC_OK;
    } catch (const std::exception &e) {
        return fail(ACC_ERR_NOMEM, std::string("loading ") + path + ": " + e.what());
    }
}
