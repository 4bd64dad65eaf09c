// internal.hpp -- host-side structures shared by the library's translation units.
#pragma once

#include "../../include/acceles.h"

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

struct acc_coo {
    int32_t rows = 0, cols = 0;
    std::vector<int32_t> row, col;
    std::vector<float> val;
};

struct acc_csr {
    int32_t rows = 0, cols = 0;
    std::vector<int64_t> ptr;  // rows + 1
    std::vector<int32_t> idx;
    std::vector<float> val;
    int64_t nnz() const { return (int64_t)idx.size(); }
};

// Low-bit sliced packets.  Rows of each SPLIT partition are ordered by kept
// entry count (descending, ties by row) and cut into 64-row slices, one row
// per wavefront lane.  Lane l of slice s owns a bit stream of entries
// (col: col_bits, value: bits), entry j at bits [j*E, (j+1)*E), E = entry_bits;
// dword d of that stream lives at words[(slice_off[s] + d) * 64 + l].
struct acc_packed {
    int32_t rows = 0, cols = 0, split = 0;
    int32_t bits = 8, col_bits = 10, entry_bits = 18, is_signed = 0;
    float scale = 0.f;
    int64_t nnz = 0, nnz_kept = 0;
    int64_t dup_rows = 0;                   // rows holding a column twice among kept entries
    uint64_t shuffle_seed = 0;              // Random-CSR: rows dealt to partitions in row_shuffle order (0: none)
    int32_t merged_max = 0, merged_min = 0; // range of the per (row, column) sums of kept values
    std::vector<int32_t> part_row_begin;    // split + 1
    std::vector<int64_t> part_slice_begin;  // split + 1
    std::vector<uint32_t> slice_off;        // in units of 64 words
    std::vector<uint32_t> slice_ndw;        // dwords per lane
    std::vector<int32_t> perm;              // nslices * 64, -1 = padding lane
    std::vector<uint32_t> words;
    int64_t nslices() const { return (int64_t)slice_off.size(); }
};

namespace acc {

int fail(int code, const std::string &msg);
void clear_error();

// ranking keys: larger key ranks first; score descending then row ascending
inline uint64_t key_i32(int32_t score, int32_t row) {
    return ((uint64_t)((uint32_t)score ^ 0x80000000u) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)row);
}
inline uint32_t ordered_f32(float f) {
    if (f == 0.0f) f = 0.0f;  // -0 == +0 ranks as a tie
    uint32_t b;
    __builtin_memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

int bits_for(int32_t n);  // ceil(log2(n)), at least 1
// Batched low-bit queries with a workspace of the caller's (acc_gpu_topk_
// lowbit_batch uses the handle's own, so one handle runs one batch at a
// time): several workspaces keep several batches of one matrix in flight on
// different streams.  The workspace grows on demand; bytes reports it.
struct BatchWorkspace;
BatchWorkspace *batch_ws_create();
void batch_ws_destroy(BatchWorkspace *w, int device);
int64_t batch_ws_bytes(const BatchWorkspace *w);
int topk_lowbit_batch_ws(acc_gpu_matrix *m, BatchWorkspace *w, const float *x_dev, int nq, const int32_t *k_host,
                         int ck, acc_element *out_dev, int out_stride, void *stream);

// Random-CSR row order: the seeded Fisher-Yates shuffle of 0 .. rows-1 that
// acc_permute_rows applies (out row i = in row order[i]).
std::vector<int32_t> row_shuffle(int32_t rows, uint64_t seed);

// Image packets: the slices of pk re-encoded for the matrix-core scores, one
// stream per slice and decode wavefront w (rows 16w .. 16w+15).  The entries
// of those rows, byte | (offset << 8) with offset = (row - 16w) * stride +
// column and byte = the LDS image's value (v - 128 with u8off, else v, as a
// byte), a row's repeated columns summed into one entry (the caller checked
// that the sums fit a byte), rows in order, are split among the 64 lanes
// (entry j of lane l = entry l*m + j modulo the list's size, m = ceil(size /
// 64)), n >= m entries of 8 + ob bits per lane (n a multiple of 4);
// lanes interleaved per dword as in pk, the four streams one after another.
// slices[s] = (offset in units of 64 dwords, n/4 of wavefront w in byte w).
// 16 * stride must fit ob bits and stride exceed cols.  Returns ACC_OK or an
// error for out-of-range sums.
int image_packets(const acc_packed *pk, int u8off, int ob, uint32_t stride, std::vector<uint32_t> &words,
                  std::vector<std::pair<uint32_t, uint32_t>> &slices);

// Device-memory admission: byte reservations against a fixed limit, each
// checked and taken under one lock, so concurrent callers cannot both pass
// against the same room.
class Budget {
  public:
    void set_limit(int64_t limit) {
        std::lock_guard<std::mutex> l(mu_);
        limit_ = limit;
    }
    // Reserves n bytes if they fit.
    bool reserve(int64_t n) {
        std::lock_guard<std::mutex> l(mu_);
        if (n < 0 || used_ + n > limit_) return false;
        used_ += n;
        return true;
    }
    // Replaces a reservation of `from` bytes by one of `to` bytes if the
    // result fits (shrinking always does).
    bool resize(int64_t from, int64_t to) {
        std::lock_guard<std::mutex> l(mu_);
        if (to > from && used_ - from + to > limit_) return false;
        used_ += to - from;
        return true;
    }
    void release(int64_t n) {
        std::lock_guard<std::mutex> l(mu_);
        used_ -= n;
    }
    int64_t used() const {
        std::lock_guard<std::mutex> l(mu_);
        return used_;
    }
    int64_t limit() const {
        std::lock_guard<std::mutex> l(mu_);
        return limit_;
    }

  private:
    mutable std::mutex mu_;
    int64_t limit_ = 0, used_ = 0;
};

}  // namespace acc
