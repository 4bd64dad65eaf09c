// k_batch.hip -- batched low-bit path, selection stages:
//   batch_select   per (query, partition): the partition's Top-M from the
//                  INT32 scores written by batch_scores (quantify_util.py:173-196)
//                  and the FP32 rescoring of those rows (quantify_util.py:354-367);
//   batch_final    per query: its Top-K of the rescored candidates.
#include "kernels_common.hpp"

__global__ __launch_bounds__(ACC_WG) void batch_select(BatchArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_n[ACC_WAVES];
    __shared__ int s_m;
    uint64_t *colbuf = (uint64_t *)smem;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int p = blockIdx.x, q = blockIdx.y;
    const int64_t b0 = a.part_slice[p] * 64, b1 = a.part_slice[p + 1] * 64;
    const int64_t chunks = (b1 - b0) / 64, per = (chunks + ACC_WAVES - 1) / ACC_WAVES;
    const int64_t w0 = b0 + wave * per * 64, w1 = min(b1, w0 + per * 64);
    const int32_t *Sq = a.S + (size_t)q * a.nslots;
    WaveTopM<2> col;
    col.init(colbuf + wave * WaveTopM<2>::cap, a.M);
    // U scores per lane per iteration; the rows of the slots that pass the
    // current threshold are fetched together, one latency per 64*U slots
    constexpr int U = 8;
    for (int64_t i0 = w0; i0 < w1; i0 += 64 * U) {
        int32_t sc[U];
#pragma unroll
        for (int u = 0; u < U; ++u)  // w1 - i0 is a multiple of 64: the guard is wave-uniform
            sc[u] = i0 + u * 64 < w1 ? __builtin_nontemporal_load(Sq + i0 + u * 64 + lane) : INT32_MIN;
        const int32_t ts = dkey_score_i32(col.thr);
        bool any = false;
#pragma unroll
        for (int u = 0; u < U; ++u) any |= i0 + u * 64 < w1 && sc[u] >= ts;
        if (__ballot(any) == 0) continue;
        int32_t row[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool maybe = i0 + u * 64 < w1 && sc[u] >= ts;
            row[u] = maybe ? a.perm[i0 + u * 64 + lane] : -1;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) col.push(row[u] >= 0, dkey_i32(sc[u], row[u]));
    }
    col.compact();
    if (lane == 0) s_n[wave] = col.n;
    __syncthreads();
    if (wave == 0) {
        wg_merge(col, colbuf, s_n, ACC_WAVES);
        if (lane == 0) s_m = col.n;
    }
    __syncthreads();
    const int mcount = s_m;
    const uint64_t *top = colbuf;
    const size_t base = ((size_t)q * a.split + p) * a.M;
    if (a.cand_rows_out) {
        for (int i = threadIdx.x; i < a.M; i += ACC_WG) {
            a.cand_rows_out[base + i] = i < mcount ? dkey_row(top[i]) : -1;
            a.cand_scores_out[base + i] = i < mcount ? dkey_score_i32(top[i]) : 0;
        }
    }
    if (!a.resc) return;
    const float *x = a.X + (size_t)q * a.cols;
    for (int b = 0; b < a.M; b += ACC_WG / a.G) {
        const int i = b + (int)threadIdx.x / a.G;
        const bool valid = i < mcount;
        const int32_t r = valid ? dkey_row(top[i]) : 0;
        const float sc = group_dot_rt(a.ptr, a.idx, a.val, x, r, valid, a.G);
        if ((threadIdx.x & (a.G - 1)) == 0 && i < a.M) a.resc[base + i] = valid ? dkey_f32(sc, r) : 0ull;
    }
}

// Per query q = blockIdx.x: its Top-K[q] of the split*M rescored keys.

template <int R>
__global__ __launch_bounds__(ACC_WG) void batch_final(BatchArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_n[ACC_WAVES];
    uint64_t *colbuf = (uint64_t *)smem;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int q = blockIdx.x;
    const int total = a.split * a.M;
    const int K = a.K[q];
    const uint64_t *keys = a.resc + (size_t)q * total;
    WaveTopM<R> col;
    col.init(colbuf + wave * WaveTopM<R>::cap, K);
    push_range(col, keys, (uint32_t)total, (uint32_t)wave * 64, ACC_WG);
    col.compact();
    if (lane == 0) s_n[wave] = col.n;
    __syncthreads();
    if (wave == 0) {
        wg_merge(col, colbuf, s_n, ACC_WAVES);
        write_results(col, K, a.out + (size_t)q * a.out_stride);
    }
}


batch_fn get_batch_select() { return batch_select; }
batch_fn get_batch_final(int R) {
    switch (R) {
        case 2: return batch_final<2>;
        case 4: return batch_final<4>;
        case 8: return batch_final<8>;
        case 16: return batch_final<16>;
        default: return nullptr;
    }
}
