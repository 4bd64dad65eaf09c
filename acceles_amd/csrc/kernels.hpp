// kernels.hpp -- HIP kernels of the Top-K SpMV path (gfx950, wave64).
//
//   lowbit_stage1   query quantization (quantify_util.py:31-49, 259) fused
//                   with the INT32 SpMV over sliced packets and a per-wavefront
//                   Top-M per partition (quantify_util.py:173-196;
//                   spmv_ucsr_top_k_multicore.hpp:195-413 on the FPGA).
//   lowbit_final    per-partition merge (host_spmv_ucsr.cpp:691-748), FP32
//                   rescoring of the candidates (quantify_util.py:354-367) and
//                   the global Top-K by the last workgroup to finish.
//   fp32_topk       plain FP32 CSR SpMV fused with the Top-K select
//                   (cpu_mkl_baseline.cpp:249-275; GPU spec
//                   host_spmv_topk_csr_gpu.cu:171-231).
//   topk_final      merge of the FP32 candidate list into the Top-K.
//   dense_topk      Top-K of a score vector (cpu_mkl_baseline.cpp:128-149).
#pragma once

#include "device_common.hpp"
#include "../../include/acceles.h"

struct Stage1Args {
    const uint32_t *words;    // lane-interleaved packet dwords
    const uint2 *slices;      // (offset in 64-word units, dwords per lane)
    const int32_t *perm;      // slot -> row (-1: padding lane)
    const int4 *items;        // (partition, first slice, end slice, 0)
    int nitems;
    const float *x;           // FP32 query (cols)
    int cols, col_bits, val_bits;
    int M;                    // rows kept per partition
    int cap;                  // collector capacity (pow2 >= M + 64)
    uint64_t *thr;            // per partition: best known lower bound of the M-th key
    uint32_t *cnt;            // per partition: candidate list length
    uint64_t *cand;           // candidate lists
    const int64_t *cand_off;  // per partition offset into cand
    int32_t *y;               // scores output (score mode)
};

// Quantize the FP32 query into LDS: q = clip(rint(127 * (x / max|x|)), -128, 127)
// (quantify_vector with B=8, symmetric; float32 division and product, half-even).
__device__ inline void quantize_query_lds(const float *x, int cols, int8_t *qx, float *red) {
    float m = 0.f;
    for (int c = threadIdx.x; c < cols; c += blockDim.x) m = fmaxf(m, fabsf(x[c]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (lane_id() == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    float s = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s = fmaxf(s, red[w]);
    for (int c = threadIdx.x; c < cols; c += blockDim.x) {
        int q = 0;
        if (s > 0.f) {
            float t = __fdiv_rn(x[c], s);
            float u = __fmul_rn(127.0f, t);
            q = (int)__builtin_rintf(u);
            q = q > 127 ? 127 : (q < -128 ? -128 : q);
        }
        qx[c] = (int8_t)q;
    }
    __syncthreads();
}

// Decodes one group of up to 32 entries (E dwords per lane) and accumulates
// value * q[col].  nd: dwords present in this group (wave-uniform).
template <int E, bool SIGNED>
__device__ __forceinline__ int32_t decode_group(const uint32_t *lanebase, int nd, const int8_t *qx, int cb, int vb,
                                                uint32_t cmask, int32_t acc) {
    uint32_t w[E + 1];
#pragma unroll
    for (int i = 0; i < E; ++i) w[i] = (i < nd) ? __builtin_nontemporal_load(lanebase + i * 64) : 0u;
    w[E] = 0u;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const int pos = j * E;
        const int wi = pos >> 5;
        const int sh = pos & 31;
        if ((j & 7) == 0 && wi >= nd) break;  // rest of the group is padding
        uint32_t bits = (sh + E <= 32) ? (w[wi] >> sh) : __builtin_amdgcn_alignbit(w[wi + 1], w[wi], sh);
        int col = (int)(bits & cmask);
        int val = SIGNED ? __builtin_amdgcn_sbfe(bits, cb, vb) : (int)__builtin_amdgcn_ubfe(bits, cb, vb);
        int q = qx[col];
        acc += val * q;
    }
    return acc;
}

// mode 0: Top-M candidates per partition; mode 1: write every row's score.
template <int E, bool SIGNED, int MODE>
__global__ __launch_bounds__(ACC_WG) void lowbit_stage1(Stage1Args a) {
    extern __shared__ __align__(16) unsigned char smem[];
    float *red = (float *)smem;                               // ACC_WAVES floats
    uint64_t *colbuf = (uint64_t *)(smem + 64);               // ACC_WAVES * cap keys
    int8_t *qx = (int8_t *)(smem + 64 + (MODE == 0 ? ACC_WAVES * a.cap * 8 : 0));
    quantize_query_lds(a.x, a.cols, qx, red);

    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int item_id = blockIdx.x * ACC_WAVES + wave;
    if (item_id >= a.nitems) return;
    const int4 item = a.items[item_id];
    const int p = item.x;
    const uint32_t cmask = (1u << a.col_bits) - 1u;

    WaveTopM col;
    col.init(colbuf + wave * a.cap, a.cap, a.M);
    uint64_t published = 0;

    for (int s = item.y; s < item.z; ++s) {
        const uint2 sl = a.slices[s];
        const uint32_t *lanebase = a.words + (size_t)sl.x * 64 + lane;
        const int ndw = (int)sl.y;
        int32_t acc = 0;
        for (int d0 = 0; d0 < ndw; d0 += E) {
            int nd = ndw - d0;
            acc = decode_group<E, SIGNED>(lanebase + (size_t)d0 * 64, nd < E ? nd : E, qx, a.col_bits, a.val_bits,
                                          cmask, acc);
        }
        if (MODE == 1) {
            int32_t row = a.perm[(size_t)s * 64 + lane];
            if (row >= 0) a.y[row] = acc;
            continue;
        }
        uint64_t g = atomic_load_u64(a.thr + p);
        uint64_t t = col.thr > g ? col.thr : g;
        int32_t tscore = dkey_score_i32(t);
        bool maybe = acc >= tscore;
        if (__ballot(maybe) == 0) continue;
        int32_t row = maybe ? a.perm[(size_t)s * 64 + lane] : -1;
        uint64_t key = dkey_i32(acc, row);
        col.push(maybe && row >= 0 && key > t, key);
        if (col.n >= a.M && col.thr > published) {
            // thr holds the M-th best of this wave's rows: a lower bound for the partition
            if (lane == 0) atomicMax((unsigned long long *)(a.thr + p), (unsigned long long)col.thr);
            published = col.thr;
        }
    }
    if (MODE == 1) return;
    col.compact();
    if (col.n >= a.M && col.thr > published && lane == 0)
        atomicMax((unsigned long long *)(a.thr + p), (unsigned long long)col.thr);
    // keep the keys that can still be in the partition's Top-M
    uint64_t g = atomic_load_u64(a.thr + p);
    int keep = 0;
    for (int i = lane; i < col.n; i += 64) keep += col.buf[i] >= g;
    for (int o = 32; o > 0; o >>= 1) keep += __shfl_xor(keep, o);
    if (keep == 0) return;
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(a.cnt + p, (uint32_t)keep);
    pos = __shfl(pos, 0);
    uint64_t *dst = a.cand + a.cand_off[p] + pos;
    for (int i = lane; i < keep; i += 64) dst[i] = col.buf[i];
}

struct FinalArgs {
    uint32_t *cnt;
    uint64_t *cand;
    const int64_t *cand_off;
    uint64_t *thr;
    const int32_t *part_row_begin;
    int split, M, cap;
    // FP32 CSR for rescoring
    const uint32_t *ptr;
    const int32_t *idx;
    const float *val;
    const float *x;
    int G;
    uint64_t *resc;   // split * M rescored keys
    uint32_t *done;   // last-workgroup counter
    int K;            // final Top-K (0: candidates only)
    acc_element *out;
    int32_t *cand_rows_out;    // optional: stage-1 candidates, p*M + i
    int32_t *cand_scores_out;
};

// Dot product of CSR row r with x by a group of G lanes: lane g sums entries
// g, g+G, ... with fmaf, then an xor butterfly of plain adds (every lane of the
// group ends with the same value).
template <int G>
__device__ __forceinline__ float group_dot(const uint32_t *ptr, const int32_t *idx, const float *val, const float *x,
                                           int r, bool valid) {
    const int g = lane_id() & (G - 1);
    float p = 0.f;
    if (valid) {
        uint32_t e = ptr[r + 1];
        for (uint32_t j = ptr[r] + g; j < e; j += G) p = __builtin_fmaf(val[j], x[idx[j]], p);
    }
#pragma unroll
    for (int s = G / 2; s >= 1; s >>= 1) p = p + __shfl_xor(p, s, G);
    return p;
}

template <int G>
__device__ inline void rescore_rows(const FinalArgs &a, const uint64_t *keys, int m, uint64_t *out_keys) {
    // keys[0, m) are stage-1 keys; writes FP32 keys (0 beyond m) for M slots
    const int per_pass = ACC_WG / G;
    for (int base = 0; base < a.M; base += per_pass) {
        int i = base + (int)threadIdx.x / G;
        bool valid = i < m;
        int32_t r = valid ? dkey_row(keys[i]) : 0;
        float s = group_dot<G>(a.ptr, a.idx, a.val, a.x, r, valid);
        if ((threadIdx.x & (G - 1)) == 0 && i < a.M) out_keys[i] = valid ? dkey_f32(s, r) : 0ull;
    }
}

template <int G>
__global__ __launch_bounds__(ACC_WG) void lowbit_final(FinalArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint64_t *colbuf = (uint64_t *)smem;                         // ACC_WAVES * cap
    uint64_t *merged = colbuf + ACC_WAVES * a.cap;                // ACC_WAVES * M
    __shared__ int s_n[ACC_WAVES];
    __shared__ int s_last;
    const int p = blockIdx.x, wave = threadIdx.x >> 6, lane = lane_id();

    // 1) Top-M of this partition's candidate list
    const uint32_t n = a.cnt[p];
    const uint64_t *keys = a.cand + a.cand_off[p];
    WaveTopM col;
    col.init(colbuf + wave * a.cap, a.cap, a.M);
    for (uint32_t b = wave * 64; b < n; b += ACC_WG) {
        uint32_t i = b + lane;
        col.push(i < n, i < n ? keys[i] : 0ull);
    }
    col.compact();
    for (int i = lane; i < a.M; i += 64) merged[wave * a.M + i] = i < col.n ? col.buf[i] : 0ull;
    if (lane == 0) s_n[wave] = col.n;
    __syncthreads();
    uint64_t *top = colbuf;  // reuse wave 0's buffer for the merged result
    if (wave == 0) {
        WaveTopM m;
        m.init(colbuf, a.cap, a.M);
        for (int b = 0; b < ACC_WAVES * a.M; b += 64) {
            int i = b + lane;
            bool ok = i < ACC_WAVES * a.M;
            uint64_t k = ok ? merged[i] : 0ull;
            m.push(ok && k != 0ull, k);
        }
        m.compact();
        if (lane == 0) s_n[0] = m.n;
    }
    __syncthreads();
    const int mcount = s_n[0];
    if (threadIdx.x == 0) {
        a.cnt[p] = 0u;  // ready for the next query
        a.thr[p] = 0ull;
    }
    if (a.cand_rows_out) {
        for (int i = threadIdx.x; i < a.M; i += ACC_WG) {
            a.cand_rows_out[(size_t)p * a.M + i] = i < mcount ? dkey_row(top[i]) : -1;
            a.cand_scores_out[(size_t)p * a.M + i] = i < mcount ? dkey_score_i32(top[i]) : 0;
        }
    }
    if (a.K <= 0) return;

    // 2) FP32 rescoring of the partition's candidates
    rescore_rows<G>(a, top, mcount, a.resc + (size_t)p * a.M);

    // 3) the last workgroup to finish selects the global Top-K
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(a.done, 1u) == (uint32_t)(a.split - 1);
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    const int total = a.split * a.M;
    const int cap2 = next_pow2(total);
    uint64_t *all = colbuf;  // needs cap2 keys: sized by the host
    for (int i = threadIdx.x; i < cap2; i += ACC_WG)
        all[i] = i < total ? __hip_atomic_load(a.resc + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
    __syncthreads();
    block_sort_desc(all, cap2);
    for (int i = threadIdx.x; i < a.K; i += ACC_WG) {
        uint64_t k = i < total ? all[i] : 0ull;
        acc_element e;
        e.value = k ? dkey_score_f32(k) : -__builtin_inff();
        e.index = k ? dkey_row(k) : -1;
        a.out[i] = e;
    }
    if (threadIdx.x == 0) *a.done = 0u;
}

struct Fp32Args {
    const uint32_t *ptr;
    const int32_t *idx;
    const float *val;
    const float *x;
    int rows, cols;
    int M, cap;
    uint64_t *thr;   // [1]
    uint32_t *cnt;   // [1]
    uint64_t *cand;  // capacity: waves * M
    float *y;        // score mode
};

// mode 0: Top-M candidates; mode 1: write y.
template <int G, int MODE>
__global__ __launch_bounds__(ACC_WG) void fp32_topk(Fp32Args a) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint64_t *colbuf = (uint64_t *)smem;
    float *xs = (float *)(smem + (MODE == 0 ? ACC_WAVES * a.cap * 8 : 0));
    for (int c = threadIdx.x; c < a.cols; c += ACC_WG) xs[c] = a.x[c];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int rows_per_iter = 64 / G;
    const int nblocks = (a.rows + rows_per_iter - 1) / rows_per_iter;
    const int gw = blockIdx.x * ACC_WAVES + wave, nw = gridDim.x * ACC_WAVES;
    WaveTopM col;
    col.init(colbuf + wave * a.cap, a.cap, a.M);
    uint64_t published = 0;
    // contiguous row range per wavefront (coalesced, cache friendly)
    const int per = (nblocks + nw - 1) / nw;
    const int b0 = gw * per, b1 = min(nblocks, b0 + per);
    for (int b = b0; b < b1; ++b) {
        const int r = b * rows_per_iter + lane / G;
        const bool valid = r < a.rows;
        float s = group_dot<G>(a.ptr, a.idx, a.val, xs, r, valid);
        if (MODE == 1) {
            if (valid && (lane & (G - 1)) == 0) a.y[r] = s;
            continue;
        }
        uint64_t g = atomic_load_u64(a.thr);
        uint64_t t = col.thr > g ? col.thr : g;
        uint64_t key = dkey_f32(s, r);
        col.push(valid && (lane & (G - 1)) == 0 && key > t, key);
        if (col.n >= a.M && col.thr > published) {
            if (lane == 0) atomicMax((unsigned long long *)a.thr, (unsigned long long)col.thr);
            published = col.thr;
        }
    }
    if (MODE == 1) return;
    col.compact();
    if (col.n >= a.M && col.thr > published && lane == 0) atomicMax((unsigned long long *)a.thr, (unsigned long long)col.thr);
    uint64_t g = atomic_load_u64(a.thr);
    int keep = 0;
    for (int i = lane; i < col.n; i += 64) keep += col.buf[i] >= g;
    for (int o = 32; o > 0; o >>= 1) keep += __shfl_xor(keep, o);
    if (keep == 0) return;
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(a.cnt, (uint32_t)keep);
    pos = __shfl(pos, 0);
    for (int i = lane; i < keep; i += 64) a.cand[pos + i] = col.buf[i];
}

// FP64 exact scores: 8 lanes per row, FP64 products and sums.
__global__ __launch_bounds__(ACC_WG) void spmv_fp64(const uint32_t *ptr, const int32_t *idx, const float *val,
                                                    const float *x, int rows, double *y) {
    const int lane8 = threadIdx.x & 7;
    for (int r = (blockIdx.x * ACC_WG + threadIdx.x) >> 3; r < ((rows + 31) & ~31); r += (gridDim.x * ACC_WG) >> 3) {
        double p = 0.0;
        if (r < rows)
            for (uint32_t j = ptr[r] + lane8; j < ptr[r + 1]; j += 8) p += (double)val[j] * (double)x[idx[j]];
        for (int s = 4; s >= 1; s >>= 1) p += __shfl_xor(p, s, 8);
        if (r < rows && lane8 == 0) y[r] = p;
    }
}

// Top-K of a candidate list (one workgroup); resets the list for the next call.
__global__ __launch_bounds__(ACC_WG) void topk_final(uint32_t *cnt, uint64_t *cand, uint64_t *thr, int K, int cap,
                                                     acc_element *out) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint64_t *colbuf = (uint64_t *)smem;             // ACC_WAVES * cap
    uint64_t *merged = colbuf + ACC_WAVES * cap;     // ACC_WAVES * K
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = *cnt;
    WaveTopM col;
    col.init(colbuf + wave * cap, cap, K);
    for (uint32_t b = wave * 64; b < n; b += ACC_WG) {
        uint32_t i = b + lane;
        col.push(i < n, i < n ? cand[i] : 0ull);
    }
    col.compact();
    for (int i = lane; i < K; i += 64) merged[wave * K + i] = i < col.n ? col.buf[i] : 0ull;
    __syncthreads();
    if (wave == 0) {
        WaveTopM m;
        m.init(colbuf, cap, K);
        for (int b = 0; b < ACC_WAVES * K; b += 64) {
            int i = b + lane;
            bool ok = i < ACC_WAVES * K;
            uint64_t k = ok ? merged[i] : 0ull;
            m.push(ok && k != 0ull, k);
        }
        m.compact();
        for (int i = lane; i < K; i += 64) {
            uint64_t k = i < m.n ? m.buf[i] : 0ull;
            acc_element e;
            e.value = k ? dkey_score_f32(k) : -__builtin_inff();
            e.index = k ? dkey_row(k) : -1;
            out[i] = e;
        }
        if (lane == 0) {
            *cnt = 0u;
            *thr = 0ull;
        }
    }
}

// Top-K of a dense FP32 score vector: per-wavefront collectors over
// contiguous ranges, appended to the candidate list like fp32_topk.
__global__ __launch_bounds__(ACC_WG) void dense_topk(const float *y, int n, int M, int cap, uint64_t *thr,
                                                     uint32_t *cnt, uint64_t *cand) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint64_t *colbuf = (uint64_t *)smem;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int gw = blockIdx.x * ACC_WAVES + wave, nw = gridDim.x * ACC_WAVES;
    const int nb = (n + 63) / 64, per = (nb + nw - 1) / nw;
    const int b0 = gw * per, b1 = min(nb, b0 + per);
    WaveTopM col;
    col.init(colbuf + wave * cap, cap, M);
    uint64_t published = 0;
    for (int b = b0; b < b1; ++b) {
        int i = b * 64 + lane;
        bool valid = i < n;
        uint64_t key = valid ? dkey_f32(y[i], i) : 0ull;
        uint64_t g = atomic_load_u64(thr);
        uint64_t t = col.thr > g ? col.thr : g;
        col.push(valid && key > t, key);
        if (col.n >= M && col.thr > published) {
            if (lane == 0) atomicMax((unsigned long long *)thr, (unsigned long long)col.thr);
            published = col.thr;
        }
    }
    col.compact();
    if (col.n >= M && col.thr > published && lane == 0) atomicMax((unsigned long long *)thr, (unsigned long long)col.thr);
    uint64_t g = atomic_load_u64(thr);
    int keep = 0;
    for (int i = lane; i < col.n; i += 64) keep += col.buf[i] >= g;
    for (int o = 32; o > 0; o >>= 1) keep += __shfl_xor(keep, o);
    if (keep == 0) return;
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(cnt, (uint32_t)keep);
    pos = __shfl(pos, 0);
    for (int i = lane; i < keep; i += 64) cand[pos + i] = col.buf[i];
}
