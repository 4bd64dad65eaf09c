// device_common.hpp -- device helpers shared by the HIP kernels (gfx950,
// wave64): ranking keys, wave primitives and the wavefront Top-M collector.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define ACC_WG 256                 // threads per workgroup (4 wavefronts)
#define ACC_WAVES (ACC_WG / 64)

// Ranking keys: a larger key ranks first.  High 32 bits: the score mapped to
// an order-preserving unsigned; low 32 bits: ~row, so on equal scores the
// lower row ranks first.  0 never occurs for a real row (row < 2^31).
__device__ __forceinline__ uint64_t dkey_i32(int32_t s, int32_t row) {
    return ((uint64_t)((uint32_t)s ^ 0x80000000u) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)row);
}
__device__ __forceinline__ uint32_t dordered_f32(float f) {
    uint32_t b = __float_as_uint(f == 0.0f ? 0.0f : f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ uint64_t dkey_f32(float s, int32_t row) {
    return ((uint64_t)dordered_f32(s) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)row);
}
__device__ __forceinline__ int32_t dkey_row(uint64_t k) { return (int32_t)(0xFFFFFFFFu - (uint32_t)k); }
__device__ __forceinline__ int32_t dkey_score_i32(uint64_t k) { return (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u); }
__device__ __forceinline__ float dkey_score_f32(uint64_t k) {
    uint32_t o = (uint32_t)(k >> 32);
    uint32_t b = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return __uint_as_float(b);
}

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Orders LDS traffic between the lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int mbcnt64(uint64_t mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t atomic_load_u64(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bitonic sort of buf[0, cap) into descending order by one wavefront.
__device__ inline void wave_sort_desc(uint64_t *buf, int cap) {
    const int lane = lane_id();
    for (int k = 2; k <= cap; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = lane; t < (cap >> 1); t += 64) {
                int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                int l = i | j;
                uint64_t a = buf[i], b = buf[l];
                bool desc_block = (i & k) == 0;
                if (desc_block ? (a < b) : (a > b)) {
                    buf[i] = b;
                    buf[l] = a;
                }
            }
            wave_sync();
        }
    }
}

// Bitonic sort of buf[0, cap) descending by the whole workgroup.
__device__ inline void block_sort_desc(uint64_t *buf, int cap) {
    for (int k = 2; k <= cap; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = (int)threadIdx.x; t < (cap >> 1); t += (int)blockDim.x) {
                int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                int l = i | j;
                uint64_t a = buf[i], b = buf[l];
                bool desc_block = (i & k) == 0;
                if (desc_block ? (a < b) : (a > b)) {
                    buf[i] = b;
                    buf[l] = a;
                }
            }
            __syncthreads();
        }
    }
}

// Wavefront Top-M collector over an LDS buffer of `cap` keys
// (cap a power of two >= M + 64).  All state is wave-uniform.  push() is
// called by all 64 lanes together; each lane offers at most one key.
struct WaveTopM {
    uint64_t *buf;
    int n, cap, M;
    uint64_t thr;  // a key must be > thr to be kept (the M-th best once full)

    __device__ void init(uint64_t *b, int cap_, int m) {
        buf = b;
        cap = cap_;
        M = m;
        n = 0;
        thr = 0;
    }
    // Sort, keep the best M, set the admission threshold.
    __device__ void compact() {
        for (int i = n + lane_id(); i < cap; i += 64) buf[i] = 0;
        wave_sync();
        wave_sort_desc(buf, cap);
        n = n < M ? n : M;
        if (n == M && M > 0) thr = buf[M - 1];
    }
    __device__ void push(bool pred, uint64_t key) {
        pred = pred && key > thr;
        uint64_t mask = __ballot(pred);
        if (mask == 0) return;
        int c = __popcll(mask);
        if (n + c > cap) {
            compact();
            pred = pred && key > thr;
            mask = __ballot(pred);
            c = __popcll(mask);
        }
        if (pred) buf[n + mbcnt64(mask)] = key;
        wave_sync();
        n += c;
    }
};

__device__ __forceinline__ int next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}
