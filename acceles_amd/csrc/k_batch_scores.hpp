// k_batch_scores.hpp -- batch_scores, included by the per-QB translation
// units k_batch16.hip and k_batch32.hip (see k_batch.hip for the pipeline).
#pragma once
#include "kernels_common.hpp"

template <int E, int QB, bool SIGNED, bool U8>
__global__ __launch_bounds__(ACC_WG) void batch_scores(BatchArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int QS = QTable<QB>::QS;
    uint8_t *qt = smem;  // cols * QS
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int q0 = blockIdx.y * QB;
    // ---- quantize the block's queries (quantify_vector, B=8 symmetric) ----
    for (int qi = wave; qi < QB; qi += ACC_WAVES) {
        const int q = q0 + qi;
        const float *x = a.X + (size_t)q * a.cols;
        float m = 0.f;
        if (q < a.nq)
            for (int c = lane; c < a.cols; c += 64) m = fmaxf(m, fabsf(x[c]));
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        for (int c = lane; c < a.cols; c += 64) {
            int qv = 0;
            if (q < a.nq && m > 0.f) {
                const float t = __fdiv_rn(x[c], m);
                qv = (int)__builtin_rintf(__fmul_rn(127.0f, t));
                qv = qv > 127 ? 127 : (qv < -128 ? -128 : qv);
            }
            qt[c * QS + qi] = (uint8_t)(U8 ? qv + 128 : qv);
        }
    }
    __syncthreads();
    const int4 item = a.items[blockIdx.x];
    const uint32_t cmask = (1u << a.col_bits) - 1u;
    for (int s = item.y + wave; s < item.z; s += ACC_WAVES) {
        const uint2 sl = a.slices[s];
        uint32_t acc[QB];
#pragma unroll
        for (int i = 0; i < QB; ++i) acc[i] = 0u;
        int32_t vsum = 0;
        const __amdgpu_buffer_rsrc_t rs = slice_rsrc(a.words, sl);
        for (int g0 = 0; g0 < (int)sl.y; g0 += E) {
            uint32_t w[E];
            const int nd = min(E, (int)sl.y - g0);
            load_group<E>(w, rs, g0);
#pragma unroll
            for (int j = 0; j < 32; ++j) {
                const int pos = j * E, wi = pos >> 5, sh = pos & 31;
                if (((((j >> 3) << 3) * E) >> 5) >= nd) continue;  // padding block (wave-uniform)
                const uint32_t hi = (wi + 1 < E) ? w[wi + 1] : 0u;
                const uint32_t bits = (sh + E <= 32) ? (w[wi] >> sh) : __builtin_amdgcn_alignbit(hi, w[wi], sh);
                const int col = (int)(bits & cmask);
                const int v = SIGNED ? __builtin_amdgcn_sbfe(bits, a.col_bits, a.val_bits)
                                     : (int)__builtin_amdgcn_ubfe(bits, a.col_bits, a.val_bits);
                if (U8) vsum += v;
                const uint32_t vb = (uint32_t)v & 0xFFu;
                const uint32_t v0 = vb, v1 = vb << 8, v2 = vb << 16, v3 = vb << 24;
                const uint2 *row = (const uint2 *)(qt + col * QS);
#pragma unroll
                for (int b = 0; b < QB / 8; ++b) {
                    const uint2 qq = row[b];
                    if (U8) {
                        acc[8 * b + 0] = __builtin_amdgcn_udot4(qq.x, v0, acc[8 * b + 0], false);
                        acc[8 * b + 1] = __builtin_amdgcn_udot4(qq.x, v1, acc[8 * b + 1], false);
                        acc[8 * b + 2] = __builtin_amdgcn_udot4(qq.x, v2, acc[8 * b + 2], false);
                        acc[8 * b + 3] = __builtin_amdgcn_udot4(qq.x, v3, acc[8 * b + 3], false);
                        acc[8 * b + 4] = __builtin_amdgcn_udot4(qq.y, v0, acc[8 * b + 4], false);
                        acc[8 * b + 5] = __builtin_amdgcn_udot4(qq.y, v1, acc[8 * b + 5], false);
                        acc[8 * b + 6] = __builtin_amdgcn_udot4(qq.y, v2, acc[8 * b + 6], false);
                        acc[8 * b + 7] = __builtin_amdgcn_udot4(qq.y, v3, acc[8 * b + 7], false);
                    } else {
                        acc[8 * b + 0] = (uint32_t)__builtin_amdgcn_sdot4((int)qq.x, (int)v0, (int)acc[8 * b + 0], false);
                        acc[8 * b + 1] = (uint32_t)__builtin_amdgcn_sdot4((int)qq.x, (int)v1, (int)acc[8 * b + 1], false);
                        acc[8 * b + 2] = (uint32_t)__builtin_amdgcn_sdot4((int)qq.x, (int)v2, (int)acc[8 * b + 2], false);
                        acc[8 * b + 3] = (uint32_t)__builtin_amdgcn_sdot4((int)qq.x, (int)v3, (int)acc[8 * b + 3], false);
                        acc[8 * b + 4] = (uint32_t)__builtin_amdgcn_sdot4((int)qq.y, (int)v0, (int)acc[8 * b + 4], false);
                        acc[8 * b + 5] = (uint32_t)__builtin_amdgcn_sdot4((int)qq.y, (int)v1, (int)acc[8 * b + 5], false);
                        acc[8 * b + 6] = (uint32_t)__builtin_amdgcn_sdot4((int)qq.y, (int)v2, (int)acc[8 * b + 6], false);
                        acc[8 * b + 7] = (uint32_t)__builtin_amdgcn_sdot4((int)qq.y, (int)v3, (int)acc[8 * b + 7], false);
                    }
                }
            }
        }
        // unsigned 8-bit: sum v*(q+128) - 128*sum v = sum v*q
        const int64_t slot = (int64_t)s * 64 + lane;
#pragma unroll
        for (int i = 0; i < QB; ++i)
            if (q0 + i < a.nq) a.S[(size_t)(q0 + i) * a.nslots + slot] = (int32_t)acc[i] - (U8 ? 128 * vsum : 0);
    }
}

// Per (partition p = blockIdx.x, query q = blockIdx.y): Top-M of S[q][slots
// of p] by (score, row), optional candidate output, FP32 rescoring.

template <int QB, bool SIGNED, bool U8>
static batch_fn bscores_for_e(int E) {
    switch (E) {
#define CASE(e) case e: return batch_scores<e, QB, SIGNED, U8>;
        CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
        CASE(17) CASE(18) CASE(19) CASE(20) CASE(21) CASE(22)
#undef CASE
        default: return nullptr;
    }
}

template <int QB>
static batch_fn bscores_for(int E, bool is_signed, int bits) {
    if (is_signed) return bscores_for_e<QB, true, false>(E);
    if (bits == 8) return bscores_for_e<QB, false, true>(E);  // unsigned 8-bit: offset query bytes
    return bscores_for_e<QB, false, false>(E);
}
