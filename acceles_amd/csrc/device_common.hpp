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

// Orders LDS traffic between the lanes of one wavefront.  The fence names
// the LDS address space only: a wavefront's LDS operations complete in order,
// and a fence over all memory would also wait for every global load in flight
// (draining the prefetch rings of the streaming kernels).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
}

// Maximum over the wavefront, valid in lane 63 (DPP steps: quad swaps,
// half-row and row mirrors, then the row 15 / row 31 broadcasts).
template <int CTRL, int ROWS>
__device__ __forceinline__ int32_t dpp_max_step(int32_t v) {
    return max(v, __builtin_amdgcn_update_dpp(v, v, CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_umax_step(uint32_t v) {
    return max(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, ROWS, 0xF, false));
}
// Unsigned maximum over the wavefront, returned to every lane.
__device__ __forceinline__ uint32_t wave_umax(uint32_t v) {
    v = dpp_umax_step<0xB1, 0xF>(v);
    v = dpp_umax_step<0x4E, 0xF>(v);
    v = dpp_umax_step<0x141, 0xF>(v);
    v = dpp_umax_step<0x140, 0xF>(v);
    v = dpp_umax_step<0x142, 0xA>(v);
    v = dpp_umax_step<0x143, 0xC>(v);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ __forceinline__ int32_t wave_max_lane63(int32_t v) {
    v = dpp_max_step<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    v = dpp_max_step<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
    v = dpp_max_step<0x141, 0xF>(v);  // row_half_mirror
    v = dpp_max_step<0x140, 0xF>(v);  // row_mirror
    v = dpp_max_step<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_max_step<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// The m-th largest (m >= 1, with repeats) of the U values per lane of the
// wavefront (0 = no value; returns 0 when fewer than m are nonzero), by
// sorting each lane's values in registers and then extracting the wavefront
// maximum m times: about 15 instructions per extracted value, against the 36
// shuffle stages of a 256-key bitonic sort.
template <int U>
__device__ __forceinline__ uint32_t wave_kth_u32(uint32_t (&v)[U], int m) {
    const int lane = lane_id();
#pragma unroll
    for (int i = 0; i < U; ++i)
#pragma unroll
        for (int j = 0; j + 1 < U - i; ++j) {
            const uint32_t x = v[j], y = v[j + 1];
            v[j] = max(x, y);
            v[j + 1] = min(x, y);
        }
    uint32_t t = 0u;
    for (int it = 0; it < m; ++it) {
        const uint32_t mx = wave_umax(v[0]);
        if (mx == 0u) return 0u;
        const int l = __builtin_ctzll(__ballot(v[0] == mx));
        if (lane == l) {
#pragma unroll
            for (int j = 0; j + 1 < U; ++j) v[j] = v[j + 1];
            v[U - 1] = 0u;
        }
        t = mx;
    }
    return t;
}

__device__ __forceinline__ int mbcnt64(uint64_t mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t atomic_load_u64(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
    uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src);
    uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m);
    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}

// Bitonic sort, descending, of 64*R keys held in registers: element
// i = r*64 + lane lives in v[r] of that lane.  Stages with j >= 64 exchange
// slots inside a lane; the others exchange lanes with shfl_xor.
template <int R>
__device__ __forceinline__ void reg_sort_desc(uint64_t (&v)[R]) {
    const int lane = lane_id();
#pragma unroll
    for (int k = 2; k <= 64 * R; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int rp = r ^ (j >> 6);
                    if (rp > r) {
                        const bool desc = ((r * 64 + lane) & k) == 0;
                        const uint64_t a = v[r], b = v[rp];
                        const bool sw = desc ? (a < b) : (a > b);
                        v[r] = sw ? b : a;
                        v[rp] = sw ? a : b;
                    }
                }
            } else {
                const bool lower = (lane & j) == 0;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint64_t o = shfl_xor_u64(v[r], j);
                    const bool desc = ((r * 64 + lane) & k) == 0;
                    const bool take_max = lower == desc;
                    v[r] = take_max ? (v[r] > o ? v[r] : o) : (v[r] < o ? v[r] : o);
                }
            }
        }
    }
}

// Wavefront Top-M collector over an LDS buffer of cap = 64*R keys
// (M <= cap - 64).  All state is wave-uniform.  push() is called by all 64
// lanes together; each lane offers at most one key.  Appends go to LDS; a
// compaction reads the buffer into registers, sorts there and writes back the
// best M.
template <int R>
struct WaveTopM {
    static constexpr int cap = 64 * R;
    uint64_t *buf;
    int n, M;
    uint64_t thr;  // a key must be > thr to be kept (the M-th best once full)

    __device__ void init(uint64_t *b, int m) {
        buf = b;
        M = m;
        n = 0;
        thr = 0;
    }
    // Sort, keep the best M, set the admission threshold.
    __device__ void compact() {
        const int lane = lane_id();
        if (n <= 32) {
            // few keys: each finds its rank by counting the keys above it (equal
            // keys, should any occur, are ordered by lane), then moves there
            const uint64_t k = lane < n ? buf[lane] : 0ull;
            const uint32_t klo = (uint32_t)k, khi = (uint32_t)(k >> 32);
            int rank = 0;
            for (int j = 0; j < n; ++j) {
                const uint64_t kj = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)khi, j) << 32) |
                                    (uint32_t)__builtin_amdgcn_readlane((int)klo, j);
                rank += (kj > k || (kj == k && j < lane)) ? 1 : 0;
            }
            const int nn = n < M ? n : M;
            wave_sync();
            if (lane < n && rank < nn) buf[rank] = k;
            if (nn == M && M > 0) {
                const uint64_t bm = __ballot(lane < n && rank == M - 1);
                thr = shfl_u64(k, __builtin_ctzll(bm));
            }
            n = nn;
            wave_sync();
            return;
        }
        if (R > 1 && n <= 64) {  // one register of keys: a 64-key sort
            uint64_t w[1] = {lane < n ? buf[lane] : 0ull};
            reg_sort_desc<1>(w);
            n = n < M ? n : M;
            if (lane < n) buf[lane] = w[0];
            if (n == M && M > 0) thr = shfl_u64(w[0], M - 1);
            wave_sync();
            return;
        }
        uint64_t v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = (r * 64 + lane < n) ? buf[r * 64 + lane] : 0ull;
        reg_sort_desc<R>(v);
        n = n < M ? n : M;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r * 64 + lane < n) buf[r * 64 + lane] = v[r];
        if (n == M && M > 0) {
            uint64_t t = 0;
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (r == (M - 1) >> 6) t = v[r];
            thr = shfl_u64(t, (M - 1) & 63);
        }
        wave_sync();
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
    // Offers U keys per lane at once: the lanes' counts are prefix-summed
    // over the wavefront and the admitted keys are written to their slots in
    // one pass.  When they do not all fit, the buffer is filled, compacted
    // (which may raise the admission threshold) and the rest offered again.
    template <int U>
    __device__ void push_many(const bool (&pred)[U], const uint64_t (&key)[U]) {
        const int lane = lane_id();
        bool pend[U];
#pragma unroll
        for (int u = 0; u < U; ++u) pend[u] = pred[u];
        for (;;) {
            int c = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                pend[u] = pend[u] && key[u] > thr;
                c += pend[u] ? 1 : 0;
            }
            int incl = c;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(incl, o);
                if (lane >= o) incl += t;
            }
            const int tot = __builtin_amdgcn_readlane(incl, 63);
            if (tot == 0) return;
            const int room = cap - n;  // >= cap - M >= 64 after a compaction
            int pos = incl - c;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (pend[u]) {
                    if (pos < room) {
                        buf[n + pos] = key[u];
                        pend[u] = false;
                    }
                    ++pos;
                }
            }
            wave_sync();
            if (tot <= room) {
                n += tot;
                return;
            }
            n = cap;
            compact();
        }
    }
};

__device__ __forceinline__ int next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}
