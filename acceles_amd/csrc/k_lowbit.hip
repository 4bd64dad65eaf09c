// k_lowbit.hip -- single-query low-bit kernels (three launches per query):
//   lowbit_slicemax  query quantization (quantify_util.py:31-49, 259) and the
//                    INT32 scores of the sliced packets (quantify_util.py:5-6,
//                    173-196; spmv_ucsr_top_k_multicore.hpp:195-355 on the
//                    FPGA); only the maximum of each 64-row slice is written;
//   lowbit_select    per partition: threshold = its M-th best slice maximum
//                    (M distinct rows reach it, so no row below it can be in
//                    the partition's Top-M), exact rescan of the slices whose
//                    maximum reaches it, Top-M (quantify_util.py:173-196,
//                    host_spmv_ucsr.cpp:691-748) and FP32 rescoring of those
//                    rows (quantify_util.py:354-367);
//   lowbit_final     the global Top-K of the rescored candidates;
//   lowbit_scores    INT32 scores of every row (parity checks).
// The streaming pass carries no selection state, so it runs at full
// occupancy with nothing but loads, decode and one wavefront max per slice.
#include "kernels_common.hpp"

template <int E, bool SIGNED>
__device__ __forceinline__ int32_t slice_score(const LowbitArgs &a, const uint2 sl, const int8_t *qx, uint32_t cmask) {
    const __amdgpu_buffer_rsrc_t rs = slice_rsrc(a.words, sl);
    int32_t acc = 0;
    for (int g0 = 0; g0 < (int)sl.y; g0 += E) {
        uint32_t w[E];
        load_group<E>(w, rs, g0);
        acc = decode_group<E, SIGNED>(w, min(E, (int)sl.y - g0), qx, a.col_bits, a.val_bits, cmask, acc);
    }
    return acc;
}

// Work item blockIdx.x (a slice range of one partition): smax[s] = the
// largest INT32 score over the 64 lanes of slice s.  Padding lanes of a
// partition's last slice score 0, so that slice's maximum may be too high;
// lowbit_select only uses full slices for its threshold.
template <int E, bool SIGNED>
__global__ __launch_bounds__(LB_WG) void lowbit_slicemax(LowbitArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint2 s_meta[LB_MAX_ITEM];
    float *red = (float *)smem;
    int8_t *qx = (int8_t *)(smem + 64);
    const int4 item = a.items[blockIdx.x];
    for (int i = threadIdx.x; i < item.z - item.y; i += LB_WG) s_meta[i] = a.slices[item.y + i];
    quantize_query_lds(a.x, a.cols, qx, red);  // its barriers also publish s_meta

    const int wave = threadIdx.x >> 6, lane = lane_id();
    const uint32_t cmask = (1u << a.col_bits) - 1u;
    // Groups (E dwords per lane) of the wave's slices item.y + wave,
    // + LB_WAVES, ... stream through a ring of three register slots, two
    // groups ahead of the decoder.  Every load of the loop is unconditional
    // (through a buffer resource bounded by the slice, so loads past its end
    // return 0 and fetch nothing): a load whose result is merged into a loop
    // variable on only some paths is copied into place after a wait for all
    // outstanding loads, which would drain the ring.  A slot records the
    // group's slice, length and whether it ends the slice.
    constexpr int RING = 3;
    uint32_t w[RING][E];
    int ds[RING], dnd[RING];
    bool dend[RING], dval[RING];
    int ls = item.y + wave, lg = 0;  // load cursor
    uint2 lsl = ls < item.z ? s_meta[ls - item.y] : make_uint2(0u, 0u);
    auto issue = [&](int slot) {
        const bool valid = ls < item.z;
        load_group<E>(w[slot], slice_rsrc(a.words, lsl), lg);
        __builtin_amdgcn_sched_barrier(0);  // keep the loads here: the scheduler would sink them to their use
        dval[slot] = valid;
        ds[slot] = ls;
        dnd[slot] = valid ? min(E, (int)lsl.y - lg) : 0;
        dend[slot] = valid && lg + E >= (int)lsl.y;
        if (valid) {
            if (dend[slot]) {
                ls += LB_WAVES;
                lg = 0;
                lsl = ls < item.z ? s_meta[ls - item.y] : make_uint2(0u, 0u);
            } else {
                lg += E;
            }
        }
    };
    issue(0);
    issue(1);
    int32_t acc = 0;
    bool done = false;
    while (!done) {
#pragma unroll
        for (int j = 0; j < RING; ++j) {
            if (done) break;
            if (!dval[j]) {
                done = true;
                break;
            }
            // Every loaded dword is read here, also on paths that skip
            // blocks: a load result that is dead on some path lets its
            // register be reused while the load is in flight, and the reuse
            // then waits for all outstanding loads.
            const int nd = dnd[j];
            uint32_t wm[E];
#pragma unroll
            for (int i = 0; i < E; ++i) wm[i] = i < nd ? w[j][i] : 0u;
            acc = decode_group<E, SIGNED>(wm, nd, qx, a.col_bits, a.val_bits, cmask, acc);
            if (dend[j]) {
                const int32_t m = wave_max_lane63(acc);
                if (lane == 63) a.smax[ds[j]] = m;
                acc = 0;
            }
            issue((j + 2) % RING);
        }
    }
}

// One workgroup per partition p = blockIdx.x, after lowbit_slicemax.
// 1. ts = the M-th largest maximum among the partition's full slices: each
//    of those M slices holds a distinct row scoring >= ts, so every row of
//    the partition's Top-M scores >= ts and lies in a slice whose maximum
//    is >= ts.
// 2. The slices whose maximum reaches ts are rescanned exactly and their
//    rows >= ts offered to the Top-M collectors.  Their indices are listed
//    in LDS (up to kList; past that, each wavefront rescans its own chunks
//    one slice at a time), the list is processed kRescan slices per
//    wavefront at a time so that their loads overlap.
template <int E, bool SIGNED>
__global__ __launch_bounds__(LB_WG) void lowbit_select(LowbitArgs a) {
    constexpr int kList = 1024, kRescan = 4;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_n[LB_WAVES];
    __shared__ int s_m, s_cnt;
    __shared__ int32_t s_ts;
    __shared__ int32_t s_list[kList];
    __shared__ uint2 s_lmeta[kList];
    float *red = (float *)smem;
    uint64_t *colbuf = (uint64_t *)(smem + 64);
    int8_t *qx = (int8_t *)(smem + 64 + 8 * LB_WAVES * WaveTopM<2>::cap);
    if (threadIdx.x == 0) s_cnt = 0;
    quantize_query_lds(a.x, a.cols, qx, red);
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int p = blockIdx.x;
    const int64_t s0 = a.part_slice[p];
    const int nsl = (int)(a.part_slice[p + 1] - s0);
    const int32_t *smax = a.smax + s0;
    const uint8_t *nvalid = a.nvalid + s0;
    const uint32_t cmask = (1u << a.col_bits) - 1u;

    // 1. threshold (keys (max, slice) are distinct and nonzero)
    WaveTopM<2> col;
    col.init(colbuf + wave * WaveTopM<2>::cap, a.M);
    {
        constexpr int U = 8;
        for (int b = wave * 64; b < nsl; b += LB_WG * U) {
            int32_t v[U];
            int nv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = min(b + u * LB_WG + lane, nsl - 1);
                v[u] = smax[i];
                nv[u] = nvalid[i];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = b + u * LB_WG + lane;
                col.push(i < nsl && nv[u] == 64, ((uint64_t)((uint32_t)v[u] ^ 0x80000000u) << 32) | (uint32_t)(i + 1));
            }
        }
    }
    col.compact();
    if (lane == 0) s_n[wave] = col.n;
    __syncthreads();
    if (wave == 0) {
        wg_merge(col, colbuf, s_n, LB_WAVES);
        if (lane == 0) s_ts = col.n >= a.M ? (int32_t)((uint32_t)(col.buf[a.M - 1] >> 32) ^ 0x80000000u) : INT32_MIN;
    }
    __syncthreads();
    const int32_t ts = s_ts;

    // 2. candidate slices
    for (int i = threadIdx.x; i < nsl; i += LB_WG) {
        if (smax[i] >= ts) {
            const int j = atomicAdd(&s_cnt, 1);
            if (j < kList) s_list[j] = i;
        }
    }
    __syncthreads();
    const int ncand = s_cnt;
    col.init(colbuf + wave * WaveTopM<2>::cap, a.M);
    if (ncand <= kList) {
        for (int j = threadIdx.x; j < ncand; j += LB_WG) s_lmeta[j] = a.slices[s0 + s_list[j]];
        __syncthreads();
        for (int j0 = wave * kRescan; j0 < ncand; j0 += LB_WAVES * kRescan) {
            uint32_t w[kRescan][E];
            int32_t row[kRescan];
            uint2 sl[kRescan];
#pragma unroll
            for (int u = 0; u < kRescan; ++u) {
                const bool ok = j0 + u < ncand;
                const int64_t s = s0 + (ok ? s_list[j0 + u] : 0);
                sl[u] = ok ? s_lmeta[j0 + u] : make_uint2(0u, 0u);
                load_group<E>(w[u], slice_rsrc(a.words, sl[u]), 0);
                row[u] = ok ? a.perm[s * 64 + lane] : -1;
            }
#pragma unroll
            for (int u = 0; u < kRescan; ++u) {
                int32_t sc = decode_group<E, SIGNED>(w[u], min(E, (int)sl[u].y), qx, a.col_bits, a.val_bits, cmask, 0);
                if ((int)sl[u].y > E) {  // slices longer than one group
                    const __amdgpu_buffer_rsrc_t rs = slice_rsrc(a.words, sl[u]);
                    for (int g0 = E; g0 < (int)sl[u].y; g0 += E) {
                        uint32_t wg[E];
                        load_group<E>(wg, rs, g0);
                        sc = decode_group<E, SIGNED>(wg, min(E, (int)sl[u].y - g0), qx, a.col_bits, a.val_bits, cmask, sc);
                    }
                }
                col.push(row[u] >= 0 && sc >= ts, dkey_i32(sc, row[u]));
            }
        }
    } else {
        for (int b = wave * 64; b < nsl; b += LB_WG) {
            uint64_t mask = __ballot(b + lane < nsl && smax[min(b + lane, nsl - 1)] >= ts);
            while (mask) {
                const int l = __builtin_ctzll(mask);
                mask &= mask - 1;
                const int64_t s = s0 + b + l;
                const int32_t sc = slice_score<E, SIGNED>(a, a.slices[s], qx, cmask);
                const int32_t row = a.perm[s * 64 + lane];
                col.push(row >= 0 && sc >= ts, dkey_i32(sc, row));
            }
        }
    }
    col.compact();
    if (lane == 0) s_n[wave] = col.n;
    __syncthreads();
    if (wave == 0) {
        wg_merge(col, colbuf, s_n, LB_WAVES);
        if (lane == 0) s_m = col.n;
    }
    __syncthreads();
    const int mcount = s_m;
    const uint64_t *top = colbuf;
    if (a.cand_rows_out) {
        for (int i = threadIdx.x; i < a.M; i += LB_WG) {
            a.cand_rows_out[(size_t)p * a.M + i] = i < mcount ? dkey_row(top[i]) : -1;
            a.cand_scores_out[(size_t)p * a.M + i] = i < mcount ? dkey_score_i32(top[i]) : 0;
        }
    }
    if (a.K <= 0) return;

    // 3. FP32 rescoring of the partition's candidates
    for (int base = 0; base < a.M; base += LB_WG / a.G) {
        const int i = base + (int)threadIdx.x / a.G;
        const bool valid = i < mcount;
        const int32_t r = valid ? dkey_row(top[i]) : 0;
        const float sc = group_dot_rt(a.ptr, a.idx, a.val, a.x, r, valid, a.G);
        if ((threadIdx.x & (a.G - 1)) == 0 && i < a.M) a.resc[(size_t)p * a.M + i] = valid ? dkey_f32(sc, r) : 0ull;
    }
}

// One workgroup: the global Top-K of the split*M rescored candidates
// (split*M <= 4*64*R).
template <int R>
__global__ __launch_bounds__(ACC_WG) void lowbit_final(LowbitArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_nz[ACC_WAVES];
    topk_by_rank<R>(a.resc, a.split * a.M, a.K, a.out, (uint64_t *)smem, s_nz);
}

// INT32 scores of every row (one work item per wavefront).

template <int E, bool SIGNED>
__global__ __launch_bounds__(ACC_WG) void lowbit_scores(ScoresArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    float *red = (float *)smem;
    int8_t *qx = (int8_t *)(smem + 64);
    quantize_query_lds(a.x, a.cols, qx, red);
    const int lane = lane_id();
    const uint32_t cmask = (1u << a.col_bits) - 1u;
    for (int64_t s = (int64_t)blockIdx.x * ACC_WAVES + (threadIdx.x >> 6); s < a.nslices;
         s += (int64_t)gridDim.x * ACC_WAVES) {
        const uint2 sl = a.slices[s];
        const __amdgpu_buffer_rsrc_t rs = slice_rsrc(a.words, sl);
        int32_t acc = 0;
        for (int g0 = 0; g0 < (int)sl.y; g0 += E) {
            uint32_t w[E];
            load_group<E>(w, rs, g0);
            acc = decode_group<E, SIGNED>(w, min(E, (int)sl.y - g0), qx, a.col_bits, a.val_bits, cmask, acc);
        }
        const int32_t row = a.perm[s * 64 + lane];
        if (row >= 0) a.y[row] = acc;
    }
}


template <bool SIGNED>
static lowbit_fn slicemax_for(int E) {
    switch (E) {
#define CASE(e) case e: return lowbit_slicemax<e, SIGNED>;
        CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
        CASE(17) CASE(18) CASE(19) CASE(20) CASE(21) CASE(22)
#undef CASE
        default: return nullptr;
    }
}

template <bool SIGNED>
static lowbit_fn select_for(int E) {
    switch (E) {
#define CASE(e) case e: return lowbit_select<e, SIGNED>;
        CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
        CASE(17) CASE(18) CASE(19) CASE(20) CASE(21) CASE(22)
#undef CASE
        default: return nullptr;
    }
}

template <bool SIGNED>
static scores_fn scores_for(int E) {
    switch (E) {
#define CASE(e) case e: return lowbit_scores<e, SIGNED>;
        CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
        CASE(17) CASE(18) CASE(19) CASE(20) CASE(21) CASE(22)
#undef CASE
        default: return nullptr;
    }
}

lowbit_fn get_lowbit_slicemax(int E, bool is_signed) {
    return is_signed ? slicemax_for<true>(E) : slicemax_for<false>(E);
}
lowbit_fn get_lowbit_select(int E, bool is_signed) { return is_signed ? select_for<true>(E) : select_for<false>(E); }
lowbit_fn get_lowbit_final(int R) {
    switch (R) {
        case 1: return lowbit_final<1>;
        case 2: return lowbit_final<2>;
        case 4: return lowbit_final<4>;
        case 8: return lowbit_final<8>;
        case 16: return lowbit_final<16>;
        default: return nullptr;
    }
}
scores_fn get_lowbit_scores(int E, bool is_signed) { return is_signed ? scores_for<true>(E) : scores_for<false>(E); }
