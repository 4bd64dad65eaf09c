// k_lowbit.hip -- single-query low-bit kernels (three launches per query):
//   lowbit_stage1   query quantization (quantify_util.py:31-49, 259) fused
//                   with the INT32 SpMV over sliced packets, a per-wavefront
//                   Top-M, the workgroup merge and the append of what can
//                   still be in the partition's Top-M (quantify_util.py:173-196;
//                   spmv_ucsr_top_k_multicore.hpp:195-413 on the FPGA);
//   lowbit_merge    per partition: its Top-M (host_spmv_ucsr.cpp:691-748) and
//                   the FP32 rescoring of those rows (quantify_util.py:354-367);
//   lowbit_final    the global Top-K of the rescored candidates;
//   lowbit_scores   INT32 scores of every row (parity checks).
#include "kernels_common.hpp"

template <int E, bool SIGNED>
__global__ __launch_bounds__(LB_WG) void lowbit_stage1(LowbitArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_n[LB_WAVES];
    __shared__ uint2 s_meta[LB_MAX_ITEM];
    float *red = (float *)smem;
    uint64_t *colbuf = (uint64_t *)(smem + 64);
    int8_t *qx = (int8_t *)(smem + 64 + 8 * LB_WAVES * WaveTopM<2>::cap);
    const int4 item = a.items[blockIdx.x];
    for (int i = threadIdx.x; i < item.z - item.y; i += LB_WG) s_meta[i] = a.slices[item.y + i];
    quantize_query_lds(a.x, a.cols, qx, red);  // its barriers also publish s_meta

    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int p = item.x;
    const uint32_t cmask = (1u << a.col_bits) - 1u;
    WaveTopM<2> col;
    col.init(colbuf + wave * WaveTopM<2>::cap, a.M);

    // Groups (E dwords per lane) of the wave's slices item.y + wave,
    // + LB_WAVES, ... stream through a ring of three register slots, two
    // groups ahead of the decoder.  Every load of the loop is unconditional
    // (the word array carries E*64 dwords of zero padding past its end;
    // dwords past a group's length are masked when it is decoded): a load
    // whose result is merged into a loop variable on only some paths is
    // copied into place after a wait for all outstanding loads, which would
    // drain the ring.  A slot records the group's slice, length and whether
    // it ends the slice.
    constexpr int RING = 3;
    uint32_t w[RING][E];
    int ds[RING], dnd[RING];
    bool dend[RING], dval[RING];
    int ls = item.y + wave, lg = 0;  // load cursor
    uint2 lsl = ls < item.z ? s_meta[ls - item.y] : make_uint2(0u, 0u);
    auto issue = [&](int slot) {
        const bool valid = ls < item.z;
        const size_t base = valid ? ((size_t)lsl.x + lg) * 64 : 0;
#pragma unroll
        for (int i = 0; i < E; ++i) w[slot][i] = __builtin_nontemporal_load(a.words + base + (size_t)i * 64 + lane);
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
    // Per step: push the rows requested at the previous slice end, decode the
    // group in slot j, at a slice end compare with the threshold and request
    // the rows of the lanes that pass (from a fixed dummy address when none
    // does, so the load is unconditional), then request the threshold for the
    // next slice end and finally the group two steps ahead.  Every load is
    // consumed one step after it is issued, so no wait reaches the ring's
    // newest groups.
    uint64_t g = atomic_load_u64(a.thr + p);
    issue(0);
    issue(1);
    int32_t acc = 0, pacc = 0, prow = -1;
    uint64_t pt = 0;
    bool pend = false, pmaybe = false;
    bool done = false;
    while (!done) {
#pragma unroll
        for (int j = 0; j < RING; ++j) {
            if (done) break;
            if (!dval[j]) {
                done = true;
                break;
            }
            if (pend) {
                const uint64_t key = dkey_i32(pacc, prow);
                col.push(pmaybe && prow >= 0 && key > pt, key);
            }
            const int nd = dnd[j];
            uint32_t wm[E];
#pragma unroll
            for (int i = 0; i < E; ++i) wm[i] = i < nd ? w[j][i] : 0u;
            acc = decode_group<E, SIGNED>(wm, nd, qx, a.col_bits, a.val_bits, cmask, acc);
            bool maybe = false;
            if (dend[j]) {
                pt = col.thr > g ? col.thr : g;
                maybe = acc >= dkey_score_i32(pt);
                pacc = acc;
                acc = 0;
            }
            pend = __ballot(maybe) != 0;
            pmaybe = maybe;
            prow = a.perm[pend ? (size_t)ds[j] * 64 + lane : (size_t)lane];
            g = atomic_load_u64(a.thr + p);
            issue((j + 2) % RING);
        }
    }
    if (pend) {
        const uint64_t key = dkey_i32(pacc, prow);
        col.push(pmaybe && prow >= 0 && key > pt, key);
    }
    col.compact();
    if (lane == 0) s_n[wave] = col.n;
    __syncthreads();
    if (wave == 0) {
        wg_merge(col, colbuf, s_n, LB_WAVES);
        publish_and_append(col, a.thr + p, a.cnt + p, a.cand + a.cand_off[p]);
    }
}

// Per partition (one workgroup each, after lowbit_stage1): the partition's
// Top-M from its candidate list, reset of the list, optional candidate output,
// FP32 rescoring of the candidates into resc.  The kernel boundary after
// lowbit_stage1 makes the lists visible (no device-scope fences, which on a
// multi-XCD part write back and invalidate whole L2 caches).

__global__ __launch_bounds__(LB_WG) void lowbit_merge(LowbitArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_n[LB_WAVES];
    __shared__ int s_flag;
    uint64_t *colbuf = (uint64_t *)(smem + 64);
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int p = blockIdx.x;
    const uint32_t n = a.cnt[p];
    const uint64_t *keys = a.cand + a.cand_off[p];
    WaveTopM<2> col;
    col.init(colbuf + wave * WaveTopM<2>::cap, a.M);
    push_range(col, keys, n, (uint32_t)wave * 64, LB_WG);
    col.compact();
    if (lane == 0) s_n[wave] = col.n;
    __syncthreads();
    if (wave == 0) {
        wg_merge(col, colbuf, s_n, LB_WAVES);
        if (lane == 0) s_flag = col.n;
    }
    __syncthreads();
    const int mcount = s_flag;
    const uint64_t *top = colbuf;
    if (threadIdx.x == 0) {  // ready for the next query
        a.cnt[p] = 0u;
        a.thr[p] = 0ull;
    }
    if (a.cand_rows_out) {
        for (int i = threadIdx.x; i < a.M; i += LB_WG) {
            a.cand_rows_out[(size_t)p * a.M + i] = i < mcount ? dkey_row(top[i]) : -1;
            a.cand_scores_out[(size_t)p * a.M + i] = i < mcount ? dkey_score_i32(top[i]) : 0;
        }
    }
    if (a.K <= 0) return;

    // ---- FP32 rescoring of the partition's candidates ----
    for (int base = 0; base < a.M; base += LB_WG / a.G) {
        const int i = base + (int)threadIdx.x / a.G;
        const bool valid = i < mcount;
        const int32_t r = valid ? dkey_row(top[i]) : 0;
        const float sc = group_dot_rt(a.ptr, a.idx, a.val, a.x, r, valid, a.G);
        if ((threadIdx.x & (a.G - 1)) == 0 && i < a.M) a.resc[(size_t)p * a.M + i] = valid ? dkey_f32(sc, r) : 0ull;
    }
}

// Writes the K best keys of a merged collector as ranked results
// (value -inf, index -1 past the available keys).

// candidates (collectors of 64*R keys, K <= 64*R - 64).
template <int R>
__global__ __launch_bounds__(ACC_WG) void lowbit_final(LowbitArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_n[ACC_WAVES];
    uint64_t *colbuf = (uint64_t *)smem;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int total = a.split * a.M;
    WaveTopM<R> col;
    col.init(colbuf + wave * WaveTopM<R>::cap, a.K);
    push_range(col, a.resc, (uint32_t)total, (uint32_t)wave * 64, ACC_WG);
    col.compact();
    if (lane == 0) s_n[wave] = col.n;
    __syncthreads();
    if (wave == 0) {
        wg_merge(col, colbuf, s_n, ACC_WAVES);
        write_results(col, a.K, a.out);
    }
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
        int32_t acc = 0;
        for (int g0 = 0; g0 < (int)sl.y; g0 += E) {
            uint32_t w[E];
            const int nd = min(E, (int)sl.y - g0);
            load_group<E>(w, a.words + ((size_t)sl.x + g0) * 64 + lane, nd);
            acc = decode_group<E, SIGNED>(w, nd, qx, a.col_bits, a.val_bits, cmask, acc);
        }
        const int32_t row = a.perm[s * 64 + lane];
        if (row >= 0) a.y[row] = acc;
    }
}


template <bool SIGNED>
static lowbit_fn stage1_for(int E) {
    switch (E) {
#define CASE(e) case e: return lowbit_stage1<e, SIGNED>;
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

lowbit_fn get_lowbit_stage1(int E, bool is_signed) { return is_signed ? stage1_for<true>(E) : stage1_for<false>(E); }
lowbit_fn get_lowbit_merge() { return lowbit_merge; }
lowbit_fn get_lowbit_final(int R) {
    switch (R) {
        case 2: return lowbit_final<2>;
        case 4: return lowbit_final<4>;
        case 8: return lowbit_final<8>;
        case 16: return lowbit_final<16>;
        default: return nullptr;
    }
}
scores_fn get_lowbit_scores(int E, bool is_signed) { return is_signed ? scores_for<true>(E) : scores_for<false>(E); }
