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
//   lowbit_final_radix / lowbit_final
//                    the global Top-K of the rescored candidates;
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

// One workgroup of SEL_WG threads per partition p = blockIdx.x, after
// lowbit_slicemax.
// 1. ts = the M-th largest maximum among the partition's full slices (radix
//    selection): each of those M slices holds a distinct row scoring >= ts,
//    so every row of the partition's Top-M scores >= ts and lies in a slice
//    whose maximum is >= ts.
// 2. The slices whose maximum reaches ts are listed in LDS and rescanned
//    exactly, CS at a time: their groups are spread over all wavefronts
//    (kU loads in flight per wavefront) and the partial scores added into
//    per-slice LDS sums, so long slices do not serialize on one wavefront.
//    Rows >= ts are appended to an LDS key list, which is cut back to its
//    best M whenever another chunk might not fit.
// 3. The list's best M by radix selection, ranked by counting.  Past kList
//    candidate slices (heavy ties) each wavefront instead rescans its share
//    slice by slice into a wavefront Top-M collector.
// 4. Optional candidate output and the FP32 rescoring of the M rows.
template <int E, bool SIGNED>
__global__ __launch_bounds__(SEL_WG) void lowbit_select(LowbitArgs a) {
    constexpr int V = 8;         // slice maxima per thread held in registers
    constexpr int kList = 2048;  // candidate slices listed in LDS
    constexpr int CS = 32;       // candidate slices per rescan chunk
    constexpr int kU = E <= 16 ? 3 : 2;  // groups in flight per wavefront (128 VGPRs at 16 waves)
    constexpr int KCAP = 4096;   // LDS key list (rows >= ts)
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t s_hist[512];
    __shared__ int32_t s_list[kList];
    __shared__ uint2 s_lmeta[kList];
    __shared__ int32_t s_sum[CS * 64];
    __shared__ int s_ustart[CS + 1];
    __shared__ int s_n[SEL_WAVES];
    __shared__ int s_m, s_cnt, s_kn, s_tn, s_need[2];
    __shared__ uint32_t s_t32[2];
    __shared__ uint64_t s_t64[2];
    __shared__ uint64_t s_tmp[64], s_top[64];
    float *red = (float *)smem;
    uint64_t *keys = (uint64_t *)(smem + 64);  // KCAP keys; the slow path's collectors reuse it
    int8_t *qx = (int8_t *)(smem + 64 + 8 * KCAP);
    float *xs = (float *)(smem + 64 + 8 * KCAP + ((a.cols + 15) & ~15));  // FP32 query (rescoring)
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int p = blockIdx.x;
    const int64_t s0 = a.part_slice[p];
    const int nsl = (int)(a.part_slice[p + 1] - s0);
    const int32_t *smax = a.smax + s0;
    const uint8_t *nvalid = a.nvalid + s0;
    const uint32_t cmask = (1u << a.col_bits) - 1u;
    const bool regs = nsl <= V * SEL_WG;
    int32_t sm[V];
    bool full[V];
    if (regs) {
        // one uniform test around loads with clamped indices: they issue
        // together (a per-element condition would wait for each load)
        if (nsl > 0) {
#pragma unroll
            for (int u = 0; u < V; ++u) {
                const int i = min(tid + u * SEL_WG, nsl - 1);
                sm[u] = smax[i];
                full[u] = nvalid[i] == 64;
            }
        } else {
#pragma unroll
            for (int u = 0; u < V; ++u) {
                sm[u] = 0;
                full[u] = false;
            }
        }
    }
    if (tid == 0) {
        s_cnt = 0;
        s_kn = 0;
        s_tn = 0;
    }
    quantize_query_lds(a.x, a.cols, qx, red, xs);

    // 1. threshold (unsigned order: score ^ 2^31; key 0 = not a full slice):
    //    up to 256 slices and M <= 32 by maximum extraction in wavefront 0,
    //    otherwise by radix selection over the workgroup
    auto mkey = [](int32_t v, bool ok) -> uint32_t { return ok ? ((uint32_t)v ^ 0x80000000u) : 0u; };
    uint32_t T;
    if (nsl <= 256 && a.M <= 32) {
        if (tid < 256) s_hist[tid] = tid < nsl ? mkey(sm[0], full[0]) : 0u;  // regs holds slice tid in sm[0]
        __syncthreads();
        if (wave == 0) {
            uint32_t v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = s_hist[u * 64 + lane];
            const uint32_t t = wave_kth_u32<4>(v, a.M);
            if (lane == 0) s_t32[0] = t;
        }
        __syncthreads();
        T = s_t32[0];
    } else {
        T = wg_radix_kth<uint32_t>(
            [&](auto fn) {
                if (regs) {
#pragma unroll
                    for (int u = 0; u < V; ++u) fn(tid + u * SEL_WG < nsl ? mkey(sm[u], full[u]) : 0u);
                } else {
                    for (int i = tid; i < nsl; i += SEL_WG) fn(mkey(smax[i], nvalid[i] == 64));
                }
            },
            a.M, s_hist, s_t32, s_need);
    }
    const int32_t ts = T == 0u ? INT32_MIN : (int32_t)(T ^ 0x80000000u);

    // 2. candidate slices, listed with their metadata
    auto offer = [&](int i, bool cand) {
        const uint64_t m = __ballot(cand);
        if (m == 0) return;
        const uint2 meta = cand ? a.slices[s0 + i] : make_uint2(0u, 0u);  // in flight during the append
        int base = 0;
        if (lane == 0) base = atomicAdd(&s_cnt, __popcll(m));
        base = __shfl(base, 0);
        const int j = base + mbcnt64(m);
        if (cand && j < kList) {
            s_list[j] = i;
            s_lmeta[j] = meta;
        }
    };
    if (regs) {
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int i = tid + u * SEL_WG;
            offer(i, i < nsl && sm[u] >= ts);
        }
    } else {
        for (int b = 0; b < nsl; b += SEL_WG) {
            const int i = b + tid;
            offer(i, i < nsl && smax[min(i, nsl - 1)] >= ts);
        }
    }
    __syncthreads();
    const int ncand = s_cnt;
    int mcount = 0;
    if (ncand <= kList) {
        // keeps the best M of the key list (all threads; the list is published)
        auto cut = [&](int n) -> int {
            const uint64_t Tk = wg_radix_kth<uint64_t>(
                [&](auto fn) {
                    for (int i = tid; i < n; i += SEL_WG) fn(keys[i]);
                },
                a.M, s_hist, s_t64, s_need);
            for (int i0 = 0; i0 < n; i0 += SEL_WG) {
                const int i = i0 + tid;
                const uint64_t k = i < n ? keys[i] : 0ull;
                const bool sel = k != 0ull && k >= Tk;
                const uint64_t m = __ballot(sel);
                if (m == 0) continue;
                int base = 0;
                if (lane == 0) base = atomicAdd(&s_tn, __popcll(m));
                base = __shfl(base, 0);
                const int j = base + mbcnt64(m);
                if (sel && j < 64) s_tmp[j] = k;  // j < M for distinct keys
            }
            __syncthreads();
            const int m = min(s_tn, a.M);
            __syncthreads();
            if (tid == 0) s_tn = 0;
            return m;
        };
        for (int cb = 0; cb < ncand; cb += CS) {
            const int cn = min(CS, ncand - cb);
            __syncthreads();  // metadata listed; the previous chunk's sums and keys in place
            if (s_kn > KCAP - CS * 64) {  // the next chunk might not fit: keep the best M
                const int m = cut(s_kn);
                if (tid < m) keys[tid] = s_tmp[tid];
                if (tid == 0) s_kn = m;
                __syncthreads();
            }
            for (int i = tid; i < CS * 64; i += SEL_WG) s_sum[i] = 0;
            if (wave == 0) {
                const int ng = lane < cn ? ((int)s_lmeta[cb + lane].y + E - 1) / E : 0;
                int inc = ng;
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(inc, o);
                    if (lane >= o) inc += t;
                }
                if (lane < CS) s_ustart[lane + 1] = inc;
                if (lane == 0) s_ustart[0] = 0;
            }
            constexpr int PR = CS / SEL_WAVES;  // slices each wavefront offers
            int32_t prow[PR];
#pragma unroll
            for (int r = 0; r < PR; ++r) {  // unconditional (clamped) lookups, masked after
                const int c = wave + r * SEL_WAVES;
                const int32_t v = a.perm[(s0 + s_list[cb + min(c, cn - 1)]) * 64 + lane];
                prow[r] = c < cn ? v : -1;
            }
            __syncthreads();
            const int nu = s_ustart[cn];
            for (int u0 = wave * kU; u0 < nu; u0 += SEL_WAVES * kU) {
                uint32_t w[kU][E];
                int cu[kU], nd[kU];
#pragma unroll
                for (int t = 0; t < kU; ++t) {
                    const int u = u0 + t;
                    int c = -1, g0 = 0;
                    uint2 sl = make_uint2(0u, 0u);
                    if (u < nu) {
                        int lo = 0, hi = cn - 1;  // last c with s_ustart[c] <= u
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (s_ustart[mid] <= u) lo = mid;
                            else hi = mid - 1;
                        }
                        c = lo;
                        g0 = (u - s_ustart[c]) * E;
                        sl = s_lmeta[cb + c];
                    }
                    load_group<E>(w[t], slice_rsrc(a.words, sl), g0);
                    cu[t] = c;
                    nd[t] = c >= 0 ? min(E, (int)sl.y - g0) : 0;
                }
#pragma unroll
                for (int t = 0; t < kU; ++t) {
                    if (cu[t] < 0) continue;
                    const int32_t part = decode_group<E, SIGNED>(w[t], nd[t], qx, a.col_bits, a.val_bits, cmask, 0);
                    atomicAdd(&s_sum[cu[t] * 64 + lane], part);
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < PR; ++r) {
                const int c = wave + r * SEL_WAVES;
                if (c < cn) {
                    const int32_t sc = s_sum[c * 64 + lane];
                    const bool ok = prow[r] >= 0 && sc >= ts;
                    const uint64_t m = __ballot(ok);
                    if (m == 0) continue;
                    int base = 0;
                    if (lane == 0) base = atomicAdd(&s_kn, __popcll(m));
                    base = __shfl(base, 0);
                    if (ok) keys[base + mbcnt64(m)] = dkey_i32(sc, prow[r]);
                }
            }
        }
        __syncthreads();
        const int nk = s_kn;
        if (nk <= 64) {  // one wavefront sorts the list (keys are distinct; 0 sorts last)
            if (wave == 0) {
                uint64_t w[1] = {lane < nk ? keys[lane] : 0ull};
                reg_sort_desc<1>(w);
                if (lane < a.M) s_top[lane] = w[0];
            }
            mcount = min(nk, a.M);
        } else {
            mcount = cut(nk);
            if (tid < mcount) {  // rank by counting (keys are distinct)
                const uint64_t k = s_tmp[tid];
                int rank = 0;
                for (int j = 0; j < mcount; ++j) rank += s_tmp[j] > k;
                s_top[rank] = k;
            }
        }
        __syncthreads();
    } else {
        uint64_t *colbuf = keys;
        WaveTopM<2> col;
        col.init(colbuf + wave * WaveTopM<2>::cap, a.M);
        for (int b = wave * 64; b < nsl; b += SEL_WG) {
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
        col.compact();
        if (lane == 0) s_n[wave] = col.n;
        __syncthreads();
        if (wave == 0) {
            wg_merge(col, colbuf, s_n, SEL_WAVES);
            if (lane < col.n) s_top[lane] = col.buf[lane];  // col.n <= M <= 64, sorted
            if (lane == 0) s_m = col.n;
        }
        __syncthreads();
        mcount = s_m;
    }
    const uint64_t *top = s_top;
    if (a.cand_rows_out) {
        for (int i = threadIdx.x; i < a.M; i += SEL_WG) {
            a.cand_rows_out[(size_t)p * a.M + i] = i < mcount ? dkey_row(top[i]) : -1;
            a.cand_scores_out[(size_t)p * a.M + i] = i < mcount ? dkey_score_i32(top[i]) : 0;
        }
    }
    if (a.K <= 0) return;

    // 4. FP32 rescoring of the partition's candidates (16 entries per lane in
    //    flight; the query from LDS)
    for (int base = 0; base < a.M; base += SEL_WG / a.G) {
        const int i = base + (int)threadIdx.x / a.G;
        const bool valid = i < mcount;
        const int32_t r = valid ? dkey_row(top[i]) : 0;
        const float sc = group_dot_rt<16>(a.ptr, a.idx, a.val, xs, r, valid, a.G);
        if ((threadIdx.x & (a.G - 1)) == 0 && i < a.M) a.resc[(size_t)p * a.M + i] = valid ? dkey_f32(sc, r) : 0ull;
    }
}

// One workgroup: the global Top-K of the split*M rescored candidates
// (split*M <= 4*64*R), ranks by merging sorted wavefront lists.
template <int R>
__global__ __launch_bounds__(ACC_WG) void lowbit_final(LowbitArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_nz[ACC_WAVES];
    topk_by_rank<R>(a.resc, a.split * a.M, a.K, a.out, (uint64_t *)smem, s_nz);
}

// The same for K <= 256 by radix selection (split*M <= 4*SEL_WG).
__global__ __launch_bounds__(SEL_WG) void lowbit_final_radix(LowbitArgs a) {
    __shared__ __align__(16) uint32_t lds[1032];
    topk_radix<4>(a.resc, a.split * a.M, a.K, a.out, lds);
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
lowbit_fn get_lowbit_final_radix() { return lowbit_final_radix; }
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
