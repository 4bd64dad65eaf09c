// k_dense.hip -- the batched low-bit path around the matrix-core scores
// (pk_scores, k_pkmfma.hpp): query preparation, per-partition selection with
// FP32 rescoring, and the per-query Top-K.
//
// The INT32 stage-1 scores (quantify_util.py:173-196: exact integer dot of
// the quantized rows with the quantized query) of a block of queries come
// from v_mfma_i32_32x32x32_i8 over the slices' packets, expanded in LDS: at
// the densities of the embedding workloads (2-15% non-zeros) the matrix cores
// multiply a dense slice image faster than the vector ALUs decode (column,
// value) pairs once a pass serves tens of queries.  Integer arithmetic is
// exact, so the scores are the same integers as the packet kernels'.  Only
// the four best (score, row) keys of each (query, 64-row slice) leave the
// scores kernel; the per-partition selection derives its threshold from the
// slice maxima and rescans exactly the rare slice whose list may have
// dropped a qualifying row.
#include "kernels_common.hpp"

// 32-bit slice keys: (score + 2^25) << 6 | row_in_slice; a larger key is a
// higher score (equal scores: any fixed order will do, the lists are only a
// filter and every candidate is re-keyed by (score, row)).  0 = no row.
// |score| < 2^25 is checked on the host before the dense layout is used.
__device__ __forceinline__ int32_t tk_score(uint32_t k) { return (int32_t)(k >> 6) - (1 << 25); }
__device__ __forceinline__ int tk_rho(uint32_t k) { return (int)(k & 63u); }

// Exact INT32 score of this lane's row (row_in_slice = lane) of slice s for
// one query, decoded from the slice's packets with the entry width known
// only at run time: the lane's dwords stream through a 64-bit bit buffer,
// eight loads in flight, and every complete E-bit entry is taken off the
// bottom (the same number of entries on every lane, so no divergence).
// qx: the query's int8 values in LDS.
__device__ __forceinline__ int32_t slice_dot_packets(const DenseArgs &a, int64_t s, const int8_t *qx) {
    const uint2 sl = a.slices[s];
    const int cb = a.col_bits, vb = a.val_bits, E = cb + vb;
    const uint64_t emask = (1ull << E) - 1ull;
    const uint32_t cmask = (1u << cb) - 1u;
    const __amdgpu_buffer_rsrc_t rs = slice_rsrc(a.words, sl);
    const int lane = lane_id();
    uint64_t buf = 0ull;
    int nb = 0;
    int32_t acc = 0;
    for (int d0 = 0; d0 < (int)sl.y; d0 += 8) {
        uint32_t w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = __builtin_amdgcn_raw_buffer_load_b32(rs, ((d0 + u) * 64 + lane) * 4, 0, 0);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            buf |= (uint64_t)w[u] << nb;
            nb += 32;
            while (nb >= E) {
                const uint32_t e = (uint32_t)(buf & emask);
                buf >>= E;
                nb -= E;
                const int v = a.is_signed ? __builtin_amdgcn_sbfe(e, cb, vb) : (int)__builtin_amdgcn_ubfe(e, cb, vb);
                acc += v * (int)qx[e & cmask];
            }
        }
    }
    return acc;
}

// One workgroup per query slot of the padded batch: int8 quantization with
// the query's own scale (quantify_util.py:31-49, 259; the rule of the packet
// kernels), written in the B-operand layout, in plain order (rescans) and as
// the unsigned-8-bit offset correction.
__global__ __launch_bounds__(ACC_WG) void dense_prep(DenseArgs a) {
    __shared__ float red[ACC_WAVES];
    __shared__ int isum[ACC_WAVES];
    const int QB = 32 * a.NT;
    const int q = blockIdx.x, qb = q / QB, ql = q % QB, t = ql >> 5, j = ql & 31;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const bool valid = q < a.nq;
    const float *x = a.X + (size_t)(valid ? q : 0) * a.cols;
    // the first XR*ACC_WG columns are read once, with loads issued together
    // (clamped indices, masked after), and kept in registers for both passes
    constexpr int XR = 4;
    float xr[XR];
#pragma unroll
    for (int u = 0; u < XR; ++u) xr[u] = x[min((int)threadIdx.x + u * ACC_WG, a.cols - 1)];
    float m = 0.f;
#pragma unroll
    for (int u = 0; u < XR; ++u)
        if (valid && (int)threadIdx.x + u * ACC_WG < a.cols) m = fmaxf(m, fabsf(xr[u]));
    if (valid)
        for (int c = threadIdx.x + XR * ACC_WG; c < a.cols; c += ACC_WG) m = fmaxf(m, fabsf(x[c]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (lane == 0) red[wave] = m;
    __syncthreads();
    float s = red[0];
    for (int w = 1; w < ACC_WAVES; ++w) s = fmaxf(s, red[w]);
    int sum = 0;
    for (int c = threadIdx.x, it = 0; c < a.cols_pad; c += ACC_WG, ++it) {
        int qv = 0;
        if (valid && c < a.cols && s > 0.f) {
            float xc = 0.f;
#pragma unroll
            for (int u = 0; u < XR; ++u)
                if (it == u) xc = xr[u];
            if (it >= XR) xc = x[c];
            const float qf = __fmul_rn(127.0f, __fdiv_rn(xc, s));
            qv = (int)__builtin_rintf(qf);
            qv = qv > 127 ? 127 : (qv < -128 ? -128 : qv);
        }
        const int kb = c >> 5, h = (c >> 4) & 1, bb = c & 15;
        a.qtab[(((size_t)qb * a.nkb + kb) * a.NT + t) * 1024 + (j + 32 * h) * 16 + bb] = (int8_t)qv;
        if (valid) a.qplain[(size_t)q * a.cols_pad + c] = (int8_t)qv;
        sum += qv;
    }
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    if (lane == 0) isum[wave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < ACC_WAVES; ++w) tot += isum[w];
        a.corr[q] = a.u8off ? 128 * tot : 0;
    }
}

// Bitonic sort, descending, of 64*R 32-bit keys in registers (element
// i = r*64 + lane in v[r] of that lane).
template <int R>
__device__ __forceinline__ void reg_sort_desc_u32(uint32_t (&v)[R]) {
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
                        const uint32_t x = v[r], y = v[rp];
                        const bool sw = desc ? (x < y) : (x > y);
                        v[r] = sw ? y : x;
                        v[rp] = sw ? x : y;
                    }
                }
            } else {
                const bool lower = (lane & j) == 0;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint32_t o = (uint32_t)__shfl_xor((int)v[r], j);
                    const bool desc = ((r * 64 + lane) & k) == 0;
                    v[r] = (lower == desc) ? max(v[r], o) : min(v[r], o);
                }
            }
        }
    }
}

// Global rows of the listed keys of 64*U slices (k4[u] = the list of slice
// s0 + b + u*64 + lane); -1 for keys of padding rows (row_in_slice >= nv[u],
// nv[u] = 0 past the partition).  Issued before the threshold is known, so
// the lookups are in flight while it is computed.
template <int U>
__device__ __forceinline__ void list_rows(const DenseArgs &a, const uint4 (&k4)[U], const int (&nv)[U], int64_t s0,
                                          int b, int32_t (&row)[U][4]) {
    const int lane = lane_id();
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t s = s0 + b + u * 64 + lane;
        const uint32_t ke[4] = {k4[u].x, k4[u].y, k4[u].z, k4[u].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            // loaded for every key (slices past the matrix clamped into it) so
            // that the loads issue together; the row is kept for real rows
            const int32_t r = a.perm[min(s, a.nslices - 1) * 64 + tk_rho(ke[e])];
            row[u][e] = tk_rho(ke[e]) < nv[u] ? r : -1;
        }
    }
}

// Offers the listed keys of those slices scoring >= ts to the collector; a
// slice whose list may have dropped such a row is rescanned exactly (from its
// packets).  Keys of padding rows are discarded, and since they may have
// displaced real rows from a list, every list whose fourth key scores >= ts
// is rescanned.  Rows are looked up for the offered keys only.
template <int U>
__device__ __forceinline__ void offer_lists(const DenseArgs &a, WaveTopM<2> &cb, const uint4 (&k4)[U], const int (&nv)[U],
                                   int64_t s0, int b, int32_t ts, int q, int8_t *qx, bool &have_qx) {
    const int lane = lane_id();
    bool rescan[U], pass[U * 4];
    uint64_t key[U * 4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        rescan[u] = nv[u] > 0 && tk_score(k4[u].w) >= ts;
        const int64_t s = min(s0 + b + u * 64 + lane, a.nslices - 1);
        const uint32_t ke[4] = {k4[u].x, k4[u].y, k4[u].z, k4[u].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            pass[u * 4 + e] = !rescan[u] && tk_rho(ke[e]) < nv[u] && tk_score(ke[e]) >= ts;
            const int32_t r = pass[u * 4 + e] ? a.perm[s * 64 + tk_rho(ke[e])] : 0;
            key[u * 4 + e] = dkey_i32(tk_score(ke[e]), r);
        }
    }
    cb.push_many<U * 4>(pass, key);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        uint64_t mask = __ballot(rescan[u]);
        while (mask) {
            const int l = __builtin_ctzll(mask);
            mask &= mask - 1;
            if (!have_qx) {
                const int8_t *src = a.qplain + (size_t)q * a.cols_pad;
                for (int c = lane; c < a.cols_pad; c += 64) qx[c] = src[c];
                wave_sync();
                have_qx = true;
            }
            const int64_t s = s0 + b + u * 64 + l;
            const int32_t sc = slice_dot_packets(a, s, qx);
            const int32_t r = a.perm[s * 64 + lane];
            cb.push(r >= 0 && sc >= ts, dkey_i32(sc, r));
        }
    }
}

// Copies the FP32 query (cols values) into this wavefront's LDS: eight
// clamped loads per lane are issued together per 512 columns (a loop of
// single loads would wait for each one before its LDS write).
__device__ __forceinline__ void stage_query(float *xs, const float *xq, int cols) {
    const int lane = lane_id();
    for (int c0 = 0; c0 < cols; c0 += 512) {
        float t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = xq[min(c0 + u * 64 + lane, cols - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (c0 + u * 64 + lane < cols) xs[c0 + u * 64 + lane] = t[u];
    }
}

// Stage-1 outputs of one (query, partition) and its FP32 rescoring
// (quantify_util.py:354-367): keys[0, nn) are the partition's Top-M in rank
// order (LDS); slot i of the candidate list is keys[i] or empty past nn.
// Each group of G lanes rescores C candidates at once in the declared order
// (strided fmaf, then the butterfly), so the scores equal those of every
// other rescoring path; xs is the FP32 query, staged in LDS; ranges
// (optional, LDS): the candidates' CSR entry ranges, already loaded.
__device__ __forceinline__ void emit_candidates(const DenseArgs &a, const uint64_t *keys, int nn, size_t base,
                                                const float *xs, const uint2 *ranges) {
    const int lane = lane_id(), M = a.M;
    if (a.cand_rows_out)
        for (int i = lane; i < M; i += 64) {
            const uint64_t key = i < nn ? keys[i] : 0ull;
            a.cand_rows_out[base + i] = key ? dkey_row(key) : -1;
            a.cand_scores_out[base + i] = key ? dkey_score_i32(key) : 0;
        }
    if (!a.out) return;
    constexpr int C = 4;
    const int G = a.G, groups = 64 / G, grp = lane / G;
    wave_sync();
    for (int i0 = 0; i0 < M; i0 += groups * C) {
        int32_t r[C];
        bool valid[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int i = i0 + c * groups + grp;
            const uint64_t key = keys[min(i, 63)];
            valid[c] = i < nn;
            r[c] = valid[c] ? dkey_row(key) : 0;
        }
        float sc[C];
        if (ranges) {
            uint32_t lo[C], hi[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const uint2 rg = ranges[min(i0 + c * groups + grp, 63)];
                lo[c] = valid[c] ? rg.x : 0u;
                hi[c] = valid[c] ? rg.y : 0u;
            }
            group_dot_ranges_multi<C, 4>(a.idx, a.val, xs, lo, hi, G, sc);
        } else {
            group_dot_rt_multi<C, 4>(a.ptr, a.idx, a.val, xs, r, valid, G, sc);
        }
        if ((lane & (G - 1)) == 0)
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int i = i0 + c * groups + grp;
                if (i < M) a.resc[base + i] = valid[c] ? dkey_f32(sc[c], r[c]) : 0ull;
            }
    }
}

// One wavefront per (partition p = blockIdx.x, query q = blockIdx.y*4 + wave):
//  1. ts = a score that M distinct rows of the partition reach: the M-th best
//     of the lanes' largest full-slice maxima on the register path, the M-th
//     best full-slice maximum otherwise (INT32_MIN when there are not M).  The
//     partition's M-th best score is >= ts, so no row below ts can be in its
//     Top-M.
//  2. Every listed key of a real row scoring >= ts is a candidate.  A list
//     whose fourth key still scores >= ts may have dropped a qualifying row:
//     that slice is rescanned exactly (all 64 rows, one dot4 pass) instead.
//     Slice maxima of partial slices (padding rows score 0) are left out of
//     step 1; the maxima of the full slices still give M distinct real rows
//     >= ts.
//  3. Top-M of the candidates by (score, row) (quantify_util.py:173-196 with
//     the lower-row tie rule), then their FP32 rescoring (out != null).
// Partitions of at most 256 slices (M <= 32, at most 64 candidates) keep
// their lists in registers: one prefix sum appends the candidates, a count
// ranks them.  In the WIDE variant, up to 1024 slices the lists are still
// read once into registers and offered to a collector; anything else streams
// them twice (threshold, then offers).
template <bool WIDE>
__global__ __launch_bounds__(ACC_WG) void dense_select(DenseArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int p = blockIdx.x, q = blockIdx.y * ACC_WAVES + wave;
    if (q >= a.nq) return;  // whole wavefront; no workgroup barrier below
    uint64_t *bufA = (uint64_t *)smem + wave * 256, *bufB = bufA + 128;
    int8_t *qx = (int8_t *)(smem + ACC_WAVES * 256 * 8) + (size_t)wave * a.cols_pad;
    float *xs = (float *)(smem + ACC_WAVES * 256 * 8 + (size_t)ACC_WAVES * a.cols_pad) + (size_t)wave * a.cols_pad;
    const float *xq = a.X + (size_t)q * a.cols;
    const int64_t s0 = a.part_slice[p];
    const int nsl = (int)(a.part_slice[p + 1] - s0);
    const uint4 *L = (const uint4 *)a.tk + (size_t)q * a.nslices + s0;
    const int M = a.M;
    const size_t base = ((size_t)q * a.split + p) * M;
    bool have_qx = false;

    if (nsl > 0 && nsl <= 256 && M <= 32) {
        // unconditional loads (clamped into the partition, masked after) so
        // that the lists and the query for the rescoring arrive in one round
        // trip; a conditional load would be waited for on its own
        const bool narrow = a.cols <= 512;
        float xt[8];
        if (a.out && narrow)
#pragma unroll
            for (int u = 0; u < 8; ++u) xt[u] = xq[min(u * 64 + lane, a.cols - 1)];
        uint4 k4[4];
        int nv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = min(u * 64 + lane, nsl - 1);
            k4[u] = L[i];
            nv[u] = (int)a.nvalid[s0 + i];
        }
        if (a.out && narrow) {
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (u * 64 + lane < a.cols) xs[u * 64 + lane] = xt[u];
        } else if (a.out) {
            stage_query(xs, xq, a.cols);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (u * 64 + lane >= nsl) {
                k4[u] = make_uint4(0u, 0u, 0u, 0u);
                nv[u] = 0;
            }
        int32_t row[4][4];
        list_rows<4>(a, k4, nv, s0, 0, row);
        int32_t ts = INT32_MIN;
        {
            uint32_t v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = nv[u] == 64 ? k4[u].x : 0u;
            // the M-th largest of the lanes' largest full-slice maxima: M lanes
            // hold a maximum >= it, i.e. M distinct rows score >= it -- a
            // valid threshold, found with one 64-key register sort
            uint32_t h[1] = {max(max(v[0], v[1]), max(v[2], v[3]))};
            reg_sort_desc_u32<1>(h);
            const uint32_t t = (uint32_t)__shfl((int)h[0], M - 1);
            if (t != 0u) ts = tk_score(t);
        }
        bool rescan[4], pass[16];
        int c = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            rescan[u] = nv[u] > 0 && tk_score(k4[u].w) >= ts;
            const uint32_t ke[4] = {k4[u].x, k4[u].y, k4[u].z, k4[u].w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                pass[u * 4 + e] = !rescan[u] && row[u][e] >= 0 && tk_score(ke[e]) >= ts;
                c += pass[u * 4 + e] ? 1 : 0;
            }
        }
        int incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        int n = __builtin_amdgcn_readlane(incl, 63);
        if (n <= 64) {
            int pos = incl - c;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t ke[4] = {k4[u].x, k4[u].y, k4[u].z, k4[u].w};
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (pass[u * 4 + e]) bufB[pos++] = dkey_i32(tk_score(ke[e]), row[u][e]);
            }
            // slices to rescan, listed in LDS so that one loop serves them all
            uint32_t *rl = (uint32_t *)bufA;
            int nr = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint64_t m = __ballot(rescan[u]);
                if (rescan[u]) rl[nr + mbcnt64(m)] = (uint32_t)(u * 64 + lane);
                nr += __popcll(m);
            }
            wave_sync();
            for (int k = 0; k < nr && n <= 64; ++k) {
                const int64_t s = s0 + rl[k];
                if (!have_qx) {
                    const int8_t *src = a.qplain + (size_t)q * a.cols_pad;
                    for (int cc = lane; cc < a.cols_pad; cc += 64) qx[cc] = src[cc];
                    wave_sync();
                    have_qx = true;
                }
                const int32_t sc = slice_dot_packets(a, s, qx);
                const int32_t r = a.perm[s * 64 + lane];
                const bool pr = r >= 0 && sc >= ts;
                const uint64_t bm = __ballot(pr);
                if (n + __popcll(bm) <= 64 && pr) bufB[n + mbcnt64(bm)] = dkey_i32(sc, r);
                n += __popcll(bm);
            }
            if (n <= 64) {
                wave_sync();
                const uint64_t k = lane < n ? bufB[lane] : 0ull;
                // the candidates' CSR entry ranges, in flight while the ranks
                // are counted (out set: the CSR is present)
                uint2 rg = make_uint2(0u, 0u);
                if (a.out) {
                    const int32_t r = lane < n ? dkey_row(k) : 0;
                    rg = make_uint2(a.ptr[r], a.ptr[r + 1]);
                }
                const uint32_t klo = (uint32_t)k, khi = (uint32_t)(k >> 32);
                int rank = 0;
                for (int j = 0; j < n; ++j) {
                    const uint64_t kj = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)khi, j) << 32) |
                                        (uint32_t)__builtin_amdgcn_readlane((int)klo, j);
                    rank += (kj > k || (kj == k && j < lane)) ? 1 : 0;
                }
                const int nn = n < M ? n : M;
                uint2 *ranges = (uint2 *)(bufA + 64);
                if (lane < n && rank < nn) {
                    bufA[rank] = k;
                    ranges[rank] = rg;
                }
                wave_sync();
                emit_candidates(a, bufA, nn, base, xs, ranges);
                return;
            }
        }
    }

    // General path: a collector of the offered keys.
    if (a.out) stage_query(xs, xq, a.cols);
    WaveTopM<2> cb;
    cb.init(bufB, M);
    constexpr int U = 4;
    int32_t ts = INT32_MIN;
    if (WIDE && nsl <= 1024 && M <= 32) {
        // every list in registers (16 per lane, loaded together, clamped into
        // the partition and masked after); the threshold is the M-th largest
        // full-slice maximum, then the lists are offered from the registers
        uint4 k4[4][U];
        int nv[4][U];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = min((j * U + u) * 64 + lane, nsl - 1);
                k4[j][u] = L[i];
                nv[j][u] = (int)a.nvalid[s0 + i];
            }
        uint32_t v[16];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if ((j * U + u) * 64 + lane >= nsl) {
                    k4[j][u] = make_uint4(0u, 0u, 0u, 0u);
                    nv[j][u] = 0;
                }
                v[j * U + u] = nv[j][u] == 64 ? k4[j][u].x : 0u;
            }
        const uint32_t t = wave_kth_u32<16>(v, M);
        if (t != 0u) ts = tk_score(t);
        offer_lists<U>(a, cb, k4[0], nv[0], s0, 0, ts, q, qx, have_qx);
        if (U * 64 < nsl) offer_lists<U>(a, cb, k4[1], nv[1], s0, U * 64, ts, q, qx, have_qx);
        if (2 * U * 64 < nsl) offer_lists<U>(a, cb, k4[2], nv[2], s0, 2 * U * 64, ts, q, qx, have_qx);
        if (3 * U * 64 < nsl) offer_lists<U>(a, cb, k4[3], nv[3], s0, 3 * U * 64, ts, q, qx, have_qx);
        cb.compact();
        emit_candidates(a, cb.buf, cb.n, base, xs, nullptr);
        return;
    }
    if (nsl <= 1024 && M <= 32) {  // the M-th largest maximum by extraction, maxima in registers
        uint32_t v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int i = u * 64 + lane;
            v[u] = i < nsl && a.nvalid[s0 + i] == 64 ? L[i].x : 0u;
        }
        const uint32_t t = wave_kth_u32<16>(v, M);
        if (t != 0u) ts = tk_score(t);
    } else {
        WaveTopM<2> ca;
        ca.init(bufA, M);
        for (int b = 0; b < nsl; b += 64 * U) {
            uint32_t k[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = b + u * 64 + lane;
                k[u] = i < nsl && a.nvalid[s0 + i] == 64 ? L[i].x : 0u;
            }
            static_assert(U == 4, "four pushes below");
            ca.push(k[0] != 0u, ((uint64_t)k[0] << 32) | (uint32_t)(b + lane));
            ca.push(k[1] != 0u, ((uint64_t)k[1] << 32) | (uint32_t)(b + 64 + lane));
            ca.push(k[2] != 0u, ((uint64_t)k[2] << 32) | (uint32_t)(b + 128 + lane));
            ca.push(k[3] != 0u, ((uint64_t)k[3] << 32) | (uint32_t)(b + 192 + lane));
        }
        ca.compact();
        if (ca.n >= M) ts = tk_score((uint32_t)(ca.buf[M - 1] >> 32));
    }
    for (int b = 0; b < nsl; b += 64 * U) {
        uint4 k4[U];
        int nv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = b + u * 64 + lane;
            k4[u] = i < nsl ? L[i] : make_uint4(0u, 0u, 0u, 0u);
            nv[u] = i < nsl ? (int)a.nvalid[s0 + i] : 0;
        }
        offer_lists<U>(a, cb, k4, nv, s0, b, ts, q, qx, have_qx);
    }
    cb.compact();
    emit_candidates(a, cb.buf, cb.n, base, xs, nullptr);
}

// One workgroup per query: each wavefront sorts a quarter of the query's
// split*M rescored keys in registers (64*R keys), the four sorted lists meet in
// LDS, and every key finds its rank in the union by binary search in the other
// three lists (keys are distinct: one per row).  Ranks below K[q] are written.
template <int R>
__global__ __launch_bounds__(ACC_WG) void batch_topk(DenseArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_nz[ACC_WAVES];
    const int q = blockIdx.x;
    const int total = a.split * a.M;
    topk_by_rank<R>(a.resc + (size_t)q * total, total, a.K[q], a.out + (size_t)q * a.out_stride, (uint64_t *)smem, s_nz);
}

// The same for K <= 256 (total <= V * ACC_WG): radix selection of the K-th
// key, then ranks by counting.
template <int V>
__global__ __launch_bounds__(ACC_WG) void batch_topk_radix(DenseArgs a) {
    __shared__ __align__(16) uint32_t lds[1032];
    const int q = blockIdx.x;
    const int total = a.split * a.M;
    topk_radix<V>(a.resc + (size_t)q * total, total, a.K[q], a.out + (size_t)q * a.out_stride, lds);
}

dense_fn get_dense_prep() { return dense_prep; }
dense_fn get_dense_select(bool wide) { return wide ? dense_select<true> : dense_select<false>; }
dense_fn get_batch_topk_radix(int V) {
    switch (V) {
        case 1: return batch_topk_radix<1>;
        case 2: return batch_topk_radix<2>;
        case 4: return batch_topk_radix<4>;
        case 8: return batch_topk_radix<8>;
        case 16: return batch_topk_radix<16>;
        default: return nullptr;
    }
}
dense_fn get_batch_topk(int R) {
    switch (R) {
        case 1: return batch_topk<1>;
        case 2: return batch_topk<2>;
        case 4: return batch_topk<4>;
        case 8: return batch_topk<8>;
        case 16: return batch_topk<16>;
        default: return nullptr;
    }
}
