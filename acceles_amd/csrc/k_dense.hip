// k_dense.hip -- batched low-bit path on dense int8 tiles and matrix cores.
//
// The INT32 stage-1 scores (quantify_util.py:173-196: exact integer dot of
// the quantized rows with the quantized query) are computed for a block of
// queries per matrix pass with v_mfma_i32_32x32x32_i8: at the densities of
// the embedding workloads (2-15% non-zeros) a dense int8 tile on the matrix
// cores costs less than decoding (column, value) packets on the vector ALUs
// once a pass serves tens of queries.  Integer arithmetic is exact, so the
// scores are the same integers as the packet kernels'.  Only the four best
// (score, row) keys of each (query, 64-row slice) leave the kernel; the
// per-partition selection derives its threshold from the slice maxima and
// rescans (exactly, from the packets) the rare slice whose list may have
// dropped a qualifying row.
#include "kernels_common.hpp"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// 32-bit slice keys: (score + 2^25) << 6 | (63 - row_in_slice); a larger key
// is a higher score, then a lower row.  0 = no row.  |score| < 2^25 is
// checked on the host before the dense layout is used.
__device__ __forceinline__ uint32_t tk_key(int32_t score, int rho) {
    return ((uint32_t)(score + (1 << 25)) << 6) | (uint32_t)(63 - rho);
}
__device__ __forceinline__ int32_t tk_score(uint32_t k) { return (int32_t)(k >> 6) - (1 << 25); }
__device__ __forceinline__ int tk_rho(uint32_t k) { return 63 - (int)(k & 63u); }

// keeps k0 >= k1 >= k2 >= k3 the four largest of what was offered
__device__ __forceinline__ void top4_insert(uint32_t &k0, uint32_t &k1, uint32_t &k2, uint32_t &k3, uint32_t x) {
    uint32_t h = max(k0, x);
    x = min(k0, x);
    k0 = h;
    h = max(k1, x);
    x = min(k1, x);
    k1 = h;
    h = max(k2, x);
    x = min(k2, x);
    k2 = h;
    k3 = max(k3, x);
}

// Exact INT32 score of this lane's row of slice s from the sliced packets
// (the entry width is a runtime value: this runs only for rescans).
__device__ inline int32_t slice_dot_rt(const DenseArgs &a, int64_t s, const int8_t *qx) {
    const int lane = lane_id();
    const uint2 sl = a.slices[s];
    const int E = a.col_bits + a.val_bits;
    const uint32_t cmask = (1u << a.col_bits) - 1u, vmask = (1u << a.val_bits) - 1u;
    int32_t acc = 0;
    for (int j = 0;; ++j) {
        const int pos = j * E, w = pos >> 5, sh = pos & 31;
        if (w >= (int)sl.y) break;
        const uint32_t lo = a.words[((size_t)sl.x + w) * 64 + lane];
        const uint32_t hi = (w + 1 < (int)sl.y) ? a.words[((size_t)sl.x + w + 1) * 64 + lane] : 0u;
        const uint32_t bits = (uint32_t)((((uint64_t)hi << 32) | lo) >> sh);
        const uint32_t v = (bits >> a.col_bits) & vmask;
        if (v == 0u) break;  // padding: entries never hold a zero value
        int sv = (int)v;
        if (a.is_signed && (v >> (a.val_bits - 1))) sv = (int)(v | ~vmask);
        acc += sv * (int)qx[bits & cmask];
    }
    return acc;
}

// One wavefront per slice: decode each lane's packet stream into a 32-row
// tile image in LDS (nkb KB), then copy it out.  Runs once per matrix.
__global__ __launch_bounds__(64) void dense_build(DenseArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = lane_id();
    const int64_t s = blockIdx.x;
    const uint2 sl = a.slices[s];
    const int E = a.col_bits + a.val_bits;
    const uint32_t cmask = (1u << a.col_bits) - 1u, vmask = (1u << a.val_bits) - 1u;
    const uint32_t fill = a.u8off ? 0x80808080u : 0u;
    const int nkb = a.nkb;
    uint4 *l4 = (uint4 *)smem;
    int8_t *l1 = (int8_t *)smem;
    for (int rt = 0; rt < 2; ++rt) {
        for (int i = lane; i < nkb * 64; i += 64) l4[i] = make_uint4(fill, fill, fill, fill);
        wave_sync();
        if ((lane >> 5) == rt) {
            const int i = lane & 31;
            for (int j = 0;; ++j) {
                const int pos = j * E, w = pos >> 5, sh = pos & 31;
                if (w >= (int)sl.y) break;
                const uint32_t lo = a.words[((size_t)sl.x + w) * 64 + lane];
                const uint32_t hi = (w + 1 < (int)sl.y) ? a.words[((size_t)sl.x + w + 1) * 64 + lane] : 0u;
                const uint32_t bits = (uint32_t)((((uint64_t)hi << 32) | lo) >> sh);
                const uint32_t v = (bits >> a.col_bits) & vmask;
                if (v == 0u) break;
                const uint32_t c = bits & cmask;
                int sv = (int)v;
                if (a.is_signed && (v >> (a.val_bits - 1))) sv = (int)(v | ~vmask);
                const int kb = (int)(c >> 5), h = (int)((c >> 4) & 1u), bb = (int)(c & 15u);
                l1[(kb * 64 + i + 32 * h) * 16 + bb] = (int8_t)(a.u8off ? sv - 128 : sv);
            }
        }
        wave_sync();
        uint4 *dst = (uint4 *)(a.tiles + ((size_t)s * 2 + rt) * nkb * 1024);
        for (int i = lane; i < nkb * 64; i += 64) dst[i] = l4[i];
        wave_sync();
    }
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
    float m = 0.f;
    if (valid)
        for (int c = threadIdx.x; c < a.cols; c += ACC_WG) m = fmaxf(m, fabsf(x[c]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (lane == 0) red[wave] = m;
    __syncthreads();
    float s = red[0];
    for (int w = 1; w < ACC_WAVES; ++w) s = fmaxf(s, red[w]);
    int sum = 0;
    for (int c = threadIdx.x; c < a.cols_pad; c += ACC_WG) {
        int qv = 0;
        if (valid && c < a.cols && s > 0.f) {
            const float u = __fmul_rn(127.0f, __fdiv_rn(x[c], s));
            qv = (int)__builtin_rintf(u);
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

// Scores of 64-row slices against a block of QB = 32*NT queries.  Workgroup
// w serves query block w % nqb (so the blocks of one slice range run side by
// side and share its tiles in L2) and slices (w / nqb)*8 + wave + k*8*(grid/nqb).
// Per slice a wavefront runs 2*NT MFMA chains over the nkb column blocks, the
// next P blocks of the matrix tiles in flight; the B operands come from the
// query table in LDS.
template <int NT>
__global__ __launch_bounds__(512) void dense_scores(DenseArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int QB = 32 * NT;
    constexpr int P = 4;
    const int nkb = a.nkb, nqb = a.nqb;
    const int qb = blockIdx.x % nqb;
    const int wx = blockIdx.x / nqb, nwx = gridDim.x / nqb;
    v4i *qt = (v4i *)smem;
    int32_t *corr = (int32_t *)(smem + (size_t)nkb * NT * 1024);
    {
        const v4i *src = (const v4i *)(a.qtab + (size_t)qb * nkb * NT * 1024);
        for (int i = threadIdx.x; i < nkb * NT * 64; i += 512) qt[i] = src[i];
        for (int i = threadIdx.x; i < QB; i += 512) corr[i] = a.corr[qb * QB + i];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = lane_id(), hi = lane >> 5;
    for (int64_t s = (int64_t)wx * 8 + wave; s < a.nslices; s += (int64_t)nwx * 8) {
        const v4i *A0 = (const v4i *)(a.tiles + (size_t)s * 2 * nkb * 1024) + lane;
        const v4i *A1 = A0 + (size_t)nkb * 64;
        v16i acc[2][NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            acc[0][t] = (v16i){0};
            acc[1][t] = (v16i){0};
        }
        v4i pa[P], pb[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            pa[p] = (v4i){0};
            pb[p] = (v4i){0};
            if (p < nkb) {
                pa[p] = __builtin_nontemporal_load(A0 + p * 64);
                pb[p] = __builtin_nontemporal_load(A1 + p * 64);
            }
        }
        for (int k0 = 0; k0 < nkb; k0 += P) {
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const int k = k0 + p;
                if (k < nkb) {
                    const v4i x0 = pa[p], x1 = pb[p];
                    if (k + P < nkb) {
                        pa[p] = __builtin_nontemporal_load(A0 + (k + P) * 64);
                        pb[p] = __builtin_nontemporal_load(A1 + (k + P) * 64);
                    }
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const v4i b = qt[(k * NT + t) * 64 + lane];
                        acc[0][t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(x0, b, acc[0][t], 0, 0, 0);
                        acc[1][t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(x1, b, acc[1][t], 0, 0, 0);
                    }
                }
            }
        }
        // D[i][j] of a 32x32 tile: lane j + 32*((i/4)%2), register (i%4) + 4*(i/8)
        const int nv = a.nvalid[s];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int ql = t * 32 + (lane & 31);
            const int32_t cr = corr[ql];
            uint32_t k0 = 0u, k1 = 0u, k2 = 0u, k3 = 0u;
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int rho = rt * 32 + (r >> 2) * 8 + hi * 4 + (r & 3);
                    const uint32_t key = rho < nv ? tk_key(acc[rt][t][r] + cr, rho) : 0u;
                    top4_insert(k0, k1, k2, k3, key);
                }
            const uint32_t o0 = (uint32_t)__shfl_xor((int)k0, 32), o1 = (uint32_t)__shfl_xor((int)k1, 32);
            const uint32_t o2 = (uint32_t)__shfl_xor((int)k2, 32), o3 = (uint32_t)__shfl_xor((int)k3, 32);
            top4_insert(k0, k1, k2, k3, o0);
            top4_insert(k0, k1, k2, k3, o1);
            top4_insert(k0, k1, k2, k3, o2);
            top4_insert(k0, k1, k2, k3, o3);
            const int q = qb * QB + ql;
            if (hi == 0 && q < a.nq) *(uint4 *)(a.tk + ((size_t)q * a.nslices + s) * 4) = make_uint4(k0, k1, k2, k3);
        }
    }
}

// One wavefront per (partition p = blockIdx.x, query q = blockIdx.y*4 + wave):
//  1. ts = the M-th best slice maximum of the partition (INT32_MIN when the
//     partition has at most M slices).  M distinct rows score >= ts, so the
//     partition's M-th best score is >= ts and no row below ts can be in its
//     Top-M.
//  2. Every listed key scoring >= ts is a candidate.  A list with four keys
//     whose fourth still scores >= ts (and a slice of more than four rows)
//     may have dropped a qualifying row: that slice is rescanned exactly from
//     the packets instead.
//  3. Top-M of the candidates by (score, row) (quantify_util.py:173-196 with
//     the lower-row tie rule), written as keys for rescoring.
__global__ __launch_bounds__(ACC_WG) void dense_select(DenseArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int p = blockIdx.x, q = blockIdx.y * ACC_WAVES + wave;
    if (q >= a.nq) return;  // whole wavefront; no workgroup barrier below
    uint64_t *bufA = (uint64_t *)smem + wave * 256, *bufB = bufA + 128;
    int8_t *qx = (int8_t *)(smem + ACC_WAVES * 256 * 8) + (size_t)wave * a.cols_pad;
    const int64_t s0 = a.part_slice[p];
    const int nsl = (int)(a.part_slice[p + 1] - s0);
    const uint4 *L = (const uint4 *)a.tk + (size_t)q * a.nslices + s0;
    const int M = a.M;
    constexpr int U = 4;

    int32_t ts = INT32_MIN;
    if (nsl > M) {
        WaveTopM<2> ca;
        ca.init(bufA, M);
        for (int b = 0; b < nsl; b += 64 * U) {
            uint32_t k[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = b + u * 64 + lane;
                k[u] = i < nsl ? L[i].x : 0u;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) ca.push(k[u] != 0u, ((uint64_t)k[u] << 32) | (uint32_t)(b + u * 64 + lane));
        }
        ca.compact();
        if (ca.n >= M) ts = tk_score((uint32_t)(ca.buf[M - 1] >> 32));
    }

    WaveTopM<2> cb;
    cb.init(bufB, M);
    bool have_qx = false;
    for (int b = 0; b < nsl; b += 64 * U) {
        uint4 k4[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = b + u * 64 + lane;
            k4[u] = i < nsl ? L[i] : make_uint4(0u, 0u, 0u, 0u);
        }
        bool rescan[U];
        int32_t row[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t s = s0 + b + u * 64 + lane;
            rescan[u] = k4[u].w != 0u && tk_score(k4[u].w) >= ts && a.nvalid[s] > 4;
            const uint32_t ke[4] = {k4[u].x, k4[u].y, k4[u].z, k4[u].w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const bool pass = !rescan[u] && ke[e] != 0u && tk_score(ke[e]) >= ts;
                row[u][e] = pass ? a.perm[s * 64 + tk_rho(ke[e])] : -1;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t ke[4] = {k4[u].x, k4[u].y, k4[u].z, k4[u].w};
#pragma unroll
            for (int e = 0; e < 4; ++e) cb.push(row[u][e] >= 0, dkey_i32(tk_score(ke[e]), row[u][e]));
        }
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
                const int32_t sc = slice_dot_rt(a, s, qx);
                const int32_t r = a.perm[s * 64 + lane];
                cb.push(r >= 0 && sc >= ts, dkey_i32(sc, r));
            }
        }
    }
    cb.compact();
    const size_t base = ((size_t)q * a.split + p) * M;
    for (int i = lane; i < M; i += 64) {
        const uint64_t key = i < cb.n ? cb.buf[i] : 0ull;
        a.ckeys[base + i] = key;
        if (a.cand_rows_out) {
            a.cand_rows_out[base + i] = key ? dkey_row(key) : -1;
            a.cand_scores_out[base + i] = key ? dkey_score_i32(key) : 0;
        }
    }
}

// FP32 rescoring of every candidate (quantify_util.py:354-367): G lanes per
// candidate in the declared strided-fmaf + butterfly order.
__global__ __launch_bounds__(ACC_WG) void batch_rescore(DenseArgs a) {
    const int G = a.G;
    const int64_t per_q = (int64_t)a.split * a.M;
    const int64_t n = (int64_t)a.nq * per_q;
    const int64_t cid = ((int64_t)blockIdx.x * ACC_WG + threadIdx.x) / G;
    const bool in = cid < n;
    const uint64_t key = in ? a.ckeys[cid] : 0ull;
    const bool valid = key != 0ull;
    const int32_t r = valid ? dkey_row(key) : 0;
    const int q = in ? (int)(cid / per_q) : 0;
    const float sc = group_dot_rt(a.ptr, a.idx, a.val, a.X + (size_t)q * a.cols, r, valid, G);
    if (in && (threadIdx.x & (G - 1)) == 0) a.resc[cid] = valid ? dkey_f32(sc, r) : 0ull;
}

// One wavefront per query: register bitonic sort of its split*M <= 64*R
// rescored keys, the first K[q] written as ranked results.
template <int R>
__global__ __launch_bounds__(ACC_WG) void batch_topk(DenseArgs a) {
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int q = blockIdx.x * ACC_WAVES + wave;
    if (q >= a.nq) return;
    const int total = a.split * a.M;
    const uint64_t *keys = a.resc + (size_t)q * total;
    uint64_t v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = r * 64 + lane;
        v[r] = i < total ? keys[i] : 0ull;
    }
    reg_sort_desc<R>(v);
    const int K = a.K[q];
    acc_element *out = a.out + (size_t)q * a.out_stride;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = r * 64 + lane;
        if (i < K) {
            acc_element e;
            e.value = v[r] ? dkey_score_f32(v[r]) : -__builtin_inff();
            e.index = v[r] ? dkey_row(v[r]) : -1;
            out[i] = e;
        }
    }
}

dense_fn get_dense_build() { return dense_build; }
dense_fn get_dense_prep() { return dense_prep; }
dense_fn get_dense_scores(int NT) {
    switch (NT) {
        case 1: return dense_scores<1>;
        case 2: return dense_scores<2>;
        case 4: return dense_scores<4>;
        default: return nullptr;
    }
}
dense_fn get_dense_select() { return dense_select; }
dense_fn get_batch_rescore() { return batch_rescore; }
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
