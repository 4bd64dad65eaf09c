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
// rescans exactly the rare slice whose list may have dropped a qualifying row.
#include "kernels_common.hpp"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// 32-bit slice keys: (score + 2^25) << 6 | row_in_slice; a larger key is a
// higher score (equal scores: any fixed order will do, the lists are only a
// filter and every candidate is re-keyed by (score, row)).  0 = no row.
// |score| < 2^25 is checked on the host before the dense layout is used.
__device__ __forceinline__ int32_t tk_score(uint32_t k) { return (int32_t)(k >> 6) - (1 << 25); }
__device__ __forceinline__ int tk_rho(uint32_t k) { return (int)(k & 63u); }

__device__ __forceinline__ uint32_t umed3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t d;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// keeps k0 >= k1 >= k2 >= k3 the four largest of what was offered: with the
// list ordered, the new i-th entry is the median of the old (i-1)-th, i-th
// and x (one v_med3 each)
__device__ __forceinline__ void top4_insert(uint32_t &k0, uint32_t &k1, uint32_t &k2, uint32_t &k3, uint32_t x) {
    const uint32_t n3 = umed3(k2, k3, x), n2 = umed3(k1, k2, x), n1 = umed3(k0, k1, x);
    k0 = max(k0, x);
    k1 = n1;
    k2 = n2;
    k3 = n3;
}

// Exact INT32 score of this lane's row (row_in_slice = lane) of slice s for
// one query, from the dense tiles with v_dot4_i32_i8: the row's 32 bytes of
// column block kb are the 16-byte registers of lanes i and i+32 of tile
// (kb, rt).  qx: the query's int8 values in LDS, plain order, cols_pad long;
// corr: its unsigned-8-bit offset term.  Integer sums in any order are the
// matrix cores' result.
__device__ __forceinline__ int32_t slice_dot_dense(const DenseArgs &a, int64_t s, const int8_t *qx, int32_t corr) {
    const int lane = lane_id(), rt = lane >> 5, i = lane & 31;
    const size_t kstr = (size_t)a.nslices * 2048;  // bytes per column block
    const int8_t *base = a.tiles + ((size_t)s * 2 + rt) * 1024 + i * 16;
    const int *q4 = (const int *)qx;
    int32_t acc = 0;
    constexpr int U = 8;
    for (int kb0 = 0; kb0 < a.nkb; kb0 += U) {
        int4 lo[U], hi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kb = min(kb0 + u, a.nkb - 1);
            lo[u] = *(const int4 *)(base + kb * kstr);
            hi[u] = *(const int4 *)(base + kb * kstr + 512);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int *qd = q4 + min(kb0 + u, a.nkb - 1) * 8;
            if (kb0 + u >= a.nkb) continue;  // wave-uniform
            acc = __builtin_amdgcn_sdot4(lo[u].x, qd[0], acc, false);
            acc = __builtin_amdgcn_sdot4(lo[u].y, qd[1], acc, false);
            acc = __builtin_amdgcn_sdot4(lo[u].z, qd[2], acc, false);
            acc = __builtin_amdgcn_sdot4(lo[u].w, qd[3], acc, false);
            acc = __builtin_amdgcn_sdot4(hi[u].x, qd[4], acc, false);
            acc = __builtin_amdgcn_sdot4(hi[u].y, qd[5], acc, false);
            acc = __builtin_amdgcn_sdot4(hi[u].z, qd[6], acc, false);
            acc = __builtin_amdgcn_sdot4(hi[u].w, qd[7], acc, false);
        }
    }
    return acc + corr;
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
                // added, not stored: a row that names a column twice keeps the
                // sum (the host checked that every sum fits the byte; the
                // cell starts at 0, or at -128 for the offset unsigned bytes)
                int8_t &cell = l1[(kb * 64 + i + 32 * h) * 16 + bb];
                cell = (int8_t)(cell + sv);
            }
        }
        wave_sync();
        uint4 *dst = (uint4 *)a.tiles + ((size_t)s * 2 + rt) * 64;
        for (int i = lane; i < nkb * 64; i += 64) dst[(size_t)(i >> 6) * a.nslices * 128 + (i & 63)] = l4[i];
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

// Scores of 64-row slices against a block of QB = 32*NT queries.  A
// workgroup serves one query block and the slices wx*8 + wave + k*8*nwx of
// its slice range wx (grid = nwx * nqb, nwx a multiple of 8).
// Per slice a wavefront runs 2*NT MFMA chains over the nkb column blocks, the
// next P blocks of the matrix tiles in flight; the B operands come from the
// query table in LDS.
template <bool STREAM>
__device__ __forceinline__ v4i tile_load(const v4i *p) {
    if (STREAM) return __builtin_nontemporal_load(p);
    return *p;
}

template <int NT, bool STREAM>
__global__ __launch_bounds__(512) void dense_scores(DenseArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int QB = 32 * NT;
    const int nkb = a.nkb, nqb = a.nqb;
    // workgroups are dealt round-robin to the 8 XCDs: the nqb query blocks of
    // one slice range (same wx) get ids b, b+8, b+16, ..., hence the same XCD,
    // and the second and later blocks find the tiles in that XCD's L2
    const int c = blockIdx.x % (8 * nqb);
    const int qb = c / 8;
    const int wx = (blockIdx.x / (8 * nqb)) * 8 + (c % 8), nwx = gridDim.x / nqb;
    v4i *qt = (v4i *)smem;
    int32_t *corr = (int32_t *)(smem + (size_t)nkb * NT * 1024);
    {
        // all loads of a round are issued before any LDS write, so the table
        // arrives in one memory latency per 64 KB instead of one per 8 KB
        const v4i *src = (const v4i *)(a.qtab + (size_t)qb * nkb * NT * 1024);
        const int n = nkb * NT * 64;
        constexpr int CU = 8;
        for (int b0 = 0; b0 < n; b0 += 512 * CU) {
            v4i tmp[CU];
#pragma unroll
            for (int u = 0; u < CU; ++u) {
                const int i = b0 + u * 512 + (int)threadIdx.x;
                tmp[u] = i < n ? src[i] : (v4i){0};
            }
#pragma unroll
            for (int u = 0; u < CU; ++u) {
                const int i = b0 + u * 512 + (int)threadIdx.x;
                if (i < n) qt[i] = tmp[u];
            }
        }
        for (int i = threadIdx.x; i < QB; i += 512) corr[i] = a.corr[qb * QB + i];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = lane_id(), hi = lane >> 5;
    const int64_t stride = (int64_t)nwx * 8;
    int64_t s = (int64_t)wx * 8 + wave;
    // Matrix blocks stream through a ring of RING register slots: the block of
    // step k sits in slot k % RING and is loaded DIST steps ahead, into the
    // slot whose MFMAs were issued RING - DIST steps before.  Loads are
    // unconditional (indices clamped into the matrix; the few past the end are
    // redundant re-reads) so each lands directly in its slot -- a conditional
    // load would be copied into place after waiting for every outstanding load.
    // The next slice's first DIST blocks are requested before the epilogue.
    constexpr int RING = 4, DIST = 2;
    v4i pa[RING], pb[RING];
#pragma unroll
    for (int j = 0; j < RING; ++j) {
        pa[j] = (v4i){0};
        pb[j] = (v4i){0};
    }
    // column block kb of all slices is contiguous: at a given step the
    // wavefronts of the whole chip read neighbouring kilobytes
    const size_t kstr = (size_t)a.nslices * 128;  // v4i per column block
    const v4i *tiles = (const v4i *)a.tiles + lane;
    if (s < a.nslices) {
        const v4i *B0 = tiles + (size_t)s * 128;
#pragma unroll
        for (int j = 0; j < DIST; ++j) {
            const size_t kj = (size_t)min(j, nkb - 1) * kstr;
            pa[j] = tile_load<STREAM>(B0 + kj);
            pb[j] = tile_load<STREAM>(B0 + kj + 64);
        }
    }
    for (; s < a.nslices; s += stride) {
        const v4i *A0 = tiles + (size_t)s * 128;
        const v4i *A1 = A0 + 64;
        v16i acc[2][NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            acc[0][t] = (v16i){0};
            acc[1][t] = (v16i){0};
        }
        for (int k0 = 0; k0 < nkb; k0 += RING) {
#pragma unroll
            for (int j = 0; j < RING; ++j) {
                const int k = k0 + j;
                const size_t kn = (size_t)min(k + DIST, nkb - 1) * kstr;
                pa[(j + DIST) % RING] = tile_load<STREAM>(A0 + kn);
                pb[(j + DIST) % RING] = tile_load<STREAM>(A1 + kn);
                if (k < nkb) {
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const v4i b = qt[(k * NT + t) * 64 + lane];
                        acc[0][t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(pa[j], b, acc[0][t], 0, 0, 0);
                        acc[1][t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(pb[j], b, acc[1][t], 0, 0, 0);
                    }
                }
            }
        }
        {
            const int64_t sn = min(s + stride, a.nslices - 1);  // past the end: a redundant in-bounds load
            const v4i *B0 = tiles + (size_t)sn * 128;
#pragma unroll
            for (int j = 0; j < DIST; ++j) {
                const size_t kj = (size_t)min(j, nkb - 1) * kstr;
                pa[j] = tile_load<STREAM>(B0 + kj);
                pb[j] = tile_load<STREAM>(B0 + kj + 64);
            }
        }
        // D[i][j] of a 32x32 tile: lane j + 32*((i/4)%2), register (i%4) + 4*(i/8);
        // row_in_slice = rt*32 + i = E(rt, r) + 4*hi.  Padding rows (the tail of a
        // partition's last slice) hold no entries and score exactly 0 in every
        // value mode; the selection discards their keys.
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int ql = t * 32 + (lane & 31);
            // key = ((score + corr + 2^25) << 6) + row_in_slice, in 32-bit wrapping
            // arithmetic (the true value is in [64, 2^32))
            const uint32_t D = ((uint32_t)(corr[ql] + (1 << 25)) << 6) + (uint32_t)(hi * 4);
            uint32_t k0 = 0u, k1 = 0u, k2 = 0u, k3 = 0u;
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const uint32_t E = (uint32_t)(rt * 32 + (r >> 2) * 8 + (r & 3));
                    top4_insert(k0, k1, k2, k3, ((uint32_t)acc[rt][t][r] << 6) + D + E);
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
// slice whose list may have dropped such a row is rescanned exactly (from the
// dense tiles).  Keys of padding rows are discarded, and since they may have
// displaced real rows from a list, every list whose fourth key scores >= ts
// is rescanned.
template <int U>
__device__ __forceinline__ void offer_lists(const DenseArgs &a, WaveTopM<2> &cb, const uint4 (&k4)[U], const int (&nv)[U],
                                   const int32_t (&row)[U][4], int64_t s0, int b, int32_t ts, int q, int8_t *qx,
                                   bool &have_qx) {
    const int lane = lane_id();
    bool rescan[U], pass[U * 4];
    uint64_t key[U * 4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        rescan[u] = nv[u] > 0 && tk_score(k4[u].w) >= ts;
        const uint32_t ke[4] = {k4[u].x, k4[u].y, k4[u].z, k4[u].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            pass[u * 4 + e] = !rescan[u] && row[u][e] >= 0 && tk_score(ke[e]) >= ts;
            key[u * 4 + e] = dkey_i32(tk_score(ke[e]), row[u][e]);
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
            const int32_t sc = slice_dot_dense(a, s, qx, a.corr[q]);
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
// ranks them.  Anything else streams the lists twice through a collector.
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
                const int32_t sc = slice_dot_dense(a, s, qx, a.corr[q]);
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

    // General path: the lists are streamed twice (threshold, then offers).
    if (a.out) stage_query(xs, xq, a.cols);
    WaveTopM<2> cb;
    cb.init(bufB, M);
    constexpr int U = 4;
    int32_t ts = INT32_MIN;
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
#pragma unroll
            for (int u = 0; u < U; ++u) ca.push(k[u] != 0u, ((uint64_t)k[u] << 32) | (uint32_t)(b + u * 64 + lane));
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
        int32_t row[U][4];
        list_rows<U>(a, k4, nv, s0, b, row);
        offer_lists<U>(a, cb, k4, nv, row, s0, b, ts, q, qx, have_qx);
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

dense_fn get_dense_build() { return dense_build; }
dense_fn get_dense_prep() { return dense_prep; }
// STREAM: non-temporal tile loads, for matrices larger than the last-level
// (MALL) cache; smaller ones are read with normal loads and stay resident
// between batches.
dense_fn get_dense_scores(int NT, bool stream) {
    switch (NT) {
        case 1: return stream ? dense_scores<1, true> : dense_scores<1, false>;
        case 2: return stream ? dense_scores<2, true> : dense_scores<2, false>;
        case 4: return stream ? dense_scores<4, true> : dense_scores<4, false>;
        default: return nullptr;
    }
}
dense_fn get_dense_select() { return dense_select; }
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
