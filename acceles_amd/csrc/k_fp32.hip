// k_fp32.hip -- plain FP32 CSR path and FP64 exact scores:
//   fp32_spmv       FP32 CSR SpMV (cpu_mkl_baseline.cpp:261);
//   select_radix_pass / select_radix_final (K <= 256),
//   select_blocks / select_final (any K <= 960)
//                   Top-K of a score vector (cpu_mkl_baseline.cpp:128-149);
//                   with fp32_spmv it is the plain FP32 CSR Top-K path
//                   (host_spmv_topk_csr_gpu.cu:171-231 is the GPU spec);
//   spmv_fp64       FP64 exact scores for precision/recall reporting.
#include "kernels_common.hpp"

// ---------------------------------------------------------------------------
// FP32 CSR path
// ---------------------------------------------------------------------------

// y = A x with G lanes per row.  Lane g of a row sums its entries
// g, g+G, g+2G, ... in that order with fmaf, then an xor butterfly of plain
// adds.  A step is B blocks of 64/G rows; the steps are dealt to the
// wavefronts in turn, so that the wavefronts in flight read neighbouring
// entries (one contiguous range of rows per wavefront, i.e. thousands of
// separate streams, measured 6% slower on a 20M-row matrix).  The step's rows
// hold one contiguous range of entries:
// for short rows (G <= 8) and when it fits kFp32Stage entries, it is first
// copied into the wavefront's LDS area with 16-byte loads (whole cache lines,
// each byte read once), one step ahead, and the G-lane order then reads LDS;
// otherwise the lanes load the entries directly, the
// first U per lane of all B blocks together (rows longer than U*G entries
// continue block by block).  The next step's bounds are fetched early.  x is
// staged in LDS.  idx and val hold 4 entries of padding past nnz (the last
// 16-byte load may reach 3 entries past the range).
static constexpr int kFp32Stage = 1024;  // entries per wavefront staged in LDS (value and column: 8 KB)

// LDS of an fp32_spmv<G> workgroup: x, then (G <= 8) each wavefront's staging area.
size_t fp32_spmv_lds(int cols, int G) {
    return (((size_t)cols * 4 + 15) & ~(size_t)15) + (G <= 8 ? (size_t)ACC_WAVES * kFp32Stage * 8 : 0);
}

template <int G, int B>
__global__ __launch_bounds__(ACC_WG) void fp32_spmv(const uint32_t *ptr, const int32_t *idx, const float *val,
                                                   const float *x, int rows, int cols, float *y) {
    extern __shared__ __align__(16) unsigned char smem[];
    float *xs = (float *)smem;
    const uint32_t xs_bytes = ((uint32_t)cols * 4u + 15u) & ~15u;
    float *sv = (float *)(smem + xs_bytes) + (threadIdx.x >> 6) * kFp32Stage;  // this wavefront's staged values
    int32_t *sc = (int32_t *)(smem + xs_bytes + ACC_WAVES * kFp32Stage * 4) + (threadIdx.x >> 6) * kFp32Stage;
    for (int c0 = 0; c0 < cols; c0 += 4 * ACC_WG) {  // four loads per thread issued together
        float t[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) t[u] = x[min(c0 + u * ACC_WG + (int)threadIdx.x, cols - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (c0 + u * ACC_WG + (int)threadIdx.x < cols) xs[c0 + u * ACC_WG + threadIdx.x] = t[u];
    }
    __syncthreads();
    constexpr int RPI = 64 / G;  // rows per block
    constexpr int U = 4;
    const int lane = lane_id(), g = lane & (G - 1);
    const int nblocks = (rows + RPI - 1) / RPI;
    const int gw = blockIdx.x * ACC_WAVES + (threadIdx.x >> 6), nw = gridDim.x * ACC_WAVES;
    // steps of B blocks dealt to the wavefronts in turn (wavefront gw takes
    // steps gw, gw + nw, ...): the wavefronts in flight read neighbouring
    // entries, not one stream each
    const int b0 = gw * B, b1 = nblocks, SB = nw * B;
    if (b0 >= b1) return;
    // row bounds are loaded unconditionally (row clamped into the matrix) and
    // masked after: a conditional load would be waited for before the step's
    // entry loads are issued
    uint32_t jb[B], je[B];
#pragma unroll
    for (int t = 0; t < B; ++t) {
        const int r = (b0 + t) * RPI + lane / G;
        const bool ok = b0 + t < b1 && r < rows;
        const int rc = min(r, rows - 1);
        const uint32_t p0 = ptr[rc], p1 = ptr[rc + 1];
        jb[t] = ok ? p0 : 0u;
        je[t] = ok ? p1 : 0u;
    }
    // the entry range of a step: its first row's start to its last row's end
    auto range = [&](int b, uint32_t &s, uint32_t &e) {
        const int r0 = b * RPI, r1 = min(min(b + B, b1) * RPI, rows);
        s = ptr[r0];
        e = ptr[r1];
    };
    // the step's entries in flight in registers: 16-byte loads over [s & ~3, e); rows of up to 8 lanes
    // only (longer rows already read whole cache lines per lane group)
    constexpr bool kStaged = G <= 8;
    constexpr int NL = kFp32Stage / 256;
    typedef float vf4 __attribute__((ext_vector_type(4)));
    typedef int vi4 __attribute__((ext_vector_type(4)));
    vf4 lv[NL];
    vi4 lc[NL];
    auto issue = [&](uint32_t s, uint32_t e) {
        const uint32_t sa = s & ~3u, cnt = e - sa;
        if (!kStaged || cnt > (uint32_t)kFp32Stage) return false;
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const uint32_t o = (uint32_t)(i * 256 + lane * 4);
            if (o < cnt) {
                lv[i] = __builtin_nontemporal_load((const vf4 *)(val + sa + o));
                lc[i] = __builtin_nontemporal_load((const vi4 *)(idx + sa + o));
            }
        }
        return true;
    };
    // ranges: this step's (rs, re), the next one's (ns, ne2); the one after is fetched in the loop
    uint32_t rs, re, ns = 0, ne2 = 0;
    range(b0, rs, re);
    if (b0 + SB < b1) range(b0 + SB, ns, ne2);
    bool have = issue(rs, re);  // this step's entries are in lv/lc
    for (int b = b0; b < b1; b += SB) {
        uint32_t nb[B], ne[B];
#pragma unroll
        for (int t = 0; t < B; ++t) {  // the next step's row bounds
            const int r = (b + SB + t) * RPI + lane / G;
            const bool ok = b + SB + t < b1 && r < rows;
            const int rc = min(r, rows - 1);
            const uint32_t p0 = ptr[rc], p1 = ptr[rc + 1];
            nb[t] = ok ? p0 : 0u;
            ne[t] = ok ? p1 : 0u;
        }
        uint32_t ms = 0, me = 0;
        if (b + 2 * SB < b1) range(b + 2 * SB, ms, me);
        if (have) {
            const uint32_t sa = rs & ~3u, cnt = re - sa;
#pragma unroll
            for (int i = 0; i < NL; ++i) {
                const uint32_t o = (uint32_t)(i * 256 + lane * 4);
                if (o < cnt) {
                    *(vf4 *)(sv + o) = lv[i];
                    *(vi4 *)(sc + o) = lc[i];
                }
            }
            wave_sync();
            // the next step's entries fly while this one is summed from LDS
            const bool next = b + SB < b1 && issue(ns, ne2);
            // the first U entries per lane of all B blocks: their LDS reads issued together (clamped
            // into the staged range, masked after), then longer rows entry by entry
            const uint32_t last = cnt ? cnt - 1u : 0u;  // clamped reads stay in the staged range
            float p[B], a[B][U], xv[B][U];
#pragma unroll
            for (int t = 0; t < B; ++t)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t o = min(jb[t] + g + u * G - sa, last);
                    a[t][u] = sv[o];
                    xv[t][u] = xs[min((uint32_t)sc[o], (uint32_t)cols - 1u)];
                }
#pragma unroll
            for (int t = 0; t < B; ++t) {
                p[t] = 0.f;
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (jb[t] + g + u * G < je[t]) p[t] = __builtin_fmaf(a[t][u], xv[t][u], p[t]);
                for (uint32_t jj = jb[t] + g + U * G; jj < je[t]; jj += G) p[t] = __builtin_fmaf(sv[jj - sa], xs[sc[jj - sa]], p[t]);
            }
#pragma unroll
            for (int t = 0; t < B; ++t) {
#pragma unroll
                for (int s2 = G / 2; s2 >= 1; s2 >>= 1) p[t] = p[t] + __shfl_xor(p[t], s2);
                const int r = (b + t) * RPI + lane / G;
                if (g == 0 && b + t < b1 && r < rows) y[r] = p[t];
            }
            wave_sync();  // the staged entries are read before the next step overwrites them
#pragma unroll
            for (int t = 0; t < B; ++t) {
                jb[t] = nb[t];
                je[t] = ne[t];
            }
            rs = ns;
            re = ne2;
            ns = ms;
            ne2 = me;
            have = next;
            continue;
        }
        float v[B][U];
        int c[B][U];
#pragma unroll
        for (int t = 0; t < B; ++t)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t jj = jb[t] + g + u * G;
                v[t][u] = jj < je[t] ? __builtin_nontemporal_load(val + jj) : 0.f;
                c[t][u] = jj < je[t] ? __builtin_nontemporal_load(idx + jj) : 0;
            }
        float p[B];
#pragma unroll
        for (int t = 0; t < B; ++t) {
            p[t] = 0.f;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (jb[t] + g + u * G < je[t]) p[t] = __builtin_fmaf(v[t][u], xs[c[t][u]], p[t]);
            for (uint32_t j = jb[t] + g + U * G; j < je[t]; j += U * G) {  // longer rows
                float w[U];
                int d[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t jj = j + u * G;
                    w[u] = jj < je[t] ? __builtin_nontemporal_load(val + jj) : 0.f;
                    d[u] = jj < je[t] ? __builtin_nontemporal_load(idx + jj) : 0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (j + u * G < je[t]) p[t] = __builtin_fmaf(w[u], xs[d[u]], p[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < B; ++t) {
#pragma unroll
            for (int s = G / 2; s >= 1; s >>= 1) p[t] = p[t] + __shfl_xor(p[t], s);
            const int r = (b + t) * RPI + lane / G;
            if (g == 0 && b + t < b1 && r < rows) y[r] = p[t];
        }
#pragma unroll
        for (int t = 0; t < B; ++t) {
            jb[t] = nb[t];
            je[t] = ne[t];
        }
        have = b + SB < b1 && issue(ns, ne2);
        rs = ns;
        re = ne2;
        ns = ms;
        ne2 = me;
    }
}


template <int R>
__global__ __launch_bounds__(ACC_WG) void select_blocks(SelectArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_n[ACC_WAVES];
    uint64_t *colbuf = (uint64_t *)smem;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int64_t b0 = (int64_t)blockIdx.x * a.chunk;
    const int64_t b1 = min((int64_t)a.n, b0 + a.chunk);
    const int64_t q = (b1 - b0 + ACC_WAVES - 1) / ACC_WAVES;
    const int64_t w0 = b0 + wave * q, w1 = min(b1, w0 + q);
    WaveTopM<R> col;
    col.init(colbuf + wave * WaveTopM<R>::cap, a.K);
    // U scores per lane are loaded together (with the shared threshold) so one
    // memory latency covers 64*U keys
    constexpr int U = 8;
    for (int64_t i0 = w0; i0 < w1; i0 += 64 * U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + u * 64 + lane;
            v[u] = i < w1 ? __builtin_nontemporal_load(a.y + i) : 0.f;
        }
        const uint64_t g = atomic_load_u64(a.thr);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + u * 64 + lane;
            const bool valid = i < w1;
            const uint64_t key = valid ? dkey_f32(v[u], (int32_t)i) : 0ull;
            const uint64_t t = col.thr > g ? col.thr : g;
            col.push(valid && key > t, key);
        }
    }
    col.compact();
    if (lane == 0) s_n[wave] = col.n;
    __syncthreads();
    if (wave == 0) {
        wg_merge(col, colbuf, s_n, ACC_WAVES);
        publish_and_append(col, a.thr, a.cnt, a.cand);
    }
}

// Second launch (one workgroup): the final K of the appended candidates; the
// state is left zeroed for the next query.
template <int R>
__global__ __launch_bounds__(ACC_WG) void select_final(SelectArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_n[ACC_WAVES];
    uint64_t *colbuf = (uint64_t *)smem;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = *a.cnt;
    WaveTopM<R> col;
    col.init(colbuf + wave * WaveTopM<R>::cap, a.K);
    push_range(col, a.cand, n, (uint32_t)wave * 64, ACC_WG);
    col.compact();
    if (lane == 0) s_n[wave] = col.n;
    __syncthreads();
    if (wave == 0) {
        wg_merge(col, colbuf, s_n, ACC_WAVES);
        write_results(col, a.K, a.out);
        if (lane == 0) {
            *a.cnt = 0u;
            *a.thr = 0ull;
        }
    }
}

// Radix Top-K (K <= 256).  Pass: workgroup b keeps the K best keys of
// elements [b*V*ACC_WG, (b+1)*V*ACC_WG) of the input (FP32 scores keyed as
// (score, row i), or keys of an earlier pass) by radix selection and writes
// them to cand[b*K, (b+1)*K), zero-padded.  Passes repeat until the survivors
// fit one workgroup; then select_radix_final ranks them.
template <bool KEYS, int V>
__global__ __launch_bounds__(ACC_WG) void select_radix_pass(const void *in, int64_t n, int K, uint64_t *cand) {
    __shared__ uint32_t hist[512];
    __shared__ uint64_t s_t[2];
    __shared__ int s_need[2], s_n;
    const int tid = threadIdx.x, lane = tid & 63;
    const int64_t b0 = (int64_t)blockIdx.x * (V * ACC_WG);
    uint64_t k[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
        const int64_t i = b0 + tid + u * ACC_WG;
        if (KEYS) k[u] = i < n ? ((const uint64_t *)in)[i] : 0ull;
        else k[u] = i < n ? dkey_f32(__builtin_nontemporal_load((const float *)in + i), (int32_t)i) : 0ull;
    }
    if (tid == 0) s_n = 0;
    const uint64_t T = wg_radix_kth<uint64_t>(
        [&](auto fn) {
#pragma unroll
            for (int u = 0; u < V; ++u) fn(k[u]);
        },
        K, hist, s_t, s_need);
    uint64_t *dst = cand + (size_t)blockIdx.x * K;
#pragma unroll
    for (int u = 0; u < V; ++u) {
        const bool sel = k[u] != 0ull && k[u] >= T;
        const uint64_t m = __ballot(sel);
        if (m == 0) continue;
        int base = 0;
        if (lane == 0) base = atomicAdd(&s_n, __popcll(m));
        base = __shfl(base, 0);
        const int j = base + mbcnt64(m);
        if (sel && j < K) dst[j] = k[u];  // j < K for distinct keys
    }
    __syncthreads();
    for (int j = min(s_n, K) + tid; j < K; j += ACC_WG) dst[j] = 0ull;
}

template <int V>
__global__ __launch_bounds__(1024) void select_radix_final(const uint64_t *cand, int total, int K, acc_element *out) {
    __shared__ __align__(16) uint32_t lds[1032];
    topk_radix<V>(cand, total, K, out, lds);
}

// Zeroes the select state (a kernel rather than a memset, so the launch
// sequence can be captured in a graph).
__global__ void select_reset(uint64_t *thr, uint32_t *cnt) {
    if (threadIdx.x == 0) {
        *thr = 0ull;
        *cnt = 0u;
    }
}

// FP64 exact scores: 8 lanes per row, FP64 products and sums.

__global__ __launch_bounds__(ACC_WG) void spmv_fp64(const uint32_t *ptr, const int32_t *idx, const float *val,
                                                    const float *x, int rows, double *y) {
    const int lane8 = threadIdx.x & 7;
    for (int r = (blockIdx.x * ACC_WG + threadIdx.x) >> 3; r < ((rows + 31) & ~31); r += (gridDim.x * ACC_WG) >> 3) {
        double p = 0.0;
        if (r < rows)
            for (uint32_t j = ptr[r] + lane8; j < ptr[r + 1]; j += 8) p += (double)val[j] * (double)x[idx[j]];
        for (int s = 4; s >= 1; s >>= 1) p += __shfl_xor(p, s);
        if (r < rows && lane8 == 0) y[r] = p;
    }
}


spmv_fn get_fp32_spmv(int G) {
    switch (G) {
        case 2: return fp32_spmv<2, 4>;
        case 4: return fp32_spmv<4, 4>;
        case 8: return fp32_spmv<8, 4>;
        case 16: return fp32_spmv<16, 2>;
        case 32: return fp32_spmv<32, 2>;
        default: return nullptr;
    }
}
select_fn get_select_blocks(int R) {
    switch (R) {
        case 2: return select_blocks<2>;
        case 4: return select_blocks<4>;
        case 8: return select_blocks<8>;
        case 16: return select_blocks<16>;
        default: return nullptr;
    }
}
select_fn get_select_final(int R) {
    switch (R) {
        case 2: return select_final<2>;
        case 4: return select_final<4>;
        case 8: return select_final<8>;
        case 16: return select_final<16>;
        default: return nullptr;
    }
}
reset_fn get_select_reset() { return select_reset; }
radix_pass_fn get_select_radix_pass(bool keys, int V) {
    if (V == 8) return keys ? select_radix_pass<true, 8> : select_radix_pass<false, 8>;
    if (V == 32) return keys ? select_radix_pass<true, 32> : select_radix_pass<false, 32>;
    return nullptr;
}
radix_final_fn get_select_radix_final(int V) {
    switch (V) {
        case 1: return select_radix_final<1>;
        case 2: return select_radix_final<2>;
        case 4: return select_radix_final<4>;
        case 8: return select_radix_final<8>;
        case 16: return select_radix_final<16>;
        default: return nullptr;
    }
}
fp64_fn get_spmv_fp64() { return spmv_fp64; }
