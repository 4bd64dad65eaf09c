// kernels_common.hpp -- device helpers, argument blocks and kernel getters
// shared by the HIP translation units of the Top-K SpMV path.
//
// A low-bit query is three launches on one stream:
//   lowbit_slicemax query quantization (quantify_util.py:31-49, 259) and the
//                   INT32 SpMV over sliced packets (quantify_util.py:5-6;
//                   spmv_ucsr_top_k_multicore.hpp:195-355 on the FPGA),
//                   keeping the maximum of each 64-row slice;
//   lowbit_select   per partition: exact threshold from the slice maxima,
//                   rescan of the slices that reach it, its Top-M
//                   (quantify_util.py:173-196, host_spmv_ucsr.cpp:691-748)
//                   and the FP32 rescoring of those rows
//                   (quantify_util.py:354-367);
//   lowbit_final    the global Top-K of the rescored candidates.
// Cross-workgroup data moves only through kernel boundaries: a device-scope
// fence on a multi-XCD part writes back and invalidates the whole L2, which
// costs microseconds per workgroup.
//   lowbit_scores   INT32 scores of every row (parity checks).
//   fp32_spmv       plain FP32 CSR SpMV (cpu_mkl_baseline.cpp:261).
//   select_blocks / select_final
//                   Top-K of a score vector (cpu_mkl_baseline.cpp:128-149);
//                   with fp32_spmv it is the plain FP32 CSR Top-K path
//                   (host_spmv_topk_csr_gpu.cu:171-231 is the GPU spec).
//   spmv_fp64       FP64 exact scores for precision/recall reporting.
#pragma once

#include "device_common.hpp"
#include "../../include/acceles.h"

#define LB_WAVES 8
#define LB_WG (LB_WAVES * 64)
#define LB_MAX_ITEM 1024  // slices per stage-1 work item (their metadata sits in LDS)
#define SEL_WAVES 16
#define SEL_WG (SEL_WAVES * 64)

// Quantize the FP32 query into LDS: q = clip(rint(127 * (x / max|x|)), -128, 127)
// (quantify_vector with B=8, symmetric; float32 division and product, half-even).
// The first 4*blockDim.x values are read once, with loads issued together
// (clamped indices, masked after), and kept in registers for both passes.
// xs (optional): the FP32 query is also copied there (LDS, cols floats).
__device__ inline void quantize_query_lds(const float *x, int cols, int8_t *qx, float *red, float *xs = nullptr) {
    constexpr int XR = 4;
    const int nt = (int)blockDim.x, tid = (int)threadIdx.x;
    float xr[XR];
#pragma unroll
    for (int u = 0; u < XR; ++u) xr[u] = x[min(tid + u * nt, cols - 1)];
    float m = 0.f;
#pragma unroll
    for (int u = 0; u < XR; ++u)
        if (tid + u * nt < cols) m = fmaxf(m, fabsf(xr[u]));
    for (int c = tid + XR * nt; c < cols; c += nt) m = fmaxf(m, fabsf(x[c]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (lane_id() == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    float s = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s = fmaxf(s, red[w]);
    for (int c = tid, it = 0; c < cols; c += nt, ++it) {
        float xc = 0.f;
#pragma unroll
        for (int u = 0; u < XR; ++u)
            if (it == u) xc = xr[u];
        if (it >= XR) xc = x[c];
        if (xs) xs[c] = xc;
        int q = 0;
        if (s > 0.f) {
            float t = __fdiv_rn(xc, s);
            float y = __fmul_rn(127.0f, t);
            q = (int)__builtin_rintf(y);
            q = q > 127 ? 127 : (q < -128 ? -128 : q);
        }
        qx[c] = (int8_t)q;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// sliced-packet decoding
// ---------------------------------------------------------------------------

// Buffer resource over one slice's packets (64 lanes x ndw dwords, lane-
// interleaved): a load past the slice's end returns 0 without touching
// memory, so group loads can be unconditional and still fetch exactly the
// slice's bytes.  off and ndw are wave-uniform.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slice_rsrc(const uint32_t *words, uint2 sl) {
    const uint32_t off = __builtin_amdgcn_readfirstlane(sl.x), ndw = __builtin_amdgcn_readfirstlane(sl.y);
    return __builtin_amdgcn_make_buffer_rsrc((void *)(words + (size_t)off * 64), (short)0, (int)(ndw * 256u),
                                             0x00020000 /* 32-bit data format */);
}

// Loads group g0 / E (E dwords per lane = up to 32 entries) of a slice,
// non-temporal; dwords past the slice's end are 0.
template <int E>
__device__ __forceinline__ void load_group(uint32_t (&w)[E], __amdgpu_buffer_rsrc_t rs, int g0) {
    const int lane = lane_id();
#pragma unroll
    for (int i = 0; i < E; ++i) w[i] = __builtin_amdgcn_raw_buffer_load_b32(rs, ((g0 + i) * 64 + lane) * 4, 0, 2 /* nt */);
}

// Accumulates value * q[col] over the entries of one group.  Padding entries
// are all-zero bits (value 0), so only whole blocks of 8 padding entries are
// skipped.
template <int E, bool SIGNED>
__device__ __forceinline__ int32_t decode_group(const uint32_t (&w)[E], int nd, const int8_t *qx, int cb, int vb,
                                                uint32_t cmask, int32_t acc) {
#pragma unroll
    for (int blk = 0; blk < 4; ++blk) {
        if (((blk * 8 * E) >> 5) >= nd) continue;  // whole block is padding (wave-uniform)
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            const int pos = (blk * 8 + jj) * E;
            const int wi = pos >> 5;
            const int sh = pos & 31;
            const uint32_t hi = (wi + 1 < E) ? w[wi + 1] : 0u;
            const uint32_t bits = (sh + E <= 32) ? (w[wi] >> sh) : __builtin_amdgcn_alignbit(hi, w[wi], sh);
            const int col = (int)(bits & cmask);
            const int val = SIGNED ? __builtin_amdgcn_sbfe(bits, cb, vb) : (int)__builtin_amdgcn_ubfe(bits, cb, vb);
            acc += __mul24(val, (int)qx[col]);  // |val|, |q| < 2^8: v_mul_i32_i24 (a full-rate op; a 32-bit multiply is quarter rate)
        }
    }
    return acc;
}

// After each wave compacted its collector (list at colbuf + w*cap, count in
// s_n[w]), wave 0 folds the other waves' lists into its own.
template <int R>
__device__ inline void wg_merge(WaveTopM<R> &col, const uint64_t *colbuf, const int *s_n, int waves) {
    const int lane = lane_id();
    for (int w = 1; w < waves; ++w) {
        const uint64_t *src = colbuf + w * WaveTopM<R>::cap;
        const int n = s_n[w];
        for (int b = 0; b < n; b += 64) {
            const int i = b + lane;
            const bool ok = i < n;
            col.push(ok, ok ? src[i] : 0ull);
        }
    }
    col.compact();
}

// Offers keys[i] for i = first + lane, first + stride + lane, ... (< n) to the
// collector; U loads per lane are issued together so one memory latency
// covers 64*U keys.  Zero keys (empty slots) are skipped.  stride is a
// multiple of 64 (the workgroup size when each wave starts at wave*64).
template <int R, int U = 8>
__device__ inline void push_range(WaveTopM<R> &col, const uint64_t *keys, uint32_t n, uint32_t first,
                                  uint32_t stride) {
    const uint32_t lane = (uint32_t)lane_id();
    for (uint32_t b = first; b < n; b += stride * U) {
        uint64_t k[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = b + u * stride + lane;
            k[u] = i < n ? keys[i] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) col.push(k[u] != 0ull, k[u]);
    }
}

// Wave 0: publish the list's M-th key as a lower bound, then append the keys
// that can still be in the final Top-M to the shared candidate list.
template <int R>
__device__ inline void publish_and_append(WaveTopM<R> &col, uint64_t *thr, uint32_t *cnt, uint64_t *dst) {
    const int lane = lane_id();
    if (col.n >= col.M && col.M > 0 && lane == 0) atomicMax((unsigned long long *)thr, (unsigned long long)col.buf[col.M - 1]);
    const uint64_t g = atomic_load_u64(thr);
    int keep = 0;
    for (int i = lane; i < col.n; i += 64) keep += col.buf[i] >= g;
    for (int o = 32; o > 0; o >>= 1) keep += __shfl_xor(keep, o);
    if (keep == 0) return;
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(cnt, (uint32_t)keep);
    pos = __shfl(pos, 0);
    for (int i = lane; i < keep; i += 64) dst[pos + i] = col.buf[i];
}

// Dot product of CSR row r with x by a group of G lanes (runtime G): lane g
// sums entries g, g+G, ... with fmaf, then an xor butterfly of plain adds.
// Entries are fetched U at a time (then their x values), so a row costs one
// memory round trip per U*G entries; the fmaf order does not depend on U.
template <int U = 4>
__device__ __forceinline__ float group_dot_rt(const uint32_t *ptr, const int32_t *idx, const float *val, const float *x,
                                              int r, bool valid, int G) {
    const int g = lane_id() & (G - 1);
    float p = 0.f;
    if (valid) {
        const uint32_t e = ptr[r + 1];
        for (uint32_t j = ptr[r] + g; j < e; j += U * G) {
            float v[U], xv[U];
            int32_t c[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // clamped into the row (j < e): unconditional loads issue
                // together, where conditional ones would each be waited for
                const uint32_t jj = min(j + u * G, e - 1);
                v[u] = val[jj];
                c[u] = idx[jj];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + u * G < e) p = __builtin_fmaf(v[u], xv[u], p);
        }
    }
    for (int s = G >> 1; s >= 1; s >>= 1) p = p + __shfl_xor(p, s);
    return p;
}

// The same dot products for C rows at once, given their entry ranges
// [lo[c], hi[c]) (lo = hi for an empty or invalid row), against x: their
// loads are issued together, so C rows cost the memory round trips of one.
// Each row is summed in the order of group_dot_rt (entries g, g+G, ... with
// fmaf, then the butterfly), so the results are the same floats.
template <int C, int U = 4>
__device__ __forceinline__ void group_dot_ranges_multi(const int32_t *idx, const float *val, const float *x,
                                                       const uint32_t (&lo)[C], const uint32_t (&hi)[C], int G,
                                                       float (&out)[C]) {
    const int g = lane_id() & (G - 1);
    uint32_t b[C], e[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        b[c] = lo[c] + g;
        e[c] = hi[c];
        out[c] = 0.f;
    }
    for (uint32_t o = 0;; o += U * G) {
        bool any = false;
#pragma unroll
        for (int c = 0; c < C; ++c) any = any || b[c] + o < e[c];
        if (!any) break;
        float v[C][U], xv[C][U];
        int32_t cl[C][U];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            // clamped into the matrix (some row has entries while the loop
            // runs): unconditional loads issue together
            const uint32_t last = e[c] > 0u ? e[c] - 1u : 0u;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t jj = min(b[c] + o + u * G, last);
                v[c][u] = val[jj];
                cl[c][u] = idx[jj];
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int u = 0; u < U; ++u) xv[c][u] = x[cl[c][u]];
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (b[c] + o + u * G < e[c]) out[c] = __builtin_fmaf(v[c][u], xv[c][u], out[c]);
    }
#pragma unroll
    for (int c = 0; c < C; ++c)
        for (int s = G >> 1; s >= 1; s >>= 1) out[c] = out[c] + __shfl_xor(out[c], s);
}

// group_dot_ranges_multi for rows r[c] (r[c] is a row of the matrix also for
// invalid candidates, so the bound loads are unconditional and issue together).
template <int C, int U = 4>
__device__ __forceinline__ void group_dot_rt_multi(const uint32_t *ptr, const int32_t *idx, const float *val,
                                                   const float *x, const int32_t (&r)[C],
                                                   const bool (&valid)[C], int G, float (&out)[C]) {
    uint32_t lo[C], hi[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const uint32_t p0 = ptr[r[c]], p1 = ptr[r[c] + 1];
        lo[c] = valid[c] ? p0 : 0u;
        hi[c] = valid[c] ? p1 : 0u;
    }
    group_dot_ranges_multi<C, U>(idx, val, x, lo, hi, G, out);
}


// One workgroup of ACC_WG threads: the K best of total <= ACC_WAVES*64*R
// ranking keys (0 = empty) into out[0, K), then (value -inf, index -1) past
// the nonzero keys.  Each wavefront sorts its 64*R keys in registers and
// stores them in LDS (lists: ACC_WAVES*64*R keys); every key finds its rank
// in the union by binary search in the other lists (keys are distinct: one
// per row).
template <int R>
__device__ inline void topk_by_rank(const uint64_t *keys, int total, int K, acc_element *out, uint64_t *lists,
                                    int *s_nz) {
    constexpr int N = 64 * R;  // keys per wavefront list
    const int wave = threadIdx.x >> 6, lane = lane_id();
    uint64_t v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = wave * N + r * 64 + lane;
        v[r] = i < total ? keys[i] : 0ull;
    }
    reg_sort_desc<R>(v);
    int nz = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        lists[wave * N + r * 64 + lane] = v[r];
        nz += v[r] != 0ull;
    }
    for (int o = 32; o > 0; o >>= 1) nz += __shfl_xor(nz, o);
    if (lane == 0) s_nz[wave] = nz;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint64_t key = v[r];
        if (key == 0ull) continue;
        int rank = r * 64 + lane;  // its place in its own (sorted) list
#pragma unroll
        for (int w = 0; w < ACC_WAVES; ++w) {
            if (w == wave) continue;
            const uint64_t *L = lists + w * N;
            int lo = 0, hi = N;  // number of keys of L above key
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (L[mid] > key) lo = mid + 1;
                else hi = mid;
            }
            rank += lo;
        }
        if (rank < K) {
            acc_element e;
            e.value = dkey_score_f32(key);
            e.index = dkey_row(key);
            out[rank] = e;
        }
    }
    int n = 0;
    for (int w = 0; w < ACC_WAVES; ++w) n += s_nz[w];
    for (int i = n + (int)threadIdx.x; i < K; i += ACC_WG) {
        acc_element e;
        e.value = -__builtin_inff();
        e.index = -1;
        out[i] = e;
    }
}

// ---------------------------------------------------------------------------
// workgroup radix selection
// ---------------------------------------------------------------------------

// The need-th largest of the nonzero keys the workgroup's threads offer
// (each thread calls fn(key) for its keys through for_each(fn); 0 = no key).
// First the keys' range: the bits above the highest bit in which the
// smallest and largest key differ are common to all keys.  Then digit by
// digit (8 bits) below them: an LDS histogram of the keys that share the
// prefix chosen so far, and wavefront 0 picks the bin that holds the need-th
// key.  Starting below the common bits spreads the first histogram over the
// keys' actual range (bins hit by many lanes at once serialize).
// Returns T with at least need keys >= T; when a bin is taken whole it
// returns its lower end, so exactly need keys are >= T when the keys are
// distinct.  Returns 0 when fewer than need keys are offered.
// hist: 512 LDS words; s_t: 2 LDS keys; s_need: 2 LDS ints.  Called by
// every thread of the workgroup.
template <typename T, class F>
__device__ T wg_radix_kth(F for_each, int need, uint32_t *hist, T *s_t, int *s_need) {
    constexpr int NB = 8 * (int)sizeof(T);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int i = tid; i < 256; i += blockDim.x) hist[i] = 0u;
    if (tid == 0) {
        s_t[0] = ~(T)0;
        s_t[1] = (T)0;
    }
    T lo = ~(T)0, hi = 0;
    for_each([&](T k) {
        if (k != 0) {
            lo = k < lo ? k : lo;
            hi = k > hi ? k : hi;
        }
    });
    for (int o = 32; o > 0; o >>= 1) {
        T ol, oh;
        if constexpr (sizeof(T) == 8) {
            ol = shfl_xor_u64(lo, o);
            oh = shfl_xor_u64(hi, o);
        } else {
            ol = (T)__shfl_xor((int)lo, o);
            oh = (T)__shfl_xor((int)hi, o);
        }
        lo = ol < lo ? ol : lo;
        hi = oh > hi ? oh : hi;
    }
    __syncthreads();
    if (lane == 0 && hi != 0) {
        atomicMin(&s_t[0], lo);
        atomicMax(&s_t[1], hi);
    }
    __syncthreads();
    lo = s_t[0];
    hi = s_t[1];
    if (hi == 0) return (T)0;
    int bits = lo == hi ? 1 : NB - (sizeof(T) == 8 ? __builtin_clzll((uint64_t)(lo ^ hi)) : __builtin_clz((uint32_t)(lo ^ hi)));
    T prefix = bits >= NB ? (T)0 : (T)(hi >> bits);  // the common high bits
    int rem = need;
    for (int pass = 0; bits > 0; ++pass) {
        const int d = bits < 8 ? bits : 8;
        const int shift = bits - d;
        uint32_t *h = hist + (pass & 1) * 256;
        for_each([&](T k) {
            if (k != 0 && (shift + d >= NB || (k >> (shift + d)) == prefix))
                atomicAdd(&h[(uint32_t)(k >> shift) & ((1u << d) - 1u)], 1u);
        });
        __syncthreads();
        if (wave == 0) {
            uint32_t c[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) c[j] = h[255 - 4 * lane - j];  // lane 0 holds the top bins
            const uint32_t sum = c[0] + c[1] + c[2] + c[3];
            uint32_t inc = sum;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(inc, o);
                if (lane >= o) inc += t;
            }
            uint32_t above = inc - sum;
            if (lane == 0) s_need[0] = -1;  // stays when fewer than rem keys share the prefix
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (above < (uint32_t)rem && (uint32_t)rem <= above + c[j]) {
                    s_t[0] = (T)((prefix << d) | (T)(255 - 4 * lane - j));
                    s_need[0] = rem - (int)above;
                    s_need[1] = c[j] == (uint32_t)rem - above;
                }
                above += c[j];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) hist[((pass + 1) & 1) * 256 + 4 * lane + j] = 0u;
        }
        __syncthreads();
        rem = s_need[0];
        if (rem < 0) return (T)0;
        prefix = s_t[0];
        bits = shift;
        if (s_need[1]) return (T)(prefix << shift);
    }
    return prefix;
}

// The K <= 256 best of total <= V * blockDim.x ranking keys (0 = empty)
// into out[0, K), then (value -inf, index -1) past the nonzero keys: radix
// selection of the K-th key, then each selected key's rank by counting the
// selected keys above it.  lds: 4128 bytes (histogram, state, 256 keys).
template <int V>
__device__ inline void topk_radix(const uint64_t *keys, int total, int K, acc_element *out, uint32_t *lds) {
    uint32_t *hist = lds;
    int *s_need = (int *)(lds + 512);
    int *s_n = (int *)(lds + 514);
    uint64_t *s_t = (uint64_t *)(lds + 516);
    uint64_t *s_sel = (uint64_t *)(lds + 520);
    const int tid = threadIdx.x, lane = tid & 63;
    uint64_t k[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
        const int i = tid + u * (int)blockDim.x;
        k[u] = i < total ? keys[i] : 0ull;
    }
    if (tid == 0) *s_n = 0;
    const uint64_t T = wg_radix_kth<uint64_t>(
        [&](auto fn) {
#pragma unroll
            for (int u = 0; u < V; ++u) fn(k[u]);
        },
        K, hist, s_t, s_need);
#pragma unroll
    for (int u = 0; u < V; ++u) {
        const bool sel = k[u] != 0ull && k[u] >= T;
        const uint64_t m = __ballot(sel);
        if (m == 0) continue;
        int base = 0;
        if (lane == 0) base = atomicAdd(s_n, __popcll(m));
        base = __shfl(base, 0);
        const int j = base + mbcnt64(m);
        if (sel && j < 256) s_sel[j] = k[u];  // j < K for distinct keys
    }
    __syncthreads();
    const int n = min(*s_n, 256);  // min(K, nonzero keys)
    if (tid < n) {
        const uint64_t key = s_sel[tid];
        int rank = 0;
        for (int j = 0; j < n; ++j) rank += s_sel[j] > key;
        acc_element e;
        e.value = dkey_score_f32(key);
        e.index = dkey_row(key);
        out[rank] = e;
    }
    for (int i = n + tid; i < K; i += blockDim.x) {
        acc_element e;
        e.value = -__builtin_inff();
        e.index = -1;
        out[i] = e;
    }
}

template <int R>
__device__ inline void write_results(const WaveTopM<R> &col, int K, acc_element *out) {
    for (int i = lane_id(); i < K; i += 64) {
        const uint64_t k = i < col.n ? col.buf[i] : 0ull;
        acc_element e;
        e.value = k ? dkey_score_f32(k) : -__builtin_inff();
        e.index = k ? dkey_row(k) : -1;
        out[i] = e;
    }
}

struct LowbitArgs {
    const uint32_t *words;      // lane-interleaved packet dwords
    const uint2 *slices;        // (offset in 64-word units, dwords per lane)
    const int32_t *perm;        // slot -> row (-1: padding lane)
    const int4 *items;          // per workgroup: (partition, first slice, end slice, 0)
    const int64_t *part_slice;  // split + 1 slice offsets
    const uint8_t *nvalid;      // rows per slice (64 except a partition's last)
    int32_t *smax;              // per slice: largest INT32 score of its lanes
    const float *x;             // FP32 query (cols)
    int cols, col_bits, val_bits;
    int split, M;               // partitions, rows kept per partition
    const uint32_t *ptr;        // FP32 CSR for rescoring
    const int32_t *idx;
    const float *val;
    int G;
    uint64_t *resc;             // split * M rescored keys
    int K;                      // final Top-K (0: candidates only)
    acc_element *out;
    int32_t *cand_rows_out;     // optional stage-1 candidates at p*M + i
    int32_t *cand_scores_out;
};

struct ScoresArgs {
    const uint32_t *words;
    const uint2 *slices;
    const int32_t *perm;
    int64_t nslices;
    const float *x;
    int cols, col_bits, val_bits;
    int32_t *y;
};


// Top-K of y[0, n) for any K <= 960 (select_blocks: workgroup b scans
// [b*chunk, (b+1)*chunk), each wavefront a contiguous quarter, with
// wavefront collectors, merges them and appends what can still be in the
// Top-K; select_final ranks the appended keys).
struct SelectArgs {
    const float *y;
    int n, chunk, K;
    uint64_t *thr;   // [1] lower bound of the K-th key
    uint32_t *cnt;   // [1]
    uint64_t *cand;  // gridDim.x * K
    acc_element *out;
};

// ---------------------------------------------------------------------------
// batched low-bit path: a block of QB queries per matrix pass
// ---------------------------------------------------------------------------
//   batch_scores   quantizes QB queries into an LDS table (per query its own
//                  scale) and computes INT32 scores of every slot for all QB
//                  queries while decoding each packet entry once; scores go to
//                  S[q][slot] (slot order, coalesced).  MACs are v_dot4 with
//                  the value placed in one byte lane: dot4(q[c][b..b+3],
//                  v << 8k) = q[c][b+k] * v.
//   batch_select   per (query, partition): the partition's Top-M from S and
//                  the FP32 rescoring of those rows.
//   batch_final    per query: its Top-K of the rescored candidates.

struct BatchArgs {
    const uint32_t *words;
    const uint2 *slices;
    const int32_t *perm;
    const int4 *items;           // batch_scores: per workgroup slice ranges
    const int64_t *part_slice;   // split + 1 slice offsets
    const float *X;              // nq x cols FP32 queries
    int nq, cols, col_bits, val_bits, QB;
    int64_t nslots;
    int32_t *S;                  // nq x nslots INT32 scores
    int split, M;
    const uint32_t *ptr;         // FP32 CSR for rescoring
    const int32_t *idx;
    const float *val;
    int G;
    uint64_t *resc;              // nq x split*M rescored keys
    const int32_t *K;            // per query
    acc_element *out;            // per query at q * out_stride
    int out_stride;
    int32_t *cand_rows_out;      // optional: nq x split*M stage-1 candidates
    int32_t *cand_scores_out;
};

// Query table layout in LDS: row c holds the QB int8 values of column c
// (offset by +128 as unsigned bytes when U8), padded to QS bytes.
template <int QB>
struct QTable {
    static constexpr int QS = QB + 8;
};


// ---------------------------------------------------------------------------
// batched low-bit path on the matrix cores, read from the packets
// ---------------------------------------------------------------------------
//   dense_prep     quantizes the batch's queries into the B-operand table;
//   pk_scores      v_mfma_i32_32x32x32_i8 over each 64-row slice and a block
//                  of queries: decode wavefronts expand the slice's image
//                  packets (the same slices re-encoded at upload: per entry
//                  the image byte and cols - column, a row's repeated columns
//                  merged) into an LDS image of the A operands, each MFMA wavefront
//                  holds its query tile's B operands in registers; per (query,
//                  slice) only the four best (score, row-in-slice) keys are
//                  kept;
//   dense_select   per (partition, query): threshold = the M-th best slice
//                  maximum, candidates from the four-key lists, exact rescan
//                  (from the packets) of any slice whose list may have dropped
//                  a candidate, Top-M, and the FP32 rescoring of those M rows;
//   batch_topk     per query: its Top-K[q] of the rescored keys.
//
// Image of a slice (LDS): row r of the slice at r * (cols_pad + 16), one
// byte per column; the padding entries write column cols.  The
// A operand (row tile rt, column block kb) of lane l is the 16 bytes of row
// rt*32 + l%32 at column kb*32 + (l/32)*16.  The query table uses the same
// lane mapping for the B operand (query l%32 of a 32-query tile), so every k
// is paired identically in both.  Unsigned 8-bit values are stored as v-128
// (absent entries as -128) and 128 * sum(query) is added back per query;
// other widths are stored as is.  Two images, then kPkLdsTail bytes that the
// chain's reads of column blocks past nkb (multiplied by zero B operands)
// may reach.
static constexpr int kPkLdsTail = 34 * 32;
// pk_scores workgroups: up to kPkMfmaWaves MFMA wavefronts (one 32-query tile
// each, or two where the B operands of two tiles fit the registers: up to 16
// column blocks), or up to kPkMfmaWaves3 of one tile each (two per SIMD,
// above 16 column blocks), and kPkDecodeWaves decode wavefronts.
static constexpr int kPkMfmaWaves = 4;
static constexpr int kPkMfmaWaves3 = 8;
static constexpr int kPkDecodeWaves = 4;

struct DenseArgs {
    const uint32_t *words;       // sliced packets
    const uint2 *slices;
    const uint32_t *pwords;      // image packets (pk_scores): (byte, offset in a decode wavefront's
                                 // 16 image rows) entries of pcb + 8 bits, one stream per wavefront
    const uint2 *pslices;        // per slice: (offset in units of 64 dwords, entries per lane / 4 of
                                 // decode wavefront w in byte w)
    int pcb;                     // offset bits
    const int32_t *perm;
    const int64_t *part_slice;   // split + 1
    const uint8_t *nvalid;       // valid rows per slice (a prefix of its lanes)
    int64_t nslices;
    int cols, cols_pad, nkb, col_bits, val_bits, is_signed, u8off;
    const float *X;              // nq x cols FP32 queries
    int nq, NT, nqb;             // queries; 32-query tiles per block; blocks
    int8_t *qtab;                // nqb x nkb x NT x 1024 B-operand bytes
    int8_t *qplain;              // nq x cols_pad quantized queries
    int32_t *corr;               // nqb*32*NT per-query offsets
    uint32_t *tk;                // nq x nslices x 4 keys
    int split, M;
    const uint32_t *ptr;         // FP32 CSR for rescoring
    const int32_t *idx;
    const float *val;
    int G;
    uint64_t *resc;              // nq x split*M rescored keys
    const int32_t *K;            // per query
    acc_element *out;
    int out_stride;
    int32_t *cand_rows_out;      // optional: nq x split*M stage-1 candidates
    int32_t *cand_scores_out;
};

// ---------------------------------------------------------------------------
// host-side getters of the kernel instantiations (one translation unit each)
// ---------------------------------------------------------------------------
typedef void (*lowbit_fn)(LowbitArgs);
typedef void (*scores_fn)(ScoresArgs);
typedef void (*select_fn)(SelectArgs);
typedef void (*batch_fn)(BatchArgs);
typedef void (*spmv_fn)(const uint32_t *, const int32_t *, const float *, const float *, int, int, float *);
typedef void (*reset_fn)(uint64_t *, uint32_t *);
typedef void (*radix_pass_fn)(const void *, int64_t, int, uint64_t *);
typedef void (*radix_final_fn)(const uint64_t *, int, int, acc_element *);
typedef void (*fp64_fn)(const uint32_t *, const int32_t *, const float *, const float *, int, double *);

lowbit_fn get_lowbit_slicemax(int E, bool is_signed);  // k_lowbit.hip
lowbit_fn get_lowbit_select(int E, bool is_signed);
lowbit_fn get_lowbit_final(int R);
lowbit_fn get_lowbit_final_radix();
scores_fn get_lowbit_scores(int E, bool is_signed);
spmv_fn get_fp32_spmv(int G);                         // k_fp32.hip
select_fn get_select_blocks(int R);
select_fn get_select_final(int R);
radix_pass_fn get_select_radix_pass(bool keys, int V);  // k_fp32.hip
radix_final_fn get_select_radix_final(int V);
reset_fn get_select_reset();
fp64_fn get_spmv_fp64();
batch_fn get_batch_scores16(int E, bool is_signed, int bits);  // k_batch16.hip
batch_fn get_batch_scores32(int E, bool is_signed, int bits);  // k_batch32.hip
batch_fn get_batch_select();                                   // k_batch.hip
batch_fn get_batch_final(int R);
typedef void (*dense_fn)(DenseArgs);
dense_fn get_dense_prep();                                     // k_dense.hip
dense_fn get_dense_select(bool wide);
dense_fn get_batch_topk_radix(int V);
dense_fn get_batch_topk(int R);
dense_fn get_pk_scores_8(int E);                               // k_pkmfma_a.hip
dense_fn get_pk_scores_16(int E);                              // k_pkmfma_b.hip
dense_fn get_pk_scores_32(int E);                              // k_pkmfma_c.hip
dense_fn get_pk_scores_24(int E);                              // k_pkmfma_d.hip
dense_fn get_pk_scores_30(int E);                              // k_pkmfma_e.hip
dense_fn get_pk2_scores_8(int E);                              // k_pkmfma_f.hip
dense_fn get_pk2_scores_16(int E);                             // k_pkmfma_g.hip
dense_fn get_pk3_scores_24(int E);                             // k_pkmfma_h.hip
dense_fn get_pk3_scores_30(int E);
dense_fn get_pk3_scores_32(int E);
