// gpu.hip -- device-resident matrices and the GPU entry points of the C ABI.
#include "kernels_common.hpp"
#include "internal.hpp"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using acc::fail;

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(ACC_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

static constexpr int kMaxM = 64;     // rows kept per partition
static constexpr int kMaxK = 960;    // Top-K limit of the FP32 path (collector cap 1024)
static constexpr int kMaxCK = 4096;  // candidates of the low-bit path
static constexpr size_t kWordPad = 32 * 64;  // dwords past the packets (>= widest group of 32 entries)
static constexpr int kPkMaxCols = 1024;  // matrix-core scores: B operands of 32 column blocks in registers
static constexpr int kPkMinE = 18, kPkMaxE = 23;  // image-packet entry widths instantiated for them
static constexpr size_t kPkLdsMax = 160 * 1024;  // LDS of a pk_scores workgroup at most

// Per-batch device workspace of the batched path.  The handle owns one;
// every plan (captured graph) owns its own, so replaying a plan never touches
// buffers that a later call on the handle may reallocate.
struct BatchWs {
    std::vector<void *> *allocs = nullptr;  // owner's allocation list
    int64_t *bytes = nullptr;               // owner's byte count
    int32_t *K = nullptr;
    size_t K_n = 0;
    int32_t *S = nullptr;        // packet layout: nq x slots INT32 scores
    size_t S_n = 0;
    uint64_t *resc = nullptr;    // nq x split*M rescored keys
    size_t resc_n = 0;
    int8_t *qtab = nullptr;      // dense layout: B-operand table
    size_t qtab_n = 0;
    int8_t *qplain = nullptr;
    size_t qplain_n = 0;
    int32_t *corr = nullptr;
    size_t corr_n = 0;
    uint32_t *tk = nullptr;      // four-key slice lists
    size_t tk_n = 0;
    int32_t *smax = nullptr;     // single-query passes: slice maxima
    size_t smax_n = 0;
    float *y = nullptr;          // FP32 plans: scores
    size_t y_n = 0;
    uint64_t *rsel = nullptr;    // FP32 plans: radix selection keys
    size_t rsel_n = 0;
    uint64_t *sthr = nullptr;    // FP32 plans, K > 256: select_blocks state (zero between launches)
    size_t sthr_n = 0;
    uint32_t *scnt = nullptr;
    size_t scnt_n = 0;
    uint64_t *scand = nullptr;
    size_t scand_n = 0;
};

struct acc_gpu_matrix {
    int device = 0;
    int cus = 256;
    int32_t rows = 0, cols = 0, split = 1, group = 8;
    int64_t nnz = 0;
    // FP32 CSR
    bool has_csr = false;
    uint32_t *d_ptr = nullptr;
    int32_t *d_idx = nullptr;
    float *d_val = nullptr;
    float *d_y = nullptr;            // FP32 scores of the plain path
    // select state of the FP32 path
    int sel_blocks = 1, sel_chunk = 8192;
    uint64_t *d_sthr = nullptr;
    uint32_t *d_scnt = nullptr;
    uint64_t *d_scand = nullptr;
    uint64_t *d_rsel = nullptr;      // radix Top-K passes (K <= 256)
    size_t rsel_cap = 0;
    // low-bit sliced packets
    bool has_packed = false;
    int bits = 8, col_bits = 10, entry_bits = 18, is_signed = 0;
    int64_t nslices = 0, stream_bytes = 0;
    uint32_t *d_words = nullptr;
    uint2 *d_slices = nullptr;
    int32_t *d_perm = nullptr;
    int4 *d_items = nullptr;         // per workgroup
    int nitems = 0;
    std::vector<int32_t> part_row_begin;
    int32_t *d_smax = nullptr;       // per slice: largest INT32 score of the last query
    uint64_t *d_resc = nullptr;      // split * kMaxM
    int32_t *d_crow = nullptr, *d_cscore = nullptr;  // split * kMaxM (candidate outputs)
    int64_t *d_part_slice = nullptr;                 // split + 1
    int64_t max_part_slices = 0;                     // slices of the largest partition
    uint32_t *d_pwords = nullptr;                    // image packets (matrix-core scores), E = 8 + pcb
    uint2 *d_pslices = nullptr;
    int pcb = 0;
    int64_t pbytes = 0;                              // image-packet bytes read per pass (words + slice table)
    BatchWs ws;                      // batched-path workspace (grown on demand)
    // batched path on the matrix cores (slices expanded from the packets in LDS)
    uint8_t *d_nvalid = nullptr;     // valid rows per slice
    bool dense_ok = false;           // the matrix cores can score this matrix exactly
    std::string dense_why;           // why not, otherwise
    int batch_mode = ACC_BATCH_AUTO;
    int dense_min_nq = 16;           // ACC_BATCH_AUTO, untuned: matrix cores from this batch size on
    // ACC_BATCH_AUTO after acc_gpu_autotune: the layout measured fastest at
    // tune_nq[i], used for batches of tune_nq[i] .. tune_nq[i+1]-1 queries
    int ntune = 0;
    int tune_nq[16] = {0};
    int tune_layout[16] = {0};
    int cols_pad = 0, nkb = 0, u8off = 0;
    // profiling events around the launches of the last query
    bool profiling = false;
    hipEvent_t ev[12] = {};  // start and end of each kernel of the last query
    int nev = 0;             // events recorded by the last query
    // host-call staging
    float *d_xstage = nullptr;
    acc_element *d_ostage = nullptr;
    int64_t device_bytes = 0;
    std::vector<void *> allocs;
};

namespace {

template <typename T>
int dalloc(acc_gpu_matrix *m, T **p, size_t count) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) return fail(ACC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    m->allocs.push_back(*p);
    m->device_bytes += (int64_t)bytes;
    return ACC_OK;
}

void free_all(acc_gpu_matrix *m) {
    if (!m) return;
    hipSetDevice(m->device);
    for (void *p : m->allocs) hipFree(p);
    m->allocs.clear();
    for (auto &e : m->ev)
        if (e) {
            hipEventDestroy(e);
            e = nullptr;
        }
}

// Grows a workspace buffer to at least `need` elements (the allocation is
// recorded in the owner's list).
template <typename T>
int grow_in(std::vector<void *> &allocs, int64_t &bytes, T **p, size_t *cap, size_t need) {
    if (*p && *cap >= need) return ACC_OK;
    if (*p) {
        hipFree(*p);
        allocs.erase(std::find(allocs.begin(), allocs.end(), (void *)*p));
        bytes -= (int64_t)(*cap * sizeof(T));
        *p = nullptr;
        *cap = 0;
    }
    const size_t n = std::max<size_t>(need, 1);
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) return fail(ACC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(*p);
    bytes += (int64_t)(n * sizeof(T));
    *cap = n;
    return ACC_OK;
}

template <typename T>
int ws_grow(BatchWs &ws, T **p, size_t *cap, size_t need) {
    return grow_in(*ws.allocs, *ws.bytes, p, cap, need);
}

// Launches kernel k as stage i of a query.  With profiling on, a pair of
// HIP events is attached to the dispatch itself (the kernel's start and end),
// so a stage's time is the kernel's own duration, without launch gaps.
#define LAUNCH(mm, i, k, grid, block, lds, st, ...)                                                              \
    do {                                                                                                       \
        acc_gpu_matrix *m_ = (mm);                                                                             \
        if (m_ && m_->profiling) {                                                                             \
            hipExtLaunchKernelGGL(k, grid, block, lds, st, m_->ev[2 * (i)], m_->ev[2 * (i) + 1], 0, __VA_ARGS__); \
            m_->nev = 2 * (i) + 2;                                                                             \
        } else {                                                                                               \
            hipLaunchKernelGGL(k, grid, block, lds, st, __VA_ARGS__);                                          \
        }                                                                                                      \
    } while (0)

int pow2_at_least(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// G lanes per row for the FP32 dot products: a power of two near a quarter of
// the mean row length, in [2, 32].
int choose_group(double avg) {
    int g = 2;
    while (g < 32 && g * 4 <= avg) g <<= 1;
    return g;
}

// Workgroup work items: contiguous slice ranges inside one partition with
// about equal work (dwords plus a fixed per-slice cost).
std::vector<int4> make_items(const acc_packed *pk, int cus, std::vector<int32_t> &part_wgs) {
    const int S = pk->split;
    const int64_t NS = pk->nslices();
    const int64_t kSliceCost = 8;
    int64_t total = 0;
    for (int64_t s = 0; s < NS; ++s) total += pk->slice_ndw[(size_t)s] + kSliceCost;
    const int64_t avg_slice = std::max<int64_t>(1, total / std::max<int64_t>(NS, 1));
    const int64_t per = std::max<int64_t>(total / ((int64_t)cus * 6), avg_slice * LB_WAVES);
    std::vector<int4> items;
    part_wgs.assign((size_t)S, 0);
    for (int p = 0; p < S; ++p) {
        int64_t s0 = pk->part_slice_begin[(size_t)p], s1 = pk->part_slice_begin[(size_t)p + 1];
        int64_t b = s0, acc = 0;
        for (int64_t s = s0; s < s1; ++s) {
            acc += pk->slice_ndw[(size_t)s] + kSliceCost;
            if (acc >= per || s + 1 == s1 || s + 1 - b >= LB_MAX_ITEM) {
                items.push_back(make_int4(p, (int)b, (int)(s + 1), 0));
                part_wgs[(size_t)p]++;
                b = s + 1;
                acc = 0;
            }
        }
    }
    return items;
}

// Collector size R (64*R keys per wavefront) for keeping the best M: M <= 64*R - 64.
int r_for(int M) { return std::max(2, pow2_at_least(M + 64) / 64); }

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ACC_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return ACC_OK;
}

int lds_limit(const void *fn, size_t bytes);

// One low-bit query (slice maxima, per-partition select + rescoring, final Top-K):
// Top-K into out (K > 0) and/or the per-partition candidates into
// crow/cscore (p*M + i).
LowbitArgs lowbit_args(acc_gpu_matrix *m, const float *x, int M, int K, acc_element *out, int32_t *crow,
                       int32_t *cscore, int32_t *smax = nullptr, uint64_t *resc = nullptr) {
    LowbitArgs a;
    a.words = m->d_words;
    a.slices = m->d_slices;
    a.perm = m->d_perm;
    a.items = m->d_items;
    a.part_slice = m->d_part_slice;
    a.nvalid = m->d_nvalid;
    a.smax = smax ? smax : m->d_smax;
    a.x = x;
    a.cols = m->cols;
    a.col_bits = m->col_bits;
    a.val_bits = m->bits;
    a.split = m->split;
    a.M = M;
    a.ptr = m->d_ptr;
    a.idx = m->d_idx;
    a.val = m->d_val;
    a.G = m->group;
    a.resc = resc ? resc : m->d_resc;
    a.K = K;
    a.out = out;
    a.cand_rows_out = crow;
    a.cand_scores_out = cscore;
    return a;
}

// Launches stage i of a low-bit query: 0 slice maxima, 1 per-partition
// select + rescoring, 2 final Top-K (a.K > 0).
int launch_lowbit_stage(acc_gpu_matrix *m, const LowbitArgs &a, int stage, hipStream_t st) {
    const size_t qlds = (size_t)((m->cols + 15) & ~15);
    if (stage == 0) {
        lowbit_fn k = get_lowbit_slicemax(m->entry_bits, m->is_signed != 0);
        if (!k) return fail(ACC_ERR_ARG, "unsupported entry width");
        LAUNCH(m, 0, k, dim3(m->nitems), dim3(LB_WG), 64 + qlds, st, a);
    } else if (stage == 1) {
        lowbit_fn ksel = get_lowbit_select(m->entry_bits, m->is_signed != 0);
        if (!ksel) return fail(ACC_ERR_ARG, "unsupported entry width");
        const size_t lds_sel = 64 + 8 * 4096 + 5 * qlds;         // key list, int8 and FP32 query
        int rc = lds_limit((const void *)ksel, lds_sel + 40960);  // plus about 36 KB static
        if (rc != ACC_OK) return rc;
        LAUNCH(m, 1, ksel, dim3(m->split), dim3(SEL_WG), lds_sel, st, a);
    } else if (a.K <= 256 && m->split * a.M <= 4 * SEL_WG) {
        LAUNCH(m, 2, get_lowbit_final_radix(), dim3(1), dim3(SEL_WG), 0, st, a);
    } else {
        const int R = pow2_at_least((m->split * a.M + ACC_WG - 1) / ACC_WG);  // split*M <= kMaxCK = 4*64*16
        lowbit_fn f = get_lowbit_final(R);
        if (!f) return fail(ACC_ERR_ARG, "too many candidates");
        LAUNCH(m, 2, f, dim3(1), dim3(ACC_WG), (size_t)ACC_WAVES * 64 * R * 8, st, a);
    }
    return check_launch();
}

// ws: a plan's or a batch's workspace for the slice maxima and rescored keys
// (null: the handle's own).
int launch_lowbit(acc_gpu_matrix *m, const float *x, int M, int K, acc_element *out, int32_t *crow, int32_t *cscore,
                  hipStream_t st, BatchWs *ws = nullptr) {
    if (ws) {
        int rc = ws_grow(*ws, &ws->smax, &ws->smax_n, (size_t)m->nslices);
        if (rc == ACC_OK) rc = ws_grow(*ws, &ws->resc, &ws->resc_n, (size_t)m->split * kMaxM);
        if (rc != ACC_OK) return rc;
    }
    const LowbitArgs a = lowbit_args(m, x, M, K, out, crow, cscore, ws ? ws->smax : nullptr, ws ? ws->resc : nullptr);
    int rc = launch_lowbit_stage(m, a, 0, st);
    if (rc == ACC_OK) rc = launch_lowbit_stage(m, a, 1, st);
    if (rc == ACC_OK && K > 0) rc = launch_lowbit_stage(m, a, 2, st);
    return rc;
}

// Batched low-bit queries on the packets: batch_scores over query blocks,
// batch_select per (partition, query), batch_final per query (skipped when out
// is null).  The workspace must already hold nq queries (S, resc, K).
int launch_batch(acc_gpu_matrix *m, BatchWs &ws, const float *X, int nq, int M, int kmax, acc_element *out,
                 int out_stride, int32_t *crow, int32_t *cscore, hipStream_t st) {
    BatchArgs a;
    a.words = m->d_words;
    a.slices = m->d_slices;
    a.perm = m->d_perm;
    a.items = m->d_items;
    a.part_slice = m->d_part_slice;
    a.X = X;
    a.nq = nq;
    a.cols = m->cols;
    a.col_bits = m->col_bits;
    a.val_bits = m->bits;
    a.nslots = m->nslices * 64;
    a.S = ws.S;
    a.split = m->split;
    a.M = M;
    a.ptr = m->d_ptr;
    a.idx = m->d_idx;
    a.val = m->d_val;
    a.G = m->group;
    a.resc = out ? ws.resc : nullptr;
    a.K = ws.K;
    a.out = out;
    a.out_stride = out_stride;
    a.cand_rows_out = crow;
    a.cand_scores_out = cscore;
    // query block: 32 queries per matrix pass, 16 for small batches
    const int QB = (nq <= 16 || (size_t)m->cols * (32 + 8) > 65536) ? 16 : 32;
    a.QB = QB;
    batch_fn fs = QB == 16 ? get_batch_scores16(m->entry_bits, m->is_signed != 0, m->bits)
                           : get_batch_scores32(m->entry_bits, m->is_signed != 0, m->bits);
    if (!fs) return fail(ACC_ERR_ARG, "unsupported entry width");
    const size_t lds_q = (size_t)m->cols * (QB + 8);
    LAUNCH(m, 0, fs, dim3(m->nitems, (nq + QB - 1) / QB), dim3(ACC_WG), lds_q, st, a);
    int rc = check_launch();
    if (rc != ACC_OK) return rc;
    LAUNCH(m, 1, get_batch_select(), dim3(m->split, nq), dim3(ACC_WG), (size_t)ACC_WAVES * 128 * 8, st, a);
    rc = check_launch();
    if (rc != ACC_OK || !out) return rc;
    const int R = r_for(kmax);
    batch_fn ff = get_batch_final(R);
    if (!ff) return fail(ACC_ERR_ARG, "k too large");
    LAUNCH(m, 2, ff, dim3(nq), dim3(ACC_WG), (size_t)ACC_WAVES * 64 * R * 8, st, a);
    return check_launch();
}

// Raises a kernel's dynamic LDS limit when it needs more than 64 KB.
// Only raised once per kernel and size (also keeps the call out of graph
// capture, where the first, eager run of a plan has already made it).
int lds_limit(const void *fn, size_t bytes) {
    if (bytes <= 65536) return ACC_OK;
    static std::mutex mu;
    static std::vector<std::pair<const void *, size_t>> done;
    std::lock_guard<std::mutex> lock(mu);
    for (auto &d : done)
        if (d.first == fn && d.second >= bytes) return ACC_OK;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return fail(ACC_ERR_DEVICE, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
    done.push_back({fn, bytes});
    return ACC_OK;
}

DenseArgs dense_args(const acc_gpu_matrix *m) {
    DenseArgs a;
    std::memset(&a, 0, sizeof(a));
    a.words = m->d_words;
    a.slices = m->d_slices;
    a.pwords = m->d_pwords;
    a.pslices = m->d_pslices;
    a.pcb = m->pcb;
    a.perm = m->d_perm;
    a.part_slice = m->d_part_slice;
    a.nvalid = m->d_nvalid;
    a.nslices = m->nslices;
    a.cols = m->cols;
    a.cols_pad = m->cols_pad;
    a.nkb = m->nkb;
    a.col_bits = m->col_bits;
    a.val_bits = m->bits;
    a.is_signed = m->is_signed;
    a.u8off = m->u8off;
    a.split = m->split;
    a.ptr = m->d_ptr;
    a.idx = m->d_idx;
    a.val = m->d_val;
    a.G = m->group;
    return a;
}

// Layout of a batch of nq queries: ACC_BATCH_SINGLE, _SPARSE or _DENSE.
int batch_layout(const acc_gpu_matrix *m, int nq) {
    if (m->batch_mode != ACC_BATCH_AUTO) return m->batch_mode;
    if (m->ntune > 0) {
        int pick = m->tune_layout[0];
        for (int i = 0; i < m->ntune; ++i)
            if (m->tune_nq[i] <= nq) pick = m->tune_layout[i];
        return pick;
    }
    return m->dense_ok && nq >= m->dense_min_nq ? ACC_BATCH_DENSE : ACC_BATCH_SPARSE;
}

// B registers for 8, 16, 24, 30 or 32 column blocks: the smallest that holds
// the matrix (the multiplications run over all of them, zero past nkb); E =
// the image packets' entry bits.  Up to 16 column blocks the chain runs in
// half-rounds: with one tile per MFMA wavefront its registers leave room for
// two workgroups per CU (k_pkmfma_{a,b}), with two tiles (tpw = 2) each A
// read feeds both (k_pkmfma_{f,g}).  Above 16 column blocks, a batch of more
// than four tiles runs eight MFMA wavefronts per workgroup, two per SIMD
// (wide, k_pkmfma_h): one decode of each slice for up to 256 queries.
dense_fn pk_kernel(const acc_gpu_matrix *m, int tpw, bool wide) {
    const int E = 8 + m->pcb;
    if (tpw == 2) return m->nkb > 8 ? get_pk2_scores_16(E) : get_pk2_scores_8(E);
    if (wide) return m->nkb > 30 ? get_pk3_scores_32(E) : m->nkb > 24 ? get_pk3_scores_30(E) : get_pk3_scores_24(E);
    if (m->nkb > 30) return get_pk_scores_32(E);
    if (m->nkb > 24) return get_pk_scores_30(E);
    if (m->nkb > 16) return get_pk_scores_24(E);
    return m->nkb > 8 ? get_pk_scores_16(E) : get_pk_scores_8(E);
}

// Batched low-bit queries on the matrix cores: dense_prep, pk_scores (slices
// expanded from the packets in LDS), dense_select (which also rescores when
// out != null), then batch_topk.
int launch_dense(acc_gpu_matrix *m, BatchWs &ws, const float *X, int nq, int M, int kmax, acc_element *out,
                 int out_stride, int32_t *crow, int32_t *cscore, hipStream_t st) {
    if (!m->dense_ok) return fail(ACC_ERR_ARG, "matrix-core scores unavailable: " + m->dense_why);
    // query blocks of at most kPkMfmaWaves MFMA wavefronts, balanced, with
    // TPW tiles of 32 queries each; every block has kPkDecodeWaves decode
    // wavefronts.  Where one tile per wavefront would take more than one
    // block, fewer blocks decode each slice fewer times: up to 16 column
    // blocks two tiles per wavefront (NT even; a wavefront's second tile may
    // be all padding; 256 queries at 500 columns: 74 -> 59 us), above that
    // (two tiles' B operands exceed the registers) up to kPkMfmaWaves3
    // one-tile wavefronts, two per SIMD.  Within one block one tile per
    // wavefront spreads the MFMAs over more SIMDs (100 queries: 41 vs 53 us)
    const int tiles = (nq + 31) / 32;
    const int tpw = m->nkb <= 16 && tiles > kPkMfmaWaves ? 2 : 1;
    const bool wide = m->nkb > 16 && tiles > kPkMfmaWaves;  // eight one-tile MFMA wavefronts
    const int per_block = wide ? kPkMfmaWaves3 : kPkMfmaWaves * tpw;
    const int nqb = (tiles + per_block - 1) / per_block;
    const int NT = ((tiles + nqb - 1) / nqb + tpw - 1) / tpw * tpw, QB = 32 * NT;
    const int W = NT / tpw + kPkDecodeWaves;
    const size_t cand = (size_t)nq * m->split * M;
    int rc = ws_grow(ws, &ws.qtab, &ws.qtab_n, (size_t)nqb * m->nkb * NT * 1024);
    if (rc == ACC_OK) rc = ws_grow(ws, &ws.qplain, &ws.qplain_n, (size_t)nq * m->cols_pad);
    if (rc == ACC_OK) rc = ws_grow(ws, &ws.corr, &ws.corr_n, (size_t)nqb * QB);
    if (rc == ACC_OK) rc = ws_grow(ws, &ws.tk, &ws.tk_n, (size_t)nq * m->nslices * 4);
    if (rc == ACC_OK) rc = ws_grow(ws, &ws.resc, &ws.resc_n, cand);
    if (rc != ACC_OK) return rc;
    DenseArgs a = dense_args(m);
    a.X = X;
    a.nq = nq;
    a.NT = NT;
    a.nqb = nqb;
    a.qtab = ws.qtab;
    a.qplain = ws.qplain;
    a.corr = ws.corr;
    a.tk = ws.tk;
    a.M = M;
    a.resc = ws.resc;
    a.K = ws.K;
    a.out = out;
    a.out_stride = out_stride;
    a.cand_rows_out = crow;
    a.cand_scores_out = cscore;

    LAUNCH(m, 0, get_dense_prep(), dim3(nqb * QB), dim3(ACC_WG), 0, st, a);
    if ((rc = check_launch()) != ACC_OK) return rc;
    dense_fn fs = pk_kernel(m, tpw, wide);
    if (!fs) return fail(ACC_ERR_ARG, "unsupported entry width");
    // two slice images (one multiplied while the other is cleared and refilled)
    const size_t img = (size_t)64 * (m->cols_pad + 16);
    const size_t lds_s = 2 * img + kPkLdsTail;
    if ((rc = lds_limit((const void *)fs, lds_s)) != ACC_OK) return rc;
    // workgroups per query block: as many as fit the chip at once, a multiple
    // of 8 (the kernel keeps a slice range's query blocks on one XCD)
    int occ = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)fs, W * 64, lds_s) != hipSuccess || occ < 1)
        occ = 1;
    const int64_t fit = (int64_t)m->cus * occ / nqb;
    const int nwx = (int)std::max<int64_t>(8, std::min<int64_t>(fit, (m->nslices + 7) / 8 * 8) / 8 * 8);
    LAUNCH(m, 1, fs, dim3(nwx * nqb), dim3(W * 64), lds_s, st, a);
    if ((rc = check_launch()) != ACC_OK) return rc;
    // per wavefront: two key buffers, the int8 query (rescans), the FP32 query (rescoring)
    const size_t lds_sel = (size_t)ACC_WAVES * 256 * 8 + (size_t)ACC_WAVES * m->cols_pad * 5;
    // partitions past the register path (more than 256 slices, or M > 32):
    // the variant that also holds up to 1024 slices' lists in registers (more
    // registers, fewer wavefronts per SIMD)
    const dense_fn fsel = get_dense_select(m->max_part_slices > 256 || M > 32);
    if ((rc = lds_limit((const void *)fsel, lds_sel)) != ACC_OK) return rc;
    LAUNCH(m, 2, fsel, dim3(m->split, (nq + ACC_WAVES - 1) / ACC_WAVES), dim3(ACC_WG), lds_sel, st, a);
    if ((rc = check_launch()) != ACC_OK) return rc;
    if (!out) return ACC_OK;
    const int total = m->split * M;  // <= kMaxCK = 4096 = 4 lists of 64*16
    const int R = pow2_at_least((total + 64 * ACC_WAVES - 1) / (64 * ACC_WAVES));
    dense_fn ft = kmax <= 256 ? get_batch_topk_radix(R) : get_batch_topk(R);
    if (!ft) return fail(ACC_ERR_ARG, "too many candidates");
    LAUNCH(m, 3, ft, dim3(nq), dim3(ACC_WG), kmax <= 256 ? 0 : (size_t)ACC_WAVES * 64 * R * 8, st, a);
    return check_launch();
}

int check_lowbit(acc_gpu_matrix *m, int ck, int k) {
    if (!m) return fail(ACC_ERR_ARG, "null matrix");
    if (!m->has_packed) return fail(ACC_ERR_ARG, "matrix was uploaded without low-bit packets");
    if (ck < m->split) return fail(ACC_ERR_ARG, "ck must be >= split");
    int M = ck / m->split;
    if (M > kMaxM) return fail(ACC_ERR_ARG, "ck/split must be <= 64");
    if ((int64_t)m->split * M > kMaxCK) return fail(ACC_ERR_ARG, "ck must be <= 4096");
    if (k < 0 || k > m->split * M || k > kMaxK) return fail(ACC_ERR_ARG, "k must be in [0, min(960, (ck/split)*split)]");
    if (k > 0 && !m->has_csr) return fail(ACC_ERR_ARG, "rescoring needs the FP32 CSR (upload it too)");
    return ACC_OK;
}

// Elements per thread of a radix Top-K pass over cnt inputs: 8 (2048 per
// workgroup) up to 2^21 inputs, so that small vectors still spread over the
// chip, else 32 (8192 per workgroup).
int radix_pass_v(int64_t cnt) { return cnt <= (int64_t(1) << 21) ? 8 : 32; }

// Keys the radix Top-K passes keep over n scores: every pass leaves K per
// workgroup chunk of its input, until at most 16384 remain (one final
// workgroup).
size_t radix_select_keys(int64_t n, int K) {
    size_t total = 0;
    int64_t cnt = n;
    do {
        const int64_t chunk = (int64_t)radix_pass_v(cnt) * ACC_WG;
        const int64_t blocks = std::max<int64_t>(1, (cnt + chunk - 1) / chunk);
        cnt = blocks * K;
        total += (size_t)cnt;
    } while (cnt > 16384);
    return total;
}

// Radix Top-K (K <= 256) of n scores; buf holds radix_select_keys(n, K) keys.
int launch_select_radix(const float *y, int64_t n, int K, uint64_t *buf, acc_element *out, hipStream_t st,
                        acc_gpu_matrix *prof) {
    const void *in = y;
    bool keys = false;
    int64_t cnt = n;
    uint64_t *dst = buf;
    do {
        const int V = radix_pass_v(cnt);
        const int64_t chunk = (int64_t)V * ACC_WG;
        const int64_t blocks = std::max<int64_t>(1, (cnt + chunk - 1) / chunk);
        radix_pass_fn f = get_select_radix_pass(keys, V);
        if (!keys) {
            LAUNCH(prof, 1, f, dim3((unsigned)blocks), dim3(ACC_WG), 0, st, in, cnt, K, dst);
        } else {
            hipLaunchKernelGGL(f, dim3((unsigned)blocks), dim3(ACC_WG), 0, st, in, cnt, K, dst);
        }
        int rc = check_launch();
        if (rc != ACC_OK) return rc;
        in = dst;
        keys = true;
        cnt = blocks * K;
        dst += cnt;
    } while (cnt > 16384);
    radix_final_fn ff = get_select_radix_final(pow2_at_least((int)((cnt + 1023) / 1024)));
    if (!ff) return fail(ACC_ERR_ARG, "too many candidates");
    LAUNCH(prof, 2, ff, dim3(1), dim3(1024), 0, st, (const uint64_t *)in, (int)cnt, K, out);
    return check_launch();
}

// Top-K of a device score vector: radix passes when K <= 256 and rbuf holds
// radix_select_keys(n, K) keys, else select_blocks then select_final (their
// state must be zero on entry and is left zero).
int launch_select(const float *y, int n, int K, int blocks, int chunk, uint64_t *thr, uint32_t *cnt, uint64_t *cand,
                  acc_element *out, hipStream_t st, acc_gpu_matrix *prof = nullptr, uint64_t *rbuf = nullptr,
                  size_t rcap = 0) {
    if (K <= 256 && rbuf && radix_select_keys(n, K) <= rcap) return launch_select_radix(y, n, K, rbuf, out, st, prof);
    SelectArgs s;
    s.y = y;
    s.n = n;
    s.chunk = chunk;
    s.K = K;
    s.thr = thr;
    s.cnt = cnt;
    s.cand = cand;
    s.out = out;
    const int R = r_for(K);
    select_fn fb = get_select_blocks(R), ff = get_select_final(R);
    if (!fb || !ff) return fail(ACC_ERR_ARG, "k too large");
    const size_t lds = (size_t)ACC_WAVES * 64 * R * 8;
    LAUNCH(prof, 1, fb, dim3(blocks), dim3(ACC_WG), lds, st, s);
    int rc = check_launch();
    if (rc != ACC_OK) return rc;
    LAUNCH(prof, 2, ff, dim3(1), dim3(ACC_WG), lds, st, s);
    return check_launch();
}

void select_geometry(int n, int cus, int &blocks, int &chunk) {
    chunk = std::max(8192, (n + 2 * cus - 1) / (2 * cus));
    blocks = std::max(1, (n + chunk - 1) / chunk);
}

}  // namespace

// ---------------------------------------------------------------------------
// device memory, streams, events
// ---------------------------------------------------------------------------

extern "C" int acc_gpu_device_count(int *count) {
    if (!count) return fail(ACC_ERR_ARG, "null argument");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
    return ACC_OK;
}

extern "C" int acc_dev_malloc(int device, size_t bytes, void **ptr) {
    if (!ptr) return fail(ACC_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(device));
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(ACC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return ACC_OK;
}

extern "C" int acc_dev_free(void *ptr) {
    HIP_TRY(hipFree(ptr));
    return ACC_OK;
}

extern "C" int acc_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return ACC_OK;
}

extern "C" int acc_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return ACC_OK;
}

extern "C" int acc_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return ACC_OK;
}

extern "C" int acc_stream_create(int device, void **stream) {
    if (!stream) return fail(ACC_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags((hipStream_t *)stream, hipStreamNonBlocking));
    return ACC_OK;
}

extern "C" int acc_stream_destroy(void *stream) {
    HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return ACC_OK;
}

extern "C" int acc_stream_synchronize(void *stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return ACC_OK;
}

extern "C" int acc_event_create(void **event) {
    if (!event) return fail(ACC_ERR_ARG, "null argument");
    HIP_TRY(hipEventCreate((hipEvent_t *)event));
    return ACC_OK;
}

extern "C" int acc_event_record(void *event, void *stream) {
    HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return ACC_OK;
}

extern "C" int acc_event_elapsed_ms(void *start, void *end, float *ms) {
    if (!ms) return fail(ACC_ERR_ARG, "null argument");
    HIP_TRY(hipEventSynchronize((hipEvent_t)end));
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return ACC_OK;
}

extern "C" int acc_event_destroy(void *event) {
    HIP_TRY(hipEventDestroy((hipEvent_t)event));
    return ACC_OK;
}


// ---------------------------------------------------------------------------
// upload / info / destroy
// ---------------------------------------------------------------------------

extern "C" int acc_gpu_upload(const acc_csr *csr, const acc_packed *pk, int device, acc_gpu_matrix **out) {
    if (!out || (!csr && !pk)) return fail(ACC_ERR_ARG, "need a CSR and/or packed matrix");
    if (csr && pk && (csr->rows != pk->rows || csr->cols != pk->cols))
        return fail(ACC_ERR_ARG, "CSR and packed matrix shapes differ");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(ACC_ERR_NODEVICE, "no GPU visible");
    if (device < 0 || device >= ndev) return fail(ACC_ERR_ARG, "bad device index");
    HIP_TRY(hipSetDevice(device));
    const int32_t cols = csr ? csr->cols : pk->cols;
    if (cols > 8192) return fail(ACC_ERR_ARG, "cols must be <= 8192");
    acc_gpu_matrix *m = new acc_gpu_matrix();
    m->device = device;
    m->ws.allocs = &m->allocs;
    m->ws.bytes = &m->device_bytes;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) m->cus = prop.multiProcessorCount;
    m->rows = csr ? csr->rows : pk->rows;
    m->cols = cols;
    int rc = ACC_OK;
    std::vector<int4> items;
    std::vector<int32_t> part_wgs;
    std::vector<uint2> sl;
#define TRY(x)                         \
    do {                               \
        rc = (x);                      \
        if (rc != ACC_OK) goto failed; \
    } while (0)
#define COPY(dst, src, bytes)                                                      \
    do {                                                                           \
        if (hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {     \
            rc = fail(ACC_ERR_DEVICE, "host to device copy failed");              \
            goto failed;                                                           \
        }                                                                          \
    } while (0)
#define ZERO(dst, bytes)                                                           \
    do {                                                                           \
        if (hipMemset(dst, 0, bytes) != hipSuccess) {                              \
            rc = fail(ACC_ERR_DEVICE, "device memset failed");                    \
            goto failed;                                                           \
        }                                                                          \
    } while (0)
    if (csr) {
        const int64_t nnz = csr->nnz();
        if (nnz >= (int64_t)UINT32_MAX) { rc = fail(ACC_ERR_RANGE, "FP32 CSR needs nnz < 2^32 on the device"); goto failed; }
        m->has_csr = true;
        m->nnz = nnz;
        m->group = choose_group(m->rows > 0 ? (double)nnz / m->rows : 1.0);
        std::vector<uint32_t> p32(csr->ptr.size());
        for (size_t i = 0; i < p32.size(); ++i) p32[i] = (uint32_t)csr->ptr[i];
        TRY(dalloc(m, &m->d_ptr, p32.size()));
        TRY(dalloc(m, &m->d_idx, (size_t)nnz + 4));  // 4 entries of padding for fp32_spmv's 16-byte loads
        TRY(dalloc(m, &m->d_val, (size_t)nnz + 4));
        TRY(dalloc(m, &m->d_y, (size_t)m->rows));
        COPY(m->d_ptr, p32.data(), p32.size() * 4);
        COPY(m->d_idx, csr->idx.data(), (size_t)nnz * 4);
        COPY(m->d_val, csr->val.data(), (size_t)nnz * 4);
        ZERO(m->d_idx + nnz, 16);
        ZERO(m->d_val + nnz, 16);
        select_geometry(m->rows, m->cus, m->sel_blocks, m->sel_chunk);
        TRY(dalloc(m, &m->d_sthr, 1));
        TRY(dalloc(m, &m->d_scnt, 1));
        TRY(dalloc(m, &m->d_scand, (size_t)m->sel_blocks * kMaxK));
        m->rsel_cap = radix_select_keys(m->rows, 256);
        TRY(dalloc(m, &m->d_rsel, m->rsel_cap));
        ZERO(m->d_sthr, 8);
        ZERO(m->d_scnt, 4);
    }
    if (pk) {
        m->has_packed = true;
        if (!csr) m->nnz = pk->nnz;
        m->split = pk->split;
        m->bits = pk->bits;
        m->col_bits = pk->col_bits;
        m->entry_bits = pk->entry_bits;
        m->is_signed = pk->is_signed;
        m->nslices = pk->nslices();
        m->part_row_begin = pk->part_row_begin;
        sl.resize((size_t)m->nslices);
        for (int64_t s = 0; s < m->nslices; ++s)
            sl[(size_t)s] = make_uint2(pk->slice_off[(size_t)s], pk->slice_ndw[(size_t)s]);
        items = make_items(pk, m->cus, part_wgs);
        m->nitems = (int)items.size();
        m->stream_bytes = (int64_t)pk->words.size() * 4 + m->nslices * 8;
        // E*64 dwords of zero padding: the packet kernels load whole groups
        // unconditionally and mask what lies past a group's length
        TRY(dalloc(m, &m->d_words, pk->words.size() + kWordPad));
        TRY(dalloc(m, &m->d_slices, sl.size()));
        TRY(dalloc(m, &m->d_perm, pk->perm.size()));
        TRY(dalloc(m, &m->d_items, items.size()));
        TRY(dalloc(m, &m->d_smax, (size_t)m->nslices));
        TRY(dalloc(m, &m->d_resc, (size_t)m->split * kMaxM));
        TRY(dalloc(m, &m->d_crow, (size_t)m->split * kMaxM));
        TRY(dalloc(m, &m->d_cscore, (size_t)m->split * kMaxM));
        TRY(dalloc(m, &m->d_part_slice, (size_t)m->split + 1));
        COPY(m->d_part_slice, pk->part_slice_begin.data(), ((size_t)m->split + 1) * 8);
        for (int p = 0; p < m->split; ++p)
            m->max_part_slices = std::max<int64_t>(m->max_part_slices,
                                                   pk->part_slice_begin[(size_t)p + 1] - pk->part_slice_begin[(size_t)p]);
        COPY(m->d_words, pk->words.data(), pk->words.size() * 4);
        ZERO(m->d_words + pk->words.size(), (size_t)kWordPad * 4);
        COPY(m->d_slices, sl.data(), sl.size() * sizeof(uint2));
        COPY(m->d_perm, pk->perm.data(), pk->perm.size() * 4);
        COPY(m->d_items, items.data(), items.size() * sizeof(int4));
        {
            // valid rows per slice: padding lanes (perm -1) close a partition's last slice
            std::vector<uint8_t> nv((size_t)m->nslices, 0);
            for (int64_t sl2 = 0; sl2 < m->nslices; ++sl2) {
                int c = 0;
                while (c < 64 && pk->perm[(size_t)sl2 * 64 + c] >= 0) ++c;
                nv[(size_t)sl2] = (uint8_t)c;
            }
            TRY(dalloc(m, &m->d_nvalid, nv.size()));
            COPY(m->d_nvalid, nv.data(), nv.size());
        }
        m->cols_pad = (m->cols + 31) & ~31;
        m->nkb = m->cols_pad / 32;
        m->u8off = (!m->is_signed && m->bits == 8) ? 1 : 0;
        {
            // a byte of the slice image holds the sum of a row's values in one
            // column (rows may name a column twice): v - 128 for unsigned
            // 8-bit, else v
            const int64_t lo = m->u8off ? 0 : -128, hi = m->u8off ? 255 : 127;
            const int64_t maxv = std::max<int64_t>({(int64_t)pk->merged_max, -(int64_t)pk->merged_min, 1});
            if (pk->merged_min < lo || pk->merged_max > hi)
                m->dense_why = "a row's values in one column sum past one byte (a row names a column twice)";
            else if (maxv * 128 * (int64_t)m->cols >= (int64_t(1) << 25))
                m->dense_why = "scores may reach 2^25 (too many columns for the value width)";
            else if (m->cols > kPkMaxCols)
                m->dense_why = "more than 1024 columns";
            else
                m->dense_ok = true;
        }
        if (m->dense_ok) {
            // the matrix-core scores' own copy of the slices: image bytes and
            // their offsets in a decode wavefront's 16 image rows, repeated
            // columns merged (k_pkmfma.hpp)
            const uint32_t stride = (uint32_t)m->cols_pad + 16u;
            m->pcb = acc::bits_for((int32_t)(16u * stride));
            if (8 + m->pcb < kPkMinE || 8 + m->pcb > kPkMaxE) {
                rc = fail(ACC_ERR_ARG, "image-packet entry width out of range");
                goto failed;
            }
            std::vector<uint32_t> pw;
            std::vector<std::pair<uint32_t, uint32_t>> ps;
            TRY(acc::image_packets(pk, m->u8off, m->pcb, stride, pw, ps));
            std::vector<uint2> psl(ps.size());
            for (size_t i = 0; i < ps.size(); ++i) psl[i] = make_uint2(ps[i].first, ps[i].second);
            TRY(dalloc(m, &m->d_pwords, pw.size() + kWordPad));
            TRY(dalloc(m, &m->d_pslices, psl.size()));
            COPY(m->d_pwords, pw.data(), pw.size() * 4);
            ZERO(m->d_pwords + pw.size(), (size_t)kWordPad * 4);
            COPY(m->d_pslices, psl.data(), psl.size() * sizeof(uint2));
            m->pbytes = (int64_t)pw.size() * 4 + (int64_t)psl.size() * 8;
        }
    }
    TRY(dalloc(m, &m->d_xstage, (size_t)m->cols));
    TRY(dalloc(m, &m->d_ostage, (size_t)kMaxCK));
    if (hipDeviceSynchronize() != hipSuccess) { rc = fail(ACC_ERR_DEVICE, "device synchronize failed"); goto failed; }
    *out = m;
    return ACC_OK;
failed:
    free_all(m);
    delete m;
    return rc;
#undef TRY
#undef COPY
#undef ZERO
}

extern "C" int acc_gpu_matrix_info(const acc_gpu_matrix *m, acc_gpu_info_t *info) {
    if (!m || !info) return fail(ACC_ERR_ARG, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->rows = m->rows;
    info->cols = m->cols;
    info->split = m->split;
    info->bits = m->bits;
    info->has_csr = m->has_csr;
    info->has_packed = m->has_packed;
    info->group = m->group;
    info->work_items = m->nitems;
    info->nnz = m->nnz;
    info->csr_bytes = m->has_csr ? m->nnz * 8 + ((int64_t)m->rows + 1) * 4 : 0;
    info->stream_bytes = m->stream_bytes;
    info->device_bytes = m->device_bytes;
    info->batch_mode = m->batch_mode;
    info->dense_ok = m->dense_ok ? 1 : 0;
    info->dense_bytes = m->pbytes;
    return ACC_OK;
}

extern "C" void acc_gpu_matrix_destroy(acc_gpu_matrix *m) {
    free_all(m);
    delete m;
}

extern "C" int acc_gpu_set_batch_mode(acc_gpu_matrix *m, int mode, int min_dense_batch) {
    if (!m) return fail(ACC_ERR_ARG, "null matrix");
    if (mode != ACC_BATCH_AUTO && mode != ACC_BATCH_SPARSE && mode != ACC_BATCH_DENSE && mode != ACC_BATCH_SINGLE)
        return fail(ACC_ERR_ARG, "mode must be ACC_BATCH_AUTO, _SINGLE, _SPARSE or _DENSE");
    if (min_dense_batch < 0) return fail(ACC_ERR_ARG, "min_dense_batch must be >= 0");
    if (mode != ACC_BATCH_SPARSE && !m->has_packed) return fail(ACC_ERR_ARG, "matrix was uploaded without low-bit packets");
    if (mode == ACC_BATCH_DENSE && !m->dense_ok) return fail(ACC_ERR_ARG, "matrix-core scores unavailable: " + m->dense_why);
    m->batch_mode = mode;
    if (min_dense_batch > 0) {
        m->dense_min_nq = min_dense_batch;
        m->ntune = 0;  // an explicit crossover replaces a tuned table
    }
    return ACC_OK;
}

// ---------------------------------------------------------------------------
// low-bit queries
// ---------------------------------------------------------------------------

extern "C" int acc_gpu_topk_lowbit(acc_gpu_matrix *m, const float *x_dev, int k, int ck, acc_element *out_dev,
                                   void *stream) {
    int rc = check_lowbit(m, ck, k);
    if (rc != ACC_OK) return rc;
    if (!x_dev || k < 1 || !out_dev) return fail(ACC_ERR_ARG, "null query or output, or k < 1");
    HIP_TRY(hipSetDevice(m->device));
    return launch_lowbit(m, x_dev, ck / m->split, k, out_dev, nullptr, nullptr, (hipStream_t)stream);
}

// Batched queries with the handle's or a plan's workspace.  k_host: per-query
// K on the host; upload_k copies it into the workspace (false under graph
// capture, where the eager run already did).
static int run_batch(acc_gpu_matrix *m, BatchWs &ws, const float *x_dev, int nq, const int32_t *k_host, bool upload_k,
                     int kmax, int ck, acc_element *out_dev, int out_stride, hipStream_t st) {
    const int M = ck / m->split;
    int layout = batch_layout(m, nq);
    // the packet kernels' query table of a 16-query block must fit in LDS
    if (layout == ACC_BATCH_SPARSE && (size_t)m->cols * (16 + 8) > 65536) layout = ACC_BATCH_SINGLE;
    if (layout == ACC_BATCH_SINGLE) {
        int rc = ACC_OK;
        for (int q = 0; q < nq && rc == ACC_OK; ++q)
            rc = launch_lowbit(m, x_dev + (size_t)q * m->cols, M, k_host[q], out_dev + (size_t)q * out_stride, nullptr,
                               nullptr, st, &ws);
        return rc;
    }
    int rc = ws_grow(ws, &ws.K, &ws.K_n, (size_t)nq);
    if (rc != ACC_OK) return rc;
    if (upload_k) HIP_TRY(hipMemcpyAsync(ws.K, k_host, (size_t)nq * 4, hipMemcpyHostToDevice, st));
    if (layout == ACC_BATCH_DENSE) return launch_dense(m, ws, x_dev, nq, M, kmax, out_dev, out_stride, nullptr, nullptr, st);
    const size_t nslots = (size_t)m->nslices * 64;
    rc = ws_grow(ws, &ws.S, &ws.S_n, (size_t)nq * nslots);
    if (rc == ACC_OK) rc = ws_grow(ws, &ws.resc, &ws.resc_n, (size_t)nq * m->split * M);
    if (rc != ACC_OK) return rc;
    return launch_batch(m, ws, x_dev, nq, M, kmax, out_dev, out_stride, nullptr, nullptr, st);
}

static int check_batch(acc_gpu_matrix *m, const float *x_dev, int nq, const int32_t *k_host, int ck,
                       acc_element *out_dev, int out_stride, int *kmax) {
    if (!m || !x_dev || !k_host || nq < 0 || !out_dev) return fail(ACC_ERR_ARG, "bad batch arguments");
    *kmax = 1;
    for (int q = 0; q < nq; ++q) {
        if (k_host[q] < 1 || k_host[q] > out_stride) return fail(ACC_ERR_ARG, "each k must be in [1, out_stride]");
        *kmax = std::max(*kmax, (int)k_host[q]);
    }
    return check_lowbit(m, ck, *kmax);
}

// A batch workspace of its own, for callers that keep several batches of one
// matrix in flight (the server's slots); like a plan's, grown on demand.
namespace acc {
struct BatchWorkspace {
    std::vector<void *> allocs;
    int64_t bytes = 0;
    BatchWs ws;
};

BatchWorkspace *batch_ws_create() {
    BatchWorkspace *w = new BatchWorkspace();
    w->ws.allocs = &w->allocs;
    w->ws.bytes = &w->bytes;
    return w;
}

void batch_ws_destroy(BatchWorkspace *w, int device) {
    if (!w) return;
    hipSetDevice(device);
    for (void *q : w->allocs) hipFree(q);
    delete w;
}

int64_t batch_ws_bytes(const BatchWorkspace *w) { return w ? w->bytes : 0; }

int topk_lowbit_batch_ws(acc_gpu_matrix *m, BatchWorkspace *w, const float *x_dev, int nq, const int32_t *k_host,
                         int ck, acc_element *out_dev, int out_stride, void *stream) {
    if (!w) return fail(ACC_ERR_ARG, "null workspace");
    int kmax = 1;
    int rc = check_batch(m, x_dev, nq, k_host, ck, out_dev, out_stride, &kmax);
    if (rc != ACC_OK || nq == 0) return rc;
    HIP_TRY(hipSetDevice(m->device));
    return run_batch(m, w->ws, x_dev, nq, k_host, true, kmax, ck, out_dev, out_stride, (hipStream_t)stream);
}
}  // namespace acc

extern "C" int acc_gpu_topk_lowbit_batch(acc_gpu_matrix *m, const float *x_dev, int nq, const int32_t *k_host, int ck,
                                         acc_element *out_dev, int out_stride, void *stream) {
    int kmax = 1;
    int rc = check_batch(m, x_dev, nq, k_host, ck, out_dev, out_stride, &kmax);
    if (rc != ACC_OK || nq == 0) return rc;
    HIP_TRY(hipSetDevice(m->device));
    return run_batch(m, m->ws, x_dev, nq, k_host, true, kmax, ck, out_dev, out_stride, (hipStream_t)stream);
}

extern "C" int acc_gpu_scores_lowbit(acc_gpu_matrix *m, const float *x_dev, int32_t *y_dev, void *stream) {
    if (!m || !m->has_packed || !x_dev || !y_dev) return fail(ACC_ERR_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    ScoresArgs a;
    a.words = m->d_words;
    a.slices = m->d_slices;
    a.perm = m->d_perm;
    a.nslices = m->nslices;
    a.x = x_dev;
    a.cols = m->cols;
    a.col_bits = m->col_bits;
    a.val_bits = m->bits;
    a.y = y_dev;
    scores_fn k = get_lowbit_scores(m->entry_bits, m->is_signed != 0);
    if (!k) return fail(ACC_ERR_ARG, "unsupported entry width");
    int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((m->nslices + ACC_WAVES - 1) / ACC_WAVES, m->cus * 8));
    hipLaunchKernelGGL(k, dim3(blocks), dim3(ACC_WG), 64 + (size_t)((m->cols + 15) & ~15), (hipStream_t)stream, a);
    return check_launch();
}

extern "C" int acc_gpu_candidates_lowbit(acc_gpu_matrix *m, const float *x_dev, int ck, int32_t *rows_dev,
                                         int32_t *scores_dev, int *count_host, void *stream) {
    int rc = check_lowbit(m, ck, 0);
    if (rc != ACC_OK) return rc;
    if (!x_dev || !rows_dev || !scores_dev) return fail(ACC_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(m->device));
    hipStream_t st = (hipStream_t)stream;
    const int M = ck / m->split;
    rc = launch_lowbit(m, x_dev, M, 0, nullptr, m->d_crow, m->d_cscore, st);
    int total = 0;
    for (int p = 0; rc == ACC_OK && p < m->split; ++p) {
        const int n = std::min(M, m->part_row_begin[(size_t)p + 1] - m->part_row_begin[(size_t)p]);
        if (n > 0 && (hipMemcpyAsync(rows_dev + total, m->d_crow + (size_t)p * M, (size_t)n * 4, hipMemcpyDeviceToDevice, st) != hipSuccess ||
                      hipMemcpyAsync(scores_dev + total, m->d_cscore + (size_t)p * M, (size_t)n * 4, hipMemcpyDeviceToDevice, st) != hipSuccess))
            rc = fail(ACC_ERR_DEVICE, "copy of candidates failed");
        total += n;
    }
    if (rc == ACC_OK && hipStreamSynchronize(st) != hipSuccess) rc = fail(ACC_ERR_DEVICE, "stream synchronize failed");
    if (count_host) *count_host = total;
    return rc;
}

extern "C" int acc_gpu_spmv_fp64(acc_gpu_matrix *m, const float *x_dev, double *y_dev, void *stream) {
    if (!m || !m->has_csr || !x_dev || !y_dev) return fail(ACC_ERR_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    int blocks = std::max(1, std::min((m->rows + 31) / 32, m->cus * 8));
    hipLaunchKernelGGL(get_spmv_fp64(), dim3(blocks), dim3(ACC_WG), 0, (hipStream_t)stream, m->d_ptr, m->d_idx, m->d_val,
                       x_dev, m->rows, y_dev);
    return check_launch();
}

// ---------------------------------------------------------------------------
// FP32 CSR path
// ---------------------------------------------------------------------------

size_t fp32_spmv_lds(int cols, int G);  // k_fp32.hip

static int launch_spmv_fp32(acc_gpu_matrix *m, const float *x_dev, float *y, hipStream_t st) {
    spmv_fn k = get_fp32_spmv(m->group);
    if (!k) return fail(ACC_ERR_ARG, "unsupported group size");
    const int rpi = 64 / m->group;
    const int nblocks = (m->rows + rpi - 1) / rpi;
    // as many workgroups as run at once (each wavefront then owns one contiguous row range, and no
    // second wave of workgroups leaves part of the chip idle at the end), fewer for small matrices
    const size_t lds = fp32_spmv_lds(m->cols, m->group);
    int occ = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)k, ACC_WG, lds) != hipSuccess || occ < 1)
        occ = 1;
    const int blocks = std::max(1, std::min((nblocks + ACC_WAVES * 4 - 1) / (ACC_WAVES * 4), m->cus * occ));
    LAUNCH(m, 0, k, dim3(blocks), dim3(ACC_WG), lds, st, m->d_ptr, m->d_idx, m->d_val, x_dev, m->rows, m->cols, y);
    return check_launch();
}

extern "C" int acc_gpu_spmv_fp32(acc_gpu_matrix *m, const float *x_dev, float *y_dev, void *stream) {
    if (!m || !m->has_csr || !x_dev || !y_dev) return fail(ACC_ERR_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    return launch_spmv_fp32(m, x_dev, y_dev, (hipStream_t)stream);
}

extern "C" int acc_gpu_topk_fp32(acc_gpu_matrix *m, const float *x_dev, int k, acc_element *out_dev, void *stream) {
    if (!m || !m->has_csr) return fail(ACC_ERR_ARG, "matrix was uploaded without the FP32 CSR");
    if (!x_dev || !out_dev || k < 1 || k > kMaxK) return fail(ACC_ERR_ARG, "k must be in [1, 960]");
    HIP_TRY(hipSetDevice(m->device));
    hipStream_t st = (hipStream_t)stream;
    int rc = launch_spmv_fp32(m, x_dev, m->d_y, st);
    if (rc != ACC_OK) return rc;
    return launch_select(m->d_y, m->rows, k, m->sel_blocks, m->sel_chunk, m->d_sthr, m->d_scnt, m->d_scand,
                         out_dev, st, m, m->d_rsel, m->rsel_cap);
}

extern "C" int acc_gpu_set_profiling(acc_gpu_matrix *m, int enable) {
    if (!m) return fail(ACC_ERR_ARG, "null matrix");
    HIP_TRY(hipSetDevice(m->device));
    if (enable)
        for (auto &e : m->ev)
            if (!e) HIP_TRY(hipEventCreate(&e));
    m->profiling = enable != 0;
    m->nev = 0;
    return ACC_OK;
}

extern "C" int acc_gpu_stage_times(acc_gpu_matrix *m, float *ms, int n, int *count) {
    if (!m || (!ms && n > 0)) return fail(ACC_ERR_ARG, "null argument");
    if (!m->profiling || m->nev < 2) return fail(ACC_ERR_ARG, "no profiled query (enable profiling first)");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipEventSynchronize(m->ev[m->nev - 1]));
    int c = 0;
    for (int i = 0; 2 * i + 1 < m->nev && c < n; ++i, ++c) HIP_TRY(hipEventElapsedTime(&ms[c], m->ev[2 * i], m->ev[2 * i + 1]));
    if (count) *count = c;
    return ACC_OK;
}

extern "C" int acc_gpu_time_lowbit_stage(acc_gpu_matrix *m, const float *x_dev, int k, int ck, int stage, int reps,
                                         void *stream, float *avg_ms) {
    int rc = check_lowbit(m, ck, k);
    if (rc != ACC_OK) return rc;
    if (!x_dev || k < 1 || stage < 0 || stage > 2 || reps < 1 || !avg_ms) return fail(ACC_ERR_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    hipStream_t st = (hipStream_t)stream;
    const bool prof = m->profiling;
    m->profiling = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const LowbitArgs a = lowbit_args(m, x_dev, ck / m->split, k, m->d_ostage, nullptr, nullptr);
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(ACC_ERR_DEVICE, "hipEventCreate");
    // one whole query leaves the state the later stages read (slice maxima, rescored candidates)
    for (int s = 0; s < 3 && rc == ACC_OK; ++s) rc = launch_lowbit_stage(m, a, s, st);
    if (rc == ACC_OK && hipEventRecord(e0, st) != hipSuccess) rc = fail(ACC_ERR_DEVICE, "hipEventRecord");
    for (int r = 0; r < reps && rc == ACC_OK; ++r) rc = launch_lowbit_stage(m, a, stage, st);
    if (rc == ACC_OK && hipEventRecord(e1, st) != hipSuccess) rc = fail(ACC_ERR_DEVICE, "hipEventRecord");
    float ms = 0.f;
    if (rc == ACC_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = fail(ACC_ERR_DEVICE, "event timing failed");
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    m->profiling = prof;
    if (rc == ACC_OK) *avg_ms = ms / reps;
    return rc;
}

extern "C" int acc_gpu_top_k_heap(const float *y_dev, int n, int k, acc_element *out_dev, void *stream) {
    if (!y_dev || !out_dev || n < 0 || k < 1 || k > kMaxK) return fail(ACC_ERR_ARG, "bad arguments (k in [1, 960])");
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    int blocks = 1, chunk = 8192;
    select_geometry(n, cus, blocks, chunk);
    uint64_t *thr = nullptr, *cand = nullptr;
    uint32_t *cnt = nullptr;
    HIP_TRY(hipMallocAsync((void **)&thr, 8, st));
    HIP_TRY(hipMallocAsync((void **)&cnt, 4, st));
    HIP_TRY(hipMallocAsync((void **)&cand, (size_t)blocks * k * 8, st));
    uint64_t *rbuf = nullptr;
    const size_t rcap = k <= 256 ? radix_select_keys(n, k) : 0;
    if (rcap) HIP_TRY(hipMallocAsync((void **)&rbuf, rcap * 8, st));
    hipLaunchKernelGGL(get_select_reset(), dim3(1), dim3(64), 0, st, thr, cnt);
    int rc = check_launch();
    if (rc == ACC_OK) rc = launch_select(y_dev, n, k, blocks, chunk, thr, cnt, cand, out_dev, st, nullptr, rbuf, rcap);
    hipFreeAsync(thr, st);
    hipFreeAsync(cnt, st);
    hipFreeAsync(cand, st);
    if (rbuf) hipFreeAsync(rbuf, st);
    return rc;
}

// ---------------------------------------------------------------------------
// top_k_heap without allocation: selection scratch in a static device array
// ---------------------------------------------------------------------------

// Keys of the static selection scratch: [0] the shared threshold, [1] the
// candidate count (low 32 bits), then the candidate keys (at most 512
// workgroups x 960 keys, or the radix passes' keys when they fit).  3.75 MiB
// per device, held until the process ends.
static constexpr int kNoAllocBlocks = 512;
static constexpr size_t kScratchKeys = 2 + (size_t)kNoAllocBlocks * kMaxK;
static __device__ __align__(256) uint64_t scratch[kScratchKeys];

// Device address of `scratch` on the current device (looked up once per
// device; the array's host-side name is not a device address).
static int scratch_address(uint64_t **out) {
    static std::mutex mu;
    static std::vector<uint64_t *> addr;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)dev >= addr.size()) addr.resize((size_t)dev + 1, nullptr);
    if (!addr[(size_t)dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(ACC_ERR_DEVICE, "hipGetSymbolAddress returned a null address");
        addr[(size_t)dev] = (uint64_t *)p;
    }
    *out = addr[(size_t)dev];
    return ACC_OK;
}

extern "C" int acc_gpu_top_k_heap_noalloc(const float *y_dev, int n, int k, acc_element *out_dev, void *stream) {
    if (!out_dev || (n > 0 && !y_dev) || n < 0 || k < 1 || k > kMaxK)
        return fail(ACC_ERR_ARG, "bad arguments (n >= 0, k in [1, 960])");
    uint64_t *s = nullptr;
    int rc = scratch_address(&s);
    if (rc != ACC_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    // at most kNoAllocBlocks workgroups, each appending at most k keys
    const int chunk = std::max(8192, (int)(((int64_t)n + kNoAllocBlocks - 1) / kNoAllocBlocks));
    const int blocks = std::max(1, (int)(((int64_t)n + chunk - 1) / chunk));
    if ((size_t)blocks * k > kScratchKeys - 2) return fail(ACC_ERR_RANGE, "selection scratch too small for n and k");
    uint64_t *thr = s, *cand = s + 2;
    uint32_t *cnt = (uint32_t *)(s + 1);
    // the previous call may have left any state: the threshold and count
    // start at zero (a kernel, not a memset, so the call can be captured)
    hipLaunchKernelGGL(get_select_reset(), dim3(1), dim3(64), 0, st, thr, cnt);
    rc = check_launch();
    if (rc != ACC_OK) return rc;
    return launch_select(y_dev, n, k, blocks, chunk, thr, cnt, cand, out_dev, st, nullptr, cand, kScratchKeys - 2);
}

// ---------------------------------------------------------------------------
// autotuner for ACC_BATCH_AUTO
// ---------------------------------------------------------------------------

extern "C" int acc_gpu_autotune(acc_gpu_matrix *m, int ck, int k, int max_batch, acc_tune_point *points, int cap,
                                int *count) {
    int rc = check_lowbit(m, ck, k);
    if (rc != ACC_OK) return rc;
    if (k < 1 || max_batch < 1 || max_batch > 4096 || cap < 0 || (cap > 0 && !points))
        return fail(ACC_ERR_ARG, "k >= 1, max_batch in [1, 4096] and room for the points are required");
    HIP_TRY(hipSetDevice(m->device));
    // the dense layout competes unless the caller pinned the packet or single-query layout
    const bool try_dense = m->dense_ok && m->batch_mode != ACC_BATCH_SPARSE && m->batch_mode != ACC_BATCH_SINGLE;
    std::vector<float> xh((size_t)max_batch * m->cols);
    acc_generate_queries(max_batch, m->cols, ACC_VAL_NORMAL, 0, 0x7E57ull, xh.data());
    float *x = nullptr;
    acc_element *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<void *> tmp;
    int64_t tmp_bytes = 0;
    size_t xn = 0, on = 0;
    rc = grow_in(tmp, tmp_bytes, &x, &xn, xh.size());
    if (rc == ACC_OK) rc = grow_in(tmp, tmp_bytes, &out, &on, (size_t)max_batch * k);
    if (rc == ACC_OK && (hipMemcpy(x, xh.data(), xh.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
                         hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                         hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(ACC_ERR_DEVICE, "autotune setup failed");
    const int saved_mode = m->batch_mode;
    const bool prof = m->profiling;
    m->profiling = false;
    m->ntune = 0;
    int n = 0;
    for (int nq = 1; rc == ACC_OK && nq <= max_batch && m->ntune < 16; nq = nq < max_batch && nq * 2 > max_batch ? max_batch : nq * 2) {
        std::vector<int32_t> ks((size_t)nq, k);
        float us[3] = {-1.f, -1.f, -1.f};
        const int layouts[3] = {ACC_BATCH_SINGLE, ACC_BATCH_SPARSE, ACC_BATCH_DENSE};
        for (int li = 0; li < 3 && rc == ACC_OK; ++li) {
            if (layouts[li] == ACC_BATCH_DENSE && !try_dense) continue;
            if (layouts[li] == ACC_BATCH_SPARSE && (size_t)m->cols * (16 + 8) > 65536) continue;
            m->batch_mode = layouts[li];
            float best = 1e30f;
            for (int r = 0; r < 4 && rc == ACC_OK; ++r) {  // the first run warms up
                hipEventRecord(e0, st);
                rc = run_batch(m, m->ws, x, nq, ks.data(), true, k, ck, out, k, st);
                hipEventRecord(e1, st);
                float ms = 0.f;
                if (rc == ACC_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                    rc = fail(ACC_ERR_DEVICE, "autotune timing failed");
                if (r > 0) best = std::min(best, ms * 1000.f);
            }
            us[li] = best;
        }
        int pick = ACC_BATCH_SINGLE;
        float t = us[0];
        for (int li = 1; li < 3; ++li)
            if (us[li] >= 0.f && us[li] < t) {
                t = us[li];
                pick = layouts[li];
            }
        m->tune_nq[m->ntune] = nq;
        m->tune_layout[m->ntune] = pick;
        ++m->ntune;
        if (n < cap) {
            points[n].nq = nq;
            points[n].layout = pick;
            points[n].us_single = us[0];
            points[n].us_sparse = us[1];
            points[n].us_dense = us[2];
            ++n;
        }
        if (nq == max_batch) break;
    }
    m->profiling = prof;
    m->batch_mode = rc == ACC_OK ? ACC_BATCH_AUTO : saved_mode;
    if (rc != ACC_OK) m->ntune = 0;
    if (count) *count = n;
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    for (void *q : tmp) hipFree(q);
    return rc;
}

// ---------------------------------------------------------------------------
// plans: a query launch sequence captured once into a HIP graph
// ---------------------------------------------------------------------------

struct acc_plan {
    acc_gpu_matrix *m = nullptr;
    int kind = 0;
    const float *x = nullptr;
    int nq = 0, ck = 0, kmax = 0, out_stride = 0;
    acc_element *out = nullptr;
    std::vector<int32_t> k;      // per-query K (host copy)
    std::vector<void *> allocs;  // the plan's own batch workspace
    int64_t bytes = 0;
    BatchWs ws;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void plan_free(acc_plan *p) {
    if (!p) return;
    if (p->m) hipSetDevice(p->m->device);
    if (p->exec) hipGraphExecDestroy(p->exec);
    if (p->graph) hipGraphDestroy(p->graph);
    if (p->cap) hipStreamDestroy(p->cap);
    for (void *q : p->allocs) hipFree(q);
    delete p;
}

// The plan's launches; upload_k is false under capture (K is already in the
// plan's workspace).
static int plan_enqueue(acc_plan *p, bool upload_k, hipStream_t st) {
    acc_gpu_matrix *m = p->m;
    switch (p->kind) {
        case ACC_PLAN_LOWBIT:
            return launch_lowbit(m, p->x, p->ck / m->split, p->kmax, p->out, nullptr, nullptr, st, &p->ws);
        case ACC_PLAN_LOWBIT_BATCH:
            return run_batch(m, p->ws, p->x, p->nq, p->k.data(), upload_k, p->kmax, p->ck, p->out, p->out_stride, st);
        case ACC_PLAN_FP32: {
            const size_t rcap = radix_select_keys(m->rows, 256);
            BatchWs &w = p->ws;
            const bool fresh = w.sthr == nullptr;
            int rc = ws_grow(w, &w.y, &w.y_n, (size_t)m->rows);
            if (rc == ACC_OK) rc = ws_grow(w, &w.rsel, &w.rsel_n, rcap);
            // the plan's own select_blocks state, so that plans of one handle
            // can run concurrently for any K; zeroed once by a kernel (the
            // kernels leave it zero), before capture
            if (rc == ACC_OK) rc = ws_grow(w, &w.sthr, &w.sthr_n, 1);
            if (rc == ACC_OK) rc = ws_grow(w, &w.scnt, &w.scnt_n, 1);
            if (rc == ACC_OK) rc = ws_grow(w, &w.scand, &w.scand_n, (size_t)m->sel_blocks * kMaxK);
            if (rc == ACC_OK && fresh) {
                hipLaunchKernelGGL(get_select_reset(), dim3(1), dim3(64), 0, st, w.sthr, w.scnt);
                rc = check_launch();
            }
            if (rc == ACC_OK) rc = launch_spmv_fp32(m, p->x, w.y, st);
            if (rc != ACC_OK) return rc;
            return launch_select(w.y, m->rows, p->kmax, m->sel_blocks, m->sel_chunk, w.sthr, w.scnt, w.scand, p->out,
                                 st, nullptr, w.rsel, rcap);
        }
        default:
            return fail(ACC_ERR_ARG, "unknown plan kind");
    }
}

extern "C" int acc_gpu_plan_create(acc_gpu_matrix *m, int kind, const float *x_dev, int nq, const int32_t *k_host,
                                   int ck, acc_element *out_dev, int out_stride, acc_plan **out) {
    if (!m || !out || !x_dev || !out_dev || !k_host || nq < 1) return fail(ACC_ERR_ARG, "bad plan arguments");
    int kmax = 1, rc = ACC_OK;
    if (kind == ACC_PLAN_LOWBIT_BATCH) {
        rc = check_batch(m, x_dev, nq, k_host, ck, out_dev, out_stride, &kmax);
    } else if (kind == ACC_PLAN_LOWBIT || kind == ACC_PLAN_FP32) {
        if (nq != 1) return fail(ACC_ERR_ARG, "single-query plans take nq = 1");
        kmax = k_host[0];
        if (kind == ACC_PLAN_LOWBIT) {
            rc = check_lowbit(m, ck, kmax);
            if (rc == ACC_OK && kmax < 1) rc = fail(ACC_ERR_ARG, "k must be >= 1");
        } else if (!m->has_csr || kmax < 1 || kmax > kMaxK) {
            rc = fail(ACC_ERR_ARG, "FP32 plans need the CSR and k in [1, 960]");
        }
    } else {
        rc = fail(ACC_ERR_ARG, "kind must be ACC_PLAN_LOWBIT, ACC_PLAN_LOWBIT_BATCH or ACC_PLAN_FP32");
    }
    if (rc != ACC_OK) return rc;
    HIP_TRY(hipSetDevice(m->device));
    acc_plan *p = new acc_plan();
    p->m = m;
    p->kind = kind;
    p->x = x_dev;
    p->nq = nq;
    p->ck = ck;
    p->kmax = kmax;
    p->out = out_dev;
    p->out_stride = out_stride;
    p->k.assign(k_host, k_host + nq);
    p->ws.allocs = &p->allocs;
    p->ws.bytes = &p->bytes;
    const bool prof = m->profiling;
    m->profiling = false;
    hipGraph_t g = nullptr;
    if (hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking) != hipSuccess) {
        rc = fail(ACC_ERR_DEVICE, "stream creation failed");
    } else {
        // an eager run first: builds what the handle makes lazily, sizes the
        // plan's workspace, uploads K and raises kernel limits outside capture
        rc = plan_enqueue(p, true, p->cap);
        if (rc == ACC_OK && hipStreamSynchronize(p->cap) != hipSuccess) rc = fail(ACC_ERR_DEVICE, "eager run failed");
        if (rc == ACC_OK) {
            if (hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                rc = fail(ACC_ERR_DEVICE, "stream capture failed to start");
            } else {
                rc = plan_enqueue(p, false, p->cap);
                const hipError_t e = hipStreamEndCapture(p->cap, &g);
                if (rc == ACC_OK && e != hipSuccess)
                    rc = fail(ACC_ERR_DEVICE, std::string("stream capture: ") + hipGetErrorString(e));
            }
        }
    }
    m->profiling = prof;
    p->graph = g;
    if (rc == ACC_OK && hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess)
        rc = fail(ACC_ERR_DEVICE, "graph instantiation failed");
    if (rc != ACC_OK) {
        plan_free(p);
        return rc;
    }
    *out = p;
    return ACC_OK;
}

extern "C" int acc_gpu_plan_launch(acc_plan *p, void *stream) {
    if (!p) return fail(ACC_ERR_ARG, "null plan");
    HIP_TRY(hipSetDevice(p->m->device));
    HIP_TRY(hipGraphLaunch(p->exec, (hipStream_t)stream));
    return ACC_OK;
}

extern "C" void acc_gpu_plan_destroy(acc_plan *p) { plan_free(p); }

// ---------------------------------------------------------------------------
// host-pointer convenience forms and the top_k_spmv drop-in
// ---------------------------------------------------------------------------

extern "C" int acc_topk_fp32_host(acc_gpu_matrix *m, const float *x_host, int k, acc_element *out_host) {
    if (!m || !x_host || !out_host || k < 1 || k > kMaxK) return fail(ACC_ERR_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipMemcpy(m->d_xstage, x_host, (size_t)m->cols * 4, hipMemcpyHostToDevice));
    int rc = acc_gpu_topk_fp32(m, m->d_xstage, k, m->d_ostage, nullptr);
    if (rc != ACC_OK) return rc;
    HIP_TRY(hipMemcpy(out_host, m->d_ostage, (size_t)k * sizeof(acc_element), hipMemcpyDeviceToHost));
    return ACC_OK;
}

extern "C" int acc_topk_lowbit_host(acc_gpu_matrix *m, const float *x_host, int k, int ck, acc_element *out_host) {
    if (!m || !x_host || !out_host || k < 1 || k > kMaxK) return fail(ACC_ERR_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipMemcpy(m->d_xstage, x_host, (size_t)m->cols * 4, hipMemcpyHostToDevice));
    int rc = acc_gpu_topk_lowbit(m, m->d_xstage, k, ck, m->d_ostage, nullptr);
    if (rc != ACC_OK) return rc;
    HIP_TRY(hipMemcpy(out_host, m->d_ostage, (size_t)k * sizeof(acc_element), hipMemcpyDeviceToHost));
    return ACC_OK;
}

// The reference's src/cpu query loop: query t is U(-1, 1) from the seed
// 0x5EED + t (generate_random_vector, cpu_mkl_baseline.cpp:111-122, is
// unseeded there), the SpMV and the Top-k are timed separately with events
// on the null stream, and the last query's results are returned.  Everything
// it touches on the device belongs to the handle (scores, selection state,
// staging), so a call allocates nothing; every runtime status is checked.
extern "C" int acc_top_k_spmv(acc_gpu_matrix *A, int row, int col, int k, int num_threads, acc_element *top_k,
                              int num_test) {
    (void)num_threads;
    if (!A || !top_k || num_test < 1) return fail(ACC_ERR_ARG, "bad arguments");
    if (!A->has_csr) return fail(ACC_ERR_ARG, "matrix was uploaded without the FP32 CSR");
    if (row != A->rows || col != A->cols) return fail(ACC_ERR_ARG, "row/col do not match the matrix");
    if (k < 1 || k > kMaxK) return fail(ACC_ERR_ARG, "k must be in [1, 960]");
    HIP_TRY(hipSetDevice(A->device));
    std::vector<float> x((size_t)col);
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    int rc = ACC_OK;
    for (auto &e : ev)
        if (rc == ACC_OK && hipEventCreate(&e) != hipSuccess) rc = fail(ACC_ERR_DEVICE, "hipEventCreate failed");
    double cpt = 0, srt = 0;
    for (int t = 0; t < num_test && rc == ACC_OK; ++t) {
        rc = acc_generate_queries(1, col, ACC_VAL_UNIFORM, 0, 0x5EEDull + (uint64_t)t, x.data());
        if (rc != ACC_OK) break;
        if (hipMemcpy(A->d_xstage, x.data(), (size_t)col * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipEventRecord(ev[0], nullptr) != hipSuccess) {
            rc = fail(ACC_ERR_DEVICE, "query upload failed");
            break;
        }
        rc = launch_spmv_fp32(A, A->d_xstage, A->d_y, nullptr);
        if (rc == ACC_OK && hipEventRecord(ev[1], nullptr) != hipSuccess) rc = fail(ACC_ERR_DEVICE, "hipEventRecord failed");
        if (rc == ACC_OK)
            rc = launch_select(A->d_y, A->rows, k, A->sel_blocks, A->sel_chunk, A->d_sthr, A->d_scnt, A->d_scand,
                               A->d_ostage, nullptr, nullptr, A->d_rsel, A->rsel_cap);
        if (rc != ACC_OK) break;
        float a = 0.f, b = 0.f;
        if (hipEventRecord(ev[2], nullptr) != hipSuccess || hipEventSynchronize(ev[2]) != hipSuccess ||
            hipEventElapsedTime(&a, ev[0], ev[1]) != hipSuccess || hipEventElapsedTime(&b, ev[1], ev[2]) != hipSuccess) {
            rc = fail(ACC_ERR_DEVICE, "query timing failed");
            break;
        }
        cpt += a;
        srt += b;
    }
    if (rc == ACC_OK && hipMemcpy(top_k, A->d_ostage, (size_t)k * sizeof(acc_element), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(ACC_ERR_DEVICE, "copy of results failed");
    if (rc == ACC_OK) {
        std::printf("compute time: %g ms\n", cpt / num_test);
        std::printf("sort time: %g ms\n", srt / num_test);
        std::printf("exec time: %gms\n", (cpt + srt) / num_test);
        std::fflush(stdout);
    }
    for (auto &e : ev)
        if (e) hipEventDestroy(e);
    return rc;
}
