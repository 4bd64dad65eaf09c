// gpu.hip -- device-resident matrices and the GPU entry points of the C ABI.
#include "kernels.hpp"
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

using acc::fail;

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(ACC_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

static constexpr int kMaxM = 64;     // rows kept per partition
static constexpr int kMaxK = 960;    // fused FP32 Top-K limit (collector cap 1024)

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
    // low-bit sliced packets
    bool has_packed = false;
    int bits = 8, col_bits = 10, entry_bits = 18, is_signed = 0;
    int64_t nslices = 0, stream_bytes = 0;
    uint32_t *d_words = nullptr;
    uint2 *d_slices = nullptr;
    int32_t *d_perm = nullptr;
    int4 *d_items = nullptr;
    int nitems = 0;
    int32_t *d_part_row_begin = nullptr;
    std::vector<int32_t> part_row_begin;
    uint64_t *d_thr = nullptr;      // split
    uint32_t *d_cnt = nullptr;      // split
    uint64_t *d_cand = nullptr;
    int64_t *d_cand_off = nullptr;  // split + 1
    uint64_t *d_resc = nullptr;     // split * kMaxM
    uint32_t *d_done = nullptr;
    // FP32 Top-K state
    int fp32_blocks = 0;
    uint64_t *d_fthr = nullptr;
    uint32_t *d_fcnt = nullptr;
    uint64_t *d_fcand = nullptr;
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
}

int choose_group(double avg) {
    int g = 2;
    while (g < 32 && g * 2 <= avg / 2) g <<= 1;
    return g;
}

// Wavefront work items: contiguous slice ranges inside one partition with
// about equal dword counts.
std::vector<int4> make_items(const acc_packed *pk, int cus) {
    const int S = pk->split;
    const int64_t NS = pk->nslices();
    const int64_t kSliceCost = 8;  // fixed per-slice work, in dwords
    int64_t total = 0;
    for (int64_t s = 0; s < NS; ++s) total += pk->slice_ndw[(size_t)s] + kSliceCost;
    int64_t target_items = (int64_t)cus * ACC_WAVES * 4;
    int64_t per = std::max<int64_t>(total / std::max<int64_t>(target_items, 1), 256);
    std::vector<int4> items;
    for (int p = 0; p < S; ++p) {
        int64_t s0 = pk->part_slice_begin[(size_t)p], s1 = pk->part_slice_begin[(size_t)p + 1];
        int64_t b = s0, acc = 0;
        for (int64_t s = s0; s < s1; ++s) {
            acc += pk->slice_ndw[(size_t)s] + kSliceCost;
            if (acc >= per || s + 1 == s1) {
                items.push_back(make_int4(p, (int)b, (int)(s + 1), 0));
                b = s + 1;
                acc = 0;
            }
        }
    }
    return items;
}

int cap_for(int M) { return std::max(128, (int)[](int v) { int p = 1; while (p < v) p <<= 1; return p; }(M + 64)); }

typedef void (*stage1_fn)(Stage1Args);

template <int E, bool SIGNED>
stage1_fn stage1_for_mode(int mode) {
    return mode == 0 ? lowbit_stage1<E, SIGNED, 0> : lowbit_stage1<E, SIGNED, 1>;
}

template <bool SIGNED>
stage1_fn stage1_for_e(int E, int mode) {
    switch (E) {
        case 10: return stage1_for_mode<10, SIGNED>(mode);
        case 11: return stage1_for_mode<11, SIGNED>(mode);
        case 12: return stage1_for_mode<12, SIGNED>(mode);
        case 13: return stage1_for_mode<13, SIGNED>(mode);
        case 14: return stage1_for_mode<14, SIGNED>(mode);
        case 15: return stage1_for_mode<15, SIGNED>(mode);
        case 16: return stage1_for_mode<16, SIGNED>(mode);
        case 17: return stage1_for_mode<17, SIGNED>(mode);
        case 18: return stage1_for_mode<18, SIGNED>(mode);
        case 19: return stage1_for_mode<19, SIGNED>(mode);
        case 20: return stage1_for_mode<20, SIGNED>(mode);
        case 21: return stage1_for_mode<21, SIGNED>(mode);
        case 22: return stage1_for_mode<22, SIGNED>(mode);
        default: return nullptr;
    }
}

stage1_fn stage1_kernel(int E, bool is_signed, int mode) {
    return is_signed ? stage1_for_e<true>(E, mode) : stage1_for_e<false>(E, mode);
}

typedef void (*final_fn)(FinalArgs);
final_fn final_kernel(int G) {
    switch (G) {
        case 2: return lowbit_final<2>;
        case 4: return lowbit_final<4>;
        case 8: return lowbit_final<8>;
        case 16: return lowbit_final<16>;
        case 32: return lowbit_final<32>;
        default: return nullptr;
    }
}

typedef void (*fp32_fn)(Fp32Args);
fp32_fn fp32_kernel(int G, int mode) {
    switch (G) {
        case 2: return mode ? fp32_topk<2, 1> : fp32_topk<2, 0>;
        case 4: return mode ? fp32_topk<4, 1> : fp32_topk<4, 0>;
        case 8: return mode ? fp32_topk<8, 1> : fp32_topk<8, 0>;
        case 16: return mode ? fp32_topk<16, 1> : fp32_topk<16, 0>;
        case 32: return mode ? fp32_topk<32, 1> : fp32_topk<32, 0>;
        default: return nullptr;
    }
}

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ACC_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return ACC_OK;
}

// Launches stage 1 for one query; mode 0 -> candidates, mode 1 -> scores.
int launch_stage1(acc_gpu_matrix *m, const float *x, int M, int mode, int32_t *y, hipStream_t st) {
    Stage1Args a;
    a.words = m->d_words;
    a.slices = m->d_slices;
    a.perm = m->d_perm;
    a.items = m->d_items;
    a.nitems = m->nitems;
    a.x = x;
    a.cols = m->cols;
    a.col_bits = m->col_bits;
    a.val_bits = m->bits;
    a.M = M;
    a.cap = cap_for(M);
    a.thr = m->d_thr;
    a.cnt = m->d_cnt;
    a.cand = m->d_cand;
    a.cand_off = m->d_cand_off;
    a.y = y;
    stage1_fn k = stage1_kernel(m->entry_bits, m->is_signed != 0, mode);
    if (!k) return fail(ACC_ERR_ARG, "unsupported entry width");
    size_t lds = 64 + (mode == 0 ? (size_t)ACC_WAVES * a.cap * 8 : 0) + (size_t)((m->cols + 15) & ~15);
    int blocks = (m->nitems + ACC_WAVES - 1) / ACC_WAVES;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(ACC_WG), lds, st, a);
    return check_launch();
}

int launch_final(acc_gpu_matrix *m, const float *x, int M, int K, acc_element *out, int32_t *crow, int32_t *cscore,
                 hipStream_t st) {
    FinalArgs f;
    f.cnt = m->d_cnt;
    f.cand = m->d_cand;
    f.cand_off = m->d_cand_off;
    f.thr = m->d_thr;
    f.part_row_begin = m->d_part_row_begin;
    f.split = m->split;
    f.M = M;
    f.cap = cap_for(M);
    f.ptr = m->d_ptr;
    f.idx = m->d_idx;
    f.val = m->d_val;
    f.x = x;
    f.G = m->group;
    f.resc = m->d_resc;
    f.done = m->d_done;
    f.K = K;
    f.out = out;
    f.cand_rows_out = crow;
    f.cand_scores_out = cscore;
    final_fn k = final_kernel(m->group);
    if (!k) return fail(ACC_ERR_ARG, "unsupported group size");
    int total = m->split * M;
    int cap2 = 1;
    while (cap2 < total) cap2 <<= 1;
    size_t lds = std::max((size_t)ACC_WAVES * f.cap * 8 + (size_t)ACC_WAVES * M * 8, (size_t)cap2 * 8);
    hipLaunchKernelGGL(k, dim3(m->split), dim3(ACC_WG), lds, st, f);
    return check_launch();
}

int check_lowbit(acc_gpu_matrix *m, int ck, int k) {
    if (!m) return fail(ACC_ERR_ARG, "null matrix");
    if (!m->has_packed) return fail(ACC_ERR_ARG, "matrix was uploaded without low-bit packets");
    if (ck < m->split) return fail(ACC_ERR_ARG, "ck must be >= split");
    int M = ck / m->split;
    if (M > kMaxM) return fail(ACC_ERR_ARG, "ck/split must be <= 64");
    if (k < 0 || k > m->split * M) return fail(ACC_ERR_ARG, "k must be in [0, (ck/split)*split]");
    if (k > 0 && !m->has_csr) return fail(ACC_ERR_ARG, "rescoring needs the FP32 CSR (upload it too)");
    if ((int64_t)m->split * M > 4096) return fail(ACC_ERR_ARG, "ck must be <= 4096");
    return ACC_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// upload / info / destroy
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

extern "C" int acc_gpu_upload(const acc_csr *csr, const acc_packed *pk, int device, acc_gpu_matrix **out) {
    if (!out || (!csr && !pk)) return fail(ACC_ERR_ARG, "need a CSR and/or packed matrix");
    if (csr && pk && (csr->rows != pk->rows || csr->cols != pk->cols))
        return fail(ACC_ERR_ARG, "CSR and packed matrix shapes differ");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(ACC_ERR_NODEVICE, "no GPU visible");
    if (device < 0 || device >= ndev) return fail(ACC_ERR_ARG, "bad device index");
    HIP_TRY(hipSetDevice(device));
    acc_gpu_matrix *m = new acc_gpu_matrix();
    m->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) m->cus = prop.multiProcessorCount;
    m->rows = csr ? csr->rows : pk->rows;
    m->cols = csr ? csr->cols : pk->cols;
    if (m->cols > 8192) { delete m; return fail(ACC_ERR_ARG, "cols must be <= 8192"); }
    int rc = ACC_OK;
#define TRY(x)                         \
    do {                               \
        rc = (x);                      \
        if (rc != ACC_OK) goto failed; \
    } while (0)
    if (csr) {
        int64_t nnz = csr->nnz();
        if (nnz >= (int64_t)UINT32_MAX) { rc = fail(ACC_ERR_RANGE, "FP32 CSR needs nnz < 2^32 on the device"); goto failed; }
        m->has_csr = true;
        m->nnz = nnz;
        m->group = choose_group(m->rows > 0 ? (double)nnz / m->rows : 1.0);
        std::vector<uint32_t> p32(csr->ptr.size());
        for (size_t i = 0; i < p32.size(); ++i) p32[i] = (uint32_t)csr->ptr[i];
        TRY(dalloc(m, &m->d_ptr, p32.size()));
        TRY(dalloc(m, &m->d_idx, (size_t)nnz));
        TRY(dalloc(m, &m->d_val, (size_t)nnz));
        if (hipMemcpy(m->d_ptr, p32.data(), p32.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->d_idx, csr->idx.data(), (size_t)nnz * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->d_val, csr->val.data(), (size_t)nnz * 4, hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(ACC_ERR_DEVICE, "copying the CSR to the device failed");
            goto failed;
        }
        // FP32 Top-K state: candidate capacity = waves * kMaxK
        m->fp32_blocks = std::max(1, std::min((m->rows + 255) / 256, m->cus * 4));
        TRY(dalloc(m, &m->d_fthr, 1));
        TRY(dalloc(m, &m->d_fcnt, 1));
        TRY(dalloc(m, &m->d_fcand, (size_t)m->fp32_blocks * ACC_WAVES * kMaxK));
        if (hipMemset(m->d_fthr, 0, 8) != hipSuccess || hipMemset(m->d_fcnt, 0, 4) != hipSuccess) {
            rc = fail(ACC_ERR_DEVICE, "hipMemset failed");
            goto failed;
        }
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
        std::vector<uint2> sl((size_t)m->nslices);
        for (int64_t s = 0; s < m->nslices; ++s) sl[(size_t)s] = make_uint2(pk->slice_off[(size_t)s], pk->slice_ndw[(size_t)s]);
        std::vector<int4> items = make_items(pk, m->cus);
        m->nitems = (int)items.size();
        std::vector<int64_t> coff((size_t)m->split + 1, 0);
        for (auto &it : items) coff[(size_t)it.x + 1] += kMaxM;
        for (int p = 0; p < m->split; ++p) coff[(size_t)p + 1] += coff[(size_t)p];
        m->stream_bytes = (int64_t)pk->words.size() * 4 + m->nslices * 8;
        TRY(dalloc(m, &m->d_words, pk->words.size()));
        TRY(dalloc(m, &m->d_slices, sl.size()));
        TRY(dalloc(m, &m->d_perm, pk->perm.size()));
        TRY(dalloc(m, &m->d_items, items.size()));
        TRY(dalloc(m, &m->d_part_row_begin, pk->part_row_begin.size()));
        TRY(dalloc(m, &m->d_thr, (size_t)m->split));
        TRY(dalloc(m, &m->d_cnt, (size_t)m->split));
        TRY(dalloc(m, &m->d_cand, (size_t)coff[(size_t)m->split]));
        TRY(dalloc(m, &m->d_cand_off, coff.size()));
        TRY(dalloc(m, &m->d_resc, (size_t)m->split * kMaxM));
        TRY(dalloc(m, &m->d_done, 1));
        if (hipMemcpy(m->d_words, pk->words.data(), pk->words.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->d_slices, sl.data(), sl.size() * sizeof(uint2), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->d_perm, pk->perm.data(), pk->perm.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->d_items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->d_part_row_begin, pk->part_row_begin.data(), pk->part_row_begin.size() * 4,
                      hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->d_cand_off, coff.data(), coff.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(m->d_thr, 0, (size_t)m->split * 8) != hipSuccess ||
            hipMemset(m->d_cnt, 0, (size_t)m->split * 4) != hipSuccess ||
            hipMemset(m->d_done, 0, 4) != hipSuccess) {
            rc = fail(ACC_ERR_DEVICE, "copying the packets to the device failed");
            goto failed;
        }
    }
    TRY(dalloc(m, &m->d_xstage, (size_t)m->cols));
    TRY(dalloc(m, &m->d_ostage, (size_t)kMaxK + 64));
    if (hipDeviceSynchronize() != hipSuccess) { rc = fail(ACC_ERR_DEVICE, "device synchronize failed"); goto failed; }
    *out = m;
    return ACC_OK;
failed:
    free_all(m);
    delete m;
    return rc;
#undef TRY
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
    return ACC_OK;
}

extern "C" void acc_gpu_matrix_destroy(acc_gpu_matrix *m) {
    free_all(m);
    delete m;
}

// ---------------------------------------------------------------------------
// queries
// ---------------------------------------------------------------------------

extern "C" int acc_gpu_topk_lowbit(acc_gpu_matrix *m, const float *x_dev, int k, int ck, acc_element *out_dev,
                                   void *stream) {
    int rc = check_lowbit(m, ck, k);
    if (rc != ACC_OK) return rc;
    if (!x_dev || (k > 0 && !out_dev)) return fail(ACC_ERR_ARG, "null query or output");
    HIP_TRY(hipSetDevice(m->device));
    hipStream_t st = (hipStream_t)stream;
    int M = ck / m->split;
    rc = launch_stage1(m, x_dev, M, 0, nullptr, st);
    if (rc != ACC_OK) return rc;
    return launch_final(m, x_dev, M, k, out_dev, nullptr, nullptr, st);
}

extern "C" int acc_gpu_lowbit_stage1(acc_gpu_matrix *m, const float *x_dev, int ck, void *stream) {
    int rc = check_lowbit(m, ck, 0);
    if (rc != ACC_OK) return rc;
    if (!x_dev) return fail(ACC_ERR_ARG, "null query");
    HIP_TRY(hipSetDevice(m->device));
    return launch_stage1(m, x_dev, ck / m->split, 0, nullptr, (hipStream_t)stream);
}

extern "C" int acc_gpu_lowbit_stage2(acc_gpu_matrix *m, const float *x_dev, int k, int ck, acc_element *out_dev,
                                     void *stream) {
    int rc = check_lowbit(m, ck, k);
    if (rc != ACC_OK) return rc;
    if (!x_dev || (k > 0 && !out_dev)) return fail(ACC_ERR_ARG, "null query or output");
    HIP_TRY(hipSetDevice(m->device));
    return launch_final(m, x_dev, ck / m->split, k, out_dev, nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int acc_gpu_spmv_fp64(acc_gpu_matrix *m, const float *x_dev, double *y_dev, void *stream) {
    if (!m || !m->has_csr || !x_dev || !y_dev) return fail(ACC_ERR_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    int blocks = std::max(1, std::min((m->rows + 31) / 32, m->cus * 8));
    hipLaunchKernelGGL(spmv_fp64, dim3(blocks), dim3(ACC_WG), 0, (hipStream_t)stream, m->d_ptr, m->d_idx, m->d_val,
                       x_dev, m->rows, y_dev);
    return check_launch();
}

extern "C" int acc_gpu_topk_lowbit_batch(acc_gpu_matrix *m, const float *x_dev, int nq, const int32_t *k_host, int ck,
                                         acc_element *out_dev, int out_stride, void *stream) {
    if (!m || !x_dev || !k_host || nq < 0) return fail(ACC_ERR_ARG, "bad batch arguments");
    for (int q = 0; q < nq; ++q) {
        if (k_host[q] > out_stride) return fail(ACC_ERR_ARG, "k exceeds out_stride");
        int rc = acc_gpu_topk_lowbit(m, x_dev + (size_t)q * m->cols, k_host[q], ck, out_dev + (size_t)q * out_stride,
                                     stream);
        if (rc != ACC_OK) return rc;
    }
    return ACC_OK;
}

extern "C" int acc_gpu_scores_lowbit(acc_gpu_matrix *m, const float *x_dev, int32_t *y_dev, void *stream) {
    if (!m || !m->has_packed || !x_dev || !y_dev) return fail(ACC_ERR_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    return launch_stage1(m, x_dev, 1, 1, y_dev, (hipStream_t)stream);
}

extern "C" int acc_gpu_candidates_lowbit(acc_gpu_matrix *m, const float *x_dev, int ck, int32_t *rows_dev,
                                         int32_t *scores_dev, int *count_host, void *stream) {
    int rc = check_lowbit(m, ck, 0);
    if (rc != ACC_OK) return rc;
    if (!x_dev || !rows_dev || !scores_dev) return fail(ACC_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(m->device));
    hipStream_t st = (hipStream_t)stream;
    int M = ck / m->split;
    // stage-1 candidates land at p*M + i; compact to the C-ABI layout afterwards
    int32_t *tmp_r = nullptr, *tmp_s = nullptr;
    HIP_TRY(hipMallocAsync((void **)&tmp_r, (size_t)m->split * M * 4, st));
    HIP_TRY(hipMallocAsync((void **)&tmp_s, (size_t)m->split * M * 4, st));
    rc = launch_stage1(m, x_dev, M, 0, nullptr, st);
    if (rc == ACC_OK) rc = launch_final(m, x_dev, M, 0, nullptr, tmp_r, tmp_s, st);
    int total = 0;
    if (rc == ACC_OK) {
        for (int p = 0; p < m->split; ++p) {
            int n = std::min(M, m->part_row_begin[(size_t)p + 1] - m->part_row_begin[(size_t)p]);
            if (n > 0) {
                if (hipMemcpyAsync(rows_dev + total, tmp_r + (size_t)p * M, (size_t)n * 4, hipMemcpyDeviceToDevice, st) != hipSuccess ||
                    hipMemcpyAsync(scores_dev + total, tmp_s + (size_t)p * M, (size_t)n * 4, hipMemcpyDeviceToDevice, st) != hipSuccess) {
                    rc = fail(ACC_ERR_DEVICE, "copy of candidates failed");
                    break;
                }
            }
            total += n;
        }
    }
    hipFreeAsync(tmp_r, st);
    hipFreeAsync(tmp_s, st);
    if (rc == ACC_OK && hipStreamSynchronize(st) != hipSuccess) rc = fail(ACC_ERR_DEVICE, "stream synchronize failed");
    if (count_host) *count_host = total;
    return rc;
}

static int launch_fp32(acc_gpu_matrix *m, const float *x_dev, int k, int mode, float *y, hipStream_t st) {
    Fp32Args a;
    a.ptr = m->d_ptr;
    a.idx = m->d_idx;
    a.val = m->d_val;
    a.x = x_dev;
    a.rows = m->rows;
    a.cols = m->cols;
    a.M = k;
    a.cap = cap_for(k);
    a.thr = m->d_fthr;
    a.cnt = m->d_fcnt;
    a.cand = m->d_fcand;
    a.y = y;
    fp32_fn kern = fp32_kernel(m->group, mode);
    if (!kern) return fail(ACC_ERR_ARG, "unsupported group size");
    size_t lds = (mode == 0 ? (size_t)ACC_WAVES * a.cap * 8 : 0) + (size_t)m->cols * 4;
    hipLaunchKernelGGL(kern, dim3(m->fp32_blocks), dim3(ACC_WG), lds, st, a);
    return check_launch();
}

extern "C" int acc_gpu_topk_fp32(acc_gpu_matrix *m, const float *x_dev, int k, acc_element *out_dev, void *stream) {
    if (!m || !m->has_csr) return fail(ACC_ERR_ARG, "matrix was uploaded without the FP32 CSR");
    if (!x_dev || !out_dev || k < 1 || k > kMaxK) return fail(ACC_ERR_ARG, "k must be in [1, 960]");
    HIP_TRY(hipSetDevice(m->device));
    hipStream_t st = (hipStream_t)stream;
    int rc = launch_fp32(m, x_dev, k, 0, nullptr, st);
    if (rc != ACC_OK) return rc;
    int cap = cap_for(k);
    size_t lds = (size_t)ACC_WAVES * cap * 8 + (size_t)ACC_WAVES * k * 8;
    hipLaunchKernelGGL(topk_final, dim3(1), dim3(ACC_WG), lds, st, m->d_fcnt, m->d_fcand, m->d_fthr, k, cap, out_dev);
    return check_launch();
}

extern "C" int acc_gpu_spmv_fp32(acc_gpu_matrix *m, const float *x_dev, float *y_dev, void *stream) {
    if (!m || !m->has_csr || !x_dev || !y_dev) return fail(ACC_ERR_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    return launch_fp32(m, x_dev, 1, 1, y_dev, (hipStream_t)stream);
}

extern "C" int acc_gpu_top_k_heap(const float *y_dev, int n, int k, acc_element *out_dev, void *stream) {
    if (!y_dev || !out_dev || n < 0 || k < 1 || k > kMaxK) return fail(ACC_ERR_ARG, "bad arguments (k in [1, 960])");
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    int blocks = std::max(1, std::min((n + 255) / 256, cus * 2));
    uint64_t *thr = nullptr, *cand = nullptr;
    uint32_t *cnt = nullptr;
    size_t ncand = (size_t)blocks * ACC_WAVES * k;
    HIP_TRY(hipMallocAsync((void **)&thr, 8, st));
    HIP_TRY(hipMallocAsync((void **)&cnt, 4, st));
    HIP_TRY(hipMallocAsync((void **)&cand, ncand * 8, st));
    HIP_TRY(hipMemsetAsync(thr, 0, 8, st));
    HIP_TRY(hipMemsetAsync(cnt, 0, 4, st));
    int cap = cap_for(k);
    hipLaunchKernelGGL(dense_topk, dim3(blocks), dim3(ACC_WG), (size_t)ACC_WAVES * cap * 8, st, y_dev, n, k, cap, thr, cnt, cand);
    int rc = check_launch();
    if (rc == ACC_OK) {
        size_t lds = (size_t)ACC_WAVES * cap * 8 + (size_t)ACC_WAVES * k * 8;
        hipLaunchKernelGGL(topk_final, dim3(1), dim3(ACC_WG), lds, st, cnt, cand, thr, k, cap, out_dev);
        rc = check_launch();
    }
    hipFreeAsync(thr, st);
    hipFreeAsync(cnt, st);
    hipFreeAsync(cand, st);
    return rc;
}

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

extern "C" int acc_top_k_spmv(acc_gpu_matrix *A, int row, int col, int k, int num_threads, acc_element *top_k,
                              int num_test) {
    (void)num_threads;
    if (!A || !top_k || num_test < 1) return fail(ACC_ERR_ARG, "bad arguments");
    if (row != A->rows || col != A->cols) return fail(ACC_ERR_ARG, "row/col do not match the matrix");
    if (k < 1 || k > kMaxK) return fail(ACC_ERR_ARG, "k must be in [1, 960]");
    HIP_TRY(hipSetDevice(A->device));
    std::vector<float> x((size_t)col);
    float *d_y = nullptr;
    HIP_TRY(hipMalloc((void **)&d_y, (size_t)row * 4 + 4));
    hipEvent_t e0, e1, e2;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventCreate(&e2);
    double cpt = 0, srt = 0;
    int rc = ACC_OK;
    for (int t = 0; t < num_test && rc == ACC_OK; ++t) {
        acc_generate_queries(1, col, ACC_VAL_UNIFORM, 0, 0x5EEDull + (uint64_t)t, x.data());
        hipMemcpy(A->d_xstage, x.data(), (size_t)col * 4, hipMemcpyHostToDevice);
        hipEventRecord(e0, nullptr);
        rc = acc_gpu_spmv_fp32(A, A->d_xstage, d_y, nullptr);
        hipEventRecord(e1, nullptr);
        if (rc == ACC_OK) rc = acc_gpu_top_k_heap(d_y, row, k, A->d_ostage, nullptr);
        hipEventRecord(e2, nullptr);
        hipEventSynchronize(e2);
        float a = 0, b = 0;
        hipEventElapsedTime(&a, e0, e1);
        hipEventElapsedTime(&b, e1, e2);
        cpt += a;
        srt += b;
    }
    if (rc == ACC_OK && hipMemcpy(top_k, A->d_ostage, (size_t)k * sizeof(acc_element), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(ACC_ERR_DEVICE, "copy of results failed");
    std::printf("compute time: %g ms\n", cpt / num_test);
    std::printf("sort time: %g ms\n", srt / num_test);
    std::printf("exec time: %gms\n", (cpt + srt) / num_test);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    hipFree(d_y);
    return rc;
}
