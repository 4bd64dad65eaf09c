/*
 * acceles.h -- C ABI of the MI355X-native AccelES Top-K SpMV library
 * (libacceles.so).  Plain pointers and sizes only.
 *
 * Conventions
 *   - Every function returns an int status: ACC_OK (0) or a negative
 *     ACC_ERR_* code; acc_last_error() gives the message of the calling
 *     thread's last failure.  The library never exits the process (the
 *     reference exits or prints and continues, utils.hpp:460-472,512-513).
 *   - Handles (acc_coo, acc_csr, acc_packed, acc_gpu_matrix) are owned by the
 *     library: create/destroy pairs.  Output arrays are owned by the caller.
 *   - A handle is not shared between threads without external locking.
 *     GPU calls enqueue on the caller's stream (hipStream_t passed as void*,
 *     NULL = the null stream) and return without synchronising, except the
 *     *_host convenience calls, which copy in, run and copy out.
 *   - Ranking order everywhere: score descending, then row index ascending.
 *
 * Each declaration cites the reference interface it replaces (path:line in
 * the AccelES tree).
 */
#ifndef ACCELES_H
#define ACCELES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ACC_OK 0
#define ACC_ERR_ARG (-1)      /* bad argument                         */
#define ACC_ERR_IO (-2)       /* file cannot be opened / read         */
#define ACC_ERR_FORMAT (-3)   /* malformed file content               */
#define ACC_ERR_RANGE (-4)    /* index or size out of range           */
#define ACC_ERR_NOMEM (-5)    /* host or device allocation failed     */
#define ACC_ERR_DEVICE (-6)   /* HIP runtime error                    */
#define ACC_ERR_NODEVICE (-7) /* no GPU visible                       */
#define ACC_ERR_BUDGET (-8)   /* serving admission: memory budget     */

/* One ranked result; same layout as Element (src/cpu/cpu_mkl_baseline.cpp:16-19). */
typedef struct {
    float value;
    int32_t index;
} acc_element;

typedef struct acc_coo acc_coo;               /* host COO triples            */
typedef struct acc_csr acc_csr;               /* host FP32 CSR               */
typedef struct acc_packed acc_packed;         /* host low-bit sliced packets */
typedef struct acc_gpu_matrix acc_gpu_matrix; /* device-resident matrix      */

const char *acc_last_error(void);
const char *acc_version(void);

/* ------------------------------------------------------------------ */
/* Loaders                                                             */
/* ------------------------------------------------------------------ */

/* read_coo_matrix (src/cpu/cpu_mkl_baseline.cpp:43-77): banner line, one
 * more header line, then "rows cols nnz" and nnz "r c v" lines; indices are
 * 1-based unless zero_index != 0.  Parsed in parallel. */
int acc_read_coo_matrix(const char *path, int zero_index, acc_coo **out);

/* readMtx (src/common/utils/utils.hpp:449-492): Matrix Market coordinate
 * file, any number of '%' comment lines; 1-based unless zero_indexed_file;
 * pattern files read values as 1.0; sort_tuples orders entries by (row, col)
 * (stable). */
int acc_read_mtx(const char *path, int zero_indexed_file, int sort_tuples, acc_coo **out);

/* write_coo_to_mtx (src/resources/python/quantify/quantify_util.py:469-477):
 * "%%MatrixMarket matrix coordinate real general", "%", size line, 1-based
 * triples with values printed to 9 significant digits. */
int acc_write_mtx(const char *path, const acc_coo *coo);

/* coo_t (src/fpga/src/ip/coo_matrix.hpp:9-24) / COOMatrix
 * (cpu_mkl_baseline.cpp:33-40) built from caller arrays (copied). */
int acc_coo_create(int32_t rows, int32_t cols, int64_t nnz, const int32_t *row,
                   const int32_t *col, const float *val, acc_coo **out);
int acc_coo_info(const acc_coo *coo, int32_t *rows, int32_t *cols, int64_t *nnz);
int acc_coo_export(const acc_coo *coo, int32_t *row, int32_t *col, float *val);
void acc_coo_destroy(acc_coo *coo);

/* ------------------------------------------------------------------ */
/* Conversion                                                          */
/* ------------------------------------------------------------------ */

/* coo2csr (src/common/utils/utils.hpp:494-547), caller-allocated outputs:
 * ptr[nrows+1], idx[nnz], val_out[nnz].  Counting sort, stable within a row;
 * sort_tuples first orders by (row, col).  Out-of-range indices return
 * ACC_ERR_RANGE instead of printing and continuing. */
int acc_coo2csr(int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *row,
                const int32_t *col, const float *val, int sort_tuples, int64_t *ptr,
                int32_t *idx, float *val_out);

/* coo_to_csr (src/cpu/cpu_mkl_baseline.cpp:79-109): library-owned CSR handle. */
int acc_coo_to_csr(const acc_coo *coo, acc_csr **out);
int acc_csr_create(int32_t rows, int32_t cols, int64_t nnz, const int64_t *ptr,
                   const int32_t *idx, const float *val, acc_csr **out);
int acc_csr_info(const acc_csr *csr, int32_t *rows, int32_t *cols, int64_t *nnz);
int acc_csr_export(const acc_csr *csr, int64_t *ptr, int32_t *idx, float *val);
void acc_csr_destroy(acc_csr *csr);

/* quantify_matrix / quantify_vector (quantify_util.py:8-49): M = 2^bits-1
 * (halved when symmetric), s = max|v|, q = clip(rint(M*(v/s)), -M-1, M) in
 * float32 with round-half-even; s is written to *scale_out. */
int acc_quantize(int64_t n, const float *v, int bits, int symmetric, float *scale_out,
                 int32_t *q);

/* Random-CSR (README.md:5; row shuffle of quantify_util.py:491 and
 * csr_util.py:101-114), seeded: out row i is input row perm[i]. */
int acc_permute_rows(const acc_csr *csr, uint64_t seed, acc_csr **out, int32_t *perm);

/* ReSparse (resparse_csr, csr_util.py:152-187; resparse_spm,
 * quantify_util.py:375-403): with avg = the mean |v| over the nonzeros
 * (mean_over_cells = 0) or over all rows*cols cells (1, the dense form), a
 * row whose smallest |v| exceeds avg is kept whole; any other row keeps its
 * ceil(n * keep) entries of largest |v| (equal |v|: the earlier entry; the
 * reference's unstable argsort leaves that order open), in column order.
 * The mean is summed in FP64.  keep in [0, 1]. */
int acc_resparse_csr(const acc_csr *csr, double keep, int mean_over_cells, acc_csr **out);

/* ---- Ranking-quality metrics ---------------------------------------- */

/* Rank NDCG of a test list against a golden list of n rows
 * (normalized_discounted_cumulative_gain, evaluation_utils.hpp:112-147):
 * relevance n - place (a row listed twice in test: the later place);
 * discount log2(place + 2). */
int acc_ndcg_rank(const int32_t *golden, const int32_t *test, int n, double *out);
/* Score NDCG (ndcg, quantify_util.py:313-329): relevance of a row = its exact
 * score in the sw list (0 when absent), dcg over the hw list's places,
 * idcg over the sw list; any output pointer may be null. */
int acc_ndcg_score(const int32_t *sw_idx, const double *sw_val, int nsw, const int32_t *hw_idx, int nhw,
                   double *ndcg, double *dcg, double *idcg);
/* Levenshtein distance of two row lists (edit_distance,
 * evaluation_utils.hpp:197-209). */
int acc_edit_distance(const int32_t *a, int na, const int32_t *b, int nb, uint32_t *out);
/* Places where the lists differ (bounded_count_errors' inner count,
 * evaluation_utils.hpp:241-252). */
int acc_count_errors(const int32_t *golden, const int32_t *test, int n, int32_t *out);

/* ---- Ultra-CSR reference packet format ---------------------------- */

/* Entries per 512-bit packet: 511 / (1 + col_bits + val_bits)
 * (src/common/types.hpp:151-162; 30 for 10-bit columns, 6-bit values). */
int acc_ultracsr_entries(int col_bits, int val_bits);

/* Q_packet_coo_partition (src/fpga/src/host_spmv_ucsr.cpp:388-425) with
 * Q_write_block_x/y/val (src/fpga/src/ip/fpga_utils.hpp:438-503): one
 * partition's COO stream -> packets of 8 little-endian 64-bit words.  Flag =
 * (row != previous entry's row), the previous row starting at last_row.
 * Returns the packet count (>= 0) or an error; cap is in packets. */
int64_t acc_pack_ultracsr(int64_t nnz, const int32_t *row, const int32_t *col,
                          const int32_t *qval, int32_t last_row, int col_bits, int val_bits,
                          uint64_t *packets, int64_t cap);

/* Q_read_block_x/y/val (fpga_utils.hpp:317-384): per entry the x-decoded row
 * counter (carried across packets from carry0), column and value. */
int acc_unpack_ultracsr(const uint64_t *packets, int64_t nnz, int col_bits, int val_bits,
                        int32_t carry0, int32_t *xrow, int32_t *col, int32_t *qval);

/* increment_array / adder (fpga_utils.hpp:279-302, x_decoder.cpp:5-22):
 * x[i] = carry + popcount(flags & ((2 << i) - 1)), i < n <= 32. */
int acc_xdecode(uint32_t flags, int n, int32_t carry, int32_t *x);

/* ---- Low-bit sliced packets (the GPU layout) ----------------------- */

typedef struct {
    int bits;      /* value bits B, 2..8 (4 and 8 are the tuned cases)       */
    int is_signed; /* 0: unsigned [0, 2^B-1] (non-negative input required),
                      1: symmetric signed [-2^(B-1), 2^(B-1)-1]              */
    int split;     /* SPLIT row partitions (quantify_util.py:184-189)        */
    int col_bits;  /* 0 = ceil(log2(cols))                                   */
    uint64_t shuffle_seed; /* Random-CSR (README.md:5; quantify_util.py:491
                      shuffles rows before the partitions of :184-189): 0 =
                      partitions of consecutive rows; else partition p holds
                      rows order[p*size ...] of acc_permute_rows(seed)'s
                      order.  Results name the original rows either way.    */
} acc_pack_opts;

typedef struct {
    int32_t rows, cols, split, bits, col_bits, entry_bits, is_signed;
    int64_t nnz;          /* stored entries of the source CSR                */
    int64_t nnz_kept;     /* entries whose quantized value is non-zero       */
    int64_t nslices;      /* 64-row slices                                   */
    int64_t stream_bytes; /* packet words + per-slice metadata (read per query) */
    int64_t perm_bytes;   /* slot -> row map (read only for candidates)      */
    double bytes_per_nnz; /* stream_bytes / nnz                               */
    float scale;          /* global FP32 scale of the values                 */
    uint64_t shuffle_seed; /* acc_pack_opts.shuffle_seed (0: no shuffle)     */
} acc_packed_info_t;

/* Quantize (acc_quantize rules) and pack the CSR (host_spmv_ucsr.cpp:278-319
 * partitions + packing, re-laid-out for wavefronts). */
int acc_pack_lowbit(const acc_csr *csr, const acc_pack_opts *opts, acc_packed **out);
int acc_packed_info(const acc_packed *pk, acc_packed_info_t *info);
/* Decode back to CSR order: ptr[rows+1], idx[nnz_kept], qval[nnz_kept]. */
int acc_unpack_lowbit(const acc_packed *pk, int64_t *ptr, int32_t *idx, int32_t *qval);
void acc_packed_destroy(acc_packed *pk);

/* Binary files of a CSR matrix or of packets, so a matrix is read and packed
 * once (no counterpart: the reference's drivers re-read .mtx text and re-pack
 * on every run, host_spmv_ucsr.cpp:278-319).  Loading checks what the kernels
 * index with (ptr monotone, columns in range, slices inside the words, lanes
 * in range) and fails with ACC_ERR_FORMAT otherwise. */
int acc_csr_save(const acc_csr *csr, const char *path);
int acc_csr_load(const char *path, acc_csr **out);
int acc_packed_save(const acc_packed *pk, const char *path);
int acc_packed_load(const char *path, acc_packed **out);

/* ---- Synthetic inputs (create_matrices.py:57-137, seeded) ---------- */

#define ACC_DEG_UNIFORM 0 /* row nnz uniform in [p0, p1]                      */
#define ACC_DEG_GAMMA 1   /* row nnz = clamp(round(Gamma(p0, p1)), 1, cols)   */
#define ACC_VAL_ABSNORMAL 0
#define ACC_VAL_NORMAL 1
#define ACC_VAL_UNIFORM 2 /* U(-1, 1)                                         */

/* Distinct uniform columns per row (sorted), values per val_kind,
 * optional L2 row normalisation.  Deterministic for a seed. */
int acc_generate_csr(int32_t rows, int32_t cols, int deg_kind, double p0, double p1,
                     int val_kind, int l2norm, uint64_t seed, acc_csr **out);
int acc_generate_queries(int32_t nq, int32_t cols, int val_kind, int l2norm, uint64_t seed,
                         float *out);

/* ------------------------------------------------------------------ */
/* GPU (HIP, gfx950)                                                   */
/* ------------------------------------------------------------------ */

int acc_gpu_device_count(int *count);

/* Device memory, streams and events for callers without a GPU framework of
 * their own (the cl::Buffer / enqueueMigrateMemObjects / profiling-event
 * plumbing of host_spmv_ucsr.cpp:501-539 and opencl_utils.hpp:335-344).
 * Copies are synchronous with respect to the host. */
int acc_dev_malloc(int device, size_t bytes, void **ptr);
int acc_dev_free(void *ptr);
int acc_memcpy_h2d(void *dst_dev, const void *src_host, size_t bytes);
int acc_memcpy_d2h(void *dst_host, const void *src_dev, size_t bytes);
int acc_memcpy_d2d(void *dst_dev, const void *src_dev, size_t bytes);
int acc_stream_create(int device, void **stream);
int acc_stream_destroy(void *stream);
int acc_stream_synchronize(void *stream);
int acc_event_create(void **event);
int acc_event_record(void *event, void *stream);
int acc_event_elapsed_ms(void *start, void *end, float *ms); /* waits for end */
int acc_event_destroy(void *event);

/* SpMV::setup (src/gpu/host_spmv_topk_csr_gpu.cu:104-169) and Q_setup
 * (host_spmv_ucsr.cpp:501-539): copy the FP32 CSR and (optionally) the
 * low-bit packets to device memory on `device`.  Either may be NULL but not
 * both; the two-stage path needs both (stage 2 rescoring reads the CSR). */
int acc_gpu_upload(const acc_csr *csr, const acc_packed *pk, int device, acc_gpu_matrix **out);

typedef struct {
    int32_t rows, cols, split, bits, has_csr, has_packed;
    int32_t group;         /* G: lanes per row in the FP32 dot products      */
    int32_t work_items;    /* stage-1 wavefront work items                   */
    int64_t nnz;
    int64_t csr_bytes;     /* FP32 CSR bytes read per query                  */
    int64_t stream_bytes;  /* low-bit bytes read per stage-1 pass            */
    int64_t device_bytes;  /* total device memory held by the handle         */
    int32_t batch_mode;    /* ACC_BATCH_* in effect                          */
    int32_t dense_ok;      /* ACC_BATCH_DENSE can score the matrix exactly   */
    int64_t dense_bytes;   /* its image packets (the packets re-encoded for the
                              matrix cores), read per pass; 0 if !dense_ok   */
} acc_gpu_info_t;

int acc_gpu_matrix_info(const acc_gpu_matrix *m, acc_gpu_info_t *info);

/* Per-launch timing with profiling events (CL_QUEUE_PROFILING_ENABLE and
 * get_event_execution_time, opencl_utils.hpp:294,335-344; the per-phase
 * clocks of host_spmv_ucsr.cpp:569-572).  While enabled, every kernel of a
 * query is launched with a pair of HIP events attached to its dispatch (the
 * kernel's own start and end); acc_gpu_stage_times waits for the last query
 * and returns up to n kernel durations in ms (low-bit: slice maxima, select
 * + rescoring, final Top-K; batched packets: scores, select, final; batched
 * matrix cores: prep, scores, select + rescoring, final; FP32: SpMV, select,
 * final select) and their count. */
int acc_gpu_set_profiling(acc_gpu_matrix *m, int enable);
int acc_gpu_stage_times(acc_gpu_matrix *m, float *ms, int n, int *count);

/* Average duration of one stage of a low-bit query (0 slice maxima, 1
 * per-partition select + rescoring, 2 final Top-K) launched reps times back
 * to back between two HIP events on the stream, after one whole query that
 * leaves the state the later stages read.  Per-dispatch events (above) add
 * the dispatch overhead to each kernel; this measures a kernel as it runs
 * inside back-to-back queries.  Waits for the stream.  No counterpart in the
 * reference (its phases are timed by host_spmv_ucsr.cpp:569-572 clocks). */
int acc_gpu_time_lowbit_stage(acc_gpu_matrix *m, const float *x_dev, int k, int ck, int stage, int reps, void *stream,
                              float *avg_ms);
void acc_gpu_matrix_destroy(acc_gpu_matrix *m);

/* Plain FP32 CSR Top-K: SpMV (cpu_mkl_baseline.cpp:261) fused with the Top-k
 * of top_k_heap (:128-149); the GPU baseline spec is
 * host_spmv_topk_csr_gpu.cu:171-231.  1 <= k <= 960.  x_dev: cols floats;
 * out_dev: k elements in ranking order (tail (-inf, -1) if k > rows). */
int acc_gpu_topk_fp32(acc_gpu_matrix *m, const float *x_dev, int k, acc_element *out_dev,
                      void *stream);

/* Low-bit two-stage Top-K: query quantized to signed 8-bit with its own scale
 * (quantify_util.py:259), INT32 scores of the packed matrix, top floor(ck/split)
 * rows per partition (quantify_util.py:173-196 / spmv_ucsr_top_k_multicore.hpp
 * :360-413), merge (host_spmv_ucsr.cpp:691-748), FP32 rescoring of the
 * candidates (quantify_util.py:354-367) and the final Top-k.
 * k <= ck, ck/split <= 64. */
int acc_gpu_topk_lowbit(acc_gpu_matrix *m, const float *x_dev, int k, int ck,
                        acc_element *out_dev, void *stream);

/* FP64 exact scores y_dev[rows] = A x with FP64 accumulation (the exact
 * reference of quantify_util.py:257-258), for precision/recall reporting. */
int acc_gpu_spmv_fp64(acc_gpu_matrix *m, const float *x_dev, double *y_dev, void *stream);

/* Batched low-bit two-stage: nq queries (row-major nq x cols), query i keeps
 * k_host[i] results in out_dev[i*out_stride ...].  One matrix pass scores a
 * block of queries. */
int acc_gpu_topk_lowbit_batch(acc_gpu_matrix *m, const float *x_dev, int nq,
                              const int32_t *k_host, int ck, acc_element *out_dev,
                              int out_stride, void *stream);

/* Scoring layout of the batched path.  The reference scores one query per
 * pass (host_spmv_ucsr.cpp:1084-1180); a batch here is one pass over the
 * sliced packets per block of queries, scored either on the vector ALUs
 * (ACC_BATCH_SPARSE, v_dot4) or on the matrix cores (ACC_BATCH_DENSE,
 * v_mfma_i32_32x32x32_i8: each 64-row slice is expanded into an int8 image
 * in LDS, one byte per (row, column), inside the scoring kernel, from the
 * handle's image packets -- the packets re-encoded at upload with each
 * entry's place in the image, 18 to 23 bits per entry,
 * acc_gpu_info_t.dense_bytes).  Both give the same integer
 * scores and the same results.  ACC_BATCH_DENSE fails with a message when
 * the matrix cores cannot score the matrix exactly (a row whose values in one
 * column sum past one byte, |score| >= 2^25, or more than 1024 columns);
 * ACC_BATCH_AUTO uses them when they can and the
 * batch has at least min_dense_batch queries (0 keeps the current value;
 * default 16). */
#define ACC_BATCH_AUTO 0
#define ACC_BATCH_SPARSE 1
#define ACC_BATCH_DENSE 2
#define ACC_BATCH_SINGLE 3 /* one single-query pass per query */
int acc_gpu_set_batch_mode(acc_gpu_matrix *m, int mode, int min_dense_batch);

/* Autotuner for ACC_BATCH_AUTO (no counterpart in the reference, whose
 * partition count and CK are fixed at build time, types.hpp:54,81): times one
 * query per pass, the packet batch and the dense batch (where the matrix
 * allows it) on synthetic queries at batch sizes 1, 2, 4, ... up to max_batch,
 * and keeps the fastest layout per batch size in the handle (a handle set to
 * ACC_BATCH_SPARSE or _SINGLE keeps the matrix cores out).  points[i] receives each measured size (up to cap entries; *count
 * gets the number written); times are per batch in microseconds, -1 where a
 * layout does not apply.  Results do not depend on the layout. */
typedef struct {
    int32_t nq, layout;   /* layout chosen: ACC_BATCH_SINGLE / _SPARSE / _DENSE */
    float us_single, us_sparse, us_dense;
} acc_tune_point;
int acc_gpu_autotune(acc_gpu_matrix *m, int ck, int k, int max_batch, acc_tune_point *points,
                     int cap, int *count);

/* Plans: one query launch sequence captured into a HIP graph and replayed
 * (the reference re-enqueues its kernel and buffer migrations for every query,
 * host_spmv_ucsr.cpp:1084-1180; a plan does that setup once).  The plan reads
 * x_dev and writes out_dev -- fixed buffers the caller refills between
 * launches -- and keeps its own batch workspace; the handle must outlive it.
 * Creation runs the sequence once eagerly (results land in out_dev), then
 * captures it.  Kinds:
 *   ACC_PLAN_LOWBIT        acc_gpu_topk_lowbit, nq = 1, k = k_host[0];
 *   ACC_PLAN_LOWBIT_BATCH  acc_gpu_topk_lowbit_batch (per-query K);
 *   ACC_PLAN_FP32          acc_gpu_topk_fp32, nq = 1, k = k_host[0].
 * Every plan owns its query workspace and selection state, so different
 * plans of one handle may run concurrently on different streams; launches of
 * one plan, and the handle's own query calls (which share the handle's
 * workspace), must not overlap in time. */
typedef struct acc_plan acc_plan;
#define ACC_PLAN_LOWBIT 1
#define ACC_PLAN_LOWBIT_BATCH 2
#define ACC_PLAN_FP32 3
int acc_gpu_plan_create(acc_gpu_matrix *m, int kind, const float *x_dev, int nq,
                        const int32_t *k_host, int ck, acc_element *out_dev, int out_stride,
                        acc_plan **plan);
int acc_gpu_plan_launch(acc_plan *plan, void *stream);
void acc_gpu_plan_destroy(acc_plan *plan);

/* ---- Serving: resident matrices, continuous batching ----------------
 * No counterpart in the reference, which answers one query at a time
 * (host_spmv_ucsr.cpp:1084-1180, cpu_mkl_baseline.cpp:254-272); for callers
 * that issue many concurrent queries.  A worker thread keeps the matrices
 * resident and forms batches from the waiting requests of each matrix (up to
 * max_batch, each request with its own K); batches of different matrices
 * run concurrently on up to `slots` streams, and requests that arrive while a
 * batch runs join the next batch of their matrix as soon as a slot frees;
 * every slot has its own batch workspace, so one matrix may have a batch on
 * every slot.  Matrices are admitted against memory_budget (resident data
 * plus every slot's workspace for a full batch of the matrix: ACC_ERR_BUDGET
 * otherwise), requests against max_request_bytes (host memory held by a
 * request until its result is read: the query, the results, bookkeeping) and
 * max_pending.  A result does not depend on the batch it lands in. */
typedef struct acc_server acc_server;
typedef struct {
    int device;
    int64_t memory_budget; /* device bytes for matrices, workspaces, staging; 0 = 90% of free */
    int max_batch;         /* queries per batch, 1..1024 (0 = 256)                 */
    int slots;             /* batches in flight, 1..8 (0 = 4)                       */
    int ck;                /* CK of every query (0 = 512)                           */
    int max_pending;       /* waiting requests before submissions are refused (0 = 65536) */
    int autotune;          /* 1: acc_gpu_autotune each matrix (K = 100) when it is added */
    int64_t max_request_bytes; /* host bytes of waiting and unread requests (0 = 1 GiB) */
} acc_server_opts;
typedef struct {
    int64_t submitted, completed, refused, batches;
    double mean_batch;         /* queries per batch                                 */
    double p50_us, p99_us, max_us; /* latency from arrival to completion            */
    double qps;                /* completed per second since the last reset         */
    int64_t device_bytes;      /* held by the server                                */
    int64_t request_bytes;     /* host bytes held by waiting and unread requests    */
    int64_t peak_inflight;     /* most batches of one matrix in flight at once      */
} acc_server_stats_t;
int acc_server_create(const acc_server_opts *opts, acc_server **out);
/* Uploads the matrix (FP32 CSR for rescoring + packets), runs a warm-up
 * batch and checks the budget. */
int acc_server_add_matrix(acc_server *s, const acc_csr *csr, const acc_packed *pk, int *matrix_id);
/* Queues one query (x: cols floats, copied); 1 <= k <= min(960, ck). */
int acc_server_submit(acc_server *s, int matrix_id, const float *x, int k, uint64_t *ticket);
/* Waits for a ticket (timeout_ms < 0: no limit) and copies its k results. */
int acc_server_wait(acc_server *s, uint64_t ticket, acc_element *out, int timeout_ms);
int acc_server_stats(acc_server *s, acc_server_stats_t *stats);
int acc_server_reset_stats(acc_server *s);
/* Open-loop load: Poisson arrivals at rate_qps for `seconds`, each query
 * N(0,1) to a uniformly chosen matrix with K uniform in [1, kmax]; latency is
 * measured from the scheduled arrival.  Returns the run's statistics. */
int acc_server_bench_poisson(acc_server *s, double rate_qps, double seconds, int kmax, uint64_t seed,
                             acc_server_stats_t *stats);
void acc_server_destroy(acc_server *s);

/* Stage outputs, for parity checks.
 *   spmv_fp32:      y_dev[rows] FP32 scores (the Top-K path's order)
 *   scores_lowbit:  y_dev[rows] INT32 stage-1 scores, original row order
 *   candidates:     per partition its top floor(ck/split) (row, INT32 score)
 *                   in ranking order, partitions in order; *count_host gets
 *                   the total (synchronises the stream). */
int acc_gpu_spmv_fp32(acc_gpu_matrix *m, const float *x_dev, float *y_dev, void *stream);
int acc_gpu_scores_lowbit(acc_gpu_matrix *m, const float *x_dev, int32_t *y_dev, void *stream);
int acc_gpu_candidates_lowbit(acc_gpu_matrix *m, const float *x_dev, int ck, int32_t *rows_dev,
                              int32_t *scores_dev, int *count_host, void *stream);

/* top_k_heap (cpu_mkl_baseline.cpp:128-149) over a device score vector. */
int acc_gpu_top_k_heap(const float *y_dev, int n, int k, acc_element *out_dev, void *stream);

/* acc_gpu_top_k_heap without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * selection scratch is a static device array of 3.75 MiB per GPU, held until
 * the process ends; limits as acc_gpu_top_k_heap (0 <= n < 2^31,
 * 1 <= k <= 960).  Every call on one GPU shares that GPU's array and resets
 * the state it needs at the start (by a kernel, so the call can be captured):
 * calls that overlap in time on one GPU race and must be serialised by the
 * caller (one stream). */
int acc_gpu_top_k_heap_noalloc(const float *y_dev, int n, int k, acc_element *out_dev, void *stream);

/* Host-pointer convenience forms: copy in, run on the handle's device,
 * copy out, synchronise. */
int acc_topk_fp32_host(acc_gpu_matrix *m, const float *x_host, int k, acc_element *out_host);
int acc_topk_lowbit_host(acc_gpu_matrix *m, const float *x_host, int k, int ck,
                         acc_element *out_host);

/* top_k_spmv (cpu_mkl_baseline.cpp:249-275) drop-in: num_test queries of
 * U(-1,1) values (generate_random_vector :111-122, seeded here), each one
 * FP32 SpMV + Top-k on the GPU; prints the reference's three timing lines
 * and leaves the last query's results in top_k.  num_threads is accepted for
 * signature compatibility and ignored. */
int acc_top_k_spmv(acc_gpu_matrix *A, int row, int col, int k, int num_threads,
                   acc_element *top_k, int num_test);

#ifdef __cplusplus
}
#endif
#endif /* ACCELES_H */
