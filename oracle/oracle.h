/*
 * oracle.h -- CPU restatement of the AccelES Top-K SpMV path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the
 * smoke() entry of __graft_entry__.py and the cpu_baseline leg of bench.py
 * may load it; the product library (acceles_amd/libacceles.so) never links,
 * loads or calls it, and has no CPU fallback for any query.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Where the reference variants disagree (tie rules, rounding), the declared
 * rule is stated next to the function; see the per-function comments.
 *
 * Declared rules used throughout:
 *   - ranking order: score descending, then row index ascending (the lower
 *     row wins a tie at the cut), i.e. the strict ">" admission of
 *     cpu_mkl_baseline.cpp:138 made total;
 *   - rounding: round half to even (numpy np.round, quantify_util.py:27,49);
 *   - FP32 dot products: a G-lane strided order (lane g sums entries
 *     g, g+G, g+2G, ... of the row with fmaf, then an xor-butterfly tree of
 *     plain adds), the order the GPU kernels use; G=1 is the sequential
 *     fmaf order.
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   orc_coo2csr            reference coo2csr built in oracle/_ref, CSR invariants
 *   orc_quantize           hand-computed half-way/clip cases, closed form
 *   orc_spmv_i32           dense int64 brute force, reference spmv_coo_gold_top_k_2
 *   orc_spmv_f32_grouped   FP64 brute force within FP32 error bound; G=1 equals
 *                          sequential fmaf; butterfly independent of lane order
 *   orc_spmv_f64           dense numpy FP64 product
 *   orc_topk               numpy lexsort brute force, reference gold Top-K
 *                          (tie-free inputs), reference heap order on the
 *                          golden fixtures
 *   orc_split_select       per-partition brute force, partition sizes of
 *                          quantify_util.py:185-189
 *   orc_split_select_order numpy lexsort brute force per partition of a
 *                          shuffled order (ties among equal scores included);
 *                          identity order equals orc_split_select
 *   orc_rescore_topk       orc_spmv_f32_grouped + orc_topk on the candidate rows
 *   orc_precision          hand-computed examples of quantify_util.py:198-241
 *   orc_ndcg_score         hand-computed lists (equal lists give 1, a
 *                          swap, a miss, a repeated row)
 *   orc_resparse_mask      hand-computed rows (kept whole, ceil rounding,
 *                          ties, empty rows, both means) and a sort-based
 *                          numpy restatement on random rows
 *   orc_ultracsr_*         x-decoder known-answer test (test_x_decoder.cpp:6),
 *                          bit positions of fpga_utils.hpp:438-503, round trip
 */
#ifndef ACCELES_ORACLE_H
#define ACCELES_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* COO -> CSR by counting sort, stable within a row (utils.hpp:494-547).
 * sort_tuples=1 first orders entries by (row, col), keeping input order for
 * equal pairs (utils.hpp:331-351 with a stable sort).  Returns 0, or -1 when
 * an index is out of range (the reference prints and continues). */
int orc_coo2csr(int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *row,
                const int32_t *col, const float *val, int sort_tuples,
                int64_t *ptr, int32_t *idx, float *vout);

/* Uniform quantization with one global scale (quantify_util.py:8-28, 31-49):
 *   M = 2^bits - 1, halved (integer) when symmetric;
 *   s = max(|max v|, |min v|);  q = clip(rint(M * (v / s)), -M-1, M)
 * evaluated in float32 as numpy does on a float32 array.  s == 0 gives q = 0.
 * Writes s to *scale_out. */
void orc_quantize(int64_t n, const float *v, int bits, int symmetric,
                  float *scale_out, int32_t *q);

/* y = A x in FP32, G-lane strided fmaf order + xor butterfly (see header). */
void orc_spmv_f32_grouped(int32_t nrows, const int64_t *ptr, const int32_t *idx,
                          const float *val, const float *x, int G, float *y);

/* y = A x in FP64 (quantify_util.py:257, matrix_vector_multiply on float64). */
void orc_spmv_f64(int32_t nrows, const int64_t *ptr, const int32_t *idx,
                  const float *val, const double *x, double *y);

/* Exact integer scores y = Aq xq (quantify_util.py:193 on int32 arrays).
 * Returns -1 if any partial sum leaves the int32 range, else 0. */
int orc_spmv_i32(int32_t nrows, const int64_t *ptr, const int32_t *idx,
                 const int32_t *qv, const int32_t *qx, int32_t *y);

/* Top-k by min-heap (cpu_mkl_baseline.cpp:128-149) over n scores given as
 * doubles (exact for int32 and float inputs); output in ranking order.
 * k > n fills the tail with (-inf, -1). */
void orc_topk(int64_t n, const double *y, int k, int32_t *out_idx, double *out_val);

/* Per-partition candidate select (quantify_util.py:173-196): partitions of
 * nrows/split rows, the remainder joining the last one; each keeps its top
 * floor(ck/split) rows.  Output grouped by partition, each group in ranking
 * order.  Returns the number of candidates, or -1 on bad arguments. */
int orc_split_select(int32_t nrows, const int32_t *y, int split, int ck,
                     int32_t *cand_rows, int32_t *cand_scores);

/* Random-CSR (README.md:5; quantify_util.py:491 shuffles the rows before the
 * partitions of :184-189): the same select with the rows dealt to partitions
 * in `order` (partition p holds rows order[p*size ...]; NULL = 0, 1, ...).
 * Rows keep their ids: ties still go to the lower row id. */
int orc_split_select_order(int32_t nrows, const int32_t *y, const int32_t *order, int split, int ck,
                           int32_t *cand_rows, int32_t *cand_scores);

/* Stage 2 (quantify_util.py:354-367, in FP32): rescore the candidate rows
 * against the unquantized query with orc_spmv_f32_grouped's order, then
 * keep the top k in ranking order. */
void orc_rescore_topk(const int64_t *ptr, const int32_t *idx, const float *val,
                      const float *x, int G, int ncand, const int32_t *cand_rows,
                      int k, int32_t *out_idx, float *out_val);

/* Whole low-bit two-stage query on a matrix whose stored values were already
 * quantized to qv (any signedness): quantize x to signed 8-bit with its own
 * scale, exact integer scores, split select, FP32 rescoring, top k.
 * cand_rows (size >= ck) receives the stage-1 candidates; returns their
 * count, or -1 on error. */
int orc_two_stage(int32_t nrows, int32_t ncols, const int64_t *ptr,
                  const int32_t *idx, const float *val, const int32_t *qv,
                  const float *x, int G, int split, int ck, int k,
                  int32_t *cand_rows, int32_t *out_idx, float *out_val);
/* orc_two_stage over the partitions of a row order (orc_split_select_order). */
int orc_two_stage_order(int32_t nrows, int32_t ncols, const int64_t *ptr, const int32_t *idx, const float *val,
                        const int32_t *qv, const float *x, int G, int split, int ck, int k, const int32_t *order,
                        int32_t *cand_rows, int32_t *out_idx, float *out_val);

/* Precision of one query (quantify_util.py:198-219): for each K in ks,
 * |set(cand) & set(exact[0:K])| / K. */
void orc_precision(const int32_t *exact, int ncand, const int32_t *cand,
                   int nks, const int32_t *ks, double *prec_out);

/* ReSparse keep mask (resparse_csr, csr_util.py:152-187; resparse_spm,
 * quantify_util.py:375-403 with mean_over_cells): avg = mean |v| over the
 * nonzeros (or all rows*cols cells); a row whose every |v| > avg is kept
 * whole, any other keeps its ceil(n*keep) entries of largest |v| (equal |v|:
 * the earlier entry).  By rank counting per entry (brute force). */
void orc_resparse_mask(int32_t rows, int32_t cols, const int64_t *ptr, const float *val, double keep,
                       int mean_over_cells, uint8_t *mask);

/* Score NDCG (quantify_util.py:313-329): relevance of a row = its sw score
 * (the last listing; 0 when absent), dcg over hw places, idcg over sw. */
double orc_ndcg_score(const int32_t *sw_idx, const double *sw_val, int nsw, const int32_t *hw_idx, int nhw);

/* x-decoder (fpga_utils.hpp:279-302): x[i] = carry + popcount(flags &
 * ((2 << i) - 1)) for i < n. */
void orc_xdecode(uint32_t flags, int n, int32_t carry, int32_t *x);

/* Ultra-CSR packet geometry (types.hpp:151-162 generalised): n entries of
 * (1-bit flag, cb-bit column, vb-bit value) per 512-bit packet,
 * n = 511 / (1 + cb + vb); flags at bits [0,n), columns at [n, n+n*cb),
 * values at [n+n*cb, n+n*(cb+vb)). */
int orc_ultracsr_entries(int cb, int vb);

/* Pack one partition's COO stream (host_spmv_ucsr.cpp:388-425): flag =
 * (row != previous entry's row), the previous row starting at last_row;
 * tail entries of the last packet are zero.  Packets are 8 little-endian
 * 64-bit words.  Returns the number of packets written (<= cap). */
int64_t orc_ultracsr_pack(int64_t nnz, const int32_t *row, const int32_t *col,
                          const int32_t *qval, int32_t last_row, int cb, int vb,
                          uint64_t *packets, int64_t cap);

/* Inverse of orc_ultracsr_pack for nnz entries: per entry the decoded row
 * counter (x-decoder carried across packets, starting at carry0), column and
 * sign-extended value (fpga_utils.hpp:317-384). */
void orc_ultracsr_unpack(const uint64_t *packets, int64_t nnz, int cb, int vb,
                         int32_t carry0, int32_t *xrow, int32_t *col, int32_t *qval);

/* The reference CPU query (cpu_mkl_baseline.cpp:249-275) for one query:
 * FP32 SpMV (sequential fmaf order, G=1) into y, then the heap Top-k.
 * Used as the timed CPU baseline. */
void orc_top_k_spmv_f32(int32_t nrows, const int64_t *ptr, const int32_t *idx,
                        const float *val, const float *x, int k, float *y,
                        int32_t *out_idx, float *out_val);

/* All-core timing legs (OpenMP), the same results as orc_top_k_spmv_f32 and
 * orc_two_stage: the CPU baselines of the benchmark beside the one-thread
 * ports.  Pinned by equality with the one-thread functions. */
void orc_top_k_spmv_f32_par(int32_t nrows, const int64_t *ptr, const int32_t *idx, const float *val,
                            const float *x, int k, int nthreads, float *y, int32_t *out_idx, float *out_val);
int orc_two_stage_par(int32_t nrows, int32_t ncols, const int64_t *ptr, const int32_t *idx, const float *val,
                      const int32_t *qv, const float *x, int G, int split, int ck, int k, int nthreads,
                      int32_t *cand_rows, int32_t *out_idx, float *out_val);

#ifdef __cplusplus
}
#endif
#endif
