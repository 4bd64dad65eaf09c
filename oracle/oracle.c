/*
 * oracle.c -- CPU restatement of the AccelES Top-K SpMV path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it.  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off; no fast-math, so the float operations below are
 * exactly the IEEE operations written).
 *
 * See oracle.h for the declared rules and the checks that pin each function.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* ranking order: score descending, row ascending                      */
/* ------------------------------------------------------------------ */

typedef struct {
    double s;
    int32_t r;
} orc_item;

/* a ranks before b */
static int better(orc_item a, orc_item b) {
    if (a.s != b.s) return a.s > b.s;
    return a.r < b.r;
}

/* min-heap on the ranking order: heap[0] is the worst kept item */
static void sift_down(orc_item *h, int n, int i) {
    for (;;) {
        int l = 2 * i + 1, r = l + 1, w = i;
        if (l < n && better(h[w], h[l])) w = l;
        if (r < n && better(h[w], h[r])) w = r;
        if (w == i) return;
        orc_item t = h[i]; h[i] = h[w]; h[w] = t;
        i = w;
    }
}

static void sift_up(orc_item *h, int i) {
    while (i > 0) {
        int p = (i - 1) / 2;
        if (!better(h[p], h[i])) return;
        orc_item t = h[i]; h[i] = h[p]; h[p] = t;
        i = p;
    }
}

static int cmp_rank(const void *pa, const void *pb) {
    orc_item a = *(const orc_item *)pa, b = *(const orc_item *)pb;
    if (better(a, b)) return -1;
    if (better(b, a)) return 1;
    return 0;
}

/* cpu_mkl_baseline.cpp:128-149: fill a heap of k, admit an item iff it ranks
 * strictly before the heap's worst, then emit best-first. */
static int heap_topk(const orc_item *items, int64_t n, int k, orc_item *out) {
    if (k <= 0) return 0;
    orc_item *h = (orc_item *)malloc(sizeof(orc_item) * (size_t)k);
    int hn = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (hn < k) {
            h[hn] = items[i];
            sift_up(h, hn);
            ++hn;
        } else if (better(items[i], h[0])) {
            h[0] = items[i];
            sift_down(h, hn, 0);
        }
    }
    qsort(h, (size_t)hn, sizeof(orc_item), cmp_rank);
    memcpy(out, h, sizeof(orc_item) * (size_t)hn);
    free(h);
    return hn;
}

void orc_topk(int64_t n, const double *y, int k, int32_t *out_idx, double *out_val) {
    orc_item *items = (orc_item *)malloc(sizeof(orc_item) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) { items[i].s = y[i]; items[i].r = (int32_t)i; }
    orc_item *out = (orc_item *)malloc(sizeof(orc_item) * (size_t)(k > 0 ? k : 1));
    int m = heap_topk(items, n, k, out);
    for (int i = 0; i < k; ++i) {
        out_idx[i] = i < m ? out[i].r : -1;
        out_val[i] = i < m ? out[i].s : -INFINITY;
    }
    free(out);
    free(items);
}

/* ------------------------------------------------------------------ */
/* COO -> CSR                                                          */
/* ------------------------------------------------------------------ */

typedef struct {
    int32_t r, c;
    int64_t i;
} rc_item;

static int cmp_rc(const void *pa, const void *pb) {
    const rc_item *a = (const rc_item *)pa, *b = (const rc_item *)pb;
    if (a->r != b->r) return a->r < b->r ? -1 : 1;
    if (a->c != b->c) return a->c < b->c ? -1 : 1;
    return a->i < b->i ? -1 : (a->i > b->i);
}

/* utils.hpp:494-547 (counting sort; the scatter walks the input in order, so
 * entries of one row keep their input order) and utils.hpp:331-351 for the
 * optional (row, col) pre-sort. */
int orc_coo2csr(int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *row,
                const int32_t *col, const float *val, int sort_tuples,
                int64_t *ptr, int32_t *idx, float *vout) {
    for (int64_t i = 0; i < nnz; ++i)
        if (row[i] < 0 || row[i] >= nrows || col[i] < 0 || col[i] >= ncols) return -1;
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
    if (sort_tuples) {
        rc_item *t = (rc_item *)malloc(sizeof(rc_item) * (size_t)(nnz > 0 ? nnz : 1));
        for (int64_t i = 0; i < nnz; ++i) { t[i].r = row[i]; t[i].c = col[i]; t[i].i = i; }
        qsort(t, (size_t)nnz, sizeof(rc_item), cmp_rc);
        for (int64_t i = 0; i < nnz; ++i) order[i] = t[i].i;
        free(t);
    } else {
        for (int64_t i = 0; i < nnz; ++i) order[i] = i;
    }
    for (int32_t r = 0; r <= nrows; ++r) ptr[r] = 0;
    for (int64_t i = 0; i < nnz; ++i) ptr[row[i] + 1]++;
    for (int32_t r = 0; r < nrows; ++r) ptr[r + 1] += ptr[r];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nrows > 0 ? nrows : 1));
    for (int32_t r = 0; r < nrows; ++r) fill[r] = ptr[r];
    for (int64_t t = 0; t < nnz; ++t) {
        int64_t i = order[t];
        int64_t d = fill[row[i]]++;
        idx[d] = col[i];
        vout[d] = val[i];
    }
    free(fill);
    free(order);
    return 0;
}

/* ------------------------------------------------------------------ */
/* quantization                                                        */
/* ------------------------------------------------------------------ */

void orc_quantize(int64_t n, const float *v, int bits, int symmetric,
                  float *scale_out, int32_t *q) {
    int32_t M = (int32_t)((1u << bits) - 1u);
    if (symmetric) M = M / 2;
    float mx = 0.0f, mn = 0.0f;
    if (n > 0) { mx = v[0]; mn = v[0]; }
    for (int64_t i = 1; i < n; ++i) {
        if (v[i] > mx) mx = v[i];
        if (v[i] < mn) mn = v[i];
    }
    float s = fabsf(mx) > fabsf(mn) ? fabsf(mx) : fabsf(mn);
    *scale_out = s;
    for (int64_t i = 0; i < n; ++i) {
        if (s == 0.0f) { q[i] = 0; continue; }
        volatile float t = v[i] / s;         /* matrix /= max(...)        */
        volatile float u = (float)M * t;     /* M * matrix                */
        float r = nearbyintf(u);             /* np.round: half to even    */
        int32_t qi = (int32_t)r;
        if (qi > M) qi = M;
        if (qi < -M - 1) qi = -M - 1;
        q[i] = qi;
    }
}

/* ------------------------------------------------------------------ */
/* SpMV                                                                */
/* ------------------------------------------------------------------ */

static float dot_grouped(int64_t b, int64_t e, const int32_t *idx, const float *val,
                         const float *x, int G) {
    float part[64];
    for (int g = 0; g < G; ++g) {
        float p = 0.0f;
        for (int64_t j = b + g; j < e; j += G) p = fmaf(val[j], x[idx[j]], p);
        part[g] = p;
    }
    for (int s = G / 2; s >= 1; s /= 2) {
        float nxt[64];
        for (int g = 0; g < G; ++g) nxt[g] = part[g] + part[g ^ s];
        memcpy(part, nxt, sizeof(float) * (size_t)G);
    }
    return part[0];
}

void orc_spmv_f32_grouped(int32_t nrows, const int64_t *ptr, const int32_t *idx,
                          const float *val, const float *x, int G, float *y) {
    for (int32_t r = 0; r < nrows; ++r) y[r] = dot_grouped(ptr[r], ptr[r + 1], idx, val, x, G);
}

void orc_spmv_f64(int32_t nrows, const int64_t *ptr, const int32_t *idx,
                  const float *val, const double *x, double *y) {
    for (int32_t r = 0; r < nrows; ++r) {
        double acc = 0.0;
        for (int64_t j = ptr[r]; j < ptr[r + 1]; ++j) acc += (double)val[j] * x[idx[j]];
        y[r] = acc;
    }
}

int orc_spmv_i32(int32_t nrows, const int64_t *ptr, const int32_t *idx,
                 const int32_t *qv, const int32_t *qx, int32_t *y) {
    int bad = 0;
    for (int32_t r = 0; r < nrows; ++r) {
        int64_t acc = 0;
        for (int64_t j = ptr[r]; j < ptr[r + 1]; ++j) {
            acc += (int64_t)qv[j] * (int64_t)qx[idx[j]];
            if (acc > INT32_MAX || acc < INT32_MIN) bad = 1;
        }
        y[r] = (int32_t)acc;
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* two-stage                                                           */
/* ------------------------------------------------------------------ */

int orc_split_select(int32_t nrows, const int32_t *y, int split, int ck,
                     int32_t *cand_rows, int32_t *cand_scores) {
    return orc_split_select_order(nrows, y, NULL, split, ck, cand_rows, cand_scores);
}

int orc_split_select_order(int32_t nrows, const int32_t *y, const int32_t *order, int split, int ck,
                           int32_t *cand_rows, int32_t *cand_scores) {
    if (split < 1 || ck < split || nrows < split) return -1;
    int32_t size = nrows / split;
    int m = ck / split;
    int n = 0;
    orc_item *items = (orc_item *)malloc(sizeof(orc_item) * (size_t)nrows);
    orc_item *out = (orc_item *)malloc(sizeof(orc_item) * (size_t)m);
    for (int p = 0; p < split; ++p) {
        int32_t b = p * size;
        int32_t e = (p == split - 1) ? nrows : b + size;
        for (int32_t i = b; i < e; ++i) {
            const int32_t r = order ? order[i] : i;
            items[i - b].s = y[r];
            items[i - b].r = r;
        }
        int got = heap_topk(items, e - b, m, out);
        for (int i = 0; i < got; ++i) {
            cand_rows[n] = out[i].r;
            cand_scores[n] = (int32_t)out[i].s;
            ++n;
        }
    }
    free(out);
    free(items);
    return n;
}

void orc_rescore_topk(const int64_t *ptr, const int32_t *idx, const float *val,
                      const float *x, int G, int ncand, const int32_t *cand_rows,
                      int k, int32_t *out_idx, float *out_val) {
    orc_item *items = (orc_item *)malloc(sizeof(orc_item) * (size_t)(ncand > 0 ? ncand : 1));
    for (int i = 0; i < ncand; ++i) {
        int32_t r = cand_rows[i];
        float s = dot_grouped(ptr[r], ptr[r + 1], idx, val, x, G);
        items[i].s = s;
        items[i].r = r;
    }
    orc_item *out = (orc_item *)malloc(sizeof(orc_item) * (size_t)(k > 0 ? k : 1));
    int m = heap_topk(items, ncand, k, out);
    for (int i = 0; i < k; ++i) {
        out_idx[i] = i < m ? out[i].r : -1;
        out_val[i] = i < m ? (float)out[i].s : -INFINITY;
    }
    free(out);
    free(items);
}

int orc_two_stage(int32_t nrows, int32_t ncols, const int64_t *ptr,
                  const int32_t *idx, const float *val, const int32_t *qv,
                  const float *x, int G, int split, int ck, int k,
                  int32_t *cand_rows, int32_t *out_idx, float *out_val) {
    return orc_two_stage_order(nrows, ncols, ptr, idx, val, qv, x, G, split, ck, k, NULL, cand_rows, out_idx,
                               out_val);
}

int orc_two_stage_order(int32_t nrows, int32_t ncols, const int64_t *ptr, const int32_t *idx, const float *val,
                        const int32_t *qv, const float *x, int G, int split, int ck, int k, const int32_t *order,
                        int32_t *cand_rows, int32_t *out_idx, float *out_val) {
    int32_t *qx = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncols);
    float sx;
    orc_quantize(ncols, x, 8, 1, &sx, qx);
    int32_t *y = (int32_t *)malloc(sizeof(int32_t) * (size_t)nrows);
    int rc = orc_spmv_i32(nrows, ptr, idx, qv, qx, y);
    int n = -1;
    if (rc == 0) {
        int32_t *cs = (int32_t *)malloc(sizeof(int32_t) * (size_t)ck);
        n = orc_split_select_order(nrows, y, order, split, ck, cand_rows, cs);
        if (n >= 0) orc_rescore_topk(ptr, idx, val, x, G, n, cand_rows, k, out_idx, out_val);
        free(cs);
    }
    free(y);
    free(qx);
    return n;
}

void orc_precision(const int32_t *exact, int ncand, const int32_t *cand,
                   int nks, const int32_t *ks, double *prec_out) {
    for (int t = 0; t < nks; ++t) {
        int K = ks[t], hit = 0;
        for (int i = 0; i < K; ++i) {
            int found = 0;
            for (int j = 0; j < i; ++j)
                if (exact[j] == exact[i]) found = -1; /* set semantics */
            if (found < 0) continue;
            for (int j = 0; j < ncand; ++j)
                if (cand[j] == exact[i]) { found = 1; break; }
            hit += found;
        }
        prec_out[t] = (double)hit / (double)K;
    }
}

/* ------------------------------------------------------------------ */
/* Ultra-CSR reference packet format                                   */
/* ------------------------------------------------------------------ */

void orc_xdecode(uint32_t flags, int n, int32_t carry, int32_t *x) {
    for (int i = 0; i < n; ++i) {
        uint32_t mask = (i >= 31) ? 0xFFFFFFFFu : ((2u << i) - 1u);
        x[i] = carry + __builtin_popcount(flags & mask);
    }
}

int orc_ultracsr_entries(int cb, int vb) { return 511 / (1 + cb + vb); }

static void put_bits(uint64_t *w, int pos, int len, uint64_t v) {
    for (int b = 0; b < len; ++b)
        if ((v >> b) & 1u) w[(pos + b) / 64] |= (uint64_t)1 << ((pos + b) % 64);
}

static uint64_t get_bits(const uint64_t *w, int pos, int len) {
    uint64_t v = 0;
    for (int b = 0; b < len; ++b) v |= ((w[(pos + b) / 64] >> ((pos + b) % 64)) & 1u) << b;
    return v;
}

int64_t orc_ultracsr_pack(int64_t nnz, const int32_t *row, const int32_t *col,
                          const int32_t *qval, int32_t last_row, int cb, int vb,
                          uint64_t *packets, int64_t cap) {
    int n = orc_ultracsr_entries(cb, vb);
    int64_t np = (nnz + n - 1) / n;
    if (np > cap) return -1;
    memset(packets, 0, sizeof(uint64_t) * 8 * (size_t)np);
    int32_t prev = last_row;
    for (int64_t i = 0; i < nnz; ++i) {
        uint64_t *w = packets + 8 * (i / n);
        int j = (int)(i % n);
        uint64_t flag = row[i] != prev;
        prev = row[i];
        put_bits(w, j, 1, flag);
        put_bits(w, n + j * cb, cb, (uint64_t)(uint32_t)col[i] & ((1ull << cb) - 1));
        put_bits(w, n + n * cb + j * vb, vb, (uint64_t)(uint32_t)qval[i] & ((1ull << vb) - 1));
    }
    return np;
}

void orc_ultracsr_unpack(const uint64_t *packets, int64_t nnz, int cb, int vb,
                         int32_t carry0, int32_t *xrow, int32_t *col, int32_t *qval) {
    int n = orc_ultracsr_entries(cb, vb);
    int32_t carry = carry0;
    int32_t x[64];
    for (int64_t p = 0; p * n < nnz; ++p) {
        const uint64_t *w = packets + 8 * p;
        uint32_t flags = (uint32_t)get_bits(w, 0, n);
        orc_xdecode(flags, n, carry, x);
        carry = x[n - 1];
        for (int j = 0; j < n && p * n + j < nnz; ++j) {
            int64_t i = p * n + j;
            xrow[i] = x[j];
            col[i] = (int32_t)get_bits(w, n + j * cb, cb);
            uint32_t v = (uint32_t)get_bits(w, n + n * cb + j * vb, vb);
            if (v & (1u << (vb - 1))) v |= ~((1u << vb) - 1u); /* two's complement */
            qval[i] = (int32_t)v;
        }
    }
}

/* ------------------------------------------------------------------ */
/* reference CPU query, timed as the CPU baseline                      */
/* ------------------------------------------------------------------ */

void orc_top_k_spmv_f32(int32_t nrows, const int64_t *ptr, const int32_t *idx,
                        const float *val, const float *x, int k, float *y,
                        int32_t *out_idx, float *out_val) {
    orc_spmv_f32_grouped(nrows, ptr, idx, val, x, 1, y);
    orc_item *h = (orc_item *)malloc(sizeof(orc_item) * (size_t)(k > 0 ? k : 1));
    int hn = 0;
    for (int32_t i = 0; i < nrows; ++i) {
        orc_item it = {y[i], i};
        if (hn < k) { h[hn] = it; sift_up(h, hn); ++hn; }
        else if (better(it, h[0])) { h[0] = it; sift_down(h, hn, 0); }
    }
    qsort(h, (size_t)hn, sizeof(orc_item), cmp_rank);
    for (int i = 0; i < k; ++i) {
        out_idx[i] = i < hn ? h[i].r : -1;
        out_val[i] = i < hn ? (float)h[i].s : -INFINITY;
    }
    free(h);
}

/* ------------------------------------------------------------------ */
/* ReSparse (csr_util.py:152-187; quantify_util.py:375-403)            */
/* ------------------------------------------------------------------ */

void orc_resparse_mask(int32_t rows, int32_t cols, const int64_t *ptr, const float *val, double keep,
                       int mean_over_cells, uint8_t *mask) {
    double sum = 0.0;
    const int64_t nnz = ptr[rows];
    for (int64_t j = 0; j < nnz; ++j) sum += fabs((double)val[j]);
    const double denom = mean_over_cells ? (double)rows * (double)cols : (double)nnz;
    const double avg = denom > 0 ? sum / denom : 0.0;
    for (int32_t r = 0; r < rows; ++r) {
        const int64_t b = ptr[r], e = ptr[r + 1], n = e - b;
        int all = n > 0;
        for (int64_t j = b; j < e; ++j)
            if (!(fabs((double)val[j]) > avg)) all = 0;
        const int64_t nk = all ? n : (int64_t)ceil((double)n * keep);
        /* entry j's place in the row's order by |v| descending, earlier first */
        for (int64_t j = b; j < e; ++j) {
            int64_t rank = 0;
            for (int64_t t = b; t < e; ++t) {
                const double a = fabs((double)val[t]), c = fabs((double)val[j]);
                if (a > c || (a == c && t < j)) ++rank;
            }
            mask[j] = rank < nk;
        }
    }
}

/* ------------------------------------------------------------------ */
/* score NDCG (quantify_util.py:313-329)                               */
/* ------------------------------------------------------------------ */

double orc_ndcg_score(const int32_t *sw_idx, const double *sw_val, int nsw, const int32_t *hw_idx, int nhw) {
    double dcg = 0.0, idcg = 0.0;
    for (int i = 0; i < nhw; ++i) {
        double rel = 0.0; /* the last listing of the row in sw, as a dict built in order keeps */
        for (int j = 0; j < nsw; ++j)
            if (sw_idx[j] == hw_idx[i]) rel = sw_val[j];
        dcg += rel / log2((double)i + 2.0);
    }
    for (int i = 0; i < nsw; ++i) idcg += sw_val[i] / log2((double)i + 2.0);
    return dcg / idcg;
}

/* ------------------------------------------------------------------ */
/* all-core CPU baselines (timing legs beside the one-thread ports)    */
/* ------------------------------------------------------------------ */

/* top_k_spmv (cpu_mkl_baseline.cpp:249-275) on nthreads cores: the SpMV
 * split into row ranges (one fmaf chain per row, as the one-thread port),
 * then the heap Top-k of each thread's rows, then the Top-k of the threads'
 * lists.  The same rows and scores as orc_top_k_spmv_f32 (the ranking order
 * is a total order).  The reference's own "32 core" mode
 * (cpu_mkl_baseline.cpp:195-245) times only thread creation and is not
 * reproduced. */
void orc_top_k_spmv_f32_par(int32_t nrows, const int64_t *ptr, const int32_t *idx, const float *val,
                            const float *x, int k, int nthreads, float *y, int32_t *out_idx, float *out_val) {
    if (nthreads < 1) nthreads = 1;
    orc_item *lists = (orc_item *)malloc(sizeof(orc_item) * (size_t)nthreads * (size_t)(k > 0 ? k : 1));
    int *counts = (int *)calloc((size_t)nthreads, sizeof(int));
#pragma omp parallel num_threads(nthreads)
    {
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
        const int32_t b = (int32_t)((int64_t)nrows * t / nt), e = (int32_t)((int64_t)nrows * (t + 1) / nt);
        for (int32_t r = b; r < e; ++r) y[r] = dot_grouped(ptr[r], ptr[r + 1], idx, val, x, 1);
        orc_item *h = lists + (size_t)t * (size_t)(k > 0 ? k : 1);
        int hn = 0;
        for (int32_t i = b; i < e; ++i) {
            orc_item it = {y[i], i};
            if (hn < k) { h[hn] = it; sift_up(h, hn); ++hn; }
            else if (better(it, h[0])) { h[0] = it; sift_down(h, hn, 0); }
        }
        counts[t] = hn;
    }
    int total = 0;
    for (int t = 0; t < nthreads; ++t) {
        memmove(lists + total, lists + (size_t)t * (size_t)(k > 0 ? k : 1), sizeof(orc_item) * (size_t)counts[t]);
        total += counts[t];
    }
    orc_item *out = (orc_item *)malloc(sizeof(orc_item) * (size_t)(k > 0 ? k : 1));
    const int m = heap_topk(lists, total, k, out);
    for (int i = 0; i < k; ++i) {
        out_idx[i] = i < m ? out[i].r : -1;
        out_val[i] = i < m ? (float)out[i].s : -INFINITY;
    }
    free(out);
    free(counts);
    free(lists);
}

/* orc_two_stage on nthreads cores: the INT32 SpMV over row ranges and the
 * per-partition selection over partitions in parallel, then the rescoring
 * of the candidates (the same results as orc_two_stage). */
int orc_two_stage_par(int32_t nrows, int32_t ncols, const int64_t *ptr, const int32_t *idx, const float *val,
                      const int32_t *qv, const float *x, int G, int split, int ck, int k, int nthreads,
                      int32_t *cand_rows, int32_t *out_idx, float *out_val) {
    if (split < 1 || ck < split || nrows < split) return -1;
    if (nthreads < 1) nthreads = 1;
    int32_t *qx = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncols);
    float sx;
    orc_quantize(ncols, x, 8, 1, &sx, qx);
    int32_t *y = (int32_t *)malloc(sizeof(int32_t) * (size_t)nrows);
    int bad = 0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(| : bad)
    for (int32_t r = 0; r < nrows; ++r) {
        int64_t acc = 0;
        for (int64_t j = ptr[r]; j < ptr[r + 1]; ++j) acc += (int64_t)qv[j] * (int64_t)qx[idx[j]];
        if (acc > INT32_MAX || acc < INT32_MIN) bad = 1;
        y[r] = (int32_t)acc;
    }
    int n = -1;
    if (!bad) {
        const int32_t size = nrows / split;
        const int m = ck / split;
        int *got = (int *)malloc(sizeof(int) * (size_t)split);
        orc_item *outs = (orc_item *)malloc(sizeof(orc_item) * (size_t)split * (size_t)m);
#pragma omp parallel num_threads(nthreads)
        {
            orc_item *items = (orc_item *)malloc(sizeof(orc_item) * (size_t)(nrows - size * (split - 1)));
#pragma omp for schedule(dynamic)
            for (int p = 0; p < split; ++p) {
                const int32_t b = p * size, e = (p == split - 1) ? nrows : b + size;
                for (int32_t r = b; r < e; ++r) { items[r - b].s = y[r]; items[r - b].r = r; }
                got[p] = heap_topk(items, e - b, m, outs + (size_t)p * (size_t)m);
            }
            free(items);
        }
        n = 0;
        for (int p = 0; p < split; ++p)
            for (int i = 0; i < got[p]; ++i) cand_rows[n++] = outs[(size_t)p * (size_t)m + (size_t)i].r;
        orc_rescore_topk(ptr, idx, val, x, G, n, cand_rows, k, out_idx, out_val);
        free(outs);
        free(got);
    }
    free(y);
    free(qx);
    return n;
}
