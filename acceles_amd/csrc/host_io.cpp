// host_io.cpp -- errors, Matrix Market loaders/writer, COO/CSR handles,
// COO->CSR conversion, quantization, row permutation and the seeded
// synthetic generators.  Host C++ (OpenMP); no GPU code here.
#include "internal.hpp"

#include <omp.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <unordered_map>

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

static thread_local std::string g_last_error;

namespace acc {
int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}
void clear_error() { g_last_error.clear(); }
int bits_for(int32_t n) {
    int b = 1;
    while (b < 31 && (int64_t(1) << b) < n) ++b;
    return b;
}
}  // namespace acc

using acc::fail;

extern "C" const char *acc_last_error(void) { return g_last_error.c_str(); }
extern "C" const char *acc_version(void) { return "acceles-mi355x 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// Matrix Market parsing
// ---------------------------------------------------------------------------

namespace {

bool read_file(const char *path, std::vector<char> &buf) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize((size_t)(n > 0 ? n : 0) + 1);
    size_t got = n > 0 ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    buf.resize(got);
    buf.push_back('\0');
    return true;
}

// position just past the end of the line starting at p
size_t next_line(const std::vector<char> &b, size_t p) {
    size_t n = b.size() - 1;
    while (p < n && b[p] != '\n') ++p;
    return p < n ? p + 1 : n;
}

bool is_blank_line(const char *s, const char *e) {
    for (; s < e; ++s)
        if (*s != ' ' && *s != '\t' && *s != '\r' && *s != '\n') return false;
    return true;
}

// Parses entry lines in [begin, end) into the output arrays in parallel.
// value_mode: 0 = strtof (istringstream >> float), 1 = strtod then cast
// (fscanf "%lf" then (T)), 2 = pattern (1.0).
int parse_entries(const std::vector<char> &b, size_t begin, int64_t nnz, int index_offset,
                  int value_mode, acc_coo *out) {
    const size_t end = b.size() - 1;
    int nth = omp_get_max_threads();
    size_t span = end > begin ? end - begin : 0;
    int nchunks = (int)std::max<size_t>(1, std::min<size_t>((size_t)nth * 4, span / (1 << 16) + 1));
    std::vector<size_t> cb(nchunks + 1);
    cb[0] = begin;
    for (int c = 1; c < nchunks; ++c) {
        size_t p = begin + span * (size_t)c / (size_t)nchunks;
        if (p > begin && b[p - 1] != '\n') p = next_line(b, p);
        cb[c] = std::max(p, cb[c - 1]);
    }
    cb[nchunks] = end;
    // count non-blank lines per chunk
    std::vector<int64_t> cnt(nchunks + 1, 0);
#pragma omp parallel for schedule(dynamic)
    for (int c = 0; c < nchunks; ++c) {
        int64_t n = 0;
        for (size_t p = cb[c]; p < cb[c + 1];) {
            size_t q = next_line(b, p);
            if (q > cb[c + 1]) q = cb[c + 1];
            if (!is_blank_line(&b[p], &b[q])) ++n;
            p = q;
        }
        cnt[c + 1] = n;
    }
    for (int c = 0; c < nchunks; ++c) cnt[c + 1] += cnt[c];
    if (cnt[nchunks] < nnz) return fail(ACC_ERR_FORMAT, "Matrix Market file has fewer entry lines than its size line states");
    out->row.resize((size_t)nnz);
    out->col.resize((size_t)nnz);
    out->val.resize((size_t)nnz);
    int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
    for (int c = 0; c < nchunks; ++c) {
        int64_t i = cnt[c];
        for (size_t p = cb[c]; p < cb[c + 1] && i < nnz;) {
            size_t q = next_line(b, p);
            if (q > cb[c + 1]) q = cb[c + 1];
            if (is_blank_line(&b[p], &b[q])) { p = q; continue; }
            const char *s = &b[p];
            char *e;
            long r = std::strtol(s, &e, 10);
            if (e == s) { bad = 1; break; }
            s = e;
            long cc = std::strtol(s, &e, 10);
            if (e == s) { bad = 1; break; }
            s = e;
            float v = 1.0f;
            if (value_mode == 0) {
                v = std::strtof(s, &e);
                if (e == s) { bad = 1; break; }
            } else if (value_mode == 1) {
                double d = std::strtod(s, &e);
                if (e == s) { bad = 1; break; }
                v = (float)d;
            }
            out->row[(size_t)i] = (int32_t)(r - index_offset);
            out->col[(size_t)i] = (int32_t)(cc - index_offset);
            out->val[(size_t)i] = v;
            ++i;
            p = q;
        }
    }
    if (bad) return fail(ACC_ERR_FORMAT, "malformed Matrix Market entry line");
    return ACC_OK;
}

int check_range(const acc_coo *c) {
    int64_t n = (int64_t)c->row.size();
    int bad = 0;
#pragma omp parallel for reduction(| : bad)
    for (int64_t i = 0; i < n; ++i)
        if (c->row[i] < 0 || c->row[i] >= c->rows || c->col[i] < 0 || c->col[i] >= c->cols) bad = 1;
    return bad ? fail(ACC_ERR_RANGE, "entry index outside the matrix dimensions (check the index base)") : ACC_OK;
}

void stable_sort_rc(acc_coo *c) {
    size_t n = c->row.size();
    std::vector<int64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        if (c->row[a] != c->row[b]) return c->row[a] < c->row[b];
        return c->col[a] < c->col[b];
    });
    std::vector<int32_t> r(n), cc(n);
    std::vector<float> v(n);
    for (size_t i = 0; i < n; ++i) {
        r[i] = c->row[order[i]];
        cc[i] = c->col[order[i]];
        v[i] = c->val[order[i]];
    }
    c->row.swap(r);
    c->col.swap(cc);
    c->val.swap(v);
}

}  // namespace

extern "C" int acc_read_coo_matrix(const char *path, int zero_index, acc_coo **out) {
    if (!path || !out) return fail(ACC_ERR_ARG, "null argument");
    std::vector<char> b;
    if (!read_file(path, b)) return fail(ACC_ERR_IO, std::string("Failed to open file: ") + path);
    // cpu_mkl_baseline.cpp:55-58: two header lines, then the size line
    size_t p = next_line(b, 0);
    p = next_line(b, p);
    long rows = 0, cols = 0, nnz = 0;
    if (std::sscanf(&b[p], "%ld %ld %ld", &rows, &cols, &nnz) != 3 || rows < 0 || cols < 0 || nnz < 0)
        return fail(ACC_ERR_FORMAT, "missing or malformed size line (third line)");
    if (rows > INT32_MAX || cols > INT32_MAX) return fail(ACC_ERR_RANGE, "dimensions exceed int32");
    p = next_line(b, p);
    acc_coo *c = new acc_coo();
    c->rows = (int32_t)rows;
    c->cols = (int32_t)cols;
    int rc = parse_entries(b, p, nnz, zero_index ? 0 : 1, 0, c);
    if (rc == ACC_OK) rc = check_range(c);
    if (rc != ACC_OK) { delete c; return rc; }
    *out = c;
    return ACC_OK;
}

extern "C" int acc_read_mtx(const char *path, int zero_indexed_file, int sort_tuples, acc_coo **out) {
    if (!path || !out) return fail(ACC_ERR_ARG, "null argument");
    std::vector<char> b;
    if (!read_file(path, b)) return fail(ACC_ERR_IO, std::string("File ") + path + " not found");
    // mm_read_banner (mmio.hpp:109): "%%MatrixMarket matrix coordinate <field> <symmetry>"
    char t0[64] = {0}, t1[64] = {0}, t2[64] = {0}, t3[64] = {0}, t4[64] = {0};
    if (std::sscanf(&b[0], "%63s %63s %63s %63s %63s", t0, t1, t2, t3, t4) != 5 ||
        std::strcmp(t0, "%%MatrixMarket") != 0)
        return fail(ACC_ERR_FORMAT, "Could not process Matrix Market banner.");
    for (char *s : {t1, t2, t3, t4})
        for (; *s; ++s) *s = (char)std::tolower(*s);
    if (std::strcmp(t1, "matrix") != 0 || std::strcmp(t2, "coordinate") != 0)
        return fail(ACC_ERR_FORMAT, "only 'matrix coordinate' Matrix Market files are supported");
    int value_mode;
    if (!std::strcmp(t3, "real") || !std::strcmp(t3, "integer") || !std::strcmp(t3, "double"))
        value_mode = 1;
    else if (!std::strcmp(t3, "pattern"))
        value_mode = 2;
    else
        return fail(ACC_ERR_FORMAT, std::string("unsupported Matrix Market field: ") + t3);
    // mm_read_mtx_crd_size (mmio.hpp:184): skip '%' comment lines
    size_t p = next_line(b, 0);
    while (p < b.size() - 1 && (b[p] == '%' || is_blank_line(&b[p], &b[next_line(b, p)]))) p = next_line(b, p);
    long rows = 0, cols = 0, nnz = 0;
    if (std::sscanf(&b[p], "%ld %ld %ld", &rows, &cols, &nnz) != 3 || rows < 0 || cols < 0 || nnz < 0)
        return fail(ACC_ERR_FORMAT, "missing or malformed size line");
    if (rows > INT32_MAX || cols > INT32_MAX) return fail(ACC_ERR_RANGE, "dimensions exceed int32");
    p = next_line(b, p);
    acc_coo *c = new acc_coo();
    c->rows = (int32_t)rows;
    c->cols = (int32_t)cols;
    int rc = parse_entries(b, p, nnz, zero_indexed_file ? 0 : 1, value_mode, c);
    if (rc == ACC_OK) rc = check_range(c);
    if (rc != ACC_OK) { delete c; return rc; }
    if (sort_tuples) stable_sort_rc(c);
    *out = c;
    return ACC_OK;
}

extern "C" int acc_write_mtx(const char *path, const acc_coo *c) {
    if (!path || !c) return fail(ACC_ERR_ARG, "null argument");
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(ACC_ERR_IO, std::string("cannot open for writing: ") + path);
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%%\n%d %d %zu\n", c->rows, c->cols,
                 c->row.size());
    for (size_t i = 0; i < c->row.size(); ++i)
        std::fprintf(f, "%d %d %.9g\n", c->row[i] + 1, c->col[i] + 1, (double)c->val[i]);
    int err = std::ferror(f);
    std::fclose(f);
    return err ? fail(ACC_ERR_IO, "write failed") : ACC_OK;
}

// ---------------------------------------------------------------------------
// COO / CSR handles
// ---------------------------------------------------------------------------

extern "C" int acc_coo_create(int32_t rows, int32_t cols, int64_t nnz, const int32_t *row,
                              const int32_t *col, const float *val, acc_coo **out) {
    if (!out || rows < 0 || cols < 0 || nnz < 0 || (nnz > 0 && (!row || !col || !val)))
        return fail(ACC_ERR_ARG, "bad COO arguments");
    acc_coo *c = new acc_coo();
    c->rows = rows;
    c->cols = cols;
    c->row.assign(row, row + nnz);
    c->col.assign(col, col + nnz);
    c->val.assign(val, val + nnz);
    int rc = check_range(c);
    if (rc != ACC_OK) { delete c; return rc; }
    *out = c;
    return ACC_OK;
}

extern "C" int acc_coo_info(const acc_coo *c, int32_t *rows, int32_t *cols, int64_t *nnz) {
    if (!c) return fail(ACC_ERR_ARG, "null handle");
    if (rows) *rows = c->rows;
    if (cols) *cols = c->cols;
    if (nnz) *nnz = (int64_t)c->row.size();
    return ACC_OK;
}

extern "C" int acc_coo_export(const acc_coo *c, int32_t *row, int32_t *col, float *val) {
    if (!c) return fail(ACC_ERR_ARG, "null handle");
    if (row) std::copy(c->row.begin(), c->row.end(), row);
    if (col) std::copy(c->col.begin(), c->col.end(), col);
    if (val) std::copy(c->val.begin(), c->val.end(), val);
    return ACC_OK;
}

extern "C" void acc_coo_destroy(acc_coo *c) { delete c; }

extern "C" int acc_coo2csr(int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *row,
                           const int32_t *col, const float *val, int sort_tuples, int64_t *ptr,
                           int32_t *idx, float *val_out) {
    if (nrows < 0 || ncols < 0 || nnz < 0 || !ptr || (nnz > 0 && (!row || !col || !val || !idx || !val_out)))
        return fail(ACC_ERR_ARG, "bad coo2csr arguments");
    int bad = 0;
#pragma omp parallel for reduction(| : bad)
    for (int64_t i = 0; i < nnz; ++i)
        if (row[i] < 0 || row[i] >= nrows || col[i] < 0 || col[i] >= ncols) bad = 1;
    if (bad) return fail(ACC_ERR_RANGE, "Error: Index out of bounds!");
    std::vector<int64_t> order;
    if (sort_tuples) {
        order.resize((size_t)nnz);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
            if (row[a] != row[b]) return row[a] < row[b];
            return col[a] < col[b];
        });
    }
    std::fill(ptr, ptr + nrows + 1, 0);
    for (int64_t i = 0; i < nnz; ++i) ptr[row[i] + 1]++;
    for (int32_t r = 0; r < nrows; ++r) ptr[r + 1] += ptr[r];
    std::vector<int64_t> fill(ptr, ptr + nrows);
    for (int64_t t = 0; t < nnz; ++t) {
        int64_t i = sort_tuples ? order[(size_t)t] : t;
        int64_t d = fill[row[i]]++;
        idx[d] = col[i];
        val_out[d] = val[i];
    }
    return ACC_OK;
}

extern "C" int acc_coo_to_csr(const acc_coo *c, acc_csr **out) {
    if (!c || !out) return fail(ACC_ERR_ARG, "null argument");
    acc_csr *m = new acc_csr();
    m->rows = c->rows;
    m->cols = c->cols;
    int64_t nnz = (int64_t)c->row.size();
    m->ptr.resize((size_t)c->rows + 1);
    m->idx.resize((size_t)nnz);
    m->val.resize((size_t)nnz);
    int rc = acc_coo2csr(c->rows, c->cols, nnz, c->row.data(), c->col.data(), c->val.data(), 0, m->ptr.data(),
                         m->idx.data(), m->val.data());
    if (rc != ACC_OK) { delete m; return rc; }
    *out = m;
    return ACC_OK;
}

extern "C" int acc_csr_create(int32_t rows, int32_t cols, int64_t nnz, const int64_t *ptr, const int32_t *idx,
                              const float *val, acc_csr **out) {
    if (!out || rows < 0 || cols < 0 || nnz < 0 || !ptr || (nnz > 0 && (!idx || !val)))
        return fail(ACC_ERR_ARG, "bad CSR arguments");
    if (ptr[0] != 0 || ptr[rows] != nnz) return fail(ACC_ERR_FORMAT, "CSR ptr must start at 0 and end at nnz");
    for (int32_t r = 0; r < rows; ++r)
        if (ptr[r + 1] < ptr[r]) return fail(ACC_ERR_FORMAT, "CSR ptr must be non-decreasing");
    int bad = 0;
#pragma omp parallel for reduction(| : bad)
    for (int64_t i = 0; i < nnz; ++i)
        if (idx[i] < 0 || idx[i] >= cols) bad = 1;
    if (bad) return fail(ACC_ERR_RANGE, "CSR column index out of range");
    acc_csr *m = new acc_csr();
    m->rows = rows;
    m->cols = cols;
    m->ptr.assign(ptr, ptr + rows + 1);
    m->idx.assign(idx, idx + nnz);
    m->val.assign(val, val + nnz);
    *out = m;
    return ACC_OK;
}

extern "C" int acc_csr_info(const acc_csr *m, int32_t *rows, int32_t *cols, int64_t *nnz) {
    if (!m) return fail(ACC_ERR_ARG, "null handle");
    if (rows) *rows = m->rows;
    if (cols) *cols = m->cols;
    if (nnz) *nnz = m->nnz();
    return ACC_OK;
}

extern "C" int acc_csr_export(const acc_csr *m, int64_t *ptr, int32_t *idx, float *val) {
    if (!m) return fail(ACC_ERR_ARG, "null handle");
    if (ptr) std::copy(m->ptr.begin(), m->ptr.end(), ptr);
    if (idx) std::copy(m->idx.begin(), m->idx.end(), idx);
    if (val) std::copy(m->val.begin(), m->val.end(), val);
    return ACC_OK;
}

extern "C" void acc_csr_destroy(acc_csr *m) { delete m; }

// ---------------------------------------------------------------------------
// quantization (quantify_util.py:8-49)
// ---------------------------------------------------------------------------

extern "C" int acc_quantize(int64_t n, const float *v, int bits, int symmetric, float *scale_out, int32_t *q) {
    if (n < 0 || bits < 1 || bits > 16 || (n > 0 && (!v || !q))) return fail(ACC_ERR_ARG, "bad quantize arguments");
    int32_t M = (int32_t)((1u << bits) - 1u);
    if (symmetric) M /= 2;
    float mx = n > 0 ? v[0] : 0.f, mn = n > 0 ? v[0] : 0.f;
#pragma omp parallel for reduction(max : mx) reduction(min : mn)
    for (int64_t i = 0; i < n; ++i) {
        mx = std::max(mx, v[i]);
        mn = std::min(mn, v[i]);
    }
    float s = std::max(std::fabs(mx), std::fabs(mn));
    if (scale_out) *scale_out = s;
    const float fM = (float)M;
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
        if (s == 0.f) { q[i] = 0; continue; }
        float t = v[i] / s;
        float u = fM * t;
        int32_t qi = (int32_t)std::nearbyint(u);
        q[i] = std::min(M, std::max(-M - 1, qi));
    }
    return ACC_OK;
}

// ---------------------------------------------------------------------------
// seeded generators (create_matrices.py:57-137 semantics, own RNG)
// ---------------------------------------------------------------------------

namespace {

inline uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** seeded per stream (seed, stream id) so results do not depend
// on the thread count.
struct Rng {
    uint64_t s[4];
    bool have_spare = false;
    double spare = 0;
    Rng(uint64_t seed, uint64_t stream) {
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull);
        for (auto &w : s) w = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    uint64_t below(uint64_t n) { return (uint64_t)(uniform() * (double)n); }
    double normal() {  // Box-Muller
        if (have_spare) { have_spare = false; return spare; }
        double u1 = 1.0 - uniform(), u2 = uniform();
        double r = std::sqrt(-2.0 * std::log(u1)), th = 6.283185307179586 * u2;
        spare = r * std::sin(th);
        have_spare = true;
        return r * std::cos(th);
    }
    double gamma(double shape, double scale) {  // Marsaglia-Tsang
        if (shape < 1.0) {
            double g = gamma(shape + 1.0, 1.0);
            return scale * g * std::pow(1.0 - uniform(), 1.0 / shape);
        }
        double d = shape - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x = normal(), v = 1.0 + c * x;
            if (v <= 0) continue;
            v = v * v * v;
            double u = 1.0 - uniform();
            if (u < 1.0 - 0.0331 * x * x * x * x) return scale * d * v;
            if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return scale * d * v;
        }
    }
    float value(int kind) {
        switch (kind) {
            case ACC_VAL_ABSNORMAL: return (float)std::fabs(normal());
            case ACC_VAL_NORMAL: return (float)normal();
            default: return (float)(uniform() * 2.0 - 1.0);
        }
    }
};

int32_t row_degree(Rng &g, int deg_kind, double p0, double p1, int32_t cols) {
    int64_t d;
    if (deg_kind == ACC_DEG_UNIFORM) {
        int64_t lo = (int64_t)p0, hi = (int64_t)p1;
        d = lo + (int64_t)g.below((uint64_t)(hi - lo + 1));
    } else {
        d = (int64_t)std::llround(g.gamma(p0, p1));
        d = std::max<int64_t>(1, d);
    }
    return (int32_t)std::min<int64_t>(std::max<int64_t>(d, 0), cols);
}

}  // namespace

extern "C" int acc_generate_csr(int32_t rows, int32_t cols, int deg_kind, double p0, double p1, int val_kind,
                                int l2norm, uint64_t seed, acc_csr **out) {
    if (!out || rows < 0 || cols < 1) return fail(ACC_ERR_ARG, "bad generator arguments");
    if (deg_kind == ACC_DEG_UNIFORM && (p0 < 0 || p1 < p0)) return fail(ACC_ERR_ARG, "bad degree range");
    if (deg_kind == ACC_DEG_GAMMA && (p0 <= 0 || p1 <= 0)) return fail(ACC_ERR_ARG, "bad gamma parameters");
    acc_csr *m = new acc_csr();
    m->rows = rows;
    m->cols = cols;
    m->ptr.assign((size_t)rows + 1, 0);
#pragma omp parallel for schedule(static, 4096)
    for (int32_t r = 0; r < rows; ++r) {
        Rng g(seed, (uint64_t)r);
        m->ptr[(size_t)r + 1] = row_degree(g, deg_kind, p0, p1, cols);
    }
    for (int32_t r = 0; r < rows; ++r) m->ptr[(size_t)r + 1] += m->ptr[(size_t)r];
    int64_t nnz = m->ptr[(size_t)rows];
    m->idx.resize((size_t)nnz);
    m->val.resize((size_t)nnz);
#pragma omp parallel
    {
        std::vector<uint8_t> used((size_t)cols, 0);
#pragma omp for schedule(static, 4096)
        for (int32_t r = 0; r < rows; ++r) {
            Rng g(seed, (uint64_t)r);
            int32_t d = row_degree(g, deg_kind, p0, p1, cols);
            int32_t *ci = m->idx.data() + m->ptr[(size_t)r];
            // Floyd's sampling of d distinct columns
            int n = 0;
            for (int32_t j = cols - d; j < cols; ++j) {
                int32_t t = (int32_t)g.below((uint64_t)j + 1);
                if (used[(size_t)t]) t = j;
                used[(size_t)t] = 1;
                ci[n++] = t;
            }
            std::sort(ci, ci + d);
            for (int k = 0; k < d; ++k) used[(size_t)ci[k]] = 0;
            float *vi = m->val.data() + m->ptr[(size_t)r];
            double ss = 0;
            for (int k = 0; k < d; ++k) {
                vi[k] = g.value(val_kind);
                ss += (double)vi[k] * (double)vi[k];
            }
            if (l2norm && ss > 0) {
                double inv = 1.0 / std::sqrt(ss);
                for (int k = 0; k < d; ++k) vi[k] = (float)((double)vi[k] * inv);
            }
        }
    }
    *out = m;
    return ACC_OK;
}

extern "C" int acc_generate_queries(int32_t nq, int32_t cols, int val_kind, int l2norm, uint64_t seed, float *out) {
    if (!out || nq < 0 || cols < 1) return fail(ACC_ERR_ARG, "bad query generator arguments");
#pragma omp parallel for
    for (int32_t q = 0; q < nq; ++q) {
        Rng g(seed ^ 0x51ED270B27B5D1F3ull, (uint64_t)q);
        float *x = out + (size_t)q * (size_t)cols;
        double ss = 0;
        for (int32_t c = 0; c < cols; ++c) {
            x[c] = g.value(val_kind);
            ss += (double)x[c] * (double)x[c];
        }
        if (l2norm && ss > 0) {
            double inv = 1.0 / std::sqrt(ss);
            for (int32_t c = 0; c < cols; ++c) x[c] = (float)((double)x[c] * inv);
        }
    }
    return ACC_OK;
}

// ---------------------------------------------------------------------------
// Random-CSR: seeded row permutation
// ---------------------------------------------------------------------------

namespace acc {
std::vector<int32_t> row_shuffle(int32_t rows, uint64_t seed) {
    std::vector<int32_t> p((size_t)rows);
    std::iota(p.begin(), p.end(), 0);
    Rng g(seed, 0xFFFFFFFFull);
    for (int32_t i = rows - 1; i > 0; --i) std::swap(p[(size_t)i], p[(size_t)g.below((uint64_t)i + 1)]);
    return p;
}
}  // namespace acc

extern "C" int acc_permute_rows(const acc_csr *m, uint64_t seed, acc_csr **out, int32_t *perm) {
    if (!m || !out) return fail(ACC_ERR_ARG, "null argument");
    const std::vector<int32_t> p = acc::row_shuffle(m->rows, seed);
    acc_csr *o = new acc_csr();
    o->rows = m->rows;
    o->cols = m->cols;
    o->ptr.assign((size_t)m->rows + 1, 0);
    for (int32_t i = 0; i < m->rows; ++i) o->ptr[(size_t)i + 1] = o->ptr[(size_t)i] + (m->ptr[(size_t)p[i] + 1] - m->ptr[(size_t)p[i]]);
    o->idx.resize(m->idx.size());
    o->val.resize(m->val.size());
#pragma omp parallel for schedule(static, 4096)
    for (int32_t i = 0; i < m->rows; ++i) {
        int64_t s = m->ptr[(size_t)p[i]], e = m->ptr[(size_t)p[i] + 1], d = o->ptr[(size_t)i];
        std::copy(m->idx.begin() + s, m->idx.begin() + e, o->idx.begin() + d);
        std::copy(m->val.begin() + s, m->val.begin() + e, o->val.begin() + d);
    }
    if (perm) std::copy(p.begin(), p.end(), perm);
    *out = o;
    return ACC_OK;
}

// ---------------------------------------------------------------------------
// ReSparse: per-row pruning of small values (csr_util.py:152-187 over the
// nonzeros; quantify_util.py:375-403 over all rows x cols cells)
// ---------------------------------------------------------------------------

extern "C" int acc_resparse_csr(const acc_csr *m, double keep, int mean_over_cells, acc_csr **out) {
    if (!m || !out) return fail(ACC_ERR_ARG, "null argument");
    if (!(keep >= 0.0 && keep <= 1.0)) return fail(ACC_ERR_ARG, "keep must be in [0, 1]");
    // the mean importance |v| (FP64 sum)
    double sum = 0.0;
    for (float v : m->val) sum += std::fabs((double)v);
    const double denom = mean_over_cells ? (double)m->rows * (double)m->cols : (double)m->val.size();
    const double avg = denom > 0 ? sum / denom : 0.0;
    std::vector<int32_t> kept((size_t)m->rows, 0);
    std::vector<uint8_t> mask(m->val.size(), 0);
#pragma omp parallel
    {
        std::vector<int64_t> order;
#pragma omp for schedule(dynamic, 1024)
        for (int32_t r = 0; r < m->rows; ++r) {
            const int64_t b = m->ptr[(size_t)r], e = m->ptr[(size_t)r + 1], n = e - b;
            if (n == 0) continue;
            double mn = INFINITY;
            for (int64_t j = b; j < e; ++j) mn = std::min(mn, std::fabs((double)m->val[(size_t)j]));
            if (mn > avg) {  // every entry matters more than the mean: keep the row
                std::fill(mask.begin() + b, mask.begin() + e, (uint8_t)1);
                kept[(size_t)r] = (int32_t)n;
                continue;
            }
            // keep ceil(n * keep) entries of largest |v| (equal |v|: the earlier entry)
            const int64_t nk = (int64_t)std::ceil((double)n * keep);
            order.resize((size_t)n);
            std::iota(order.begin(), order.end(), b);
            std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
                return std::fabs(m->val[(size_t)x]) > std::fabs(m->val[(size_t)y]);
            });
            for (int64_t t = 0; t < nk; ++t) mask[(size_t)order[(size_t)t]] = 1;
            kept[(size_t)r] = (int32_t)nk;
        }
    }
    acc_csr *o = new acc_csr();
    o->rows = m->rows;
    o->cols = m->cols;
    o->ptr.assign((size_t)m->rows + 1, 0);
    for (int32_t r = 0; r < m->rows; ++r) o->ptr[(size_t)r + 1] = o->ptr[(size_t)r] + kept[(size_t)r];
    o->idx.resize((size_t)o->ptr[(size_t)m->rows]);
    o->val.resize(o->idx.size());
#pragma omp parallel for schedule(static, 4096)
    for (int32_t r = 0; r < m->rows; ++r) {
        int64_t d = o->ptr[(size_t)r];
        for (int64_t j = m->ptr[(size_t)r]; j < m->ptr[(size_t)r + 1]; ++j)
            if (mask[(size_t)j]) {  // kept entries stay in their column order
                o->idx[(size_t)d] = m->idx[(size_t)j];
                o->val[(size_t)d] = m->val[(size_t)j];
                ++d;
            }
    }
    *out = o;
    return ACC_OK;
}

// ---------------------------------------------------------------------------
// Ranking-quality metrics (evaluation_utils.hpp:112-267, quantify_util.py:313-329)
// ---------------------------------------------------------------------------

extern "C" int acc_ndcg_rank(const int32_t *golden, const int32_t *test, int n, double *out) {
    if ((n > 0 && (!golden || !test)) || !out || n < 0) return fail(ACC_ERR_ARG, "bad arguments");
    // relevance of a row: n - (its place in the list); for a row that appears
    // twice in test, the later place counts
    std::unordered_map<int32_t, int> rel;
    for (int i = 0; i < n; ++i) rel[test[i]] = n - i;
    double dcg = 0.0, idcg = 0.0;
    for (int i = 0; i < n; ++i) {
        const int g = n - i;
        const auto it = rel.find(golden[i]);
        const int t = it == rel.end() ? 0 : it->second;
        // the argument is a float (place + 2) and the logarithm a double one,
        // as in the reference's expression
        const double disc = std::log2((double)((float)std::abs(g - n) + 2));
        dcg += (double)t / disc;
        idcg += (double)g / disc;
    }
    *out = dcg / idcg;
    return ACC_OK;
}

extern "C" int acc_ndcg_score(const int32_t *sw_idx, const double *sw_val, int nsw, const int32_t *hw_idx, int nhw,
                              double *ndcg, double *dcg_out, double *idcg_out) {
    if (nsw < 0 || nhw < 0 || (nsw > 0 && (!sw_idx || !sw_val)) || (nhw > 0 && !hw_idx))
        return fail(ACC_ERR_ARG, "bad arguments");
    std::unordered_map<int32_t, double> rel;  // a row listed twice: the later value counts
    for (int i = 0; i < nsw; ++i) rel[sw_idx[i]] = sw_val[i];
    double dcg = 0.0, idcg = 0.0;
    for (int i = 0; i < nhw; ++i) {
        const auto it = rel.find(hw_idx[i]);
        dcg += (it == rel.end() ? 0.0 : it->second) / std::log2((double)(i + 2));
    }
    for (int i = 0; i < nsw; ++i) idcg += sw_val[i] / std::log2((double)(i + 2));
    if (ndcg) *ndcg = dcg / idcg;
    if (dcg_out) *dcg_out = dcg;
    if (idcg_out) *idcg_out = idcg;
    return ACC_OK;
}

extern "C" int acc_edit_distance(const int32_t *a, int na, const int32_t *b, int nb, uint32_t *out) {
    if (na < 0 || nb < 0 || (na > 0 && !a) || (nb > 0 && !b) || !out) return fail(ACC_ERR_ARG, "bad arguments");
    std::vector<uint32_t> prev((size_t)nb + 1), cur((size_t)nb + 1);  // one row of the Levenshtein table
    for (int j = 0; j <= nb; ++j) prev[(size_t)j] = (uint32_t)j;
    for (int i = 1; i <= na; ++i) {
        cur[0] = (uint32_t)i;
        for (int j = 1; j <= nb; ++j)
            cur[(size_t)j] = std::min({prev[(size_t)j] + 1, cur[(size_t)j - 1] + 1,
                                       prev[(size_t)j - 1] + (a[i - 1] == b[j - 1] ? 0u : 1u)});
        std::swap(prev, cur);
    }
    *out = prev[(size_t)nb];
    return ACC_OK;
}

extern "C" int acc_count_errors(const int32_t *golden, const int32_t *test, int n, int32_t *out) {
    if (n < 0 || (n > 0 && (!golden || !test)) || !out) return fail(ACC_ERR_ARG, "bad arguments");
    int32_t e = 0;
    for (int i = 0; i < n; ++i) e += golden[i] != test[i];
    *out = e;
    return ACC_OK;
}
