// ref_gold.cpp -- C entry points onto the reference's own header-only gold
// code, compiled from the headers where they lie under the reference tree
// (oracle/Makefile, target `ref`; output oracle/_ref/libref_gold.so).
//
// TEST INFRASTRUCTURE ONLY: used by tests/ and tests/golden/make_golden.py to
// pin the oracle; nothing in the product loads it.
//
// Wrapped reference functions:
//   readMtx                    src/common/utils/utils.hpp:449-492
//   coo2csr                    src/common/utils/utils.hpp:494-547
//   spmv_gold                  src/fpga/src/gold_algorithms/gold_algorithms.hpp:5-18
//   spmv_coo_gold_top_k        gold_algorithms.hpp:189-247
//   spmv_coo_gold_top_k_2      gold_algorithms.hpp:249-308
//   sort_tuples                src/common/utils/evaluation_utils.hpp:40-62
//   normalized_discounted_cumulative_gain, edit_distance, bounded_count_errors
//                              evaluation_utils.hpp:112-252
#include "types.hpp"
#include "csc_matrix/csc_matrix.hpp"
#include "utils/mmio.hpp"
#include "utils/utils.hpp"
#include "utils/evaluation_utils.hpp"
#include "gold_algorithms/gold_algorithms.hpp"

#include <cstdint>
#include <vector>

extern "C" {

// Returns nnz (>= 0) and fills dims[3] = {rows, cols, nnz}; fills the arrays
// only when cap >= nnz.
long long ref_readMtx(const char *path, int zero_indexed_file, int sort_tuples, int *dims,
                      int *row, int *col, float *val, long long cap) {
    std::vector<unsigned> r, c;
    std::vector<float> v;
    unsigned nr = 0, nc = 0, nz = 0;
    readMtx<unsigned, float>(path, &r, &c, &v, &nr, &nc, &nz, 1, true, false,
                             zero_indexed_file != 0, sort_tuples != 0);
    dims[0] = (int)nr;
    dims[1] = (int)nc;
    dims[2] = (int)nz;
    if ((long long)r.size() <= cap) {
        for (size_t i = 0; i < r.size(); ++i) {
            row[i] = (int)r[i];
            col[i] = (int)c[i];
            val[i] = v[i];
        }
    }
    return (long long)r.size();
}

void ref_coo2csr(unsigned nrows, unsigned ncols, unsigned nnz, const unsigned *row,
                 const unsigned *col, const float *val, int sort_tuples, unsigned *ptr,
                 unsigned *idx, float *vout) {
    std::vector<unsigned> r(row, row + nnz), c(col, col + nnz);
    std::vector<float> v(val, val + nnz);
    coo2csr<unsigned, float>(ptr, idx, vout, r, c, v, nrows, ncols, sort_tuples != 0);
}

void ref_spmv_gold_f32(unsigned n, unsigned *ptr, unsigned *idx, float *val, float *vec,
                       float *result) {
    spmv_gold<unsigned, float>(ptr, idx, val, n, result, vec);
}

void ref_spmv_coo_gold_top_k_f32(unsigned nnz, const unsigned *row, const unsigned *col,
                                 const float *val, float *vec, int k, unsigned *res_idx,
                                 float *res_val) {
    std::vector<unsigned> r(row, row + nnz), c(col, col + nnz);
    std::vector<float> v(val, val + nnz);
    coo_t<unsigned, float> coo(r, c, v);
    spmv_coo_gold_top_k<unsigned, float>(coo, vec, k, res_idx, res_val);
}

void ref_spmv_coo_gold_top_k_2_i32(unsigned nnz, const unsigned *row, const unsigned *col,
                                   const int *val, int *vec, int k, unsigned *res_idx,
                                   int *res_val) {
    std::vector<unsigned> r(row, row + nnz), c(col, col + nnz);
    std::vector<int> v(val, val + nnz);
    coo_t<unsigned, int> coo(r, c, v);
    spmv_coo_gold_top_k_2<unsigned, int, int>(coo, vec, k, res_idx, res_val);
}

void ref_sort_tuples_i32(unsigned n, unsigned *idx, int *val) {
    sort_tuples<unsigned, int>(n, idx, val);
}

void ref_sort_tuples_f32(unsigned n, unsigned *idx, float *val) {
    sort_tuples<unsigned, float>(n, idx, val);
}


double ref_ndcg(const int *golden, const int *test, int n) {
    return normalized_discounted_cumulative_gain(std::vector<int>(golden, golden + n), std::vector<int>(test, test + n));
}

unsigned ref_edit_distance(const int *a, int na, const int *b, int nb) {
    return edit_distance(std::vector<int>(a, a + na), std::vector<int>(b, b + nb));
}

int ref_count_errors(const int *golden, const int *test, int n) {
    std::vector<int> bounds{n};
    return bounded_count_errors(std::vector<int>(golden, golden + n), std::vector<int>(test, test + n), bounds)[0];
}

}  // extern "C"
