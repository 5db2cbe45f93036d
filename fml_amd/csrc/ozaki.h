// FP64 trailing updates of the Cholesky factorisation on INT8 matrix cores
// (the error-free splitting of Ozaki, Ogita, Oishi and Rump, "Error-free
// transformations of matrix multiplication by using fast routines of matrix
// multiplication and its applications", Numer. Algorithms 59 (2012), with
// integer slices as in Mukunoki, Ozaki, Ogita and Imamura, "DGEMM using
// Tensor Cores, and Its Accurate and Reproducible Versions", ISC 2020).
#pragma once

#include "common.h"

namespace fmlh {

constexpr int kOzPanel = 1024;  // panel width the slices are made for (k <= 1024)
constexpr int kOzSlices = 6;    // slices per entry (the one instantiated form)

// Bytes of the slice planes and row exponents for rows [row0, n) of an
// n-row matrix with `slices` planes of kOzPanel bytes a row.
size_t oz_ws_bytes(int n, int slices);

// Rows [row0, row0 + nrows), panel columns [kb, kb + kw) of the row-major
// view M[r][c] = a[r lda + c]: each row r is scaled by 2^-e_r (e_r the
// exponent of its largest |entry|, so |M[r][k]| 2^-e_r < 1) and split
// exactly into `slices` signed 7-bit integers,
//   M[r][k] = 2^e_r (sum_t q_t[r][k] 2^-7t + rest), |q_t| <= 127,
// stored in planes of kOzPanel bytes a row (zero beyond kw).
int oz_split(const double *a, int64_t lda, int row0, int nrows, int kb, int kw, int slices,
             void *ws, const int32_t *info, hipStream_t st);

// C[r][c] -= sum_k M[r][k] M[c][k] over the panel split by oz_split (rows
// split from srow0), for r in [r0, r1), c in [c0, c1), c <= r: the products
// of slices t, u with t + u <= slices + 1 on v_mfma_i32_16x16x64_i8
// (exact int32 sums), combined per level in FP64.  tri: the region is the
// square [r0, r1)^2 (lower-triangle tiles only).
int oz_update(double *a, int64_t lda, int r0, int r1, int c0, int c1, int srow0, int nsplit,
              int slices, const void *ws, bool tri, const int32_t *info, hipStream_t st);

}  // namespace fmlh
