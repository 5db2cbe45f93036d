// FP64 trailing updates of the Cholesky factorisation on INT8 matrix cores
// (the error-free splitting of Ozaki, Ogita, Oishi and Rump, "Error-free
// transformations of matrix multiplication by using fast routines of matrix
// multiplication and its applications", Numer. Algorithms 59 (2012), with
// integer slices as in Mukunoki, Ozaki, Ogita and Imamura, "DGEMM using
// Tensor Cores, and Its Accurate and Reproducible Versions", ISC 2020).
#pragma once

#include "common.h"

namespace fmlh {

// outer panel width of the Ozaki factorisation, the k the slices are made
// for (36k: 1024 -> 145 ms a factorisation, 1536 -> 156 ms, 2048 -> 168 ms
// and one more refinement step)
constexpr int kOzPanel = 1024;
constexpr int kOzSlices = 5;    // slices per entry (the one instantiated form)

// Bytes of the slice planes and row exponents for rows [row0, n) of an
// n-row matrix with `slices` planes of kOzPanel bytes a row.
size_t oz_ws_bytes(int n, int slices);

// Rows [row0, row0 + nrows), panel columns [kb, kb + kw) of the row-major
// view M[r][c] = a[r lda + c]: each row r gets a power-of-two scale 2^e_r
// with |M[r][k]| 2^-e_r <= 0.498, the scaled entries are truncated to
// integers of 8 * slices bits, X = trunc(M 2^(8 slices - e_r)), and X is
// written in signed base-256 digits,
//   M[r][k] = 2^e_r (sum_t q_t[r][k] 2^-8(t+1) + rest),  q_t in [-128, 127],
//   |rest| < 2^-8 slices,
// stored in `slices` planes of kOzPanel bytes a row (zero beyond kw),
// k-step major: byte (r, k) of a plane at (k / 32) * 32 nrows + 32 r + k % 32.
int oz_split(const double *a, int64_t lda, int row0, int nrows, int kb, int kw, int slices,
             void *ws, const int32_t *info, hipStream_t st);

// C[r][c] -= sum_k M[r][k] M[c][k] over the panel split by oz_split (rows
// split from srow0, nsplit of them), for r in [r0, r1), c in [c0, c1),
// c <= r: the digit products q_t q_u with t + u <= slices - 1 on
// v_mfma_i32_32x32x32_i8 (exact int32 sums per level t + u), combined in
// FP64.  tri: the region is the square [r0, r1)^2 and only its
// lower-triangle tiles are launched.
int oz_update(double *a, int64_t lda, int r0, int r1, int c0, int c1, int srow0, int nsplit,
              int slices, const void *ws, bool tri, const int32_t *info, hipStream_t st);

}  // namespace fmlh
