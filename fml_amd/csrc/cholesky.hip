// (e) Blocked Cholesky factorisation and solve of a symmetric positive
// definite matrix, LAPACK dpotrf('U') / dpotrs('U') semantics
// (fml/math/fcho_solve.f90:35,42).
//
// Storage: the reference's Fortran-order (n, n) array whose UPPER triangle
// holds A.  Read as row-major with the same leading dimension, that upper
// triangle is the LOWER triangle of M (M[r][c] = a[r*lda + c], r >= c), and
// A = U^T U with U upper becomes M = L L^T with L = U^T lower.  Everything
// below works on that row-major lower view, so the strictly upper part of M
// (the strictly lower part of the caller's array) is never touched.
//
// Factorisation (right-looking, two block levels):
//   for each 256-column outer panel:
//     for each 64-column inner block inside it:
//       k_potf2   diagonal 64x64 block in registers (one wave, lane = row)
//       k_trsm    rows below: M[r][k] <- M[r][k] L_kk^{-T} (one lane/row)
//       k_update  panel columns right of the block: C -= P Q^T, 64 tiles
//     k_update    trailing matrix right of the outer panel, K = 256, on
//                 128x128 MFMA tiles over its lower triangle only
// A non-positive (or NaN) pivot sets *info to its 1-based index, as LAPACK;
// every later kernel sees *info != 0 and returns, so the factor is left
// undefined but no kernel waits on anything.
//
// Solve: blocked forward (L w = b) and backward (L^T x = w) substitution
// over 64-row blocks: a multi-workgroup partial product of the already
// solved part followed by a one-workgroup triangular solve of the block.
#include <math.h>

#include "common.h"
#include "gemm_nt.h"

namespace {

using namespace fmlh;

constexpr int kNB = 64;     // inner block
constexpr int kNBO = 256;   // outer panel

// Value of v in lane l (l uniform) for every lane.
__device__ __forceinline__ double lane_value(double v, int l) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, l);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// One wave factors the diagonal block: lane r holds row r in registers.
// Step j takes the square root of the pivot (lane j's diagonal, read by
// every lane), scales column j below it, then every lane r updates its row
// M[r][c] -= L[r][j] L[c][j] for c > j, L[c][j] read from lane c (entries
// above the diagonal are updated too and never stored).  Fully unrolled,
// so the row stays in registers.
template <typename T>
__global__ __launch_bounds__(64) void k_potf2(T *__restrict__ a, int64_t lda, int k0, int b,
                                              int32_t *__restrict__ info) {
  if (*info != 0) return;
  const int r = threadIdx.x;
  T *grow = a + (int64_t)(k0 + r) * lda + k0;
  T row[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) row[c] = (r < b && c <= r) ? grow[c] : T(0);
#pragma unroll
  for (int j = 0; j < kNB; ++j) {
    if (j < b) {
      const T d = lane_value(row[j], j);
      if (!(d > T(0))) {  // uniform: every lane read the same pivot
        if (r == 0) *info = k0 + j + 1;
        return;
      }
      const T piv = sqrt(d);
      const T lrj = (r > j && r < b) ? row[j] / piv : T(0);
      row[j] = (r == j) ? piv : (r > j ? lrj : row[j]);
#pragma unroll
      for (int c = j + 1; c < kNB; ++c) row[c] -= lrj * lane_value(lrj, c);
    }
  }
  if (r < b) {
#pragma unroll
    for (int c = 0; c < kNB; ++c)
      if (c <= r) grow[c] = row[c];
  }
}

// Rows r in [r0, n): x = M[r][k0:k0+b] solves x L^T = m_r against the
// factored diagonal block (LDS, loaded by all 256 threads), one row per
// thread held in registers; column by column, x_c = m_c / L[c][c] then
// m_t -= x_c L[t][c] for t > c (independent updates, so the row's
// dependency chain is short).
template <typename T>
__global__ __launch_bounds__(256) void k_trsm(T *__restrict__ a, int64_t lda, int n, int k0, int b,
                                              int r0, const int32_t *__restrict__ info) {
  __shared__ T L[kNB][kNB + 1];
  if (*info != 0) return;
  for (int idx = threadIdx.x; idx < b * b; idx += 256) {
    const int r = idx / b, c = idx % b;
    L[r][c] = (c <= r) ? a[(int64_t)(k0 + r) * lda + k0 + c] : T(0);
  }
  __syncthreads();
  const int r = r0 + blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  T *row = a + (int64_t)r * lda + k0;
  T x[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) x[c] = (c < b) ? row[c] : T(0);
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    if (c < b) {
      x[c] = x[c] / L[c][c];
#pragma unroll
      for (int t = c + 1; t < kNB; ++t) x[t] -= x[c] * L[t][c];
    }
  }
#pragma unroll
  for (int c = 0; c < kNB; ++c)
    if (c < b) row[c] = x[c];
}

// Tile (tx, ty), tx <= ty < g, of the lower triangle of a g x g tile grid
// at position `local` of this order: groups of kGroup tile rows, each group
// walked column by column, so a group's row tiles stay in L2 while its
// column tiles stream past once (the row-major triangle would stream every
// column tile once per row).  Returns false past the end.
__device__ __forceinline__ bool tri_tile(int g, int64_t local, int &tx, int &ty) {
  if (local >= (int64_t)g * (g + 1) / 2) return false;
  // row of `local` in row-major order: T(r) <= local < T(r + 1), T(r) = r(r+1)/2
  int r = (int)((sqrt(8.0 * (double)local + 1.0) - 1.0) * 0.5);
  while ((int64_t)(r + 1) * (r + 2) / 2 <= local) ++r;
  while ((int64_t)r * (r + 1) / 2 > local) --r;
  const int r0 = (r / kGroup) * kGroup;  // first row of the group (it starts at T(r0))
  const int R = min(kGroup, g - r0);
  int64_t o = local - (int64_t)r0 * (r0 + 1) / 2;
  const int64_t full = (int64_t)(r0 + 1) * R;  // columns 0..r0 hold R tiles each
  if (o < full) {
    tx = (int)(o / R);
    ty = r0 + (int)(o % R);
    return true;
  }
  o -= full;
  for (int c = r0 + 1; c < r0 + R; ++c) {  // column c holds rows c..r0+R-1
    const int cnt = r0 + R - c;
    if (o < cnt) {
      tx = c;
      ty = c + (int)o;
      return true;
    }
    o -= cnt;
  }
  return false;
}

// C[r][c] -= sum_{t in [kb, ke)} M[r][t] M[c][t] for r in [r0, r1),
// c in [c0, c1), c <= r, on MT x MT tiles.  kTri: the region is the square
// [r0, r1)^2 and only its lower-triangle tiles are launched (tri_tile order,
// in XCD-contiguous runs); otherwise a rectangular gx x gy grid in which
// tiles above the diagonal exit.
template <typename T, int MT, bool kTri>
__global__ __launch_bounds__(256, (MT == 128 ? 2 : 1)) void k_update(T *__restrict__ a, int64_t lda, int r0, int r1,
                                                int c0, int c1, int kb, int ke, int gx, int gy,
                                                const int32_t *__restrict__ info) {
  __shared__ TileSmem<T, MT> s;
  if (*info != 0) return;
  int tx, ty;
  if (kTri) {
    const int64_t t = (int64_t)gy * (gy + 1) / 2;
    const int64_t per = (t + kXcds - 1) / kXcds;
    const int64_t L = blockIdx.x;
    if (!tri_tile(gy, (L % kXcds) * per + L / kXcds, tx, ty)) return;
  } else if (!xcd_tile(gx, gy, tx, ty)) {
    return;
  }
  const int tr = r0 + ty * MT;
  const int tc = c0 + tx * MT;
  if (tc > tr + MT - 1) return;  // tile entirely above the diagonal
  TileAcc<T, MT> acc;
  tile_nt<T, false, MT>(a + (int64_t)tr * lda + kb, lda, r1 - tr, a + (int64_t)tc * lda + kb, lda,
                        c1 - tc, ke - kb, s, acc);
  constexpr int NB = MT / 32;
#pragma unroll
  for (int im = 0; im < NB; ++im)
#pragma unroll
    for (int in = 0; in < NB; ++in)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = tr + tile_row<T, MT>(im, r);
        const int gc = tc + tile_col<MT>(in);
        if (gr < r1 && gc < c1 && gc <= gr) a[(int64_t)gr * lda + gc] -= acc.v[im][in][r];
      }
}

__global__ void k_set_int_impl(int32_t *p, int v) { *p = v; }

template <typename T>
int potrf_impl(T *a, int64_t lda, int n, int32_t *info, hipStream_t st) {
  hipLaunchKernelGGL(k_set_int_impl, dim3(1), dim3(1), 0, st, info, 0);
  FMLH_LAUNCH_CHECK("k_set_int");
  for (int K0 = 0; K0 < n; K0 += kNBO) {
    const int kend = min(K0 + kNBO, n);
    for (int k0 = K0; k0 < kend; k0 += kNB) {
      const int b = min(kNB, kend - k0);
      hipLaunchKernelGGL(k_potf2<T>, dim3(1), dim3(64), 0, st, a, lda, k0, b, info);
      FMLH_LAUNCH_CHECK("k_potf2");
      const int r0 = k0 + b;
      if (r0 < n) {
        hipLaunchKernelGGL(k_trsm<T>, dim3(ceil_div(n - r0, 256)), dim3(256), 0, st, a, lda, n, k0,
                           b, r0, info);
        FMLH_LAUNCH_CHECK("k_trsm");
        if (r0 < kend) {  // rest of the outer panel
          const int gx = ceil_div(kend - r0, kTile), gy = ceil_div(n - r0, kTile);
          hipLaunchKernelGGL((k_update<T, kTile, false>), dim3((unsigned)xcd_grid(gx, gy)),
                             dim3(256), 0, st, a, lda, r0, n, r0, kend, k0, k0 + b, gx, gy, info);
          FMLH_LAUNCH_CHECK("k_update(panel)");
        }
      }
    }
    if (kend < n) {  // trailing matrix: lower-triangle tiles only
      const int m = n - kend;
      if (m >= 4 * 128) {
        const int g = ceil_div(m, 128);
        const int64_t t = (int64_t)g * (g + 1) / 2;
        hipLaunchKernelGGL((k_update<T, 128, true>), dim3((unsigned)xcd_grid((int)t, 1)), dim3(256),
                           0, st, a, lda, kend, n, kend, n, K0, kend, g, g, info);
      } else {
        const int g = ceil_div(m, kTile);
        const int64_t t = (int64_t)g * (g + 1) / 2;
        hipLaunchKernelGGL((k_update<T, kTile, true>), dim3((unsigned)xcd_grid((int)t, 1)),
                           dim3(256), 0, st, a, lda, kend, n, kend, n, K0, kend, g, g, info);
      }
      FMLH_LAUNCH_CHECK("k_update(trailing)");
    }
  }
  return 0;
}

// ---- triangular solves -------------------------------------------------

constexpr int kChunk = 256;

// Forward partials: part[ch][r*nrhs + j] = sum_{c in chunk ch, c < k0}
// M[k0+r][c] x[c + j*ldx].  Each wave owns 16 rows, lanes span columns.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_partial(const T *__restrict__ a, int64_t lda, int k0,
                                                     int b, const T *__restrict__ x, int64_t ldx,
                                                     int nrhs, T *__restrict__ part) {
  const int ch = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cbeg = ch * kChunk;
  const int cend = min(cbeg + kChunk, k0);
  for (int j = 0; j < nrhs; ++j) {
    for (int rr = 0; rr < 16; ++rr) {
      const int r = wave * 16 + rr;
      T acc = T(0);
      if (r < b) {
        const T *row = a + (int64_t)(k0 + r) * lda;
        for (int c = cbeg + lane; c < cend; c += 64) acc += row[c] * x[c + (int64_t)j * ldx];
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
      if (lane == 0 && r < b) part[((int64_t)ch * kNB + r) * nrhs + j] = acc;
    }
  }
}

// Backward partials: part[ch][r*nrhs + j] = sum_{t in chunk ch, t >= k0+b}
// M[t][k0+r] x[t + j*ldx].  Lanes span r; waves split t, combined in order.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_partial(const T *__restrict__ a, int64_t lda, int n,
                                                     int k0, int b, const T *__restrict__ x,
                                                     int64_t ldx, int nrhs, T *__restrict__ part) {
  __shared__ T red[4][kNB];
  const int ch = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tbeg = k0 + b + ch * kChunk;
  const int tend = min(tbeg + kChunk, n);
  for (int j = 0; j < nrhs; ++j) {
    T acc = T(0);
    if (lane < b)
      for (int t = tbeg + wave; t < tend; t += 4)
        acc += a[(int64_t)t * lda + k0 + lane] * x[t + (int64_t)j * ldx];
    red[wave][lane] = acc;
    __syncthreads();
    if (threadIdx.x < 64 && lane < b) {
      const T v = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
      part[((int64_t)ch * kNB + lane) * nrhs + j] = v;
    }
    __syncthreads();
  }
}

// Triangular solve of one diagonal block; forward (lower) or backward
// (transposed).  rhs = x_block - sum_ch part[ch] (ascending ch).
template <typename T, bool kForward>
__global__ __launch_bounds__(256) void k_block_solve(const T *__restrict__ a, int64_t lda, int k0,
                                                     int b, T *__restrict__ x, int64_t ldx,
                                                     int nrhs, const T *__restrict__ part,
                                                     int nch) {
  __shared__ T L[kNB][kNB + 1];
  __shared__ T rhs[kNB];
  for (int idx = threadIdx.x; idx < b * b; idx += 256) {
    const int r = idx / b, c = idx % b;
    L[r][c] = (c <= r) ? a[(int64_t)(k0 + r) * lda + k0 + c] : T(0);
  }
  for (int j = 0; j < nrhs; ++j) {
    __syncthreads();
    if ((int)threadIdx.x < b) {
      const int r = threadIdx.x;
      T v = x[k0 + r + (int64_t)j * ldx];
      for (int ch = 0; ch < nch; ++ch) v -= part[((int64_t)ch * kNB + r) * nrhs + j];
      rhs[r] = v;
    }
    __syncthreads();
    if (kForward) {
      for (int c = 0; c < b; ++c) {
        if (threadIdx.x == 0) rhs[c] = rhs[c] / L[c][c];
        __syncthreads();
        for (int r = c + 1 + threadIdx.x; r < b; r += 256) rhs[r] -= L[r][c] * rhs[c];
        __syncthreads();
      }
    } else {
      for (int c = b - 1; c >= 0; --c) {
        if (threadIdx.x == 0) rhs[c] = rhs[c] / L[c][c];
        __syncthreads();
        for (int r = threadIdx.x; r < c; r += 256) rhs[r] -= L[c][r] * rhs[c];
        __syncthreads();
      }
    }
    if ((int)threadIdx.x < b) x[k0 + threadIdx.x + (int64_t)j * ldx] = rhs[threadIdx.x];
  }
}

template <typename T>
size_t potrs_ws(int n, int nrhs) {
  return WsCarver::need(sizeof(T) * (size_t)ceil_div(n > 0 ? n : 1, kChunk) * kNB *
                        (nrhs > 0 ? nrhs : 1));
}

template <typename T>
int potrs_impl(const T *a, int64_t lda, int n, T *x, int64_t ldx, int nrhs, void *ws,
               hipStream_t st) {
  T *part = static_cast<T *>(ws);
  const int nblk = ceil_div(n, kNB);
  for (int kb = 0; kb < nblk; ++kb) {
    const int k0 = kb * kNB, b = min(kNB, n - k0);
    const int nch = ceil_div(k0, kChunk);
    if (nch > 0) {
      hipLaunchKernelGGL(k_fwd_partial<T>, dim3(nch), dim3(256), 0, st, a, lda, k0, b, x, ldx,
                         nrhs, part);
      FMLH_LAUNCH_CHECK("k_fwd_partial");
    }
    hipLaunchKernelGGL((k_block_solve<T, true>), dim3(1), dim3(256), 0, st, a, lda, k0, b, x, ldx,
                       nrhs, part, nch);
    FMLH_LAUNCH_CHECK("k_block_solve<fwd>");
  }
  for (int kb = nblk - 1; kb >= 0; --kb) {
    const int k0 = kb * kNB, b = min(kNB, n - k0);
    const int nch = ceil_div(n - (k0 + b), kChunk);
    if (nch > 0) {
      hipLaunchKernelGGL(k_bwd_partial<T>, dim3(nch), dim3(256), 0, st, a, lda, n, k0, b, x, ldx,
                         nrhs, part);
      FMLH_LAUNCH_CHECK("k_bwd_partial");
    }
    hipLaunchKernelGGL((k_block_solve<T, false>), dim3(1), dim3(256), 0, st, a, lda, k0, b, x,
                       ldx, nrhs, part, nch);
    FMLH_LAUNCH_CHECK("k_block_solve<bwd>");
  }
  return 0;
}

}  // namespace

extern "C" {

size_t fmlh_potrf_ws_bytes(int n) { return n >= 0 ? 0 : 0; }
size_t fmlh_spotrf_ws_bytes(int n) { return n >= 0 ? 0 : 0; }
size_t fmlh_potrs_ws_bytes(int n, int nrhs) { return potrs_ws<double>(n, nrhs); }
size_t fmlh_spotrs_ws_bytes(int n, int nrhs) { return potrs_ws<float>(n, nrhs); }

int fmlh_potrf(double *a, int64_t lda, int n, int32_t *info, void *ws, size_t ws_bytes,
               void *stream) {
  (void)ws;
  (void)ws_bytes;
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(lda >= (n > 1 ? n : 1), 2);
  FMLH_ARG_CHECK(info != nullptr, 4);
  if (n > 0) FMLH_ARG_CHECK(a != nullptr, 1);
  if (n == 0) {
    hipLaunchKernelGGL(k_set_int_impl, dim3(1), dim3(1), 0, as_stream(stream), info, 0);
    FMLH_LAUNCH_CHECK("k_set_int");
    return 0;
  }
  timing_begin(as_stream(stream));
  int rc = potrf_impl<double>(a, lda, n, info, as_stream(stream));
  timing_end(as_stream(stream));
  return rc;
}

int fmlh_spotrf(float *a, int64_t lda, int n, int32_t *info, void *ws, size_t ws_bytes,
                void *stream) {
  (void)ws;
  (void)ws_bytes;
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(lda >= (n > 1 ? n : 1), 2);
  FMLH_ARG_CHECK(info != nullptr, 4);
  if (n > 0) FMLH_ARG_CHECK(a != nullptr, 1);
  if (n == 0) {
    hipLaunchKernelGGL(k_set_int_impl, dim3(1), dim3(1), 0, as_stream(stream), info, 0);
    FMLH_LAUNCH_CHECK("k_set_int");
    return 0;
  }
  timing_begin(as_stream(stream));
  int rc = potrf_impl<float>(a, lda, n, info, as_stream(stream));
  timing_end(as_stream(stream));
  return rc;
}

int fmlh_potrs(const double *a, int64_t lda, int n, double *x, int64_t ldx, int nrhs, void *ws,
               size_t ws_bytes, void *stream) {
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(lda >= (n > 1 ? n : 1), 2);
  FMLH_ARG_CHECK(nrhs >= 0, 6);
  FMLH_ARG_CHECK(ldx >= (n > 1 ? n : 1), 5);
  if (n == 0 || nrhs == 0) return 0;
  FMLH_ARG_CHECK(a != nullptr, 1);
  FMLH_ARG_CHECK(x != nullptr, 4);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= potrs_ws<double>(n, nrhs), 7);
  return potrs_impl<double>(a, lda, n, x, ldx, nrhs, ws, as_stream(stream));
}

int fmlh_spotrs(const float *a, int64_t lda, int n, float *x, int64_t ldx, int nrhs, void *ws,
                size_t ws_bytes, void *stream) {
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(lda >= (n > 1 ? n : 1), 2);
  FMLH_ARG_CHECK(nrhs >= 0, 6);
  FMLH_ARG_CHECK(ldx >= (n > 1 ? n : 1), 5);
  if (n == 0 || nrhs == 0) return 0;
  FMLH_ARG_CHECK(a != nullptr, 1);
  FMLH_ARG_CHECK(x != nullptr, 4);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= potrs_ws<float>(n, nrhs), 7);
  return potrs_impl<float>(a, lda, n, x, ldx, nrhs, ws, as_stream(stream));
}

}  // extern "C"
