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
//   for each 1024-column outer panel:
//     for each 64-column inner block inside it:
//       k_potf2   diagonal 64x64 block in registers (one wave, lane = row)
//       k_trsm    rows below: M[r][k] <- M[r][k] L_kk^{-T} (one thread/row)
//       k_update  panel columns right of the block: C -= P Q^T, 64 tiles
//     k_update    the next outer panel's columns, K = 1024, 64 tiles; then
//                 the rest of the trailing matrix on 128x128 MFMA tiles over
//                 its lower triangle only, while the next panel is factored
//                 on an auxiliary stream (one panel of look-ahead)
// A non-positive (or NaN) pivot sets *info to its 1-based index, as LAPACK;
// every later kernel sees *info != 0 and returns, so the factor is left
// undefined but no kernel waits on anything.
//
// Solve: blocked forward (L w = b) and backward (L^T x = w) substitution
// over 64-row blocks.  One right-hand side: one launch per direction, the
// blocks chained through ready flags (k_trsv_chain).  Several: per block a
// multi-workgroup partial product of the already solved part followed by a
// one-workgroup triangular solve of the block.
#include <math.h>
#include <stdio.h>

#include <mutex>

#include <type_traits>

#include "common.h"
#include "gemm_nt.h"
#include "ozaki.h"

namespace {

using namespace fmlh;

constexpr int kNB = 64;     // inner block
constexpr int kNBO = 1024;  // outer panel

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
// every lane) and scales column j below it; every lane r then updates its
// row, M[r][c] -= L[r][j] L[c][j] for c > j (entries above the diagonal are
// updated too and never stored).  The steps go kPB columns at a time: inside
// a block, L[c][j] of the block's own later columns comes from lane c by
// readlane; the block's kPB columns then go through LDS once, and the rows'
// columns right of the block subtract their kPB products in step order.
// Every entry sees the same operations in the same order as one step at a
// time, with a quarter of the LDS round trips.  Fully unrolled, so the row
// stays in registers.
constexpr int kPB = 4;
constexpr int kPC = 8;

template <typename T>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_potf2(
    T *__restrict__ a, int64_t lda, int k0, int b, int32_t *__restrict__ info) {
  __shared__ T colb[kNB][kPB];  // colb[c][jj] = L[c][j0 + jj]
  if (*info != 0) return;
  const int r = threadIdx.x;
  T *grow = a + (int64_t)(k0 + r) * lda + k0;
  T row[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) row[c] = (r < b && c <= r) ? grow[c] : T(0);
#pragma unroll
  for (int j0 = 0; j0 < kNB; j0 += kPB) {
    if (j0 < b) {
      T lr[kPB];
#pragma unroll
      for (int jj = 0; jj < kPB; ++jj) {
        const int j = j0 + jj;
        lr[jj] = T(0);
        if (j < b) {
          const T d = lane_value(row[j], j);
          if (!(d > T(0))) {  // uniform: every lane read the same pivot
            if (r == 0) *info = k0 + j + 1;
            return;
          }
          const T piv = sqrt(d);
          const T lrj = (r > j && r < b) ? row[j] / piv : T(0);
          row[j] = (r == j) ? piv : (r > j ? lrj : row[j]);
          lr[jj] = lrj;
#pragma unroll
          for (int cc = jj + 1; cc < kPB; ++cc) row[j0 + cc] -= lrj * lane_value(lrj, j0 + cc);
        }
      }
#pragma unroll
      for (int jj = 0; jj < kPB; ++jj) colb[r][jj] = lr[jj];
      __syncthreads();
      // kPC columns' values are read before their products (the scheduling
      // barrier keeps the reads together), so one LDS wait covers kPC reads
#pragma unroll
      for (int c0 = j0 + kPB; c0 < kNB; c0 += kPC) {
        T v[kPC][kPB];
#pragma unroll
        for (int cc = 0; cc < kPC; ++cc)
#pragma unroll
          for (int jj = 0; jj < kPB; ++jj) v[cc][jj] = c0 + cc < kNB ? colb[c0 + cc][jj] : T(0);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int cc = 0; cc < kPC; ++cc)
          if (c0 + cc < kNB) {
#pragma unroll
            for (int jj = 0; jj < kPB; ++jj) row[c0 + cc] -= lr[jj] * v[cc][jj];
          }
      }
      __syncthreads();
    }
  }
  if (r < b) {
#pragma unroll
    for (int c = 0; c < kNB; ++c)
      if (c <= r) grow[c] = row[c];
  }
}

// Rows r in [r0, n): x = M[r][k0:k0+b] solves x L^T = m_r (forward
// substitution against the factored diagonal block, held in LDS).  Each
// entry is x[c] = (m[c] - sum_{t < c} x[t] L[c][t]) / L[c][c] with the
// products subtracted in t order; four entries' chains run interleaved up
// to the first of them, then the four's own triangle in order.  A full
// block (kFull) stages the factor transposed (LT[t][c] = L[c][t], so the
// four entries' factor values at one t are contiguous) and reads them for
// eight t's at a time before their products: the substitution is bound by
// LDS round trips, and reading one value per product took 31 of the kernel's
// 35 us (20 us this way, the same operations in the same order).  Rows move
// in 16-byte pieces when aligned.
template <typename T>
struct TrsmLds {
  // row strides keeping every row 16-byte aligned
  static constexpr int kStride = sizeof(T) == 8 ? kNB + 2 : kNB + 4;
};

template <typename T, bool kFull>
__global__ __launch_bounds__(256) void k_trsm(T *__restrict__ a, int64_t lda, int n, int k0, int b,
                                              int r0, const int32_t *__restrict__ info) {
  constexpr int kS = TrsmLds<T>::kStride;
  __shared__ __attribute__((aligned(16))) T L[kNB * kS];  // full: LT[t][c]; ragged: L[r][c]
  if (*info != 0) return;
  if (kFull) {
#pragma unroll
    for (int idx = threadIdx.x; idx < kNB * kNB; idx += 256) {
      const int r = idx >> 6, c = idx & 63;
      L[c * kS + r] = (c <= r) ? a[(int64_t)(k0 + r) * lda + k0 + c] : T(0);
    }
  } else {
    for (int idx = threadIdx.x; idx < b * b; idx += 256) {
      const int r = idx / b, c = idx % b;
      L[r * kS + c] = (c <= r) ? a[(int64_t)(k0 + r) * lda + k0 + c] : T(0);
    }
  }
  __syncthreads();
  const int r = r0 + blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  T *row = a + (int64_t)r * lda + k0;
  T x[kNB];
  typedef T T2 __attribute__((ext_vector_type(2)));
  // 16-byte pieces when the row's start is 16-byte aligned (lda even, k0 a
  // multiple of 64)
  const bool vec = kFull && sizeof(T) == 8 && ((reinterpret_cast<uintptr_t>(row) & 15) == 0);
  if (vec) {
#pragma unroll
    for (int c = 0; c < kNB; c += 2) {
      const T2 p = *reinterpret_cast<const T2 *>(row + c);
      x[c] = p.x;
      x[c + 1] = p.y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < kNB; ++c) x[c] = (kFull || c < b) ? row[c] : T(0);
  }
  if (kFull) {
    constexpr int kCh = 8;
#pragma unroll
    for (int c0 = 0; c0 < kNB; c0 += 4) {
      T v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = x[c0 + i];
#pragma unroll
      for (int t0 = 0; t0 < c0; t0 += kCh) {
        T l[kCh][4];
#pragma unroll
        for (int tt = 0; tt < kCh; ++tt)
          if (t0 + tt < c0) {
            const T *src = L + (t0 + tt) * kS + c0;
            if (sizeof(T) == 8) {
              const T2 p0 = *reinterpret_cast<const T2 *>(src);
              const T2 p1 = *reinterpret_cast<const T2 *>(src + 2);
              l[tt][0] = p0.x;
              l[tt][1] = p0.y;
              l[tt][2] = p1.x;
              l[tt][3] = p1.y;
            } else {
              typedef T T4 __attribute__((ext_vector_type(4)));
              const T4 p = *reinterpret_cast<const T4 *>(src);
              l[tt][0] = p.x;
              l[tt][1] = p.y;
              l[tt][2] = p.z;
              l[tt][3] = p.w;
            }
          }
        // the chunk's reads issue together, one wait before its products
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int tt = 0; tt < kCh; ++tt)
          if (t0 + tt < c0) {
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] -= x[t0 + tt] * l[tt][i];
          }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = c0 + i;
#pragma unroll
        for (int t = c0; t < c; ++t) v[i] -= x[t] * L[t * kS + c];
        x[c] = v[i] / L[c * kS + c];
      }
    }
  } else {
#pragma unroll
    for (int c0 = 0; c0 < kNB; c0 += 4) {
      if (c0 < b) {
        T v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = x[c0 + i];
#pragma unroll
        for (int t = 0; t < c0; ++t)
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] -= x[t] * L[(c0 + i) * kS + t];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int c = c0 + i;
          if (c < b) {
#pragma unroll
            for (int t = c0; t < c; ++t) v[i] -= x[t] * L[c * kS + t];
            x[c] = v[i] / L[c * kS + c];
          } else {
            x[c] = T(0);
          }
        }
      }
    }
  }
  if (vec) {
#pragma unroll
    for (int c = 0; c < kNB; c += 2) *reinterpret_cast<T2 *>(row + c) = T2{x[c], x[c + 1]};
  } else {
#pragma unroll
    for (int c = 0; c < kNB; ++c)
      if (kFull || c < b) row[c] = x[c];
  }
}

// C[r][c] -= sum_{t in [kb, ke)} M[r][t] M[c][t] for r in [r0, r1),
// c in [c0, c1), c <= r, on MT x MT tiles.  kTri: the region is the square
// [r0, r1)^2 and only its lower-triangle tiles are launched (tri_tile order,
// in XCD-contiguous runs); otherwise a rectangular gx x gy grid in which
// tiles above the diagonal exit.
template <typename T, int MT, bool kTri>
__global__ __launch_bounds__(MT == 128 ? 512 : 256, (MT == 128 ? 2 : 1)) void k_update(T *__restrict__ a, int64_t lda, int r0, int r1,
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
  constexpr int WR = MT == 128 ? 4 : 2;
  TileAcc<T, MT, WR> acc;
  tile_nt<T, false, MT, WR>(a + (int64_t)tr * lda + kb, lda, r1 - tr, a + (int64_t)tc * lda + kb,
                            lda, c1 - tc, ke - kb, s, acc);
#pragma unroll
  for (int im = 0; im < TileAcc<T, MT, WR>::NBM; ++im)
#pragma unroll
    for (int in = 0; in < TileAcc<T, MT, WR>::NBN; ++in)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = tr + tile_row<T, MT, WR>(im, r);
        const int gc = tc + tile_col<MT>(in);
        if (gr < r1 && gc < c1 && gc <= gr) a[(int64_t)gr * lda + gc] -= acc.v[im][in][r];
      }
}

__global__ void k_set_int_impl(int32_t *p, int v) { *p = v; }

// The inner blocks of outer panel [K0, kend), in middle blocks of kMid
// columns: each 64-column block is factored (k_potf2), the rows below it
// solved (k_trsm) and the rest of its middle block's columns updated with it
// (k_update, 64 tiles); then the rest of the outer panel's columns are
// updated with the whole middle block at once, so that they are read and
// written once per kMid columns of k instead of once per 64 (36k, Ozaki /
// FP64 factorisation: 64 -> 151.4 / 306.7 ms, 128 -> 146.0 / 300.2,
// 256 -> 144 / 297.9, 512 -> 145.1 / 300.0).
#ifndef FMLH_PANEL_MID
#define FMLH_PANEL_MID 256
#endif
constexpr int kMid = FMLH_PANEL_MID;
static_assert(kMid % kNB == 0, "middle blocks hold whole inner blocks");

template <typename T>
int panel_factor(T *a, int64_t lda, int n, int K0, int kend, int32_t *info, hipStream_t st) {
  for (int m0 = K0; m0 < kend; m0 += kMid) {
    const int mend = min(m0 + kMid, kend);
    for (int k0 = m0; k0 < mend; k0 += kNB) {
      const int b = min(kNB, mend - k0);
      hipLaunchKernelGGL(k_potf2<T>, dim3(1), dim3(64), 0, st, a, lda, k0, b, info);
      FMLH_LAUNCH_CHECK("k_potf2");
      const int r0 = k0 + b;
      if (r0 >= n) continue;
      if (b == kNB)
        hipLaunchKernelGGL((k_trsm<T, true>), dim3(ceil_div(n - r0, 256)), dim3(256), 0, st, a, lda, n,
                           k0, b, r0, info);
      else
        hipLaunchKernelGGL((k_trsm<T, false>), dim3(ceil_div(n - r0, 256)), dim3(256), 0, st, a, lda,
                           n, k0, b, r0, info);
      FMLH_LAUNCH_CHECK("k_trsm");
      if (r0 < mend) {  // rest of the middle block
        const int gx = ceil_div(mend - r0, kTile), gy = ceil_div(n - r0, kTile);
        hipLaunchKernelGGL((k_update<T, kTile, false>), dim3((unsigned)xcd_grid(gx, gy)), dim3(256),
                           0, st, a, lda, r0, n, r0, mend, k0, k0 + b, gx, gy, info);
        FMLH_LAUNCH_CHECK("k_update(panel)");
      }
    }
    if (mend < kend) {  // rest of the outer panel, all rows below mend, with this middle block
      if (n - mend >= 4 * 128) {
        const int gx = ceil_div(kend - mend, 128), gy = ceil_div(n - mend, 128);
        hipLaunchKernelGGL((k_update<T, 128, false>), dim3((unsigned)xcd_grid(gx, gy)), dim3(512),
                           0, st, a, lda, mend, n, mend, kend, m0, mend, gx, gy, info);
      } else {
        const int gx = ceil_div(kend - mend, kTile), gy = ceil_div(n - mend, kTile);
        hipLaunchKernelGGL((k_update<T, kTile, false>), dim3((unsigned)xcd_grid(gx, gy)), dim3(256),
                           0, st, a, lda, mend, n, mend, kend, m0, mend, gx, gy, info);
      }
      FMLH_LAUNCH_CHECK("k_update(panel rest)");
    }
  }
  return 0;
}

// C -= P P^T over the lower triangle of the square [r0, n)^2, P the columns
// [kb, ke) of those rows.
template <typename T>
int trailing_update(T *a, int64_t lda, int n, int r0, int kb, int ke, int32_t *info, hipStream_t st) {
  const int m = n - r0;
  if (m <= 0) return 0;
  if (m >= 4 * 128) {
    const int g = ceil_div(m, 128);
    const int64_t t = (int64_t)g * (g + 1) / 2;
    hipLaunchKernelGGL((k_update<T, 128, true>), dim3((unsigned)xcd_grid((int)t, 1)), dim3(512), 0,
                       st, a, lda, r0, n, r0, n, kb, ke, g, g, info);
  } else {
    const int g = ceil_div(m, kTile);
    const int64_t t = (int64_t)g * (g + 1) / 2;
    hipLaunchKernelGGL((k_update<T, kTile, true>), dim3((unsigned)xcd_grid((int)t, 1)), dim3(256),
                       0, st, a, lda, r0, n, r0, n, kb, ke, g, g, info);
  }
  FMLH_LAUNCH_CHECK("k_update(trailing)");
  return 0;
}

// One auxiliary stream (non-blocking, default priority) and two events per
// device, created on first use, for the factorisation's look-ahead.  A
// caller on a CU-masked stream made by fmlh_stream_create (the serving
// loop's fit stream) gets no look-ahead: a stream of the library's own would
// run the panel work outside the caller's CUs, and every further stream
// takes one of the process's few hardware queues, which the serving loop's
// prediction stream must not share.
struct Aux {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  std::mutex *enqueue = nullptr;  // held while one factorisation enqueues on s
};

bool aux_for_stream(hipStream_t caller, Aux &out) {
  static std::mutex mu;
  static std::mutex enqueue_mu[64];
  static Aux table[64];
  std::vector<uint32_t> mask;
  if (stream_cu_mask(caller, mask)) return false;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    return false;
  }
  std::lock_guard<std::mutex> lk(mu);
  Aux &x = table[dev];
  if (x.s == nullptr) {
    Aux y;
    if (hipStreamCreateWithFlags(&y.s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&y.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&y.join, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    y.enqueue = &enqueue_mu[dev];
    x = y;
  }
  out = x;
  return true;
}

// Right-looking factorisation over 1024-column outer panels with one panel
// of look-ahead: after panel p is factored, the auxiliary stream updates the
// columns of panel p + 1 with it and factors panel p + 1, while the caller's
// stream updates the rest of the trailing matrix (disjoint columns); the
// next update waits for that factor.  Inside a
// graph capture, or without the auxiliary stream, everything runs on the
// caller's stream in the same order.
//
// oz_ws (FP64 only, fmlh_potrf_ozaki): the next panel's columns and the
// trailing matrix are updated on INT8 matrix cores from the Ozaki split of
// the finished panel (ozaki.h) instead of FP64 MFMA tiles; the panels
// themselves are factored in FP64 as always.
template <typename T>
int potrf_impl(T *a, int64_t lda, int n, int32_t *info, hipStream_t st, void *oz_ws = nullptr) {
  hipLaunchKernelGGL(k_set_int_impl, dim3(1), dim3(1), 0, st, info, 0);
  FMLH_LAUNCH_CHECK("k_set_int");
  Aux aux;
  bool ahead = true;
  {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
      (void)hipGetLastError();
      ahead = false;   // nothing is created or forked while a graph is captured
    }
  }
  if (ahead) ahead = aux_for_stream(st, aux);
  // Calls from several host threads share the device's auxiliary stream and
  // events: one factorisation enqueues at a time (each event wait takes the
  // event's latest record, so serialised enqueues keep every fork and join
  // paired; the device work of the calls still overlaps on their streams).
  std::unique_lock<std::mutex> lk;
  if (ahead) lk = std::unique_lock<std::mutex>(*aux.enqueue);
  const bool oz = std::is_same<T, double>::value && oz_ws != nullptr;
  const int nbo = oz ? kOzPanel : kNBO;  // outer panel width
  int rc = panel_factor(a, lda, n, 0, min(nbo, n), info, st);
  if (rc) return rc;
  for (int K0 = 0; K0 < n; K0 += nbo) {
    const int kend = min(K0 + nbo, n);
    if (kend >= n) break;
    const int nend = min(kend + nbo, n);  // the next panel is [kend, nend)
    if (oz) {  // the finished panel's rows below it, split once for both updates
      rc = oz_split(reinterpret_cast<double *>(a), lda, kend, n - kend, K0, kend - K0, kOzSlices,
                    oz_ws, info, st);
      if (rc) return rc;
    }
    // the next panel's columns, all rows below kend, with this panel
    auto next_update = [&](hipStream_t s) -> int {
      if (oz)
        return oz_update(reinterpret_cast<double *>(a), lda, kend, n, kend, nend, kend, n - kend,
                         kOzSlices, oz_ws, false, info, s);
      if (n - kend >= 4 * 128) {
        const int gx = ceil_div(nend - kend, 128), gy = ceil_div(n - kend, 128);
        hipLaunchKernelGGL((k_update<T, 128, false>), dim3((unsigned)xcd_grid(gx, gy)), dim3(512),
                           0, s, a, lda, kend, n, kend, nend, K0, kend, gx, gy, info);
      } else {
        const int gx = ceil_div(nend - kend, kTile), gy = ceil_div(n - kend, kTile);
        hipLaunchKernelGGL((k_update<T, kTile, false>), dim3((unsigned)xcd_grid(gx, gy)),
                           dim3(256), 0, s, a, lda, kend, n, kend, nend, K0, kend, gx, gy, info);
      }
      FMLH_LAUNCH_CHECK("k_update(next panel)");
      return 0;
    };
    // the rest of the trailing matrix (columns from nend), with this panel
    auto rest_update = [&]() -> int {
      return oz ? oz_update(reinterpret_cast<double *>(a), lda, nend, n, nend, n, kend, n - kend,
                            kOzSlices, oz_ws, true, info, st)
                : trailing_update(a, lda, n, nend, K0, kend, info, st);
    };
    if (ahead) {
      // the auxiliary stream updates the next panel's columns and factors
      // them while the caller's stream updates the rest (disjoint columns,
      // both reading only this panel)
      hipError_t e = hipEventRecord(aux.fork, st);
      if (e == hipSuccess) e = hipStreamWaitEvent(aux.s, aux.fork, 0);
      if (e != hipSuccess) return record_error(e, "potrf fork");
      rc = next_update(aux.s);
      if (rc) return rc;
      rc = panel_factor(a, lda, n, kend, nend, info, aux.s);
      if (rc) return rc;
      e = hipEventRecord(aux.join, aux.s);
      if (e != hipSuccess) return record_error(e, "potrf join record");
      rc = rest_update();
      if (rc) return rc;
      e = hipStreamWaitEvent(st, aux.join, 0);
      if (e != hipSuccess) return record_error(e, "potrf join");
    } else {
      rc = next_update(st);
      if (rc) return rc;
      rc = panel_factor(a, lda, n, kend, nend, info, st);
      if (rc) return rc;
      rc = rest_update();
      if (rc) return rc;
    }
  }
  return 0;
}

// ---- inverse from the factor (LAPACK dpotri('U') semantics) ------------
//
// inv(M) = L^{-T} L^{-1}.  X = L^{-1} is formed by blocked forward
// substitution with the identity as right-hand side, stored transposed in
// the workspace (XT[j][k] = X[k][j], row-major), so that both of its
// products are "NT" MFMA tiles over contiguous rows:
//   block row b (rows r0..r0+bs): C = L[b rows][0:r0] X[0:r0][j] for the
//     columns j < r0 (k_potri_gemm; X is lower triangular, so columns
//     j >= r0 have no entries above block b), then every column j < r0+bs
//     solves L_bb x = e_j - C (k_potri_diag, one column per thread);
//   inv(M)[r][c] = sum_k XT[r][k] XT[c][k] over the lower-triangle tiles
//     (k_potri_syrk), written over the factor.

// XT[j][r0 + m] = -sum_{k < r0} M[r0 + m][k] XT[j][k], m < bs, j < r0.
// XT[j][k] is 0 for k < j (X = L^{-1} is lower triangular; the workspace is
// zero-filled and the diagonal solves leave exact zeros there), so the tile
// of columns j0.. sums k from j0: the products dropped are exact zeros, on
// the same 4-aligned k steps.
template <typename T>
__global__ __launch_bounds__(256) void k_potri_gemm(const T *__restrict__ a, int64_t lda, int r0,
                                                    int bs, T *__restrict__ xt, int64_t ldx,
                                                    const int32_t *__restrict__ info) {
  __shared__ TileSmem<T, kTile> s;
  if (*info != 0) return;
  const int j0 = blockIdx.x * kTile;
  TileAcc<T, kTile> acc;
  tile_nt<T, false, kTile>(a + (int64_t)r0 * lda + j0, lda, bs, xt + (int64_t)j0 * ldx + j0, ldx,
                           r0 - j0, r0 - j0, s, acc);
#pragma unroll
  for (int im = 0; im < 2; ++im)
#pragma unroll
    for (int in = 0; in < 2; ++in)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = tile_row<T, kTile>(im, r);
        const int j = j0 + tile_col<kTile>(in);
        if (m < bs && j < r0) xt[(int64_t)j * ldx + r0 + m] = -acc.v[im][in][r];
      }
}

// Column j < r0 + bs: x = L_bb^{-1} (e_j - C) for the block rows r0..r0+bs,
// forward substitution in registers, L_bb in LDS.  A zero pivot sets *info.
// The same product split over k: block (tile, c) of a 2-D grid sums
// k in [j0 + c*kPotriKC, min(j0 + (c+1)*kPotriKC, r0)) into its own 64 x 64
// partial; k_potri_sum adds a tile's partials in chunk order.  Used when
// the deepest tile (j0 = 0, depth r0) spans several chunks, so that a
// block row's few long tiles become many short ones.
constexpr int kPotriKC = 2048;

template <typename T>
__global__ __launch_bounds__(256) void k_potri_gemm_part(const T *__restrict__ a, int64_t lda,
                                                         int r0, int bs, const T *__restrict__ xt,
                                                         int64_t ldx, int cmax,
                                                         T *__restrict__ part,
                                                         const int32_t *__restrict__ info) {
  __shared__ TileSmem<T, kTile> s;
  if (*info != 0) return;
  const int tile = blockIdx.x, c = blockIdx.y;
  const int j0 = tile * kTile;
  const int kb = j0 + c * kPotriKC;
  if (kb >= r0) return;
  const int ke = min(kb + kPotriKC, r0);
  TileAcc<T, kTile> acc;
  tile_nt<T, false, kTile>(a + (int64_t)r0 * lda + kb, lda, bs, xt + (int64_t)j0 * ldx + kb, ldx,
                           r0 - j0, ke - kb, s, acc);
  T *pt = part + ((int64_t)tile * cmax + c) * kTile * kTile;
#pragma unroll
  for (int im = 0; im < 2; ++im)
#pragma unroll
    for (int in = 0; in < 2; ++in)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        pt[tile_row<T, kTile>(im, r) * kTile + tile_col<kTile>(in)] = acc.v[im][in][r];
}

template <typename T>
__global__ __launch_bounds__(256) void k_potri_sum(int r0, int bs, T *__restrict__ xt, int64_t ldx,
                                                   int cmax, const T *__restrict__ part,
                                                   const int32_t *__restrict__ info) {
  if (*info != 0) return;
  const int tile = blockIdx.x;
  const int j0 = tile * kTile;
  const int nc = (r0 - j0 + kPotriKC - 1) / kPotriKC;
  for (int e = threadIdx.x; e < kTile * kTile; e += 256) {
    const int m = e / kTile, jj = e % kTile;
    const int j = j0 + jj;
    if (m >= bs || j >= r0) continue;
    T v = T(0);
    for (int c = 0; c < nc; ++c) v = v + part[((int64_t)tile * cmax + c) * kTile * kTile + e];
    xt[(int64_t)j * ldx + r0 + m] = -v;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_potri_diag(const T *__restrict__ a, int64_t lda, int r0,
                                                    int bs, T *__restrict__ xt, int64_t ldx,
                                                    int32_t *__restrict__ info) {
  __shared__ T L[kNB][kNB + 1];
  if (*info != 0) return;
  for (int idx = threadIdx.x; idx < bs * bs; idx += 256) {
    const int r = idx / bs, c = idx % bs;
    L[r][c] = (c <= r) ? a[(int64_t)(r0 + r) * lda + r0 + c] : T(0);
  }
  __syncthreads();
  if (threadIdx.x < bs && L[threadIdx.x][threadIdx.x] == T(0)) atomicCAS(info, 0, r0 + threadIdx.x + 1);
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= r0 + bs) return;
  T *row = xt + (int64_t)j * ldx + r0;
  T x[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    T v = T(0);
    if (c < bs) v = (j < r0) ? row[c] : T(j - r0 == c ? 1 : 0);
    x[c] = v;
  }
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    if (c < bs) {
      x[c] = x[c] / L[c][c];
#pragma unroll
      for (int t = c + 1; t < kNB; ++t) x[t] -= x[c] * L[t][c];
    }
  }
#pragma unroll
  for (int c = 0; c < kNB; ++c)
    if (c < bs) row[c] = x[c];
}

// M[r][c] = sum_{k >= tile row start} XT[r][k] XT[c][k] for c <= r < n,
// lower-triangle tiles in tri_tile order (XT[r][k] = 0 for k < r).
// Runs of consecutive tri_tile tiles, one per XCD (workgroup L runs on XCD
// L % 8 and takes tile runs.start[L % 8] + L / 8), cut so that each run
// holds about the same work.  For the inverse product a tile's depth is
// n - ty*MT, so equal tile counts would give the first XCD the deepest
// tiles and the last the shallowest.
struct XcdRuns {
  int64_t start[kXcds + 1];
};

inline XcdRuns balanced_runs(int g, int n, int mt, int64_t *max_len) {
  const int64_t t = (int64_t)g * (g + 1) / 2;
  const int64_t overhead = 4 * mt;  // staging and epilogue, in k columns
  double total = 0.0;
  for (int64_t l = 0; l < t; ++l) {
    int tx = 0, ty = 0;
    tri_tile(g, l, tx, ty);
    total += (double)((n - ty * mt > 0 ? n - ty * mt : 0) + overhead);
  }
  XcdRuns r;
  r.start[0] = 0;
  int x = 1;
  double acc = 0.0;
  for (int64_t l = 0; l < t && x < kXcds; ++l) {
    int tx = 0, ty = 0;
    tri_tile(g, l, tx, ty);
    acc += (double)((n - ty * mt > 0 ? n - ty * mt : 0) + overhead);
    while (x < kXcds && acc >= total * x / kXcds) r.start[x++] = l + 1;
  }
  while (x <= kXcds) r.start[x++] = t;
  int64_t m = 0;
  for (int k = 0; k < kXcds; ++k) m = r.start[k + 1] - r.start[k] > m ? r.start[k + 1] - r.start[k] : m;
  *max_len = m;
  return r;
}

template <typename T, int MT>
__global__ __launch_bounds__(MT == 128 ? 512 : 256, (MT == 128 ? 2 : 1)) void k_potri_syrk(
    const T *__restrict__ xt, int64_t ldx, int n, int g, T *__restrict__ a, int64_t lda,
    XcdRuns runs, const int32_t *__restrict__ info) {
  __shared__ TileSmem<T, MT> s;
  if (*info != 0) return;
  int tx, ty;
  const int64_t L = blockIdx.x;
  const int x = (int)(L % kXcds);
  const int64_t local = runs.start[x] + L / kXcds;
  if (local >= runs.start[x + 1] || !tri_tile(g, local, tx, ty)) return;
  const int tr = ty * MT, tc = tx * MT;
  constexpr int WR = MT == 128 ? 4 : 2;
  TileAcc<T, MT, WR> acc;
  tile_nt<T, false, MT, WR>(xt + (int64_t)tr * ldx + tr, ldx, n - tr, xt + (int64_t)tc * ldx + tr,
                            ldx, n - tc, n - tr, s, acc);
#pragma unroll
  for (int im = 0; im < TileAcc<T, MT, WR>::NBM; ++im)
#pragma unroll
    for (int in = 0; in < TileAcc<T, MT, WR>::NBN; ++in)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = tr + tile_row<T, MT, WR>(im, r);
        const int gc = tc + tile_col<MT>(in);
        if (gr < n && gc <= gr) a[(int64_t)gr * lda + gc] = acc.v[im][in][r];
      }
}

__global__ void k_fill_rows(double *p, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0.0;
}

int potri_impl(double *a, int64_t lda, int n, int32_t *info, double *xt, double *part,
               hipStream_t st) {
  const int64_t ldx = n;
  hipLaunchKernelGGL(k_set_int_impl, dim3(1), dim3(1), 0, st, info, 0);
  FMLH_LAUNCH_CHECK("k_set_int");
  const int64_t cells = (int64_t)n * ldx;
  hipLaunchKernelGGL(k_fill_rows, dim3((unsigned)ceil_div64(cells, 256)), dim3(256), 0, st, xt, cells);
  FMLH_LAUNCH_CHECK("k_fill_rows");
  for (int r0 = 0; r0 < n; r0 += kNB) {
    const int bs = min(kNB, n - r0);
    if (r0 > kPotriKC) {
      const int tiles = ceil_div(r0, kTile), cmax = ceil_div(r0, kPotriKC);
      hipLaunchKernelGGL(k_potri_gemm_part<double>, dim3(tiles, cmax), dim3(256), 0, st, a, lda, r0,
                         bs, xt, ldx, cmax, part, info);
      FMLH_LAUNCH_CHECK("k_potri_gemm_part");
      hipLaunchKernelGGL(k_potri_sum<double>, dim3(tiles), dim3(256), 0, st, r0, bs, xt, ldx, cmax,
                         part, info);
      FMLH_LAUNCH_CHECK("k_potri_sum");
    } else if (r0 > 0) {
      hipLaunchKernelGGL(k_potri_gemm<double>, dim3(ceil_div(r0, kTile)), dim3(256), 0, st, a, lda,
                         r0, bs, xt, ldx, info);
      FMLH_LAUNCH_CHECK("k_potri_gemm");
    }
    hipLaunchKernelGGL(k_potri_diag<double>, dim3(ceil_div(r0 + bs, 256)), dim3(256), 0, st, a, lda,
                       r0, bs, xt, ldx, info);
    FMLH_LAUNCH_CHECK("k_potri_diag");
  }
  if (n >= 4 * 128) {
    const int g = ceil_div(n, 128);
    int64_t len = 0;
    const XcdRuns runs = balanced_runs(g, n, 128, &len);
    hipLaunchKernelGGL((k_potri_syrk<double, 128>), dim3((unsigned)(len * kXcds)), dim3(512), 0, st,
                       xt, ldx, n, g, a, lda, runs, info);
  } else {
    const int g = ceil_div(n, kTile);
    int64_t len = 0;
    const XcdRuns runs = balanced_runs(g, n, kTile, &len);
    hipLaunchKernelGGL((k_potri_syrk<double, kTile>), dim3((unsigned)(len * kXcds)), dim3(256), 0, st,
                       xt, ldx, n, g, a, lda, runs, info);
  }
  FMLH_LAUNCH_CHECK("k_potri_syrk");
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

// One-launch triangular solves for a single right-hand side.  Workgroups
// take 64-row blocks in ticket order (forward: block = ticket; backward:
// block = last - ticket), accumulate the products of their rows with each
// solved block as its values appear in the publish buffer (blocks in solve
// order, each tile loaded before its values are awaited), solve their
// diagonal block in one wave (multiplying by the pivots' reciprocals) and
// publish it.  A ticket waits only on smaller tickets, whose workgroups are
// already running, so every wait ends.  ctl[0]: the ticket counter (zero
// before the launch); pub: n values, the sentinel before the launch.

// Publication of solved values between workgroups without fences: the
// publish buffer starts filled with a signalling-NaN pattern that no
// arithmetic produces; each solved value is written once with a relaxed
// device-scope atomic store, and a reader polls its value with relaxed
// device-scope atomic loads until the pattern is gone.  Each value carries
// its own readiness, so no ordering between values (and no cache
// write-back or invalidation) is needed.
template <typename T> struct PubBits;
template <> struct PubBits<double> {
  typedef unsigned long long U;
  static constexpr U kSentinel = 0x7FF4000000000DEAull;
};
template <> struct PubBits<float> {
  typedef unsigned int U;
  static constexpr U kSentinel = 0x7FA00DEAu;
};

template <typename T>
__device__ __forceinline__ void publish(T *pub, int i, T v) {
  typedef typename PubBits<T>::U U;
  U u;
  __builtin_memcpy(&u, &v, sizeof(u));
  __hip_atomic_store(reinterpret_cast<U *>(pub) + i, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ T await_value(const T *pub, int i) {
  typedef typename PubBits<T>::U U;
  const U *p = reinterpret_cast<const U *>(pub) + i;
  U u = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (u == PubBits<T>::kSentinel) {
    __builtin_amdgcn_s_sleep(1);
    u = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  T v;
  __builtin_memcpy(&v, &u, sizeof(v));
  return v;
}

// Before a chained solve: the ticket counter to zero, the publish buffer to
// the sentinel (a kernel, not a memset, so the call can be captured).
template <typename T>
__global__ void k_trsv_prep(int32_t *ticket, T *pub, int n) {
  typedef typename PubBits<T>::U U;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *ticket = 0;
  if (i < n) reinterpret_cast<U *>(pub)[i] = PubBits<T>::kSentinel;
}

// kInv: a full diagonal block is applied as its inverse W = L_bb^-1,
// formed by wave 0 while the workgroup waits for the blocks before it, so
// the block's own step on the chain is a 64 x 64 product over all four
// waves instead of 64 dependent substitution steps.  Used by FP32 (the
// mixed form's preconditioner solve) and by the FP64 correction solve of
// iterative refinement.  The plain FP64 solve keeps the substitution: its
// rounding meets a triangular solve's backward-error bound, the inverse's
// only a bound that grows with the diagonal block's condition number.
// Two workgroups per CU (at most 256 registers a lane, LDS for two): with
// one, only 256 of the 563 blocks of a 36,000-row solve are resident and
// the later ones start too late to clear their backlog of products before
// the chain reaches them.
template <typename T, bool kForward, bool kInv>
__global__ __launch_bounds__(256, 2) void k_trsv_chain(const T *__restrict__ a, int64_t lda, int n,
                                                    T *x, int32_t *ctl, T *pub) {
  constexpr int kWRows = kInv ? kNB : 1;
  __shared__ T L[kNB][kNB + 1];
  __shared__ T W[kWRows][kNB + 1];
  __shared__ T part[4][kNB];
  __shared__ T rhs[kNB];
  __shared__ T rinv[kNB];
  __shared__ int s_ticket;
  const int nblk = (n + kNB - 1) / kNB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_ticket = atomicAdd(&ctl[0], 1);
  __syncthreads();
  const int ticket = s_ticket;
  if (ticket >= nblk) return;
  const int b = kForward ? ticket : nblk - 1 - ticket;
  const int r0 = b * kNB, bs = min(kNB, n - r0);
  for (int idx = tid; idx < bs * bs; idx += 256) {
    const int r = idx / bs, c = idx % bs;
    L[r][c] = (c <= r) ? a[(int64_t)(r0 + r) * lda + r0 + c] : T(0);
  }
  // the block's own right-hand side, and the reciprocals of its pivots (the
  // substitution multiplies by them instead of dividing), before any wait
  const T xown = tid < bs ? x[r0 + tid] : T(0);
  __syncthreads();
  if (tid < bs) rinv[tid] = T(1) / L[tid][tid];
  const bool use_inv = kInv && bs == kNB;
  if (use_inv) {
    __syncthreads();  // rinv
    if (wave == 0) {
      // lane j: column j of W by right-looking substitution of e_j (the
      // entries above the diagonal stay zero)
      T w[kNB];
#pragma unroll
      for (int r = 0; r < kNB; ++r) w[r] = r == lane ? T(1) : T(0);
#pragma unroll
      for (int c = 0; c < kNB; ++c) {
        const T wc = w[c] * rinv[c];
        w[c] = wc;
#pragma unroll
        for (int r = c + 1; r < kNB; ++r) w[r] = fma(-L[r][c], wc, w[r]);
      }
#pragma unroll
      for (int r = 0; r < kNB; ++r) W[kInv ? r : 0][lane] = w[r];
    }
  }
  // The products with the solved blocks: the tile of block j is loaded one
  // step ahead (it does not depend on x), so each step waits only for the
  // flag and the 64 values of x_j.
  T ca[16], cb[16];
  if (kForward) {
    // rows wave*16 .. wave*16+15 of the block, lanes over block j's columns
    T acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = T(0);
    auto load_tile = [&](int j, T *dst) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int r = wave * 16 + i;
        dst[i] = r < bs ? a[(int64_t)(r0 + r) * lda + j * kNB + lane] : T(0);
      }
    };
    auto step = [&](int j, const T *t) {
      const T xv = await_value(pub, j * kNB + lane);  // blocks before the last are full
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] = fma(t[i], xv, acc[i]);
    };
    if (b > 0) load_tile(0, ca);
    int j = 0;
    for (; j + 1 < b; j += 2) {
      load_tile(j + 1, cb);
      step(j, ca);
      if (j + 2 < b) load_tile(j + 2, ca);
      step(j + 1, cb);
    }
    if (j < b) step(j, ca);
    // the 16 row sums over the wave's lanes by a transposing butterfly: at
    // exchange distance 32, 16, 8, 4 a lane keeps half of its partial rows
    // and adds its partner's copy of them (8 + 4 + 2 + 1 exchanges), then
    // two plain exchanges; lane l ends with row (l / 4) % 16
    const bool up32 = lane & 32, up16 = lane & 16, up8 = lane & 8, up4 = lane & 4;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      acc[k] = (up32 ? acc[k + 8] : acc[k]) + __shfl_xor(up32 ? acc[k] : acc[k + 8], 32, 64);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      acc[k] = (up16 ? acc[k + 4] : acc[k]) + __shfl_xor(up16 ? acc[k] : acc[k + 4], 16, 64);
#pragma unroll
    for (int k = 0; k < 2; ++k)
      acc[k] = (up8 ? acc[k + 2] : acc[k]) + __shfl_xor(up8 ? acc[k] : acc[k + 2], 8, 64);
    T rsum = (up4 ? acc[1] : acc[0]) + __shfl_xor(up4 ? acc[0] : acc[1], 4, 64);
    rsum += __shfl_xor(rsum, 2, 64);
    rsum += __shfl_xor(rsum, 1, 64);
    if ((lane & 3) == 0) part[0][wave * 16 + ((lane >> 2) & 15)] = rsum;
    __syncthreads();
    if (tid < bs) rhs[tid] = xown - part[0][tid];
  } else {
    // lanes over the block's columns, waves over the rows of each solved
    // block (row wave + 4 i of block j), in ascending order
    T acc = T(0);
    auto load_tile = [&](int j, T *dst) {
      const int jb = min(kNB, n - j * kNB);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int tt = wave + 4 * i;
        dst[i] = (lane < bs && tt < jb) ? a[(int64_t)(j * kNB + tt) * lda + r0 + lane] : T(0);
      }
    };
    auto step = [&](int j, const T *t) {
      const int jb = min(kNB, n - j * kNB);
      const T xl = lane < jb ? await_value(pub, j * kNB + lane) : T(0);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int tt = wave + 4 * i;
        const T xv = __shfl(xl, tt, 64);
        if (tt < jb) acc = fma(t[i], xv, acc);
      }
    };
    if (b < nblk - 1) load_tile(nblk - 1, ca);
    int j = nblk - 1;
    for (; j - 1 > b; j -= 2) {
      load_tile(j - 1, cb);
      step(j, ca);
      if (j - 2 > b) load_tile(j - 2, ca);
      step(j - 1, cb);
    }
    if (j > b) step(j, ca);
    part[wave][lane] = acc;
    __syncthreads();
    if (tid < bs) rhs[tid] = xown - (((part[0][tid] + part[1][tid]) + part[2][tid]) + part[3][tid]);
  }
  __syncthreads();
  if (use_inv) {
    // x = W rhs (forward) or W^T rhs (backward): thread (row lane, quarter
    // wave) sums 16 columns, the quarters are added in order
    T sp = T(0);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int c = 16 * wave + k;
      sp = fma(kForward ? W[kInv ? lane : 0][c] : W[kInv ? c : 0][lane], rhs[c], sp);
    }
    part[wave][lane] = sp;
    __syncthreads();
    if (wave == 0) {
      const T v = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
      x[r0 + lane] = v;
      publish(pub, r0 + lane, v);
    }
    return;
  }
  if (wave == 0) {
    // the block's substitution in one wave: lane r holds row r; each pivot
    // value is broadcast with readlane (a register read, no LDS round trip)
    T v = lane < bs ? rhs[lane] : T(0);
    if (bs == kNB) {
      // full block: fully unrolled, the factor entries this lane needs
      // (forward L[lane][c], backward L[c][lane]) read from LDS 16 steps
      // ahead, so a step's chain is readlane, multiply, fma; lane c scales
      // its own value by its own reciprocal pivot.  Entries outside the
      // triangle are zero in L, so a lane the step does not update keeps
      // its value; the select keeps it exact for any x_c.
      const T myr = rinv[lane];
      T cur[16], nxt[16];
      auto fetch = [&](int q, T *dst) {
#pragma unroll
        for (int k = 0; k < 16; ++k)
          dst[k] = kForward ? L[lane][16 * q + k] : L[kNB - 1 - 16 * q - k][lane];
      };
      fetch(0, cur);
#pragma unroll
      for (int q = 0; q < kNB / 16; ++q) {
        if (q + 1 < kNB / 16) fetch(q + 1, nxt);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          const int c = kForward ? 16 * q + k : kNB - 1 - 16 * q - k;
          const T xc = lane_value(v * myr, c);
          const T upd = fma(-cur[k], xc, v);
          v = lane == c ? xc : ((kForward ? lane > c : lane < c) ? upd : v);
        }
        if (q + 1 < kNB / 16) {
#pragma unroll
          for (int k = 0; k < 16; ++k) cur[k] = nxt[k];
        }
      }
    } else if (kForward) {
      T lnext = L[lane][0];
      for (int c = 0; c < bs; ++c) {
        const T lc = lnext;
        if (c + 1 < bs) lnext = L[lane][c + 1];
        const T xc = lane_value(v, c) * rinv[c];
        if (lane == c) v = xc;
        if (lane > c) v -= lc * xc;
      }
    } else {
      T lnext = L[bs - 1][lane];
      for (int c = bs - 1; c >= 0; --c) {
        const T lc = lnext;
        if (c > 0) lnext = L[c - 1][lane];
        const T xc = lane_value(v, c) * rinv[c];
        if (lane == c) v = xc;
        if (lane < c) v -= lc * xc;
      }
    }
    if (lane < bs) {
      x[r0 + lane] = v;
      publish(pub, r0 + lane, v);
    }
  }
}

template <typename T>
size_t potrs_ws(int n, int nrhs) {
  const size_t blocked = WsCarver::need(sizeof(T) * (size_t)ceil_div(n > 0 ? n : 1, kChunk) * kNB *
                                        (nrhs > 0 ? nrhs : 1));
  // ticket counters, then the publish buffer of solved values
  const size_t chain = WsCarver::need(sizeof(int32_t) * 2) + WsCarver::need(sizeof(T) * (size_t)(n > 0 ? n : 1));
  return blocked > chain ? blocked : chain;
}

// inv_diag: the one-launch solve applies full diagonal blocks as their
// inverses (always for FP32; for FP64 only the correction solve).
template <typename T>
int potrs_impl(const T *a, int64_t lda, int n, T *x, int64_t ldx, int nrhs, void *ws,
               hipStream_t st, bool inv_diag = sizeof(T) == 4) {
  const int nblk = ceil_div(n, kNB);
  if (nrhs == 1) {  // one launch per direction
    WsCarver w(ws, potrs_ws<T>(n, nrhs));
    int32_t *ctl = w.take<int32_t>(2);
    T *pub = w.take<T>((size_t)n);
    if (!ctl || !pub) return -7;
    const unsigned g = (unsigned)ceil_div(n, 256);
    hipLaunchKernelGGL(k_trsv_prep<T>, dim3(g), dim3(256), 0, st, ctl, pub, n);
    FMLH_LAUNCH_CHECK("k_trsv_prep");
    if (inv_diag)
      hipLaunchKernelGGL((k_trsv_chain<T, true, true>), dim3(nblk), dim3(256), 0, st, a, lda, n, x,
                         ctl, pub);
    else
      hipLaunchKernelGGL((k_trsv_chain<T, true, false>), dim3(nblk), dim3(256), 0, st, a, lda, n,
                         x, ctl, pub);
    FMLH_LAUNCH_CHECK("k_trsv_chain<fwd>");
    hipLaunchKernelGGL(k_trsv_prep<T>, dim3(g), dim3(256), 0, st, ctl + 1, pub, n);
    FMLH_LAUNCH_CHECK("k_trsv_prep");
    if (inv_diag)
      hipLaunchKernelGGL((k_trsv_chain<T, false, true>), dim3(nblk), dim3(256), 0, st, a, lda, n,
                         x, ctl + 1, pub);
    else
      hipLaunchKernelGGL((k_trsv_chain<T, false, false>), dim3(nblk), dim3(256), 0, st, a, lda, n,
                         x, ctl + 1, pub);
    FMLH_LAUNCH_CHECK("k_trsv_chain<bwd>");
    return 0;
  }
  T *part = static_cast<T *>(ws);
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

// ---- the allocation-free solve ----------------------------------------
//
// fmlh_potrs_noalloc takes its scratch (the chained solve's ready flags and
// tickets, or the blocked solve's partial products) from this file-scope
// device array instead of a caller workspace.  Host code never passes the
// array's name to a kernel: on each device the address comes from
// hipGetSymbolAddress, taken once per device (every device holds its own
// copy of the array) and kept below.
constexpr size_t kStaticScratchBytes = size_t(1) << 20;
__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The address of this device's copy of scratch (nullptr and the HIP error
// recorded on failure).
void *device_scratch(int &rc) {
  rc = 0;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    rc = record_error(e, "hipGetDevice(potrs scratch)");
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    rc = record_message(FMLH_EHIP, "fmlh_potrs_noalloc: device index beyond the library's table");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      rc = record_error(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace

namespace fmlh {
// Takes this device's scratch address ahead of the first solve (fmlh_init).
int potrs_scratch_init() {
  int rc = 0;
  device_scratch(rc);
  return rc;
}
}  // namespace fmlh

extern "C" {

size_t fmlh_potrs_noalloc_max_bytes(void) { return kStaticScratchBytes; }

int fmlh_potrs_noalloc(const double *a, int64_t lda, int n, double *x, int64_t ldx, int nrhs,
                       void *stream) {
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(lda >= (n > 1 ? n : 1), 2);
  FMLH_ARG_CHECK(nrhs >= 0, 6);
  FMLH_ARG_CHECK(ldx >= (n > 1 ? n : 1), 5);
  if (n == 0 || nrhs == 0) return 0;
  FMLH_ARG_CHECK(a != nullptr, 1);
  FMLH_ARG_CHECK(x != nullptr, 4);
  const size_t need = potrs_ws<double>(n, nrhs);
  if (need > kStaticScratchBytes) {
    char msg[200];
    snprintf(msg, sizeof(msg),
             "fmlh_potrs_noalloc: n = %d with nrhs = %d needs %zu bytes of scratch, more than its "
             "%zu-byte static array (use fmlh_potrs with a workspace)",
             n, nrhs, need, kStaticScratchBytes);
    return record_message(-3, msg);
  }
  int rc = 0;
  void *ws = device_scratch(rc);
  if (!ws) return rc;
  return potrs_impl<double>(a, lda, n, x, ldx, nrhs, ws, as_stream(stream));
}

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

size_t fmlh_potrf_ozaki_ws_bytes(int n) { return oz_ws_bytes(n, kOzSlices); }

int fmlh_potrf_ozaki(double *a, int64_t lda, int n, int32_t *info, void *ws, size_t ws_bytes,
                     void *stream) {
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(lda >= (n > 1 ? n : 1), 2);
  FMLH_ARG_CHECK(info != nullptr, 4);
  if (n > 0) FMLH_ARG_CHECK(a != nullptr, 1);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= oz_ws_bytes(n, kOzSlices), 5);
  if (n == 0) {
    hipLaunchKernelGGL(k_set_int_impl, dim3(1), dim3(1), 0, as_stream(stream), info, 0);
    FMLH_LAUNCH_CHECK("k_set_int");
    return 0;
  }
  timing_begin(as_stream(stream));
  int rc = potrf_impl<double>(a, lda, n, info, as_stream(stream), ws);
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

// X^T (n x n) and the split-k partials of the largest block row:
// ceil(n/64) tiles x ceil(n/2048) chunks of 64 x 64.
size_t fmlh_potri_ws_bytes(int n) {
  if (n <= 0) return 0;
  const size_t tiles = (size_t)ceil_div(n, kTile), chunks = (size_t)ceil_div(n, kPotriKC);
  return WsCarver::need(sizeof(double) * (size_t)n * (size_t)n) +
         WsCarver::need(sizeof(double) * tiles * chunks * kTile * kTile);
}

int fmlh_potri(double *a, int64_t lda, int n, int32_t *info, void *ws, size_t ws_bytes,
               void *stream) {
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(lda >= (n > 1 ? n : 1), 2);
  FMLH_ARG_CHECK(info != nullptr, 4);
  if (n == 0) {
    hipLaunchKernelGGL(k_set_int_impl, dim3(1), dim3(1), 0, as_stream(stream), info, 0);
    FMLH_LAUNCH_CHECK("k_set_int");
    return 0;
  }
  FMLH_ARG_CHECK(a != nullptr, 1);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= fmlh_potri_ws_bytes(n), 5);
  timing_begin(as_stream(stream));
  WsCarver wc(ws, ws_bytes);
  double *xt = wc.take<double>((size_t)n * (size_t)n);
  double *part = wc.take<double>((size_t)ceil_div(n, kTile) * ceil_div(n, kPotriKC) * kTile * kTile);
  if (xt == nullptr || part == nullptr) return -5;
  const int rc = potri_impl(a, lda, n, info, xt, part, as_stream(stream));
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
  timing_begin(as_stream(stream));
  const int rc = potrs_impl<double>(a, lda, n, x, ldx, nrhs, ws, as_stream(stream));
  timing_end(as_stream(stream));
  return rc;
}

int fmlh_potrs_correction(const double *a, int64_t lda, int n, double *x, void *ws,
                          size_t ws_bytes, void *stream) {
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(lda >= (n > 1 ? n : 1), 2);
  if (n == 0) return 0;
  FMLH_ARG_CHECK(a != nullptr, 1);
  FMLH_ARG_CHECK(x != nullptr, 4);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= potrs_ws<double>(n, 1), 5);
  timing_begin(as_stream(stream));
  const int rc = potrs_impl<double>(a, lda, n, x, n, 1, ws, as_stream(stream), true);
  timing_end(as_stream(stream));
  return rc;
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
  timing_begin(as_stream(stream));
  const int rc = potrs_impl<float>(a, lda, n, x, ldx, nrhs, ws, as_stream(stream));
  timing_end(as_stream(stream));
  return rc;
}

}  // extern "C"
