// Ozaki-split FP64 updates on INT8 MFMA for the Cholesky factorisation
// (the slicing and its error bound: ozaki.h).
//
// k_oz_split: one workgroup of 256 threads per row; each thread holds four
// consecutive panel entries, the row maximum comes from a wave butterfly and
// an LDS exchange, e_r from frexp, and each entry is peeled into `S` signed
// 7-bit slices by exact scaling and truncation: R = M 2^-e_r in (-1, 1),
// then S times R <- 128 R, q = trunc(R), R <- R - q.  Each step is exact
// (a power-of-two scaling, a truncation, and a subtraction of the integer
// part), so M = 2^e_r (sum_t q_t 2^-7t + R_S 2^-7S) holds exactly.
//
// k_oz_update: 64 x 64 output tiles, 256 threads as 2 x 2 waves of 2 x 2
// blocks of v_mfma_i32_16x16x64_i8 (230 registers: two waves per SIMD, the
// fewest with which this MFMA reaches its sustained rate).  A k step stages the S planes of the
// tile's 64 rows and 64 columns in LDS, 64 bytes a row with the four 16-byte
// chunks of row r permuted by c -> c ^ swz(r) (lds_swizzle, gemm_nt.h), so
// neither the row-wise staging writes nor the 16-byte fragment reads of 16
// rows collide in a bank; the next step's planes are loaded into registers
// meanwhile.  Products q_t q_u
// with t + u <= S + 1 go to one int32 accumulator per level l = t + u
// (S(S+1)/2 MFMAs per block and step; each level sums at most S products
// of 1024 terms of |q_t q_u| <= 127^2, below 2^31).  The epilogue forms
// sum_l acc_l 2^-7l in FP64 (each term exact, smallest level first) and
// subtracts it, scaled by 2^(e_r + e_c), from the FP64 entry.  Dropped
// terms (t + u > S + 1 and the remainders) are below about
// (S + 2) 2^-7S 2^(e_r + e_c) per product, summed over the panel.
//
// int8 16x16x64 MFMA lane layout on gfx950 (checked with an in-bounds
// probe):  A lane l -> A[l%16][16(l/16) + e], B lane l -> B[16(l/16) + e][l%16]
// (e = 0..15, byte e of the lane's 16), D lane l, reg r -> D[4(l/16) + r][l%16].
#include <math.h>

#include "common.h"
#include "gemm_nt.h"
#include "ozaki.h"

namespace {

using namespace fmlh;

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kOzT = 64;    // tile edge
constexpr int kOzK = 64;    // k step (bytes of a slice row per step)
constexpr int kOzRow = 64;  // LDS bytes per staged row

__host__ __device__ inline size_t oz_plane_bytes(int nrows) {
  return (size_t)nrows * kOzPanel;
}

template <int S>
__global__ __launch_bounds__(256) void k_oz_split(const double *__restrict__ a, int64_t lda,
                                                  int row0, int kb, int kw,
                                                  int8_t *__restrict__ planes, int64_t plane,
                                                  int32_t *__restrict__ expo,
                                                  const int32_t *__restrict__ info) {
  if (*info != 0) return;
  __shared__ double wmax[4];
  const int r = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double *row = a + (int64_t)(row0 + r) * lda + kb;
  const int k0 = tid * 4;
  double v[4];
  double m = 0.0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    v[e] = k0 + e < kw ? row[k0 + e] : 0.0;
    m = fmax(m, fabs(v[e]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if (lane == 0) wmax[wave] = m;
  __syncthreads();
  m = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
  int ex = 0;
  if (m > 0.0) (void)frexp(m, &ex);  // m = f 2^ex, f in [0.5, 1): |v| < 2^ex
  if (tid == 0) expo[r] = ex;
  double rem[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) rem[e] = ldexp(v[e], -ex);
#pragma unroll
  for (int t = 0; t < S; ++t) {
    uint32_t packed = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double x = rem[e] * 128.0;
      const double q = trunc(x);
      rem[e] = x - q;
      packed |= (uint32_t)(uint8_t)(int8_t)(int)q << (8 * e);
    }
    *reinterpret_cast<uint32_t *>(planes + t * plane + (int64_t)r * kOzPanel + k0) = packed;
  }
}

template <int S>
struct OzSmem {
  uint8_t a[S][kOzT * kOzRow];
  uint8_t b[S][kOzT * kOzRow];
};

template <int S>
__global__ __launch_bounds__(256, 2) void k_oz_update(double *__restrict__ c, int64_t lda, int r0,
                                                      int r1, int c0, int c1, int srow0,
                                                      const int8_t *__restrict__ planes,
                                                      int64_t plane,
                                                      const int32_t *__restrict__ expo, int kp,
                                                      int gx, int gy,
                                                      const int32_t *__restrict__ info) {
  if (*info != 0) return;
  __shared__ __attribute__((aligned(16))) OzSmem<S> s;
  int tx, ty;
  if (!xcd_tile(gx, gy, tx, ty)) return;
  const int tr = r0 + ty * kOzT;  // output rows (A operand)
  const int tc = c0 + tx * kOzT;  // output columns (B operand)
  if (tc > tr + kOzT - 1) return;  // tile above the diagonal
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // staging: thread -> (row tid/4, 16-byte chunk tid%4) of every plane
  const int srow = tid >> 2, schunk = (tid & 3) * 16;
  const int ar = min(tr + srow, r1 - 1) - srow0;
  const int br = min(tc + srow, c1 - 1) - srow0;
  u32x4 ra[S], rb[S];
  auto gload = [&](int k0) {
#pragma unroll
    for (int t = 0; t < S; ++t) {
      ra[t] = *reinterpret_cast<const u32x4 *>(planes + t * plane + (int64_t)ar * kOzPanel + k0 + schunk);
      rb[t] = *reinterpret_cast<const u32x4 *>(planes + t * plane + (int64_t)br * kOzPanel + k0 + schunk);
    }
  };
  auto lstore = [&]() {
    const int o = srow * kOzRow + 16 * ((tid & 3) ^ lds_swizzle(srow));
#pragma unroll
    for (int t = 0; t < S; ++t) {
      *reinterpret_cast<u32x4 *>(&s.a[t][o]) = ra[t];
      *reinterpret_cast<u32x4 *>(&s.b[t][o]) = rb[t];
    }
  };

  i32x4 acc[S][2][2];
#pragma unroll
  for (int l = 0; l < S; ++l)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) acc[l][j][i] = i32x4{0, 0, 0, 0};

  const int nk = kp / kOzK;
  gload(0);
  for (int kt = 0; kt < nk; ++kt) {
    __syncthreads();  // the previous step's readers are done
    lstore();
    __syncthreads();
    if (kt + 1 < nk) gload((kt + 1) * kOzK);
    // fragment of rows r0f .. r0f + 15, chunk lane / 16 (swizzled)
    auto frag_off = [&](int r0f) {
      const int row = r0f + (lane & 15);
      return row * kOzRow + 16 * ((lane >> 4) ^ lds_swizzle(row));
    };
    i32x4 bf[S][2];
#pragma unroll
    for (int u = 0; u < S; ++u)
#pragma unroll
      for (int i = 0; i < 2; ++i)
        bf[u][i] = *reinterpret_cast<const i32x4 *>(&s.b[u][frag_off(wn * 32 + i * 16)]);
#pragma unroll
    for (int t = 0; t < S; ++t) {
      i32x4 af[2];
#pragma unroll
      for (int j = 0; j < 2; ++j)
        af[j] = *reinterpret_cast<const i32x4 *>(&s.a[t][frag_off(wm * 32 + j * 16)]);
#pragma unroll
      for (int u = 0; u + t < S; ++u)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int i = 0; i < 2; ++i)
            acc[t + u][j][i] =
                __builtin_amdgcn_mfma_i32_16x16x64_i8(af[j], bf[u][i], acc[t + u][j][i], 0, 0, 0);
    }
  }
  // epilogue: level l (0-based t + u) carries 2^-7(l + 2)
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int gc = tc + wn * 32 + i * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = tr + wm * 32 + j * 16 + 4 * (lane >> 4) + r;
        if (gr >= r1 || gc >= c1 || gc > gr) continue;
        double v = 0.0;
#pragma unroll
        for (int l = S - 1; l >= 0; --l) v += ldexp((double)acc[l][j][i][r], -7 * (l + 2));
        const int e = expo[gr - srow0] + expo[gc - srow0];
        c[(int64_t)gr * lda + gc] -= ldexp(v, e);
      }
    }
}

template <int S>
int split_launch(const double *a, int64_t lda, int row0, int nrows, int kb, int kw, void *ws,
                 const int32_t *info, hipStream_t st) {
  int8_t *planes = static_cast<int8_t *>(ws);
  const int64_t plane = (int64_t)oz_plane_bytes(nrows);
  int32_t *expo = reinterpret_cast<int32_t *>(static_cast<char *>(ws) + WsCarver::need(S * plane));
  hipLaunchKernelGGL(k_oz_split<S>, dim3(nrows), dim3(256), 0, st, a, lda, row0, kb, kw, planes,
                     plane, expo, info);
  FMLH_LAUNCH_CHECK("k_oz_split");
  return 0;
}

template <int S>
int update_launch(double *a, int64_t lda, int r0, int r1, int c0, int c1, int srow0, int nsplit,
                  const void *ws, const int32_t *info, hipStream_t st) {
  const int8_t *planes = static_cast<const int8_t *>(ws);
  const int64_t plane = (int64_t)oz_plane_bytes(nsplit);
  const int32_t *expo = reinterpret_cast<const int32_t *>(static_cast<const char *>(ws) +
                                                         WsCarver::need(S * plane));
  const int gx = ceil_div(c1 - c0, kOzT), gy = ceil_div(r1 - r0, kOzT);
  hipLaunchKernelGGL(k_oz_update<S>, dim3((unsigned)xcd_grid(gx, gy)), dim3(256), 0, st, a, lda,
                     r0, r1, c0, c1, srow0, planes, plane, expo, kOzPanel, gx, gy, info);
  FMLH_LAUNCH_CHECK("k_oz_update");
  return 0;
}

}  // namespace

namespace fmlh {

size_t oz_ws_bytes(int n, int slices) {
  const size_t rows = (size_t)(n > 0 ? n : 1);
  return WsCarver::need((size_t)slices * oz_plane_bytes((int)rows)) + WsCarver::need(4 * rows);
}

int oz_split(const double *a, int64_t lda, int row0, int nrows, int kb, int kw, int slices,
             void *ws, const int32_t *info, hipStream_t st) {
  if (nrows <= 0) return 0;
  if (kw > kOzPanel) return -6;
  if (slices == 6) return split_launch<6>(a, lda, row0, nrows, kb, kw, ws, info, st);
  return -7;
}

int oz_update(double *a, int64_t lda, int r0, int r1, int c0, int c1, int srow0, int nsplit,
              int slices, const void *ws, bool tri, const int32_t *info, hipStream_t st) {
  (void)tri;  // the rectangular grid's tiles above the diagonal exit at once
  if (r1 <= r0 || c1 <= c0) return 0;
  if (slices == 6) return update_launch<6>(a, lda, r0, r1, c0, c1, srow0, nsplit, ws, info, st);
  return -7;
}

}  // namespace fmlh
