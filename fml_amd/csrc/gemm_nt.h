// 64x64 "NT" MFMA tile: C[m][n] = sum_k P[m][k] * Q[n][k] for row-major
// P (rows m, k contiguous) and Q (rows n, k contiguous), 256 threads.
//
// Waves are laid out 2x2; each wave owns a 32x32 sub-tile built from 2x2
// 16x16 MFMA blocks (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).
// Operand tiles of BK=16 columns are staged global -> registers -> LDS with
// two LDS buffers (one barrier per K step).  LDS rows are padded to 17
// elements so the 16 rows a 16-lane group reads fall in distinct banks.
//
// MFMA 16x16x4 lane layout on gfx950 (single-block form; checked on the
// device with an in-bounds probe):
//   A operand  lane l -> A[l%16][l/16]          (16 x 4)
//   B operand  lane l -> B[l/16][l%16]          (4 x 16)
//   D (f32)    lane l, reg r -> D[4*(l/16) + r][l%16]
//   D (f64)    lane l, reg r -> D[l/16 + 4*r][l%16]
#pragma once

#include "common.h"

namespace fmlh {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MfmaOp;

template <>
struct MfmaOp<double> {
  typedef dbl4 acc_t;
  static __device__ __forceinline__ acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

template <>
struct MfmaOp<float> {
  typedef flt4 acc_t;
  static __device__ __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

constexpr int kTile = 64;
constexpr int kBK = 16;
constexpr int kLdsStride = kBK + 1;
constexpr int kTileThreads = 256;

template <typename T>
struct TileSmem {
  T p[2][kTile * kLdsStride];
  T q[2][kTile * kLdsStride];
  T pn[kTile];  // optional squared row norms of the P tile
  T qn[kTile];  // optional squared row norms of the Q tile
};

// Result of one tile: acc[im][in][r] is C at m = tile_row<T>(im, r),
// n = tile_col(in).
template <typename T>
struct TileAcc {
  typename MfmaOp<T>::acc_t v[2][2];
};

__device__ __forceinline__ int tile_lane() { return threadIdx.x & 63; }
__device__ __forceinline__ int tile_wm() { return (threadIdx.x >> 6) >> 1; }
__device__ __forceinline__ int tile_wn() { return (threadIdx.x >> 6) & 1; }
template <typename T>
__device__ __forceinline__ int tile_row(int im, int r) {
  const int q = tile_lane() >> 4;
  return tile_wm() * 32 + im * 16 + (sizeof(T) == 8 ? q + 4 * r : 4 * q + r);
}
__device__ __forceinline__ int tile_col(int in) {
  return tile_wn() * 32 + in * 16 + (tile_lane() & 15);
}

// Computes the 64x64 tile for P rows [0, mvalid) and Q rows [0, nvalid)
// (pointers already offset to the tile origin and to column kbeg), over
// columns [0, klen).  Out-of-range rows/columns read as zero.  When
// kNorms is set, s.pn / s.qn receive sequential sums of squares of the
// staged rows (ascending k).
template <typename T, bool kNorms>
__device__ __forceinline__ void tile_nt(const T *__restrict__ P, int64_t ldp, int mvalid,
                                        const T *__restrict__ Q, int64_t ldq, int nvalid,
                                        int klen, TileSmem<T> &s, TileAcc<T> &acc) {
  typedef MfmaOp<T> Op;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wm = tile_wm(), wn = tile_wn();
  const int lrow = tid >> 2;        // 0..63
  const int lcol = (tid & 3) * 4;   // 0,4,8,12

  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc.v[i][j] = Op::zero();

  T pn_acc = T(0), qn_acc = T(0);
  T rp[4], rq[4];
  const bool prow_ok = lrow < mvalid;
  const bool qrow_ok = lrow < nvalid;
  const T *prow = P + (int64_t)lrow * ldp;
  const T *qrow = Q + (int64_t)lrow * ldq;

  auto gload = [&](int k0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int k = k0 + lcol + c;
      rp[c] = (prow_ok && k < klen) ? prow[k] : T(0);
      rq[c] = (qrow_ok && k < klen) ? qrow[k] : T(0);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      s.p[buf][lrow * kLdsStride + lcol + c] = rp[c];
      s.q[buf][lrow * kLdsStride + lcol + c] = rq[c];
    }
  };

  const int nk = (klen + kBK - 1) / kBK;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * kBK);
    const T *sp = s.p[buf];
    const T *sq = s.q[buf];
#pragma unroll
    for (int kq = 0; kq < kBK / 4; ++kq) {
      const int kk = kq * 4 + (lane >> 4);
      T a0 = sp[(wm * 32 + 0 + (lane & 15)) * kLdsStride + kk];
      T a1 = sp[(wm * 32 + 16 + (lane & 15)) * kLdsStride + kk];
      T b0 = sq[(wn * 32 + 0 + (lane & 15)) * kLdsStride + kk];
      T b1 = sq[(wn * 32 + 16 + (lane & 15)) * kLdsStride + kk];
      acc.v[0][0] = Op::mma(a0, b0, acc.v[0][0]);
      acc.v[0][1] = Op::mma(a0, b1, acc.v[0][1]);
      acc.v[1][0] = Op::mma(a1, b0, acc.v[1][0]);
      acc.v[1][1] = Op::mma(a1, b1, acc.v[1][1]);
    }
    if (kNorms) {
      if (tid < 64) {
#pragma unroll
        for (int k = 0; k < kBK; ++k) {
          T v = sp[tid * kLdsStride + k];
          pn_acc = pn_acc + v * v;
        }
      } else if (tid < 128) {
#pragma unroll
        for (int k = 0; k < kBK; ++k) {
          T v = sq[(tid - 64) * kLdsStride + k];
          qn_acc = qn_acc + v * v;
        }
      }
    }
    if (kt + 1 < nk) lstore(buf ^ 1);
    __syncthreads();
  }
  if (kNorms) {
    if (tid < 64)
      s.pn[tid] = pn_acc;
    else if (tid < 128)
      s.qn[tid - 64] = qn_acc;
    __syncthreads();
  }
}

}  // namespace fmlh
