// "NT" MFMA tiles: C[m][n] = sum_k P[m][k] * Q[n][k] for row-major P (rows m,
// k contiguous) and Q (rows n, k contiguous), 256 threads per tile.
//
// Tile MT x MT with MT = 64 or 128.  Waves are laid out 2x2; each wave owns
// an (MT/2) x (MT/2) sub-tile built from NB x NB 16x16 MFMA blocks
// (NB = MT/32; v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).  Operand
// tiles of BK columns (16; 8 for the FP64 128 tile, which keeps its two LDS
// buffers at 37 KB so four workgroups share a CU) are staged global ->
// registers -> LDS with two LDS buffers (one barrier per K step).  LDS rows
// are padded to BK + 1 elements so the 16 rows a 16-lane group reads fall in
// distinct banks.  The 128 tile halves the operand traffic per flop of the
// 64 tile; it is the shape for large kernels and trailing updates, the 64
// tile for small ones.
//
// MFMA 16x16x4 lane layout on gfx950 (single-block form; checked on the
// device with an in-bounds probe):
//   A operand  lane l -> A[l%16][l/16]          (16 x 4)
//   B operand  lane l -> B[l/16][l%16]          (4 x 16)
//   D (f32)    lane l, reg r -> D[4*(l/16) + r][l%16]
//   D (f64)    lane l, reg r -> D[l/16 + 4*r][l%16]
#pragma once

#include <math.h>

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

constexpr int kTile = 64;  // the small tile (also the Cholesky block size)
constexpr int kTileThreads = 256;

template <typename T, int MT>
constexpr int tile_bk() { return (sizeof(T) == 8 && MT == 128) ? 8 : 16; }

template <typename T, int MT>
struct TileSmem {
  static constexpr int BK = tile_bk<T, MT>();
  static constexpr int LS = BK + 1;  // padded LDS row
  T p[2][MT * LS];
  T q[2][MT * LS];
  T pn[MT];  // optional squared row norms of the P tile
  T qn[MT];  // optional squared row norms of the Q tile
};

// Result of one tile: v[im][in][r] is C at m = tile_row<T, MT, WR>(im, r),
// n = tile_col<MT>(in).  WR: wave rows (2: 2 x 2 waves, 256 threads, each
// wave (MT/2) x (MT/2); 4: 4 x 2 waves, 512 threads, each wave
// (MT/4) x (MT/2) -- half the accumulator registers per wave, so twice the
// waves per SIMD).
template <typename T, int MT, int WR = 2>
struct TileAcc {
  static constexpr int NBM = MT / (16 * WR);
  static constexpr int NBN = MT / 32;
  typename MfmaOp<T>::acc_t v[NBM][NBN];
};

__device__ __forceinline__ int tile_lane() { return threadIdx.x & 63; }
__device__ __forceinline__ int tile_wm() { return (threadIdx.x >> 6) >> 1; }
__device__ __forceinline__ int tile_wn() { return (threadIdx.x >> 6) & 1; }
template <typename T, int MT = kTile, int WR = 2>
__device__ __forceinline__ int tile_row(int im, int r) {
  const int q = tile_lane() >> 4;
  return tile_wm() * (MT / WR) + im * 16 + (sizeof(T) == 8 ? q + 4 * r : 4 * q + r);
}
template <int MT = kTile>
__device__ __forceinline__ int tile_col(int in) {
  return tile_wn() * (MT / 2) + in * 16 + (tile_lane() & 15);
}

// Computes the MT x MT tile for P rows [0, mvalid) and Q rows [0, nvalid)
// (pointers already offset to the tile origin and to column kbeg), over
// columns [0, klen).  Out-of-range rows/columns read as zero.  When kNorms
// is set, s.pn / s.qn receive sequential fused sums of squares of the staged
// rows (ascending k), the same for either tile size.
template <typename T, bool kNorms, int MT = kTile, int WR = 2>
__device__ __forceinline__ void tile_nt(const T *__restrict__ P, int64_t ldp, int mvalid,
                                        const T *__restrict__ Q, int64_t ldq, int nvalid,
                                        int klen, TileSmem<T, MT> &s, TileAcc<T, MT, WR> &acc) {
  typedef MfmaOp<T> Op;
  constexpr int NBM = TileAcc<T, MT, WR>::NBM;  // MFMA blocks per wave, rows
  constexpr int NBN = TileAcc<T, MT, WR>::NBN;  // and columns
  constexpr int kThreads = 128 * WR;
  constexpr int kBK = TileSmem<T, MT>::BK;
  constexpr int kLdsStride = TileSmem<T, MT>::LS;
  constexpr int EPT = MT * kBK / kThreads;  // staged elements per thread and operand
  constexpr int TPR = kBK / EPT;        // threads per staged row
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wm = tile_wm(), wn = tile_wn();
  const int lrow = tid / TPR;
  const int lcol = (tid % TPR) * EPT;

#pragma unroll
  for (int i = 0; i < NBM; ++i)
#pragma unroll
    for (int j = 0; j < NBN; ++j) acc.v[i][j] = Op::zero();

  T pn_acc = T(0), qn_acc = T(0);
  T rp[EPT], rq[EPT];
  const bool prow_ok = lrow < mvalid;
  const bool qrow_ok = lrow < nvalid;
  const T *prow = P + (int64_t)lrow * ldp;
  const T *qrow = Q + (int64_t)lrow * ldq;

  auto gload = [&](int k0) {
#pragma unroll
    for (int c = 0; c < EPT; ++c) {
      const int k = k0 + lcol + c;
      rp[c] = (prow_ok && k < klen) ? prow[k] : T(0);
      rq[c] = (qrow_ok && k < klen) ? qrow[k] : T(0);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int c = 0; c < EPT; ++c) {
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
      T a[NBM], b[NBN];
#pragma unroll
      for (int i = 0; i < NBM; ++i) a[i] = sp[(wm * (MT / WR) + i * 16 + (lane & 15)) * kLdsStride + kk];
#pragma unroll
      for (int j = 0; j < NBN; ++j) b[j] = sq[(wn * (MT / 2) + j * 16 + (lane & 15)) * kLdsStride + kk];
#pragma unroll
      for (int i = 0; i < NBM; ++i)
#pragma unroll
        for (int j = 0; j < NBN; ++j) acc.v[i][j] = Op::mma(a[i], b[j], acc.v[i][j]);
    }
    if (kNorms) {
      if (tid < MT) {
#pragma unroll
        for (int k = 0; k < kBK; ++k) {
          T v = sp[tid * kLdsStride + k];
          pn_acc = fma(v, v, pn_acc);
        }
      } else if (tid < 2 * MT) {
#pragma unroll
        for (int k = 0; k < kBK; ++k) {
          T v = sq[(tid - MT) * kLdsStride + k];
          qn_acc = fma(v, v, qn_acc);
        }
      }
    }
    if (kt + 1 < nk) lstore(buf ^ 1);
    __syncthreads();
  }
  if (kNorms) {
    if (tid < MT)
      s.pn[tid] = pn_acc;
    else if (tid < 2 * MT)
      s.qn[tid - MT] = qn_acc;
    __syncthreads();
  }
}

// LDS swizzle for tiles staged as 64-byte rows read in 16-byte chunks: chunk
// c of row r is stored at chunk c ^ lds_swizzle(r).  A row-wise staging
// write covers whole rows (consecutive banks), and the fragment read of 16
// consecutive rows at one logical chunk then touches 16 distinct 4-bank
// groups for 64 banks (and each aligned 8 rows distinct groups for 32
// banks), where an unswizzled or padded row stride repeats a bank group.
__host__ __device__ inline int lds_swizzle(int r) { return ((r >> 1) ^ (r >> 2)) & 3; }

// XCD-aware tile order.  Workgroups are dispatched round-robin over the
// chip's 8 XCDs (workgroup L runs on XCD L % 8), each XCD with its own L2.
// The grid is padded to a multiple of 8; workgroup L takes tile
// (L % 8) * per + L / 8 of a grouped raster order, so the workgroups of one
// XCD cover a contiguous run of tiles that share operand tiles, and within
// the run kGroup consecutive tile rows are walked column by column.  Returns
// false for the padding workgroups (their tile index is past the end).
constexpr int kXcds = 8;
constexpr int kGroup = 8;

__host__ __device__ inline int64_t xcd_grid(int gx, int gy) {
  const int64_t t = (int64_t)gx * gy;
  return ((t + kXcds - 1) / kXcds) * kXcds;
}

__device__ __forceinline__ bool xcd_tile(int gx, int gy, int &tx, int &ty) {
  const int64_t t = (int64_t)gx * gy;
  const int64_t per = (t + kXcds - 1) / kXcds;
  const int64_t L = blockIdx.x;
  const int64_t local = (L % kXcds) * per + L / kXcds;
  if (local >= t) return false;
  const int64_t span = (int64_t)kGroup * gx;
  const int64_t g = local / span;
  const int first = (int)(g * kGroup);
  const int rows = min(kGroup, gy - first);
  const int64_t in = local - g * span;
  ty = first + (int)(in % rows);
  tx = (int)(in / rows);
  return true;
}

// Tile (tx, ty), tx <= ty < g, of the lower triangle of a g x g tile grid
// at position `local` of this order: groups of kGroup tile rows, each group
// walked column by column, so a group's row tiles stay in L2 while its
// column tiles stream past once (the row-major triangle would stream every
// column tile once per row).  Returns false past the end.
__host__ __device__ inline bool tri_tile(int g, int64_t local, int &tx, int &ty) {
  if (local >= (int64_t)g * (g + 1) / 2) return false;
  // row of `local` in row-major order: T(r) <= local < T(r + 1), T(r) = r(r+1)/2
  int r = (int)((sqrt(8.0 * (double)local + 1.0) - 1.0) * 0.5);
  while ((int64_t)(r + 1) * (r + 2) / 2 <= local) ++r;
  while ((int64_t)r * (r + 1) / 2 > local) --r;
  const int r0 = (r / kGroup) * kGroup;  // first row of the group (it starts at T(r0))
  const int R = kGroup < g - r0 ? kGroup : g - r0;
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

}  // namespace fmlh
