// Ozaki-split FP64 updates on INT8 MFMA for the Cholesky factorisation
// (the slicing and its error bound: ozaki.h).
//
// k_oz_split: a workgroup of 256 threads takes 8 rows.  Each wave finds the
// largest |entry| of two rows (a butterfly over 16 entries a lane) and
// picks the row's exponent e_r; then each thread takes the 32-entry chunks
// c, c + 32, ... of one row (one chunk at the 1024-column panel width): X =
// trunc(M 2^(8S - e_r)) is exact in int64 (|X| < 2^39 for
// S = 5), and its signed base-256 digits come from the low byte, b - 256
// when b >= 128, and an exact shift of X - b.  The chunk's digits of plane t
// are one 32-byte store, and the 8 rows of a chunk are 256 contiguous bytes.
//
// Dropped terms.  With digits bounded by 2^7 and level l = t + u (0-based)
// weighted 2^-8(l+2), the products with t + u >= S have |q_t q_u| 2^-8(l+2)
// <= 2^14 2^-8(S+2) each, summed over the panel: the update of entry
// (r, c) is exact to about S 2^(e_r + e_c) 2^(14 - 8(S+2)) per k (2^-39.7
// for S = 5), plus the truncation of X (2^-40 of the scale per factor).
// Balanced digits make both errors zero-mean; a truncated sign-magnitude
// split with 7-bit slices is biased and needs six slices (21 products) for
// the accuracy five balanced ones (15 products) give.
//
// k_oz_update: 128 x 128 output tiles, 512 threads as 4 x 2 waves, each
// wave 32 rows x 64 columns (two 32 x 32 blocks of v_mfma_i32_32x32x32_i8,
// S levels of int32 accumulators: 16 S registers per block).  A k step is
// 32 bytes of every plane for the tile's 128 rows and 128 columns (8 S KB),
// moved global -> LDS by the DMA form of the load (global_load_lds_dwordx4,
// 16 bytes a lane, the wave's 64 lanes filling 1 KB) into one of three
// stages, two steps ahead: no registers are spent on staging, and a step's
// data has two steps of MFMA work to arrive.  Within a 32-byte LDS row the
// two 16-byte halves of row r are exchanged when bit 2 xor bit 3 of r is
// set, so the 16 (or 8) rows a ds_read_b128 serves at once fall in distinct
// bank groups for 64 (or 32) banks; the DMA realises the permutation by the
// source address each lane loads.  One barrier a step: a wave waits for its
// own loads of the step (s_waitcnt vmcnt leaving the next step's S loads in
// flight), and the barrier makes every wave's loads visible and frees the
// stage the next loads overwrite; a step's loads are issued one per group
// of MFMAs, and each B fragment is read one group ahead of its MFMAs.  The
// epilogue forms sum_l acc_l 2^-8(l+2) in FP64 (each term exact, smallest
// level first) and subtracts it, scaled by 2^(e_r + e_c), from the FP64
// entry.  Measured alone on a 32768 x 32768 lower triangle: 7.5 ms (2.2
// POP/s of tile work); without the loads 5.5 ms, the MFMAs and barriers
// alone 4.1 ms (4.0 POP/s) -- the INT8 MFMA pipe is not the limit.
//
// int8 32x32x32 MFMA lane layout on gfx950 (checked with an in-bounds
// probe): A lane l -> row l % 32, k bytes 16 (l / 32) + e; B lane l ->
// column l % 32, the same k bytes; D lane l, reg r -> row 8 (r / 4) +
// 4 (l / 32) + r % 4, column l % 32.  global_load_lds_dwordx4 writes lane
// l's 16 bytes at the wave's LDS base + 16 l.
#include <math.h>

#include "common.h"
#include "gemm_nt.h"
#include "ozaki.h"

namespace {

using namespace fmlh;

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kOzT = 128;     // tile edge
constexpr int kOzKS = 32;     // k step: bytes of a slice row per MFMA
// LDS stages, loads kOzStages - 1 steps ahead (at 36k: 3 stages 169 ms per
// factorisation, 4 stages 190 ms -- four leave no LDS for the exponents and
// spill more of the epilogue)
constexpr int kOzStages = 3;
// the row and column exponents of a tile are staged in LDS when they fit
constexpr bool kOzExpLds = kOzStages <= 3;
constexpr int kOzSplitRows = 8;

__host__ __device__ inline size_t oz_plane_bytes(int nrows) {
  return (size_t)nrows * kOzPanel;
}

// largest |M| 2^-e for which every S-digit signed base-256 number holds
// trunc(M 2^(8S - e)): (127 (256^S - 1) / 255) / 256^S > 0.498
constexpr double kOzLim = 0.498;

template <int S>
__global__ __launch_bounds__(256) void k_oz_split(const double *__restrict__ a, int64_t lda,
                                                  int row0, int nrows, int kb, int kw,
                                                  int8_t *__restrict__ planes, int64_t plane,
                                                  int32_t *__restrict__ expo,
                                                  const int32_t *__restrict__ info) {
  if (*info != 0) return;
  __shared__ int sh_e[kOzSplitRows];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int rb = blockIdx.x * kOzSplitRows;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int rl = 2 * wave + j;
    const int r = rb + rl;
    double m = 0.0;
    if (r < nrows) {
      const double *row = a + (int64_t)(row0 + r) * lda + kb;
#pragma unroll 4
      for (int k = lane; k < kw; k += 64) m = fmax(m, fabs(row[k]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (lane == 0) {
      int e = 0;
      if (m > 0.0) {
        (void)frexp(m, &e);  // m = f 2^e, f in [0.5, 1)
        e += 1;              // m 2^-e in [0.25, 0.5)
        if (ldexp(m, -e) > kOzLim) e += 1;
      }
      sh_e[rl] = e;
      if (r < nrows) expo[r] = e;
    }
  }
  __syncthreads();
  const int rl = tid & (kOzSplitRows - 1);
  const int r = rb + rl;
  if (r >= nrows) return;
  const int e = sh_e[rl];
  // 32 threads a row, each taking every 32nd chunk of 32 entries
  for (int chunk = tid >> 3; chunk < kOzPanel / kOzKS; chunk += 32) {
    const double *src = a + (int64_t)(row0 + r) * lda + kb + chunk * kOzKS;
    uint32_t packed[S][kOzKS / 4];
#pragma unroll
    for (int t = 0; t < S; ++t)
#pragma unroll
      for (int w = 0; w < kOzKS / 4; ++w) packed[t][w] = 0;
#pragma unroll
    for (int i = 0; i < kOzKS; ++i) {
      const int k = chunk * kOzKS + i;
      const double x = src[min(k, kw - 1) - chunk * kOzKS];  // a load on every path
      const double v = k < kw ? x : 0.0;
      long long X = (long long)ldexp(v, 8 * S - e);  // exact scaling, truncation toward zero
#pragma unroll
      for (int t = S - 1; t >= 0; --t) {
        int b = (int)(X & 255);
        b = b >= 128 ? b - 256 : b;
        X = (X - b) >> 8;  // exact: X - b is a multiple of 256
        packed[t][i >> 2] |= (uint32_t)(uint8_t)(int8_t)b << (8 * (i & 3));
      }
    }
    const int64_t off = (int64_t)chunk * ((int64_t)nrows * kOzKS) + (int64_t)r * kOzKS;
#pragma unroll
    for (int t = 0; t < S; ++t) {
      u32x4 *dst = reinterpret_cast<u32x4 *>(planes + t * plane + off);
      dst[0] = u32x4{packed[t][0], packed[t][1], packed[t][2], packed[t][3]};
      dst[1] = u32x4{packed[t][4], packed[t][5], packed[t][6], packed[t][7]};
    }
  }
}

template <int S>
struct OzStage {
  uint8_t a[S][kOzT * kOzKS];  // the tile's 128 rows, 32 bytes each, per plane
  uint8_t b[S][kOzT * kOzKS];  // the tile's 128 columns
};

// half of LDS row r that holds logical half h
__device__ __forceinline__ int oz_half(int r, int h) { return h ^ (((r >> 2) ^ (r >> 3)) & 1); }

// 16 bytes a lane from src into LDS at the wave-uniform address lds + 16 lane.
// Issued from asm: the compiler, which cannot tell the three stages apart,
// would otherwise wait for every DMA in flight before each LDS read; the
// kernel waits for its own loads with s_waitcnt vmcnt instead.  The
// instruction has no result registers (its data goes to LDS).  M0 holds the
// LDS address; it is declared clobbered, and no other code of this file
// uses M0 (the compiler sets M0 only right before its own uses, of which
// k_oz_update has none), so the reserved-register warning is silenced here.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void oz_dma16(const void *src, uint32_t lds) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "s"(lds)
               : "memory", "m0");
}
#pragma clang diagnostic pop

template <int S>
__global__ __launch_bounds__(512, 1) void k_oz_update(double *__restrict__ c, int64_t lda, int r0,
                                                      int r1, int c0, int c1, int srow0, int nsplit,
                                                      const int8_t *__restrict__ planes,
                                                      int64_t plane,
                                                      const int32_t *__restrict__ expo, int nk,
                                                      int gx, int gy, int tri,
                                                      const int32_t *__restrict__ info) {
  if (*info != 0) return;
  __shared__ __attribute__((aligned(16))) OzStage<S> st[kOzStages];
  __shared__ int ea[kOzExpLds ? kOzT : 1], eb[kOzExpLds ? kOzT : 1];
  int tx, ty;
  if (tri) {
    const int64_t t = (int64_t)gy * (gy + 1) / 2;
    const int64_t per = (t + kXcds - 1) / kXcds;
    const int64_t L = blockIdx.x;
    if (!tri_tile(gy, (L % kXcds) * per + L / kXcds, tx, ty)) return;
  } else if (!xcd_tile(gx, gy, tx, ty)) {
    return;
  }
  const int tr = r0 + ty * kOzT;  // output rows (A operand)
  const int tc = c0 + tx * kOzT;  // output columns (B operand)
  if (tc > tr + kOzT - 1) return;  // tile above the diagonal
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 1, wn = wave & 1;
  if (kOzExpLds) {
    if (tid < kOzT)
      ea[tid] = expo[min(tr + tid, r1 - 1) - srow0];
    else if (tid < 2 * kOzT)
      eb[tid - kOzT] = expo[min(tc + tid - kOzT, c1 - 1) - srow0];
  }

  // this wave's S of the stage's 8 S one-KB blocks: block j = (operand,
  // plane, 32-row group); lane l fills LDS slot l (row 32 q + l / 2, half
  // l % 2) from the logical half that slot holds
  const int8_t *src[S];
  int dst[S];
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const int j = wave * S + i;
    const int o = j / (4 * S), t = (j % (4 * S)) >> 2, q = j & 3;
    const int r = 32 * q + (lane >> 1);
    const int h = oz_half(r, lane & 1);
    const int grow = (o == 0 ? min(tr + r, r1 - 1) : min(tc + r, c1 - 1)) - srow0;
    src[i] = planes + t * plane + (int64_t)grow * kOzKS + 16 * h;
    dst[i] = (o == 0 ? (int)offsetof(OzStage<S>, a) : (int)offsetof(OzStage<S>, b)) +
             t * (kOzT * kOzKS) + q * 1024;
  }
  const int64_t kstride = (int64_t)nsplit * kOzKS;
  const uint32_t lds0 = (uint32_t)(size_t)(__attribute__((address_space(3))) uint8_t *)(
      reinterpret_cast<uint8_t *>(&st[0]));
  auto issue = [&](int kt, int stage) {
    const uint32_t base = lds0 + (uint32_t)(stage * (int)sizeof(OzStage<S>));
#pragma unroll
    for (int i = 0; i < S; ++i) oz_dma16(src[i] + kt * kstride, base + (uint32_t)dst[i]);
  };
  // fragment of rows rb .. rb + 31 (lane % 32), k bytes 16 (lane / 32)..
  auto frag = [&](const uint8_t *p, int rb) {
    const int row = rb + (lane & 31);
    return *reinterpret_cast<const i32x4 *>(p + row * kOzKS + 16 * oz_half(row, lane >> 5));
  };

  i32x16 acc[S][2];
#pragma unroll
  for (int l = 0; l < S; ++l)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[l][i][r] = 0;

  constexpr int ahead = kOzStages - 1;
#pragma unroll
  for (int k = 0; k < ahead; ++k)
    if (k < nk) issue(k, k);
  for (int kt = 0; kt < nk; ++kt) {
    // this wave's loads of step kt are done (the S loads of each later step
    // already issued may still be in flight); after the barrier every wave's
    // are, and every wave has finished step kt - 1, whose stage the loads of
    // step kt + ahead overwrite
    const int later = min(ahead - 1, nk - 1 - kt);
    if (later >= 2)
      asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(2 * S) : "memory");
    else if (later == 1)
      asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(S) : "memory");
    else
      asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    // this step's S loads of step kt + ahead are issued one per u step below,
    // between the MFMA groups, rather than in a burst
    const bool more = kt + ahead < nk;
    const uint32_t nbase = lds0 + (uint32_t)(((kt + ahead) % kOzStages) * (int)sizeof(OzStage<S>));
    const int64_t koff = (int64_t)(kt + ahead) * kstride;
    const OzStage<S> &s = st[kt % kOzStages];
    i32x4 af[S];
#pragma unroll
    for (int t = 0; t < S; ++t) af[t] = frag(s.a[t], wm * 32);
    // B fragments one u step ahead of their MFMAs
    i32x4 bq0 = frag(s.b[0], wn * 64), bq1 = frag(s.b[0], wn * 64 + 32);
#pragma unroll
    for (int u = 0; u < S; ++u) {
      const i32x4 b0 = bq0, b1 = bq1;
      if (u + 1 < S) {
        bq0 = frag(s.b[u + 1], wn * 64);
        bq1 = frag(s.b[u + 1], wn * 64 + 32);
      }
      if (more) oz_dma16(src[u] + koff, nbase + (uint32_t)dst[u]);
      // levels t + u < S; t descending keeps two MFMAs on one accumulator
      // a whole u step apart
#pragma unroll
      for (int t = S - 1 - u; t >= 0; --t) {
        acc[t + u][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[t], b0, acc[t + u][0], 0, 0, 0);
        acc[t + u][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[t], b1, acc[t + u][1], 0, 0, 0);
      }
    }
  }
  // epilogue: level l (0-based t + u) carries 2^-8(l + 2).  First every
  // value of the lane is formed (freeing all accumulators), then per 32 x 32
  // block its 16 old entries are loaded before the first use and the
  // entries inside the region stored, non-temporal (C is streamed once per
  // update).  Byte offsets from the tile's first entry are 32-bit (128 rows
  // of lda) and clamped into the region.  The scheduling barrier keeps the
  // loads from being hoisted among the conversions, where they would spill.
  double *tile = c + (int64_t)tr * lda + tc;
  const int rmax = r1 - 1 - tr, cmax = c1 - 1 - tc;
  double v[2][16];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int cl = wn * 64 + i * 32 + (lane & 31);
    const int gc = tc + cl;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int rl = wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      double x = 0.0;
#pragma unroll
      for (int l = S - 1; l >= 0; --l) x += ldexp((double)acc[l][i][r], -8 * (l + 2));
      const int e2 = kOzExpLds ? ea[rl] + eb[cl]
                               : expo[min(tr + rl, r1 - 1) - srow0] + expo[min(gc, c1 - 1) - srow0];
      v[i][r] = ldexp(x, e2);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  const int rl0 = wm * 32 + 4 * (lane >> 5);
  auto offs = [&](int i, int r) -> uint32_t {
    const int rl = rl0 + 8 * (r >> 2) + (r & 3);
    const int cl = wn * 64 + i * 32 + (lane & 31);
    return (uint32_t)(min(rl, rmax) * (int)lda + min(cl, cmax)) * 8u;
  };
  auto at = [&](uint32_t o) { return reinterpret_cast<double *>(reinterpret_cast<char *>(tile) + o); };
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    double old[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) old[r] = __builtin_nontemporal_load(at(offs(i, r)));
    const int cl = wn * 64 + i * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int rl = rl0 + 8 * (r >> 2) + (r & 3);
      if (rl <= rmax && cl <= cmax && tc + cl <= tr + rl)
        __builtin_nontemporal_store(old[r] - v[i][r], at(offs(i, r)));
    }
  }
}

template <int S>
int split_launch(const double *a, int64_t lda, int row0, int nrows, int kb, int kw, void *ws,
                 const int32_t *info, hipStream_t st) {
  int8_t *planes = static_cast<int8_t *>(ws);
  const int64_t plane = (int64_t)oz_plane_bytes(nrows);
  int32_t *expo = reinterpret_cast<int32_t *>(static_cast<char *>(ws) + WsCarver::need(S * plane));
  hipLaunchKernelGGL(k_oz_split<S>, dim3(ceil_div(nrows, kOzSplitRows)), dim3(256), 0, st, a, lda,
                     row0, nrows, kb, kw, planes, plane, expo, info);
  FMLH_LAUNCH_CHECK("k_oz_split");
  return 0;
}

template <int S>
int update_launch(double *a, int64_t lda, int r0, int r1, int c0, int c1, int srow0, int nsplit,
                  const void *ws, bool tri, const int32_t *info, hipStream_t st) {
  const int8_t *planes = static_cast<const int8_t *>(ws);
  const int64_t plane = (int64_t)oz_plane_bytes(nsplit);
  const int32_t *expo = reinterpret_cast<const int32_t *>(static_cast<const char *>(ws) +
                                                         WsCarver::need(S * plane));
  const int gx = ceil_div(c1 - c0, kOzT), gy = ceil_div(r1 - r0, kOzT);
  const int64_t tiles = tri ? (int64_t)gy * (gy + 1) / 2 : (int64_t)gx * gy;
  hipLaunchKernelGGL(k_oz_update<S>, dim3((unsigned)xcd_grid((int)tiles, 1)), dim3(512), 0, st, a,
                     lda, r0, r1, c0, c1, srow0, nsplit, planes, plane, expo, kOzPanel / kOzKS, gx,
                     gy, tri ? 1 : 0, info);
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
  if (slices == kOzSlices) return split_launch<kOzSlices>(a, lda, row0, nrows, kb, kw, ws, info, st);
  return -7;
}

int oz_update(double *a, int64_t lda, int r0, int r1, int c0, int c1, int srow0, int nsplit,
              int slices, const void *ws, bool tri, const int32_t *info, hipStream_t st) {
  if (r1 <= r0 || c1 <= c0) return 0;
  if (tri && (r0 != c0 || r1 != c1)) return -11;
  if (slices == kOzSlices)
    return update_launch<kOzSlices>(a, lda, r0, r1, c0, c1, srow0, nsplit, ws, tri, info, st);
  return -7;
}

}  // namespace fmlh
