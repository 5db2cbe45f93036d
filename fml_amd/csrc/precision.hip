// Split-BF16x3 Gaussian kernel (precision form of fgaussian_kernel,
// fml/kernels/fkernels.f90:317-349).
//
// Each float operand row is split once, by k_split_bf16, into two bf16
// planes in the workspace: hi = bf16(x), lo = bf16(x - hi) (round to nearest
// even), columns zero-padded to a multiple of 32, with the row's FP32
// squared norm (sequential fused sum of the original floats, so a row's norm
// does not depend on the tile that uses it).  The dot products
//   a.b ~ hi_a.hi_b + hi_a.lo_b + lo_a.hi_b
// are three v_mfma_f32_16x16x32_bf16 per 16x16 block and 32-wide k step,
// accumulated in FP32.  The dropped lo.lo term and the bf16 rounding of lo
// leave ~2^-16 relative error per product, so |a|^2 + |b|^2 - 2 a.b loses
// accuracy when d2 is small against the norms: entries with d2 <= tau
// (|a|^2 + |b|^2) are recomputed from direct FP32 differences after the
// tile's other entries are stored, four at a time per wavefront
// (recompute_near_deferred, near_pairs.h), and counted when the caller asks.
// FP32 outputs are exp2(d2 * (log2(e) inv)) with the scale rounded once.
//
// Tile: 128 x 128 outputs per 256 threads, waves 2 x 2, each wave 64 x 64
// from 4 x 4 16x16 blocks (48 MFMAs per k step).  A k step stages 128 rows x
// 32 columns of the four planes (32 KB) in LDS, 64 bytes a row with the
// 16-byte chunks of row r permuted by lds_swizzle(r) (gemm_nt.h), so
// neither the staging writes nor the 16-byte fragment reads of 16
// consecutive rows collide in a bank; the next step's planes are loaded into
// registers while the current one is multiplied.  The grid is 1-D in XCD-grouped tile order
// (gemm_nt.h), so the tiles one XCD runs share operand rows in its L2.
//
// bf16 16x16x32 MFMA lane layout on gfx950 (checked with an in-bounds probe):
//   A lane l -> A[l%16][8*(l/16) + e], e = 0..7
//   B lane l -> B[8*(l/16) + e][l%16]
//   D lane l, reg r -> D[4*(l/16) + r][l%16]
#include <math.h>

#include "common.h"
#include "gemm_nt.h"
#include "near_pairs.h"

#pragma clang fp contract(off)

namespace {

using namespace fmlh;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

constexpr int kBK = 32;     // k step
constexpr int kRow = 32;    // bf16 per LDS row (64 bytes; 16-byte chunks swizzled)
constexpr int kMT = 128;    // tile edge
constexpr float kTau = 2e-2f;

__device__ __forceinline__ uint16_t bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__host__ __device__ inline int padded_cols(int d) { return (d + kBK - 1) / kBK * kBK; }

// One thread per (row, 8-column chunk) for the planes; the chunk-0 thread of
// each row also forms the row's squared norm.
__global__ __launch_bounds__(256) void k_split_bf16(const float *__restrict__ X, int64_t ld, int n,
                                                    int d, int dp, uint16_t *__restrict__ hi,
                                                    uint16_t *__restrict__ lo,
                                                    float *__restrict__ nrm) {
  const int chunks = dp / 8;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * chunks) return;
  const int row = (int)(t / chunks), c0 = (int)(t % chunks) * 8;
  const float *x = X + (int64_t)row * ld;
  uint16_t h[8], l[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const float v = c0 + e < d ? x[c0 + e] : 0.f;
    h[e] = bf16_rne(v);
    l[e] = bf16_rne(v - bf16_to_f(h[e]));
  }
  u4v hv, lv;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    hv[e] = (uint32_t)h[2 * e] | ((uint32_t)h[2 * e + 1] << 16);
    lv[e] = (uint32_t)l[2 * e] | ((uint32_t)l[2 * e + 1] << 16);
  }
  *reinterpret_cast<u4v *>(hi + (int64_t)row * dp + c0) = hv;
  *reinterpret_cast<u4v *>(lo + (int64_t)row * dp + c0) = lv;
  if (c0 == 0) {
    float s = 0.f;
    for (int k = 0; k < d; ++k) s = fmaf(x[k], x[k], s);
    nrm[row] = s;
  }
}

struct Plane {
  const uint16_t *hi, *lo;
  const float *nrm;
};

// Tile row held by LDS row s of the A operand.  LDS row blk*64 + 16 i + c
// (the column c of MFMA block i) holds tile row blk*64 + 4 c + i, so the lane
// that owns output column c of the four blocks i = 0..3 owns four
// consecutive columns of K and stores them with one vector store.
__device__ __forceinline__ int a_row(int s) {
  return (s & ~63) | ((s & 15) << 2) | ((s >> 4) & 3);
}

struct B3Smem {
  uint16_t ah[kMT * kRow];
  uint16_t al[kMT * kRow];
  uint16_t bh[kMT * kRow];
  uint16_t bl[kMT * kRow];
};

template <typename OutT>
__device__ __forceinline__ OutT gauss_value(float d2, double inv, float inv_log2) {
  if (sizeof(OutT) == 8) return (OutT)exp(inv * (double)d2);
  return (OutT)exp2f(inv_log2 * d2);
}

// K[gm][gn] for B rows gm (output rows) and A rows gn (output columns).
// kSym (A and B the same rows): only the lower-triangle tiles run (tri_tile
// order); each entry with gn <= gm is stored at (gm, gn) and, when gn < gm,
// at (gn, gm) too, so the kernel is exactly symmetric.  (A full-grid entry
// above the diagonal may differ from its mirror in the last FP32 bits: its
// tile adds the two cross products in the other order.)
template <typename OutT, bool kSym = false>
__global__ __launch_bounds__(256, 2) void k_gauss_b3(Plane pa, int na, Plane pb, int nb, int dp,
                                                     const float *__restrict__ A, int64_t lda,
                                                     const float *__restrict__ B, int64_t ldb,
                                                     int d, double inv, float inv_log2,
                                                     OutT *__restrict__ K,
                                                     int64_t ldk, unsigned long long *nrecomp,
                                                     int gx, int gy) {
  __shared__ __attribute__((aligned(16))) B3Smem s;
  __shared__ unsigned long long s_near[4][64];
  int tx, ty;
  if (kSym) {
    const int64_t t = (int64_t)gy * (gy + 1) / 2;
    const int64_t per = (t + kXcds - 1) / kXcds;
    const int64_t L = blockIdx.x;
    if (!tri_tile(gy, (L % kXcds) * per + L / kXcds, tx, ty)) return;
  } else if (!xcd_tile(gx, gy, tx, ty)) {
    return;
  }
  const bool diag = kSym && tx == ty;  // stores only gn <= gm
  const bool mirror = kSym;            // entries with gn < gm also at (gn, gm)
  const int n0 = tx * kMT;  // A rows (output columns)
  const int m0 = ty * kMT;  // B rows (output rows)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // staging: each thread moves 2 chunks of 8 bf16 per plane and k step
  // (rows tid / 4 and 64 + tid / 4, columns 8 (tid % 4))
  const int srow = tid >> 2, scol = (tid & 3) * 8;
  u4v rah[2], ral[2], rbh[2], rbl[2];
  auto gload = [&](int k0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = srow + 64 * h;
      const int ga = n0 + a_row(r), gb = m0 + r;
      const int64_t oa = (int64_t)min(ga, na - 1) * dp + k0 + scol;
      const int64_t ob = (int64_t)min(gb, nb - 1) * dp + k0 + scol;
      rah[h] = *reinterpret_cast<const u4v *>(pa.hi + oa);
      ral[h] = *reinterpret_cast<const u4v *>(pa.lo + oa);
      rbh[h] = *reinterpret_cast<const u4v *>(pb.hi + ob);
      rbl[h] = *reinterpret_cast<const u4v *>(pb.lo + ob);
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int row = srow + 64 * h;
      const int o = row * kRow + 8 * ((scol >> 3) ^ lds_swizzle(row));
      *reinterpret_cast<u4v *>(&s.ah[o]) = rah[h];
      *reinterpret_cast<u4v *>(&s.al[o]) = ral[h];
      *reinterpret_cast<u4v *>(&s.bh[o]) = rbh[h];
      *reinterpret_cast<u4v *>(&s.bl[o]) = rbl[h];
    }
  };

  f4v acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};

  const int nk = dp / kBK;
  gload(0);
  for (int kt = 0; kt < nk; ++kt) {
    __syncthreads();  // the previous step's readers are done
    lstore();
    __syncthreads();
    if (kt + 1 < nk) gload((kt + 1) * kBK);
    const int ko = 8 * (lane >> 4);
    bf16x8 bh[4], bl[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = wm * 64 + j * 16 + (lane & 15);
      const int r = row * kRow + 8 * ((ko >> 3) ^ lds_swizzle(row));
      bh[j] = *reinterpret_cast<const bf16x8 *>(&s.bh[r]);
      bl[j] = *reinterpret_cast<const bf16x8 *>(&s.bl[r]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = wn * 64 + i * 16 + (lane & 15);
      const int r = row * kRow + 8 * ((ko >> 3) ^ lds_swizzle(row));
      const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(&s.ah[r]);
      const bf16x8 al = *reinterpret_cast<const bf16x8 *>(&s.al[r]);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // block (j, i): rows from B (A operand of the MFMA), columns from A
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah, acc[i][j], 0, 0, 0);
      }
    }
  }
  // epilogue: d2 = |b|^2 + |a|^2 - 2 a.b, near pairs recomputed, exp.
  // This lane's output columns are gn0 + i (i = 0..3, from block i) and its
  // rows m0 + wm*64 + 16 j + 4 (lane / 16) + r.
  const int gn0 = n0 + wn * 64 + 4 * (lane & 15);
  float qn[4], pn[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i) qn[i] = gn0 + i < na ? pa.nrm[gn0 + i] : 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gm = m0 + wm * 64 + j * 16 + 4 * (lane >> 4) + r;
      pn[j][r] = gm < nb ? pb.nrm[gm] : 0.f;
    }
  // whole tile inside K and 4-element stores aligned: one vector store per row
  const bool vec = !diag && n0 + kMT <= na && m0 + kMT <= nb && (ldk & 3) == 0 &&
                   (reinterpret_cast<uintptr_t>(K) & (4 * sizeof(OutT) - 1)) == 0;
  // entries in the cancellation band are stored after the tile, from direct
  // differences: step b = 16 j + 4 r + i records its ballot in masks[b]
  unsigned long long *masks = s_near[wave];
  unsigned long long steps = 0;
  unsigned long long mynear = 0;  // this lane's band entries, bit 16 j + 4 r + i
  unsigned int near_count = 0;    // this wave's entries of K written from direct differences
  OutT v[4][4][4];                // [j][r][i]
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gm = m0 + wm * 64 + j * 16 + 4 * (lane >> 4) + r;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gn = gn0 + i;
        const bool valid = gm < nb && gn < na && !(diag && gn > gm);
        const float nsum = pn[j][r] + qn[i];
        float d2 = fmaf(-2.f, acc[i][j][r], nsum);
        const bool nr = valid && d2 <= kTau * nsum;
        const unsigned long long mk = __ballot(nr);
        if (mk) {
          masks[16 * j + 4 * r + i] = mk;
          steps |= 1ull << (16 * j + 4 * r + i);
          // a mirrored entry is written twice
          near_count += (unsigned int)__popcll(mk) + (unsigned int)__popcll(__ballot(mirror && nr && gn < gm));
        }
        if (nr) mynear |= 1ull << (16 * j + 4 * r + i);
        if (d2 < 0.f) d2 = 0.f;
        v[j][r][i] = gauss_value<OutT>(d2, inv, inv_log2);
      }
    }
  typedef OutT o4v __attribute__((ext_vector_type(4)));
  // row gm, columns gn0 .. gn0 + 3: one vector store when whole
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gm = m0 + wm * 64 + j * 16 + 4 * (lane >> 4) + r;
      const unsigned long long nb4 = (mynear >> (16 * j + 4 * r)) & 15ull;
      OutT *row = K + (int64_t)gm * ldk + gn0;
      if (vec && nb4 == 0) {
        __builtin_nontemporal_store(o4v{v[j][r][0], v[j][r][1], v[j][r][2], v[j][r][3]},
                                    reinterpret_cast<o4v *>(row));
      } else if (gm < nb) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (gn0 + i < na && !((nb4 >> i) & 1ull) && !(diag && gn0 + i > gm))
            __builtin_nontemporal_store(v[j][r][i], row + i);
      }
    }
  // the mirror: row gn0 + i, columns gm .. gm + 3 (this lane's four r), one
  // vector store when whole (every entry of an off-diagonal tile has gn < gm)
  if (mirror) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gmb = m0 + wm * 64 + j * 16 + 4 * (lane >> 4);
        const int gn = gn0 + i;
        bool band = false;
#pragma unroll
        for (int r = 0; r < 4; ++r) band = band || ((mynear >> (16 * j + 4 * r + i)) & 1ull);
        OutT *col = K + (int64_t)gn * ldk + gmb;
        if (vec && !band) {
          __builtin_nontemporal_store(o4v{v[j][0][i], v[j][1][i], v[j][2][i], v[j][3][i]},
                                      reinterpret_cast<o4v *>(col));
        } else if (gn < na) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (gmb + r < nb && gn < gmb + r && !((mynear >> (16 * j + 4 * r + i)) & 1ull))
              __builtin_nontemporal_store(v[j][r][i], col + r);
        }
      }
  }
  if (steps) {
    auto idx = [&](int b, int L, int &gn, int &gm) {
      gn = n0 + wn * 64 + 4 * (L & 15) + (b & 3);
      gm = m0 + wm * 64 + 16 * (b >> 4) + 4 * (L >> 4) + ((b >> 2) & 3);
    };
    auto put = [&](int gn, int gm, float d2) {
      if (diag && gn > gm) return;
      const OutT kv = gauss_value<OutT>(d2, inv, inv_log2);
      __builtin_nontemporal_store(kv, K + (int64_t)gm * ldk + gn);
      if (mirror && gn < gm) __builtin_nontemporal_store(kv, K + (int64_t)gn * ldk + gm);
    };
    recompute_near_deferred<float>(steps, masks, idx, A, lda, B, ldb, d, put);
  }
  if (nrecomp && lane == 0 && near_count) atomicAdd(nrecomp, (unsigned long long)near_count);
}

size_t plane_bytes(int n, int d) {
  const size_t dp = (size_t)padded_cols(d);
  return 2 * WsCarver::need(sizeof(uint16_t) * (size_t)n * dp) + WsCarver::need(sizeof(float) * (size_t)n);
}

bool same_operand(const float *a, int64_t lda, int na, const float *b, int64_t ldb, int nb) {
  return a == b && lda == ldb && na == nb;
}

int split(WsCarver &w, const float *X, int64_t ld, int n, int d, Plane &p, hipStream_t st) {
  const int dp = padded_cols(d);
  uint16_t *hi = w.take<uint16_t>((size_t)n * dp);
  uint16_t *lo = w.take<uint16_t>((size_t)n * dp);
  float *nrm = w.take<float>((size_t)n);
  if (!hi || !lo || !nrm) return -1;
  const int64_t work = (int64_t)n * (dp / 8);
  hipLaunchKernelGGL(k_split_bf16, dim3((unsigned)ceil_div64(work, 256)), dim3(256), 0, st, X, ld,
                     n, d, dp, hi, lo, nrm);
  FMLH_LAUNCH_CHECK("k_split_bf16");
  p = Plane{hi, lo, nrm};
  return 0;
}

}  // namespace

namespace fmlh {

size_t gaussian_bf16x3_ws_bytes(const float *a, int64_t lda, int na, const float *b, int64_t ldb,
                                int nb, int d) {
  size_t bytes = plane_bytes(na, d);
  if (!same_operand(a, lda, na, b, ldb, nb)) bytes += plane_bytes(nb, d);
  return bytes;
}

int gaussian_bf16x3(const float *a, int64_t lda, int na, const float *b, int64_t ldb, int nb,
                    int d, double sigma, float *k32, double *k64, int64_t ldk, void *ws,
                    size_t ws_bytes, unsigned long long *recomputed, hipStream_t st) {
  const double inv = -0.5 / (sigma * sigma);
  // exp(inv d2) = exp2(inv log2(e) d2) for the FP32 outputs
  const float inv_log2 = (float)(inv * 1.4426950408889634);
  WsCarver w(ws, ws_bytes);
  Plane pa, pb;
  int rc = split(w, a, lda, na, d, pa, st);
  if (rc) return rc == -1 ? -12 : rc;
  if (same_operand(a, lda, na, b, ldb, nb)) {
    pb = pa;
  } else {
    rc = split(w, b, ldb, nb, d, pb, st);
    if (rc) return rc == -1 ? -12 : rc;
  }
  const int gx = ceil_div(na, kMT), gy = ceil_div(nb, kMT);
  const int dp = padded_cols(d);
  timing_begin(st);
  if (same_operand(a, lda, na, b, ldb, nb)) {
    // a kernel of a set with itself: lower-triangle tiles, mirrored stores
    const unsigned grid = (unsigned)xcd_grid((int)((int64_t)gy * (gy + 1) / 2), 1);
    if (k32)
      hipLaunchKernelGGL((k_gauss_b3<float, true>), dim3(grid), dim3(256), 0, st, pa, na, pb, nb, dp,
                         a, lda, b, ldb, d, inv, inv_log2, k32, ldk, recomputed, gx, gy);
    else
      hipLaunchKernelGGL((k_gauss_b3<double, true>), dim3(grid), dim3(256), 0, st, pa, na, pb, nb, dp,
                         a, lda, b, ldb, d, inv, inv_log2, k64, ldk, recomputed, gx, gy);
  } else {
    const unsigned grid = (unsigned)xcd_grid(gx, gy);
    if (k32)
      hipLaunchKernelGGL(k_gauss_b3<float>, dim3(grid), dim3(256), 0, st, pa, na, pb, nb, dp, a, lda,
                         b, ldb, d, inv, inv_log2, k32, ldk, recomputed, gx, gy);
    else
      hipLaunchKernelGGL(k_gauss_b3<double>, dim3(grid), dim3(256), 0, st, pa, na, pb, nb, dp, a, lda,
                         b, ldb, d, inv, inv_log2, k64, ldk, recomputed, gx, gy);
  }
  timing_end(st);
  FMLH_LAUNCH_CHECK("k_gauss_b3");
  return 0;
}

}  // namespace fmlh
