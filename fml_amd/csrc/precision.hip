// Split-BF16x3 Gaussian kernel (precision form of fgaussian_kernel,
// fml/kernels/fkernels.f90:317-349).
//
// Each float operand x is split into hi = bf16(x) and lo = bf16(x - hi)
// (round to nearest even) while it is staged into LDS; the dot products
// a.b ~ hi_a.hi_b + hi_a.lo_b + lo_a.hi_b are three v_mfma_f32_16x16x32_bf16
// per 16x16 block and 32-wide k step, accumulated in FP32.  The dropped
// lo.lo term and the bf16 rounding of lo leave ~2^-16 relative error per
// product, so |a|^2 + |b|^2 - 2 a.b loses accuracy when d2 is small against
// the norms: entries with d2 <= tau (|a|^2 + |b|^2) are recomputed from direct
// FP32 differences by their wavefront (near_pairs.h).  Norms are FP32 sums of the original floats.
//
// Tile: 64 x 64 outputs per 256 threads, waves 2 x 2, each 32 x 32 from
// 2 x 2 16x16 blocks; 32-wide k stages double buffered in LDS (hi and lo
// planes of both operands, rows padded to 40 bf16 = 80 bytes so the
// 16-byte reads of 8 consecutive rows hit distinct banks).
//
// bf16 16x16x32 MFMA lane layout on gfx950 (checked with an in-bounds probe):
//   A lane l -> A[l%16][8*(l/16) + e], e = 0..7
//   B lane l -> B[8*(l/16) + e][l%16]
//   D lane l, reg r -> D[4*(l/16) + r][l%16]
#include <math.h>

#include "common.h"
#include "near_pairs.h"

namespace {

using namespace fmlh;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int kBK = 32;
constexpr int kRow = 40;  // bf16 per LDS row

struct B3Smem {
  uint16_t ph[2][64 * kRow];
  uint16_t pl[2][64 * kRow];
  uint16_t qh[2][64 * kRow];
  uint16_t ql[2][64 * kRow];
  float pn[64];
  float qn[64];
};

__device__ __forceinline__ uint16_t bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <typename OutT>
__global__ __launch_bounds__(256) void k_gaussian_bf16x3(const float *__restrict__ A, int64_t lda,
                                                         int na, const float *__restrict__ B,
                                                         int64_t ldb, int nb, int d, double inv,
                                                         OutT *__restrict__ K, int64_t ldk,
                                                         float tau) {
  __shared__ __attribute__((aligned(16))) B3Smem s;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int n0 = blockIdx.x * 64;  // A rows (output columns)
  const int m0 = blockIdx.y * 64;  // B rows (output rows)
  const int lrow = tid >> 2;       // staged row
  const int lk = (tid & 3) * 8;    // staged k offset
  const bool prow_ok = m0 + lrow < nb, qrow_ok = n0 + lrow < na;
  const float *prow = B + (int64_t)(m0 + lrow) * ldb;
  const float *qrow = A + (int64_t)(n0 + lrow) * lda;

  float rp[8], rq[8];
  float pn_acc = 0.f, qn_acc = 0.f;
  auto gload = [&](int k0) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + lk + e;
      rp[e] = (prow_ok && k < d) ? prow[k] : 0.f;
      rq[e] = (qrow_ok && k < d) ? qrow[k] : 0.f;
    }
  };
  auto lstore = [&](int buf) {
    uint16_t hp[8], lp[8], hq[8], lq[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      pn_acc = pn_acc + rp[e] * rp[e];
      qn_acc = qn_acc + rq[e] * rq[e];
      hp[e] = bf16_rne(rp[e]);
      lp[e] = bf16_rne(rp[e] - bf16_to_f(hp[e]));
      hq[e] = bf16_rne(rq[e]);
      lq[e] = bf16_rne(rq[e] - bf16_to_f(hq[e]));
    }
    const int o = lrow * kRow + lk;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      s.ph[buf][o + e] = hp[e];
      s.pl[buf][o + e] = lp[e];
      s.qh[buf][o + e] = hq[e];
      s.ql[buf][o + e] = lq[e];
    }
  };

  f4v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};

  const int nk = (d + kBK - 1) / kBK;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * kBK);
    const int ko = 8 * (lane >> 4);
    bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int pr = (wm * 32 + i * 16 + (lane & 15)) * kRow + ko;
      const int qr = (wn * 32 + i * 16 + (lane & 15)) * kRow + ko;
      ah[i] = *reinterpret_cast<const bf16x8 *>(&s.ph[buf][pr]);
      al[i] = *reinterpret_cast<const bf16x8 *>(&s.pl[buf][pr]);
      bh[i] = *reinterpret_cast<const bf16x8 *>(&s.qh[buf][qr]);
      bl[i] = *reinterpret_cast<const bf16x8 *>(&s.ql[buf][qr]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    if (kt + 1 < nk) lstore(buf ^ 1);
    __syncthreads();
  }
  // row norms: the four threads of a row sit in consecutive lanes
  pn_acc += __shfl_xor(pn_acc, 1, 64);
  pn_acc += __shfl_xor(pn_acc, 2, 64);
  qn_acc += __shfl_xor(qn_acc, 1, 64);
  qn_acc += __shfl_xor(qn_acc, 2, 64);
  if ((tid & 3) == 0) {
    s.pn[lrow] = pn_acc;
    s.qn[lrow] = qn_acc;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        const int n = wn * 32 + j * 16 + (lane & 15);
        const int gm = m0 + m, gn = n0 + n;
        const bool valid = gm < nb && gn < na;
        const float pn = s.pn[m], qn = s.qn[n];
        float d2 = pn + qn - 2.f * acc[i][j][r];
        const bool near = valid && d2 <= tau * (pn + qn);
        d2 = recompute_near(near, d2, A, lda, gn, B, ldb, gm, d);
        if (!valid) continue;
        if (d2 < 0.f) d2 = 0.f;
        OutT v;
        if (sizeof(OutT) == 8)
          v = (OutT)exp(inv * (double)d2);
        else
          v = (OutT)expf((float)inv * d2);
        __builtin_nontemporal_store(v, K + (int64_t)gn + (int64_t)gm * ldk);
      }
}

}  // namespace

extern "C" int fmlh_gaussian_kernel_bf16x3(const float *a, int64_t lda, int na, const float *b,
                                           int64_t ldb, int nb, int d, double sigma, float *k32,
                                           double *k64, int64_t ldk, void *stream) {
  const double inv = -0.5 / (sigma * sigma);
  dim3 grid(ceil_div(na, 64), ceil_div(nb, 64));
  hipStream_t st = as_stream(stream);
  timing_begin(st);
  if (k32)
    hipLaunchKernelGGL(k_gaussian_bf16x3<float>, grid, dim3(256), 0, st, a, lda, na, b, ldb, nb, d,
                       inv, k32, ldk, 2e-2f);
  else
    hipLaunchKernelGGL(k_gaussian_bf16x3<double>, grid, dim3(256), 0, st, a, lda, na, b, ldb, nb, d,
                       inv, k64, ldk, 2e-2f);
  timing_end(st);
  FMLH_LAUNCH_CHECK("k_gaussian_bf16x3");
  return 0;
}
