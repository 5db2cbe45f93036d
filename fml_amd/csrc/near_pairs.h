// Cancellation-band recompute shared by the Gaussian kernels: entries whose
// GEMM-form distance |a|^2 + |b|^2 - 2 a.b lies in the band where it has lost
// accuracy are recomputed as sum_t (a_t - b_t)^2.
#pragma once

#include "common.h"

namespace fmlh {

// sum_t (a_t - b_t)^2 over one row pair by the 16 lanes of a quarter wave
// (lane l16 takes t = l16 + 16 u); every load of a 16 U-column chunk is
// issued before the first use, so a pair costs one memory round trip per
// chunk.  Returns the lane's partial sum.
template <typename T>
__device__ __forceinline__ T quarter_row_dist(const T *__restrict__ a, const T *__restrict__ b,
                                              int d, int l16) {
  constexpr int U = sizeof(T) == 4 ? 28 : 14;
  T part = T(0);
  for (int t0 = 0; t0 < d; t0 += 16 * U) {
    T av[U], bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = t0 + l16 + 16 * u;
      av[u] = t < d ? a[t] : T(0);
      bv[u] = t < d ? b[t] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T df = av[u] - bv[u];
      part = fma(df, df, part);
    }
  }
  return part;
}

__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long x) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)x);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(x >> 32));
  return (unsigned long long)lo | ((unsigned long long)hi << 32);
}

// Deferred form for an epilogue that has already stored every entry outside
// the band.  The epilogue records, for each of its (at most 64) steps b, the
// ballot of its in-band lanes in masks[b] (wave-private LDS, written with the
// same value by every lane) and sets bit b of the wave-uniform `steps`.
// idx(b, L, gn, gm) names the entry that lane L held at step b.  The flagged
// entries are taken four at a time, one per quarter wave, and for each
// store(gn, gm, d2) runs on the quarter's first lane with the direct-
// difference distance.
template <typename T, typename IdxF, typename StoreF>
__device__ void recompute_near_deferred(unsigned long long steps, const unsigned long long *masks,
                                        IdxF idx, const T *__restrict__ A, int64_t lda,
                                        const T *__restrict__ B, int64_t ldb, int d,
                                        StoreF store) {
  const int lane = threadIdx.x & 63, q = lane >> 4, l16 = lane & 15;
  unsigned long long m = 0;
  int b = 0;
  for (;;) {
    int en = 0, em = 0, cnt = 0;
    bool have = false;
    while (cnt < 4) {
      if (m == 0) {
        if (steps == 0) break;
        b = __ffsll((long long)steps) - 1;
        steps &= steps - 1;
        m = uniform_u64(masks[b]);
      }
      const int L = __ffsll((long long)m) - 1;
      m &= m - 1;
      int gn, gm;
      idx(b, L, gn, gm);
      if (q == cnt) {
        en = gn;
        em = gm;
        have = true;
      }
      ++cnt;
    }
    if (cnt == 0) break;
    T part = T(0);
    if (have) part = quarter_row_dist(A + (int64_t)en * lda, B + (int64_t)em * ldb, d, l16);
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) part += __shfl_xor(part, off, 16);
    if (have && l16 == 0) store(en, em, part);
    if (cnt < 4) break;
  }
}

}  // namespace fmlh
