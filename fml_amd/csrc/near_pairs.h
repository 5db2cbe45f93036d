// Cancellation-band recompute shared by the Gaussian kernels.
#pragma once

#include "common.h"

namespace fmlh {

// Entries in the cancellation band are recomputed as sum_t (a_t - b_t)^2 by
// the whole wavefront, one flagged entry at a time: lanes stride over t with
// coalesced loads and a butterfly reduction combines them.  Every lane
// executes this (it contains cross-lane operations); `near` marks the
// lanes whose entry needs it, and each gets its own recomputed value back.
template <typename T>
__device__ T recompute_near(bool near, T d2, const T *__restrict__ A, int64_t lda, int gn,
                            const T *__restrict__ B, int64_t ldb, int gm, int d) {
  const int lane = threadIdx.x & 63;
  unsigned long long mask = __ballot(near);
  while (mask) {
    const int L = __ffsll((long long)mask) - 1;
    mask &= mask - 1;
    const int rn = __shfl(gn, L, 64), rm = __shfl(gm, L, 64);
    const T *a = A + (int64_t)rn * lda;
    const T *b = B + (int64_t)rm * ldb;
    T part = T(0);
    for (int t = lane; t < d; t += 64) {
      const T diff = a[t] - b[t];
      part += diff * diff;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if (lane == L) d2 = part;
  }
  return d2;
}

}  // namespace fmlh
