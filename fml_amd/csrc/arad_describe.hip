// ARAD descriptors on the device: ARAD.describe (fml/arad.py:70-145, the
// non-periodic form; Python entry Molecule.generate_arad_descriptor,
// fml/fml.py:80-86), one wavefront per atom row (molecule, atom i).
//
// For atom i of a molecule with L atoms, every atom k (i itself included)
// is a neighbour: v_k = R_k - R_i, d_k = sqrt((x^2 + y^2) + z^2), weight
// w_k = 1 - sin((pi d_k) / (2 cut)) (arad.py:98-108).  The angle at i
// between neighbours m and k is arccos(sp / (d_m d_k)), sp = the dot product
// ((x_m x_k + y_m y_k) + z_m z_k), and 0 where |sp - d_m d_k| <= eps or
// |d_m d_k| <= eps, eps = 100 DBL_EPSILON (arad.py:29-33, 42); arccos is not
// clipped, as in the reference.  Neighbours are ordered by distance (ties in
// index order) and kept while d < cut (arad.py:110-121); the five channels
// of row i are d, the PTP row and column of the neighbour, and
// sum_k cos(angle) w_k / W and sum_k sin(angle) w_k / W over the kept
// neighbours k in that order, W the sum of the kept weights (arad.py:
// 137-142).  Empty distance slots hold 1e100 (arad.py:93).
//
// Every sum is evaluated in the order numpy's np.sum takes along a
// contiguous axis of up to 128 values (pairwise_sum): fewer than 8 values
// left to right from 0; otherwise eight running sums over blocks of 8,
// combined ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then the rest left to
// right.  With contraction off, distances, weights' arguments and the order
// of neighbours are the host computation's bit for bit; the angular
// channels differ from it only by the device sin/cos/acos rounding.
#include <float.h>
#include <math.h>

#include "common.h"

#pragma clang fp contract(off)

namespace {

using namespace fmlh;

constexpr int kDescMaxAtoms = 64;  // atoms per molecule (one lane each)

__global__ __launch_bounds__(64) void k_arad_describe(const double *__restrict__ z,
                                                      const double *__restrict__ r,
                                                      const int32_t *__restrict__ offsets, int nmol,
                                                      int amax, double cut,
                                                      const double *__restrict__ ptp, int ptp_rows,
                                                      double *__restrict__ q, double *__restrict__ zt,
                                                      int32_t *__restrict__ n_out,
                                                      int32_t *__restrict__ status) {
  __shared__ double xs[kDescMaxAtoms], ys[kDescMaxAtoms], zs[kDescMaxAtoms];
  __shared__ double vx[kDescMaxAtoms], vy[kDescMaxAtoms], vz[kDescMaxAtoms];
  __shared__ double dist[kDescMaxAtoms], wgt[kDescMaxAtoms];
  __shared__ double trow[kDescMaxAtoms], tcol[kDescMaxAtoms];
  __shared__ int order[kDescMaxAtoms];
  const int mol = blockIdx.x;
  const int i = blockIdx.y;   // the atom whose row this workgroup writes
  const int lane = threadIdx.x;
  const int a0 = offsets[mol];
  const int L = offsets[mol + 1] - a0;
  const int64_t blk = (int64_t)5 * amax;
  double *qm = q + (int64_t)mol * amax * blk;
  // each charge is a key of the PTP dictionary (arad.py:98): an integral
  // value with an entry; rows without an entry hold 0
  bool bad = L < 0 || L > amax || L > kDescMaxAtoms;
  if (!bad && lane < L) {
    const double zc = z[a0 + lane];
    bad = !(zc >= 0.0 && zc < (double)ptp_rows);
    if (!bad) bad = (double)(int)zc != zc || !(ptp[2 * (int)zc] > 0.0);
  }
  bad = __any(bad);
  if (i == 0) {
    if (status && lane == 0) status[mol] = bad ? 1 : 0;
    if (lane == 0) n_out[mol] = bad ? 0 : L;
    for (int a = lane; a < amax; a += 64) {
      const int zi = (!bad && a < L) ? (int)z[a0 + a] : 0;
      zt[((int64_t)mol * amax + a) * 2 + 0] = (!bad && a < L) ? ptp[2 * zi] : 0.0;
      zt[((int64_t)mol * amax + a) * 2 + 1] = (!bad && a < L) ? ptp[2 * zi + 1] : 0.0;
    }
  }
  double *Mi = qm + (int64_t)i * blk;
  if (bad || i >= L) {  // a padding row (all rows of a molecule flagged in status)
    for (int t = lane; t < (int)blk; t += 64) Mi[t] = t < amax ? 1e100 : 0.0;
    return;
  }
  if (lane < L) {
    xs[lane] = r[3 * (int64_t)(a0 + lane) + 0];
    ys[lane] = r[3 * (int64_t)(a0 + lane) + 1];
    zs[lane] = r[3 * (int64_t)(a0 + lane) + 2];
    const int zi = (int)z[a0 + lane];
    trow[lane] = ptp[2 * zi];
    tcol[lane] = ptp[2 * zi + 1];
  }
  __syncthreads();
  const double eps = 100.0 * DBL_EPSILON;
  const double pi = 3.141592653589793;
  {
    double d = 0.0;
    if (lane < L) {
      const double x = xs[lane] - xs[i], y = ys[lane] - ys[i], w = zs[lane] - zs[i];
      vx[lane] = x;
      vy[lane] = y;
      vz[lane] = w;
      d = sqrt((x * x + y * y) + w * w);
      dist[lane] = d;
      wgt[lane] = 1.0 - sin(pi * d / (2.0 * cut));
    }
    __syncthreads();
    // stable ascending order of the distances; kept: d < cut
    if (lane < L) {
      int rank = 0;
      for (int j = 0; j < L; ++j) {
        const double dj = dist[j];
        rank += (dj < d || (dj == d && j < lane)) ? 1 : 0;
      }
      order[rank] = lane;
    }
    const int K = __popcll(__ballot(lane < L && d < cut));
    __syncthreads();
    // W: the kept weights in sorted order (every lane the same value)
    double W;
    {
      auto val = [&](int t) { return wgt[order[t]]; };  // 1 - sin(pi d / (2 cut)), arad.py:137
      if (K < 8) {
        W = 0.0;
        for (int t = 0; t < K; ++t) W += val(t);
      } else {
        double s8[8];
        for (int j = 0; j < 8; ++j) s8[j] = val(j);
        int t = 8;
        for (; t < K - (K % 8); t += 8)
          for (int j = 0; j < 8; ++j) s8[j] += val(t + j);
        W = ((s8[0] + s8[1]) + (s8[2] + s8[3])) + ((s8[4] + s8[5]) + (s8[6] + s8[7]));
        for (; t < K; ++t) W += val(t);
      }
    }
    // lane p: the row of the p-th kept neighbour m, summed over kept k
    double sc = 0.0, ss = 0.0;
    if (lane < K) {
      const int m = order[lane];
      const double mx = vx[m], my = vy[m], mz = vz[m], dm = dist[m];
      auto terms = [&](int t, double &c, double &s) {
        const int k = order[t];
        const double sp = (mx * vx[k] + my * vy[k]) + mz * vz[k];
        const double nrm = dm * dist[k];
        double ang = 0.0;
        if (fabs(sp - nrm) > eps && fabs(nrm) > eps) ang = acos(sp / nrm);
        c = cos(ang) * wgt[k];
        s = sin(ang) * wgt[k];
      };
      if (K < 8) {
        for (int t = 0; t < K; ++t) {
          double c, s;
          terms(t, c, s);
          sc += c;
          ss += s;
        }
      } else {
        double c8[8], s8[8];
        for (int j = 0; j < 8; ++j) terms(j, c8[j], s8[j]);
        int t = 8;
        for (; t < K - (K % 8); t += 8)
          for (int j = 0; j < 8; ++j) {
            double c, s;
            terms(t + j, c, s);
            c8[j] += c;
            s8[j] += s;
          }
        sc = ((c8[0] + c8[1]) + (c8[2] + c8[3])) + ((c8[4] + c8[5]) + (c8[6] + c8[7]));
        ss = ((s8[0] + s8[1]) + (s8[2] + s8[3])) + ((s8[4] + s8[5]) + (s8[6] + s8[7]));
        for (; t < K; ++t) {
          double c, s;
          terms(t, c, s);
          sc += c;
          ss += s;
        }
      }
    }
    for (int p = lane; p < amax; p += 64) {
      const bool kept = p < K;
      const int m = kept ? order[p] : 0;
      Mi[p] = kept ? dist[m] : 1e100;
      Mi[amax + p] = kept ? trow[m] : 0.0;
      Mi[2 * amax + p] = kept ? tcol[m] : 0.0;
      Mi[3 * amax + p] = kept ? sc / W : 0.0;
      Mi[4 * amax + p] = kept ? ss / W : 0.0;
    }
  }
}

}  // namespace

extern "C" int fmlh_arad_describe(const double *z, const double *r, const int32_t *offsets,
                                  int nmol, int amax, double cut, const double *ptp, int ptp_rows,
                                  double *q, double *zt, int32_t *n, int32_t *status,
                                  void *stream) {
  FMLH_ARG_CHECK(nmol >= 0, 4);
  FMLH_ARG_CHECK(amax >= 1 && amax <= kDescMaxAtoms, 5);
  FMLH_ARG_CHECK(cut > 0.0, 6);
  FMLH_ARG_CHECK(ptp_rows >= 1, 8);
  if (nmol == 0) return 0;
  FMLH_ARG_CHECK(z && r && offsets, 1);
  FMLH_ARG_CHECK(ptp != nullptr, 7);
  FMLH_ARG_CHECK(q && zt && n, 9);
  hipLaunchKernelGGL(k_arad_describe, dim3(nmol, amax), dim3(64), 0, as_stream(stream), z, r, offsets,
                     nmol, amax, cut, ptp, ptp_rows, q, zt, n, status);
  FMLH_LAUNCH_CHECK("k_arad_describe");
  return 0;
}
