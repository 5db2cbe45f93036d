// ARAS kernels: fget_kernels_aras and fget_symmetric_kernels_aras
// (fml/kernels/faras_kernels.f90:150-723, atom distance aras_utils::
// atomic_distl2 :7-147; Python entries get_atomic_kernels_aras and
// get_atomic_symmetric_kernels_aras, fml/kernels/kernels.py:408-560).
//
// Descriptors x (nm, amax, 3 + amax, amax) as ARAS.describe makes them:
// per atom, channel 0 the neighbour distances (ascending, 1e100 past the
// last), channel 1 their nuclear charges, channel 3 + m the angles between
// neighbour m and every other neighbour.
//
// Three launches per set and one for the pairs:
//   k_aras_prep (one wavefront per molecule, atom by atom): the neighbour
//     count nb (distances < cut, self included), the atom's distinct
//     neighbour elements in ascending order, ksi_m = 1/d_m^6 q0 q_m, and per
//     order o and element p the projections
//       cp[o][j][p] = sum_k (cos(o th) - cos((th + pi) o)) ksi3(j, k),
//       sp[o][j][p] = sum_k (sin(o th) - sin((th + pi) o)) ksi3(j, k)
//     over neighbours k != j of charge p (k ascending), th the j-k angle and
//     ksi3 = q0 qj qk (1 + cos th cos_i cos_j) / (di dj dij)^3 from the law
//     of cosines (faras_kernels.f90:258-361), each lane a neighbour j;
//   k_aras_self: atomic_distl2 of every atom with itself;
//   k_aras_pairs (one wavefront per molecule pair): per atom pair
//     l2 = self1 + self2 - 2 pd(q1, q2) D(a, b), zeroed below 5e-12, and
//     K_s(i, j) = sum over the atom pairs of exp(-l2 / sigma_s^2), the sum
//     in a fixed order (lane partials in atom-pair order, then a fixed
//     wavefront tree), so a pair's value does not depend on the batch.
// D(a, b) = pd(q0a, q0b) + sum over neighbours m1 of a, m2 of b (both past
// self, inside the cut) with (d2 - d1)^2 < (8 d_width)^2 of
//   exp((d2 - d1)^2 (-1/(4 d_width^2))) pd(q_m1, q_m2)
//   * (ksi_m1 ksi_m2 + scale sum_o s_o sum_{p1} sum_{p2}
//        (cp1[o][m1][p1] cp2[o][m2][p2] + sp1 sp2) pd(p2, p1)),
// p1, p2 over each atom's own neighbour elements in ascending order, every
// expression in the reference's order with contraction off.
#include <math.h>

#include "common.h"

#pragma clang fp contract(off)

namespace {

using namespace fmlh;

constexpr int kMaxAmax = 64;    // atoms per molecule (one lane each in the prep)
constexpr int kMaxTypes = 16;   // distinct elements within one atom's cut
constexpr int kMaxOrder = 8;
constexpr int kMaxSig = 8;

struct ArasParams {
  double s[kMaxOrder];   // g1 exp(-(t_width o)^2 / 2), o = 1..order
  int order;
  double inv_width;      // -1 / (4 d_width^2)
  double maxg2;          // (8 d_width)^2
  double cut;
  double scale;          // angular_scale
  int amax;
  int emax;
  const double *pd;      // Fortran (emax, emax)
  int nsig;
  double inv_sig[kMaxSig];  // -1 / sigma^2
};

struct ArasSet {
  const double *x;
  const int32_t *n;
  const int32_t *nn;  // optional neighbour counts (nm, amax); null: count d < cut
  int nm;
  int32_t *nb;      // (nm, amax)
  int32_t *nt;      // (nm, amax)
  int32_t *types;   // (nm, amax, kMaxTypes)
  double *ksi;      // (nm, amax, amax)
  double *cp, *sp;  // (nm, amax, order, amax, kMaxTypes)
  double *self;     // (nm, amax)
};

__device__ __forceinline__ double xval(const ArasSet &S, int amax, int a, int i, int c, int m) {
  return S.x[((((int64_t)a * amax + i) * (3 + amax) + c) * amax) + m];
}

// Fortran pd(p, q); NaN for a charge outside the table (flagged by the prep)
__device__ __forceinline__ double pdv(const ArasParams &P, int p, int q) {
  if (p < 1 || q < 1 || p > P.emax || q > P.emax) return __builtin_nan("");
  return P.pd[(int64_t)(q - 1) * P.emax + (p - 1)];
}

// Atoms of molecule a that the descriptor can hold.
__device__ __forceinline__ int mol_atoms(const int32_t *n, int a, int amax) {
  return max(0, min(n[a], amax));
}

__device__ __forceinline__ int64_t atom_index(const ArasParams &P, int a, int i) {
  return (int64_t)a * P.amax + i;
}

__global__ __launch_bounds__(64) void k_aras_prep(ArasSet S, ArasParams P, int32_t *status) {
  __shared__ double d_s[kMaxAmax], q_s[kMaxAmax];
  __shared__ int tix[kMaxAmax];          // element index of each neighbour
  __shared__ int types[kMaxTypes];
  __shared__ int ntypes_s;
  __shared__ double cacc[kMaxAmax * kMaxTypes], sacc[kMaxAmax * kMaxTypes];
  const int a = blockIdx.x;
  const int lane = threadIdx.x;
  const int amax = P.amax;
  const int n = mol_atoms(S.n, a, amax);
  if (lane == 0 && status && (S.n[a] > amax || S.n[a] < 0)) status[0] = 1;
  const double pi = 3.141592653589793;
  for (int i = 0; i < n; ++i) {
    const int64_t ai = atom_index(P, a, i);
    double d = 1e300;
    double q = 0.0;
    if (lane < amax) {
      d = xval(S, amax, a, i, 0, lane);
      q = xval(S, amax, a, i, 1, lane);
      d_s[lane] = d;
      q_s[lane] = q;
    }
    const int nb = S.nn ? max(0, min((int)S.nn[ai], amax))
                        : __popcll(__ballot(lane < amax && d < P.cut));
    __syncthreads();
    if (lane == 0) {
      // distinct neighbour elements, ascending (faras_kernels.f90:64-75)
      int nt = 0;
      bool bad = false;
      for (int m = 0; m < nb; ++m) {
        const double qm = q_s[m];
        const int p = (int)qm;
        if (!((double)p == qm && p >= 1 && p <= P.emax)) {
          bad = true;
          continue;
        }
        int pos = 0;
        while (pos < nt && types[pos] < p) ++pos;
        if (pos < nt && types[pos] == p) continue;
        if (nt == kMaxTypes) {
          bad = true;
          continue;
        }
        for (int t = nt; t > pos; --t) types[t] = types[t - 1];
        types[pos] = p;
        ++nt;
      }
      if (bad && status) status[0] = 1;
      ntypes_s = nt;
      S.nb[ai] = nb;
      S.nt[ai] = nt;
      for (int t = 0; t < kMaxTypes; ++t) S.types[ai * kMaxTypes + t] = t < nt ? types[t] : 0;
    }
    __syncthreads();
    const int nt = ntypes_s;
    if (lane < nb) {
      const int p = (int)q_s[lane];
      int t = 0;
      while (t < nt && types[t] != p) ++t;
      tix[lane] = t;  // nt for an invalid charge: never matched
    }
    // ksi (faras_kernels.f90:228-236)
    const double q0 = q_s[0];
    if (lane < amax) {
      double k = 0.0;
      if (lane >= 1 && lane < nb && d < P.cut) {
        const double d2 = d * d;
        const double d6 = d2 * d * d * d * d;
        k = 1.0 / d6 * q0 * q;
      }
      S.ksi[ai * amax + lane] = k;
    }
    __syncthreads();
    // projections, lane j (faras_kernels.f90:288-361)
    for (int o = 0; o < P.order; ++o) {
      const double mo = (double)(o + 1);
      for (int t = 0; t < kMaxTypes; ++t) {
        cacc[lane * kMaxTypes + t] = 0.0;
        sacc[lane * kMaxTypes + t] = 0.0;
      }
      if (lane >= 1 && lane < nb) {
        const int j = lane;
        const double dj0 = d_s[j];
        for (int k = 1; k < nb; ++k) {
          if (k == j) continue;
          // ksi3(j, k): di = d_j, dj = d_k, angle x(3 + j, k)
          const double di = dj0, dk = d_s[k];
          const double ang = xval(S, amax, a, i, 3 + j, k);
          const double cang = cos(ang);
          const double dij = sqrt(di * di + dk * dk - 2.0 * di * dk * cang);
          const double cosi = (dk * dk + dij * dij - di * di) / (2.0 * dk * dij);
          const double cosj = (di * di + dij * dij - dk * dk) / (2.0 * di * dij);
          const double p3 = di * dk * dij;
          const double k3 = q0 * q_s[j] * q_s[k] * (1.0 + cang * cosi * cosj) / (p3 * p3 * p3);
          const double th = ang;
          const int t = tix[k];
          if (t < nt) {
            cacc[j * kMaxTypes + t] = cacc[j * kMaxTypes + t] + (cos(mo * th) - cos((th + pi) * mo)) * k3;
            sacc[j * kMaxTypes + t] = sacc[j * kMaxTypes + t] + (sin(mo * th) - sin((th + pi) * mo)) * k3;
          }
        }
      }
      __syncthreads();
      if (lane < amax) {
        const int64_t base = ((ai * P.order + o) * amax + lane) * kMaxTypes;
        for (int t = 0; t < kMaxTypes; ++t) {
          S.cp[base + t] = cacc[lane * kMaxTypes + t];
          S.sp[base + t] = sacc[lane * kMaxTypes + t];
        }
      }
      __syncthreads();
    }
  }
}

// aras_utils::atomic_distl2 (faras_kernels.f90:7-147) for atom (a1, i1) of
// set A against atom (a2, i2) of set B.
__device__ double aras_distl2(const ArasSet &A, int a1, int i1, const ArasSet &B, int a2, int i2,
                              const ArasParams &P) {
  const int amax = P.amax;
  const int64_t x1 = atom_index(P, a1, i1), x2 = atom_index(P, a2, i2);
  const int nb1 = A.nb[x1], nb2 = B.nb[x2];
  const int nt1 = A.nt[x1], nt2 = B.nt[x2];
  const int32_t *T1 = A.types + x1 * kMaxTypes;
  const int32_t *T2 = B.types + x2 * kMaxTypes;
  const double *ks1 = A.ksi + x1 * amax;
  const double *ks2 = B.ksi + x2 * amax;
  double aad = pdv(P, (int)xval(A, amax, a1, i1, 1, 0), (int)xval(B, amax, a2, i2, 1, 0));
  for (int m1 = 1; m1 < nb1; ++m1) {
    const double d1 = xval(A, amax, a1, i1, 0, m1);
    if (d1 > P.cut) break;
    const int q1 = (int)xval(A, amax, a1, i1, 1, m1);
    for (int m2 = 1; m2 < nb2; ++m2) {
      const double d2 = xval(B, amax, a2, i2, 0, m2);
      if (d2 > P.cut) break;
      const double dd = d2 - d1;
      const double r2 = dd * dd;
      if (!(r2 < P.maxg2)) continue;
      const int q2 = (int)xval(B, amax, a2, i2, 1, m2);
      const double dw = exp(r2 * P.inv_width) * pdv(P, q1, q2);
      double angular = 0.0 * 0.0;
      for (int o = 0; o < P.order; ++o) {
        const double *c1 = A.cp + ((x1 * P.order + o) * amax + m1) * kMaxTypes;
        const double *s1 = A.sp + ((x1 * P.order + o) * amax + m1) * kMaxTypes;
        const double *c2 = B.cp + ((x2 * P.order + o) * amax + m2) * kMaxTypes;
        const double *s2 = B.sp + ((x2 * P.order + o) * amax + m2) * kMaxTypes;
        double temp = 0.0;
        for (int t1 = 0; t1 < nt1; ++t1) {
          const double cc = c1[t1], ss = s1[t1];
          const int p1 = T1[t1];
          for (int t2 = 0; t2 < nt2; ++t2)
            temp = temp + (cc * c2[t2] + ss * s2[t2]) * pdv(P, T2[t2], p1);
        }
        angular = angular + temp * P.s[o];
      }
      aad = aad + dw * (ks1[m1] * ks2[m2] + angular * P.scale);
    }
  }
  return aad;
}

__global__ void k_aras_self(ArasSet S, ArasParams P) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)S.nm * P.amax) return;
  const int a = (int)(t / P.amax), i = (int)(t % P.amax);
  if (i >= mol_atoms(S.n, a, P.amax)) return;
  S.self[t] = aras_distl2(S, a, i, S, a, i, P);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
  return v;
}

// One wavefront per molecule pair (i of set A, j of set B); symmetric:
// pairs i >= j of one set, written to both halves.
__global__ __launch_bounds__(64) void k_aras_pairs(ArasSet A, ArasSet B, ArasParams P, int sym,
                                                   double *__restrict__ out) {
  __shared__ double l2s[kMaxAmax * kMaxAmax];
  int i, j;
  const int64_t w = blockIdx.x;
  if (sym) {
    // w = i (i + 1) / 2 + j, j <= i
    i = (int)((sqrt(8.0 * (double)w + 1.0) - 1.0) * 0.5);
    while ((int64_t)(i + 1) * (i + 2) / 2 <= w) ++i;
    while ((int64_t)i * (i + 1) / 2 > w) --i;
    j = (int)(w - (int64_t)i * (i + 1) / 2);
  } else {
    i = (int)(w % A.nm);
    j = (int)(w / A.nm);
  }
  const int lane = threadIdx.x;
  const int ni = mol_atoms(A.n, i, P.amax), nj = mol_atoms(B.n, j, P.amax);
  const int np = ni * nj;
  for (int t = lane; t < np; t += 64) {
    const int i1 = t % ni, j1 = t / ni;  // the reference's column-major atom block
    const double l = aras_distl2(A, i, i1, B, j, j1, P);
    const double q1 = xval(A, P.amax, i, i1, 1, 0), q2 = xval(B, P.amax, j, j1, 1, 0);
    double l2 = A.self[atom_index(P, i, i1)] + B.self[atom_index(P, j, j1)] -
                2.0 * l * pdv(P, (int)q1, (int)q2);
    if (fabs(l2) < 5.0e-12) l2 = 0.0;
    l2s[t] = l2;
  }
  __syncthreads();
  for (int s = 0; s < P.nsig; ++s) {
    double v = 0.0;
    for (int t = lane; t < np; t += 64) v = v + exp(l2s[t] * P.inv_sig[s]);
    v = wave_sum(v);
    if (lane == 0) {
      out[((int64_t)j * A.nm + i) * P.nsig + s] = v;
      if (sym) out[((int64_t)i * A.nm + j) * P.nsig + s] = v;
    }
  }
}

__global__ void k_zero_status1(int32_t *s) { s[0] = 0; }

// Workspace of one set's tables.
size_t set_bytes(int nm, int amax, int order) {
  const size_t na = (size_t)nm * amax;
  size_t b = 0;
  b += WsCarver::need(sizeof(int32_t) * na) * 2;
  b += WsCarver::need(sizeof(int32_t) * na * kMaxTypes);
  b += WsCarver::need(sizeof(double) * na * amax);
  b += WsCarver::need(sizeof(double) * na * order * amax * kMaxTypes) * 2;
  b += WsCarver::need(sizeof(double) * na);
  return b;
}

bool carve_set(WsCarver &wc, ArasSet &S, int nm, int amax, int order) {
  const size_t na = (size_t)nm * amax;
  S.nb = wc.take<int32_t>(na);
  S.nt = wc.take<int32_t>(na);
  S.types = wc.take<int32_t>(na * kMaxTypes);
  S.ksi = wc.take<double>(na * amax);
  S.cp = wc.take<double>(na * order * amax * kMaxTypes);
  S.sp = wc.take<double>(na * order * amax * kMaxTypes);
  S.self = wc.take<double>(na);
  return S.nb && S.nt && S.types && S.ksi && S.cp && S.sp && S.self;
}

}  // namespace

extern "C" size_t fmlh_aras_ws_bytes(int nm1, int nm2, int amax, int order, int mode) {
  if (nm1 < 0 || nm2 < 0 || amax < 1 || order < 1) return 0;
  return set_bytes(nm1, amax, order) + (mode == 1 ? 0 : set_bytes(nm2, amax, order));
}

extern "C" int fmlh_aras_kernels(int mode, const double *x1, const int32_t *n1,
                                 const int32_t *nn1, int nm1, const double *x2,
                                 const int32_t *n2, const int32_t *nn2, int nm2, int amax,
                                 const double *sigmas, int nsig, double t_width, double d_width,
                                 double cut, int order, const double *pd, int emax,
                                 double angular_scale, double *out, int32_t *status, void *ws,
                                 size_t ws_bytes, void *stream) {
  FMLH_ARG_CHECK(mode == 0 || mode == 1, 1);
  FMLH_ARG_CHECK(nm1 >= 0, 5);
  FMLH_ARG_CHECK(mode == 1 || nm2 >= 0, 9);
  FMLH_ARG_CHECK(amax >= 1 && amax <= kMaxAmax, 10);
  FMLH_ARG_CHECK(nsig >= 1 && nsig <= kMaxSig && sigmas != nullptr, 12);
  for (int s = 0; s < nsig; ++s) FMLH_ARG_CHECK(sigmas[s] > 0.0, 11);
  FMLH_ARG_CHECK(t_width > 0.0, 13);
  FMLH_ARG_CHECK(d_width > 0.0, 14);
  FMLH_ARG_CHECK(cut > 0.0, 15);
  FMLH_ARG_CHECK(order >= 1 && order <= kMaxOrder, 16);
  FMLH_ARG_CHECK(emax >= 1, 18);
  if (mode == 1) nm2 = nm1;
  if (nm1 == 0 || nm2 == 0) return 0;
  FMLH_ARG_CHECK(x1 != nullptr && n1 != nullptr, 2);
  FMLH_ARG_CHECK(mode == 1 || (x2 != nullptr && n2 != nullptr), 6);
  FMLH_ARG_CHECK(pd != nullptr, 17);
  FMLH_ARG_CHECK(out != nullptr, 20);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= fmlh_aras_ws_bytes(nm1, nm2, amax, order, mode), 22);

  // angular normalisation and order weights (faras_kernels.f90:198-207, 84-88)
  const double pi = 3.141592653589793;
  double an = 0.0;
  for (int k = -10000; k <= 10000; ++k) {
    const double tn = t_width * k;
    an = an + exp(-(tn * tn)) * (2.0 - 2.0 * cos(k * pi));
  }
  an = sqrt(an * pi) * 2.0;
  ArasParams P = {};
  const double g1 = sqrt(2.0 * pi) / an;
  for (int o = 0; o < order; ++o) {
    const double tm = t_width * (o + 1);
    P.s[o] = g1 * exp(-(tm * tm) / 2.0);
  }
  P.order = order;
  P.inv_width = -1.0 / (4.0 * (d_width * d_width));
  const double g8 = 8.0 * d_width;
  P.maxg2 = g8 * g8;
  P.cut = cut;
  P.scale = angular_scale;
  P.amax = amax;
  P.emax = emax;
  P.pd = pd;
  P.nsig = nsig;
  for (int s = 0; s < nsig; ++s) P.inv_sig[s] = -1.0 / (sigmas[s] * sigmas[s]);

  WsCarver wc(ws, ws_bytes);
  ArasSet A = {}, B = {};
  A.x = x1;
  A.n = n1;
  A.nn = nn1;
  A.nm = nm1;
  if (!carve_set(wc, A, nm1, amax, order)) return -22;
  if (mode == 0) {
    B.x = x2;
    B.n = n2;
    B.nn = nn2;
    B.nm = nm2;
    if (!carve_set(wc, B, nm2, amax, order)) return -22;
  } else {
    B = A;
  }
  const hipStream_t st = as_stream(stream);
  if (status) {
    hipLaunchKernelGGL(k_zero_status1, dim3(1), dim3(1), 0, st, status);
    FMLH_LAUNCH_CHECK("k_zero_status1");
  }
  hipLaunchKernelGGL(k_aras_prep, dim3(nm1), dim3(64), 0, st, A, P, status);
  FMLH_LAUNCH_CHECK("k_aras_prep");
  hipLaunchKernelGGL(k_aras_self, dim3(ceil_div(nm1 * amax, 128)), dim3(128), 0, st, A, P);
  FMLH_LAUNCH_CHECK("k_aras_self");
  if (mode == 0) {
    hipLaunchKernelGGL(k_aras_prep, dim3(nm2), dim3(64), 0, st, B, P, status);
    FMLH_LAUNCH_CHECK("k_aras_prep");
    hipLaunchKernelGGL(k_aras_self, dim3(ceil_div(nm2 * amax, 128)), dim3(128), 0, st, B, P);
    FMLH_LAUNCH_CHECK("k_aras_self");
  }
  const int64_t nw = mode == 1 ? (int64_t)nm1 * (nm1 + 1) / 2 : (int64_t)nm1 * nm2;
  if (nw > 0x7fffffff) return -4;
  timing_begin(st);
  hipLaunchKernelGGL(k_aras_pairs, dim3((unsigned)nw), dim3(64), 0, st, A, B, P, mode, out);
  timing_end(st);
  FMLH_LAUNCH_CHECK("k_aras_pairs");
  return 0;
}
