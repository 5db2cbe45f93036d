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
// Launches: k_aras_init; per set k_aras_prep (one wavefront per atom): the
// neighbour count nb (distances < cut, self included),
// the atom's distinct neighbour elements in ascending order, ksi_m =
// 1/d_m^6 q0 q_m, and per order o and element p the projections
//   cp[o][j][p] = sum_k (cos(o th) - cos((th + pi) o)) ksi3(j, k),
//   sp[o][j][p] = sum_k (sin(o th) - sin((th + pi) o)) ksi3(j, k)
// over neighbours k != j of charge p (k ascending), th the j-k angle and
// ksi3 = q0 qj qk (1 + cos th cos_i cos_j) / (di dj dij)^3 from the law of
// cosines (faras_kernels.f90:258-361), each lane a neighbour j;
// k_aras_elems: the elements of both sets as one ascending list; per set
// k_aras_proj (element indices into that list, the projections by list
// index) and k_aras_self (atomic_distl2 of every atom with itself);
// k_aras_tiles: per atom pair l2 = self1 + self2 - 2 pd(q1, q2) D(a, b),
// zeroed below 5e-12, and K_s(i, j) = sum over the atom pairs of
// exp(-l2 / sigma_s^2) in a fixed order (per set-1 atom over the set-2
// atoms in order, then over the set-1 atoms in order), so a pair's value
// does not depend on the batch.
// D(a, b) = pd(q0a, q0b) + sum over neighbours m1 of a, m2 of b (both past
// self, inside the cut) with (d2 - d1)^2 < (8 d_width)^2 of
//   exp((d2 - d1)^2 (-1/(4 d_width^2))) pd(q_m1, q_m2)
//   * (ksi_m1 ksi_m2 + scale sum_o s_o sum_{p1} sum_{p2}
//        (cp1[o][m1][p1] cp2[o][m2][p2] + sp1 sp2) pd(p2, p1)),
// p1, p2 over each atom's own neighbour elements in ascending order, every
// expression in the reference's order with contraction off.
#include <math.h>

#include "common.h"
#include "exp_neg.h"

#pragma clang fp contract(off)

namespace {

using namespace fmlh;

__constant__ ExpPair c_exp_pairs[128] = {FMLH_EXP2_128_PAIRS};

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
  double *cs;       // (nm, amax, order, amax, kMaxTypes, 2): (cp, sp) per the atom's own elements
  double *csg;      // the same per element of the data set's list, zero where absent
  double *self;     // (nm, amax)
  double *dist;     // (nm, amax, amax) neighbour distances
  int32_t *gq;      // (nm, amax, amax) neighbour element as an index into the
                    // data set's element list (G: a charge outside it)
  int32_t *tg;      // (nm, amax, kMaxTypes) the atom's elements as indices into that list
};

// Elements present in either set (bitmask, then an ascending list).
struct ArasElems {
  uint32_t bits[4];
  int32_t count;
  int32_t list[kMaxTypes];
  int32_t gidx[128];   // element -> index into list, count for none
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

__global__ __launch_bounds__(64) void k_aras_prep(ArasSet S, ArasParams P, ArasElems *E,
                                                  int32_t *status) {
  __shared__ double d_s[kMaxAmax], q_s[kMaxAmax];
  __shared__ int tix[kMaxAmax];          // element index of each neighbour
  __shared__ int types[kMaxTypes];
  __shared__ int ntypes_s;
  __shared__ double cacc[kMaxAmax * kMaxTypes], sacc[kMaxAmax * kMaxTypes];
  const int lane = threadIdx.x;
  const int amax = P.amax;
  const int a = (int)(blockIdx.x / amax), i = (int)(blockIdx.x % amax);
  const int n = mol_atoms(S.n, a, amax);
  if (i == 0 && lane == 0 && status && (S.n[a] > amax || S.n[a] < 0)) status[0] = 1;
  if (i >= n) return;
  const double pi = 3.141592653589793;
  {
    const int64_t ai = atom_index(P, a, i);
    double d = 1e300;
    double q = 0.0;
    if (lane < amax) {
      d = xval(S, amax, a, i, 0, lane);
      q = xval(S, amax, a, i, 1, lane);
      d_s[lane] = d;
      q_s[lane] = q;
      S.dist[ai * amax + lane] = d;
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
        if (p < 128)
          atomicOr(&E->bits[p >> 5], 1u << (p & 31));
        else
          bad = true;
      }
      // the atom's own charge (used even when no neighbour is inside the cut)
      const int p0 = (int)q_s[0];
      if ((double)p0 == q_s[0] && p0 >= 1 && p0 <= P.emax && p0 < 128)
        atomicOr(&E->bits[p0 >> 5], 1u << (p0 & 31));
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
          S.cs[2 * (base + t)] = cacc[lane * kMaxTypes + t];
          S.cs[2 * (base + t) + 1] = sacc[lane * kMaxTypes + t];
        }
      }
      __syncthreads();
    }
  }
}

// The data set's element list from the bitmask (one thread).
__global__ void k_aras_elems(ArasElems *E, int32_t *status) {
  int c = 0;
  for (int p = 1; p < 128; ++p)
    if (E->bits[p >> 5] & (1u << (p & 31))) {
      if (c < kMaxTypes)
        E->list[c] = p;
      else if (status)
        status[0] = 1;
      ++c;
    }
  E->count = min(c, kMaxTypes);
  for (int p = 0; p < 128; ++p) E->gidx[p] = E->count;
  for (int g = 0; g < E->count; ++g) E->gidx[E->list[g]] = g;
}

// Per atom: each neighbour's element (lane m) and each of the atom's own
// elements (lane t) as an index into the data set's element list, for the
// element similarity table that the pair kernels keep in LDS.
__global__ __launch_bounds__(64) void k_aras_proj(ArasSet S, ArasParams P, const ArasElems *E) {
  const int a = blockIdx.x;
  const int lane = threadIdx.x;
  const int amax = P.amax;
  const int n = mol_atoms(S.n, a, amax);
  const int G = E->count;
  for (int i = 0; i < n; ++i) {
    const int64_t ai = atom_index(P, a, i);
    if (lane < amax) {
      const int nb = S.nb[ai];
      int g = G;
      if (lane < nb || lane == 0) {
        const double qm = xval(S, amax, a, i, 1, lane);
        const int p = (int)qm;
        if ((double)p == qm && p >= 1 && p < 128) g = E->gidx[p];
      }
      S.gq[ai * amax + lane] = g;
    }
    if (lane < kMaxTypes) {
      const int p = S.types[ai * kMaxTypes + lane];
      S.tg[ai * kMaxTypes + lane] = (lane < S.nt[ai] && p >= 1 && p < 128) ? E->gidx[p] : G;
    }
    // (cp, sp) by the data set's element index, lane m a neighbour
    if (lane < amax) {
      const int nt = S.nt[ai];
      const int32_t *T = S.types + ai * kMaxTypes;
      for (int o = 0; o < P.order; ++o) {
        const int64_t base = 2 * (((ai * P.order + o) * amax + lane) * kMaxTypes);
        double c[kMaxTypes], sn[kMaxTypes];
#pragma unroll
        for (int g = 0; g < kMaxTypes; ++g) c[g] = sn[g] = 0.0;
        for (int t = 0; t < nt; ++t) {
          const int p = T[t];
          const int g = (p >= 1 && p < 128) ? E->gidx[p] : G;
#pragma unroll
          for (int gg = 0; gg < kMaxTypes; ++gg)
            if (gg == g) {
              c[gg] = S.cs[base + 2 * t];
              sn[gg] = S.cs[base + 2 * t + 1];
            }
        }
#pragma unroll
        for (int g = 0; g < kMaxTypes; ++g) {
          S.csg[base + 2 * g] = c[g];
          S.csg[base + 2 * g + 1] = sn[g];
        }
      }
    }
  }
}

// Element similarity between the data set's elements, pdg[a (G + 1) + b] =
// pd(list[a], list[b]); row and column G (padding, and a charge outside the
// list, which the status flags) and the rest of the kMaxTypes + 1 square
// hold 0.
__device__ void load_pdg(double *pdg, const ArasParams &P, const ArasElems *E) {
  const int G = E->count;
  for (int t = threadIdx.x; t < (kMaxTypes + 1) * (kMaxTypes + 1); t += blockDim.x) {
    const int a = t / (G + 1), b = t % (G + 1);
    pdg[t] = (a < G && b < G) ? pdv(P, E->list[a], E->list[b]) : 0.0;
  }
}

// aras_utils::atomic_distl2 (faras_kernels.f90:7-147) of every atom with
// itself, one wavefront per atom.  The reference's loops: m1 and m2 over the
// neighbours past self, each loop left at its first distance above the cut
// (the same index mc for both, one list), pairs with (d2 - d1)^2 <
// (8 d_width)^2 adding dw (ksi_m1 ksi_m2 + scale sum_o s_o sum_{p1} sum_{p2}
// (cp[p1] cp[p2] + sp[p1] sp[p2]) pd(p2, p1)) in m1-major order.  Lanes
// evaluate 64 consecutive (m1, m2) pairs of that order at a time (0 for a
// pair outside the window: adding +0 leaves the sum unchanged), then lane 0
// adds them in order.
__global__ __launch_bounds__(64) void k_aras_self(ArasSet S, ArasParams P, const ArasElems *E) {
  __shared__ double pdg[(kMaxTypes + 1) * (kMaxTypes + 1)];
  __shared__ double term[64];
  __shared__ ExpPair etab[128];
  load_pdg(pdg, P, E);
  for (int t = threadIdx.x; t < 128; t += blockDim.x) etab[t] = c_exp_pairs[t];
  __syncthreads();
  const int amax = P.amax;
  const int a = (int)(blockIdx.x / amax), i = (int)(blockIdx.x % amax);
  if (i >= mol_atoms(S.n, a, amax)) return;
  const int lane = threadIdx.x;
  const int G1 = E->count + 1;
  const int64_t x = atom_index(P, a, i);
  const int nb = S.nb[x];
  const double *dx = S.dist + x * amax;
  const int32_t *gx = S.gq + x * amax;
  const double *kx = S.ksi + x * amax;
  const int nt = S.nt[x];
  const int32_t *tg = S.tg + x * kMaxTypes;
  // first neighbour past self beyond the cut
  const bool beyond = lane >= 1 && lane < nb && dx[lane] > P.cut;
  const uint64_t bb = __ballot(beyond);
  const int mc = bb ? (int)__ffsll((long long)bb) - 1 : max(nb, 1);
  const int K = max(mc - 1, 0);
  double aad = pdg[gx[0] * G1 + gx[0]];
  for (int base = 0; base < K * K; base += 64) {
    const int k = base + lane;
    double v = 0.0;
    if (k < K * K) {
      const int m1 = 1 + k / K, m2 = 1 + k % K;
      const double d1 = dx[m1], d2 = dx[m2];
      const double dd = d2 - d1;
      const double r2 = dd * dd;
      if (r2 < P.maxg2) {
        const double dw = exp_neg(r2 * P.inv_width, etab) * pdg[gx[m1] * G1 + gx[m2]];
        double angular = 0.0 * 0.0;
        for (int o = 0; o < P.order; ++o) {
          const double *c1 = S.cs + 2 * ((x * P.order + o) * amax + m1) * kMaxTypes;
          const double *c2 = S.cs + 2 * ((x * P.order + o) * amax + m2) * kMaxTypes;
          double temp = 0.0;
          for (int t1 = 0; t1 < nt; ++t1) {
            const double cc = c1[2 * t1], ss = c1[2 * t1 + 1];
            for (int t2 = 0; t2 < nt; ++t2)
              temp = temp + (cc * c2[2 * t2] + ss * c2[2 * t2 + 1]) * pdg[tg[t2] * G1 + tg[t1]];
          }
          angular = angular + temp * P.s[o];
        }
        v = dw * (kx[m1] * kx[m2] + angular * P.scale);
      }
    }
    term[lane] = v;
    __syncthreads();
    if (lane == 0) {
      const int cnt = min(64, K * K - base);
      for (int l = 0; l < cnt; ++l) aad = aad + term[l];
    }
    __syncthreads();
  }
  if (lane == 0) S.self[x] = aad;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
  return v;
}

constexpr int kJB = 1;   // set-B molecules per workgroup

// Set-B tables the tile kernel reads at wave-uniform addresses.
struct ArasUniform {
  const int32_t *__restrict__ n;
  const int32_t *__restrict__ nb;
  const int32_t *__restrict__ nt;
  const int32_t *__restrict__ tg;
  const int32_t *__restrict__ gq;
  const double *__restrict__ dist;
  const double *__restrict__ ksi;
  const double *__restrict__ cs;
  const double *__restrict__ self;
};

// One lane's set-A atom (per-lane values) for the tile kernel.
struct LaneAtom {
  int64_t x1;
  int nb1;
  bool valid;
};

// atomic_distl2 of the lane's set-A atom against set-B atom x2 (the same
// for every lane), accumulated on aad in the reference's order: m1 over
// the lane's neighbours, m2 over x2's, and for the angular term p1 over
// the data set's elements g (ascending; the lane's projections are zero
// for elements its atom lacks, and elements no lane of the block has --
// bit clear in gmask -- are skipped: both only drop exact zeros) and p2
// over x2's own elements t (ascending).  NT2 > 0: x2 has exactly NT2
// elements, whose similarity row pd(p2_t, g) sits in registers; NT2 == 0:
// any count nt2, similarities read from LDS.  The window factor's exponent
// lies in (-16, 0] ((8 d_width)^2 (-1/(4 d_width^2)) = -16), where exp_neg
// (exp_neg.h: Tang's table-driven method, a head/tail 2^(j/128) table in
// LDS, rounding like a correctly rounded exp but for ~0.004 ulp around
// midpoints) evaluates it.
template <int GB, int NT2>
__device__ __forceinline__ double atom_pair(const ArasSet &A, const ArasUniform &B, const ArasParams &P,
                                            const double *pdg, const ExpPair *etab, int G1,
                                            const LaneAtom &L, unsigned gmask, int64_t x2,
                                            int nt2, double aad) {
  const int amax = P.amax;
  const int nb2 = B.nb[x2];
  constexpr int NT = NT2 > 0 ? NT2 : GB;
  int tg2[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) tg2[t] = (NT2 > 0 || t < nt2) ? B.tg[x2 * kMaxTypes + t] : G1 - 1;
  double pw[NT2 > 0 ? NT2 : 1][GB];
  if (NT2 > 0) {
#pragma unroll
    for (int t = 0; t < (NT2 > 0 ? NT2 : 1); ++t)
#pragma unroll
      for (int g = 0; g < GB; ++g) pw[t][g] = pdg[tg2[t] * G1 + g];
  }
  const double *d2p = B.dist + x2 * amax;
  const int32_t *g2p = B.gq + x2 * amax;
  const double *k2p = B.ksi + x2 * amax;
  bool alive = L.valid;
  for (int m1 = 1;; ++m1) {
    bool act = alive && m1 < L.nb1;
    if (__ballot(act) == 0) break;
    double d1 = 0.0;
    if (act) {
      d1 = A.dist[L.x1 * amax + m1];
      if (d1 > P.cut) {   // the reference leaves the m1 loop here
        alive = false;
        act = false;
      }
    }
    int g1 = G1 - 1;
    double k1 = 0.0;
    double c1[GB], s1[GB];
#pragma unroll
    for (int g = 0; g < GB; ++g) c1[g] = s1[g] = 0.0;
    if (act) {
      g1 = A.gq[L.x1 * amax + m1];
      k1 = A.ksi[L.x1 * amax + m1];
      const double *cp = A.csg + 2 * ((L.x1 * P.order) * amax + m1) * kMaxTypes;
#pragma unroll
      for (int g = 0; g < GB; ++g) {
        c1[g] = cp[2 * g];
        s1[g] = cp[2 * g + 1];
      }
    }
    for (int m2 = 1; m2 < nb2; ++m2) {
      const double d2 = d2p[m2];
      if (d2 > P.cut) break;
      const double dd = d2 - d1;
      const double r2 = dd * dd;
      if (act && r2 < P.maxg2) {
        const double dw = exp_neg(r2 * P.inv_width, etab) * pdg[g1 * G1 + g2p[m2]];
        double angular = 0.0 * 0.0;
        for (int o = 0; o < P.order; ++o) {
          const double *c2 = B.cs + 2 * ((x2 * P.order + o) * amax + m2) * kMaxTypes;
          const double *cq = A.csg + 2 * ((L.x1 * P.order + o) * amax + m1) * kMaxTypes;
          double c2r[NT], s2r[NT];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            c2r[t] = (NT2 > 0 || t < nt2) ? c2[2 * t] : 0.0;
            s2r[t] = (NT2 > 0 || t < nt2) ? c2[2 * t + 1] : 0.0;
          }
          double temp = 0.0;
#pragma unroll
          for (int g = 0; g < GB; ++g) {
            if (!((gmask >> g) & 1u)) continue;
            const double cc = o == 0 ? c1[g] : cq[2 * g];
            const double ss = o == 0 ? s1[g] : cq[2 * g + 1];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
              if (NT2 > 0) {
                temp = temp + (cc * c2r[t] + ss * s2r[t]) * pw[t][g];
              } else if (t < nt2) {
                temp = temp + (cc * c2r[t] + ss * s2r[t]) * pdg[tg2[t] * G1 + g];
              }
            }
          }
          angular = angular + temp * P.s[o];
        }
        aad = aad + dw * (k1 * k2p[m2] + angular * P.scale);
      }
    }
  }
  return aad;
}

// Molecule-pair kernels over tiles: a workgroup (one wavefront) takes a
// block of set-A molecules, one atom per lane (at most 64 atoms: molecules
// ord[blk[b]] .. ord[blk[b + 1] - 1], ord the identity when NULL), against
// kJB molecules of set B, whose atoms it walks one at a time, so every
// set-B value is the same for all lanes.  Each lane forms atomic_distl2 of
// its atom against the current set-B atom (atom_pair), then adds
// exp(-l2 / sigma^2) over the set-B atoms in the molecule's atom order; the
// molecule pair's value is the sum of its lanes in atom order.  Symmetric:
// only pairs i >= j, written to both halves.  GB is the smallest of
// {4, 5, 6, 8, 16} that holds the data set's element count (the other
// instantiations return at once).
template <int GB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_aras_tiles(ArasSet A, ArasUniform B, int nm2, ArasParams P,
                                                   const int32_t *__restrict__ blk,
                                                   const int32_t *__restrict__ ord, int nblk,
                                                   int per_blk, int sym, int g_lo,
                                                   const ArasElems *__restrict__ E,
                                                   double *__restrict__ out,
                                                   int32_t *__restrict__ status) {
  const int G = E->count;
  if (G > GB || G <= g_lo) return;
  __shared__ double pdg[(kMaxTypes + 1) * (kMaxTypes + 1)];
  __shared__ double red[64 * kMaxSig];
  __shared__ int lane_mol[64], lane_atom[64], natoms_s, maxmol_s;
  __shared__ ExpPair etab[128];
  const int lane = threadIdx.x;
  const int64_t w = blockIdx.x;
  const int ib = (int)(w % nblk);
  const int j0 = (int)(w / nblk) * kJB;
  const int b0 = blk ? blk[ib] : ib * per_blk;
  const int b1 = blk ? blk[ib + 1] : min(A.nm, (ib + 1) * per_blk);
  load_pdg(pdg, P, E);
  for (int t = lane; t < 128; t += 64) etab[t] = c_exp_pairs[t];
  if (lane == 0) {
    int c = 0, mx = -1;
    for (int b = b0; b < b1; ++b) {
      const int m = ord ? ord[b] : b;
      mx = max(mx, m);
      const int nm_ = mol_atoms(A.n, m, P.amax);
      for (int i1 = 0; i1 < nm_; ++i1) {
        if (c < 64) {
          lane_mol[c] = m;
          lane_atom[c] = i1;
        }
        ++c;
      }
    }
    if (c > 64 && status) status[0] = 1;
    natoms_s = min(c, 64);
    maxmol_s = mx;
  }
  __syncthreads();
  const int maxmol = maxmol_s;
  if (sym && j0 > maxmol) return;   // every pair of the tile has i < j
  const int amax = P.amax;
  const int G1 = G + 1;
  LaneAtom L;
  L.valid = lane < natoms_s;
  const int mi = L.valid ? lane_mol[lane] : 0;
  L.x1 = L.valid ? atom_index(P, mi, lane_atom[lane]) : 0;
  L.nb1 = L.valid ? A.nb[L.x1] : 0;
  const int g01 = L.valid ? A.gq[L.x1 * amax] : G;
  const double self1 = L.valid ? A.self[L.x1] : 0.0;
  const bool first = L.valid && lane_atom[lane] == 0;
  const int ni = L.valid ? mol_atoms(A.n, mi, amax) : 0;
  // elements held by some lane's atom
  unsigned lm = 0u;
  if (L.valid)
    for (int t = 0; t < A.nt[L.x1]; ++t) lm |= 1u << A.tg[L.x1 * kMaxTypes + t];
  unsigned gmask = 0u;
#pragma unroll
  for (int g = 0; g < GB; ++g)
    if (__ballot((lm >> g) & 1u) != 0) gmask |= 1u << g;

  for (int jj = 0; jj < kJB; ++jj) {
    const int j = j0 + jj;
    if (j >= nm2) break;
    if (sym && j > maxmol) break;
    LaneAtom Lj = L;
    if (sym && mi < j) Lj.valid = false;   // this pair is the other orientation's
    // the lane's sums over the set-B atoms, per sigma, in LDS (registers go
    // to the atom-pair walk)
    double *e = red + lane * kMaxSig;
    for (int s = 0; s < P.nsig; ++s) e[s] = 0.0;
    const int nj = mol_atoms(B.n, j, amax);
    for (int j1 = 0; j1 < nj; ++j1) {
      const int64_t x2 = atom_index(P, j, j1);
      const int g02 = B.gq[x2 * amax];
      double aad = pdg[g01 * G1 + g02];
      const int nt2 = B.nt[x2];
      if constexpr (GB <= 6) {
        switch (nt2) {
          case 1: aad = atom_pair<GB, 1>(A, B, P, pdg, etab, G1, Lj, gmask, x2, nt2, aad); break;
          case 2: aad = atom_pair<GB, 2>(A, B, P, pdg, etab, G1, Lj, gmask, x2, nt2, aad); break;
          case 3: aad = atom_pair<GB, 3>(A, B, P, pdg, etab, G1, Lj, gmask, x2, nt2, aad); break;
          case 4: aad = atom_pair<GB, (GB >= 4 ? 4 : 0)>(A, B, P, pdg, etab, G1, Lj, gmask, x2, nt2, aad); break;
          case 5: aad = atom_pair<GB, (GB >= 5 ? 5 : 0)>(A, B, P, pdg, etab, G1, Lj, gmask, x2, nt2, aad); break;
          default: aad = atom_pair<GB, 0>(A, B, P, pdg, etab, G1, Lj, gmask, x2, nt2, aad); break;
        }
      } else {
        aad = atom_pair<GB, 0>(A, B, P, pdg, etab, G1, Lj, gmask, x2, nt2, aad);
      }
      if (Lj.valid) {
        double l2 = self1 + B.self[x2] - 2.0 * aad * pdg[g01 * G1 + g02];
        if (fabs(l2) < 5.0e-12) l2 = 0.0;
        for (int s = 0; s < P.nsig; ++s) e[s] = e[s] + exp(l2 * P.inv_sig[s]);
      }
    }
    __syncthreads();
    if (first && (!sym || mi >= j)) {
      for (int s = 0; s < P.nsig; ++s) {
        double v = red[lane * kMaxSig + s];
        for (int k = 1; k < ni && lane + k < 64; ++k) v = v + red[(lane + k) * kMaxSig + s];
        out[((int64_t)j * A.nm + mi) * P.nsig + s] = v;
        if (sym) out[((int64_t)mi * A.nm + j) * P.nsig + s] = v;
      }
    }
    __syncthreads();
  }
}

__global__ void k_aras_init(int32_t *status, ArasElems *E) {
  if (status) status[0] = 0;
  for (int k = 0; k < 4; ++k) E->bits[k] = 0u;
}

// Workspace of one set's tables.
size_t set_bytes(int nm, int amax, int order) {
  const size_t na = (size_t)nm * amax;
  size_t b = 0;
  b += WsCarver::need(sizeof(int32_t) * na) * 2;
  b += WsCarver::need(sizeof(int32_t) * na * kMaxTypes);
  b += WsCarver::need(sizeof(double) * na * amax) * 2;
  b += WsCarver::need(sizeof(int32_t) * na * amax);
  b += WsCarver::need(sizeof(int32_t) * na * kMaxTypes);
  b += WsCarver::need(sizeof(double) * 2 * na * order * amax * kMaxTypes) * 2;
  b += WsCarver::need(sizeof(double) * na);
  return b;
}

bool carve_set(WsCarver &wc, ArasSet &S, int nm, int amax, int order) {
  const size_t na = (size_t)nm * amax;
  S.nb = wc.take<int32_t>(na);
  S.nt = wc.take<int32_t>(na);
  S.types = wc.take<int32_t>(na * kMaxTypes);
  S.ksi = wc.take<double>(na * amax);
  S.cs = wc.take<double>(2 * na * order * amax * kMaxTypes);
  S.csg = wc.take<double>(2 * na * order * amax * kMaxTypes);
  S.self = wc.take<double>(na);
  S.dist = wc.take<double>(na * amax);
  S.gq = wc.take<int32_t>(na * amax);
  S.tg = wc.take<int32_t>(na * kMaxTypes);
  return S.nb && S.nt && S.types && S.ksi && S.cs && S.csg && S.self && S.dist && S.gq && S.tg;
}

}  // namespace

extern "C" size_t fmlh_aras_ws_bytes(int nm1, int nm2, int amax, int order, int mode) {
  if (nm1 < 0 || nm2 < 0 || amax < 1 || order < 1) return 0;
  return WsCarver::need(sizeof(ArasElems)) + set_bytes(nm1, amax, order) +
         (mode == 1 ? 0 : set_bytes(nm2, amax, order));
}

extern "C" int fmlh_aras_kernels(int mode, const double *x1, const int32_t *n1,
                                 const int32_t *nn1, int nm1, const double *x2,
                                 const int32_t *n2, const int32_t *nn2, int nm2, int amax,
                                 const double *sigmas, int nsig, double t_width, double d_width,
                                 double cut, int order, const double *pd, int emax,
                                 double angular_scale, const int32_t *blk1, const int32_t *ord1,
                                 int nblk1, double *out, int32_t *status, void *ws, size_t ws_bytes,
                                 void *stream) {
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
  FMLH_ARG_CHECK(blk1 == nullptr || nblk1 >= 1, 22);
  FMLH_ARG_CHECK(out != nullptr, 23);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= fmlh_aras_ws_bytes(nm1, nm2, amax, order, mode), 25);

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
  ArasElems *E = wc.take<ArasElems>(1);
  if (E == nullptr) return -25;
  ArasSet A = {}, B = {};
  A.x = x1;
  A.n = n1;
  A.nn = nn1;
  A.nm = nm1;
  if (!carve_set(wc, A, nm1, amax, order)) return -25;
  if (mode == 0) {
    B.x = x2;
    B.n = n2;
    B.nn = nn2;
    B.nm = nm2;
    if (!carve_set(wc, B, nm2, amax, order)) return -25;
  } else {
    B = A;
  }
  if ((int64_t)max(nm1, nm2) * amax > 0x7fffffff) return -5;
  const hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_aras_init, dim3(1), dim3(1), 0, st, status, E);
  FMLH_LAUNCH_CHECK("k_aras_init");
  hipLaunchKernelGGL(k_aras_prep, dim3(nm1 * amax), dim3(64), 0, st, A, P, E, status);
  FMLH_LAUNCH_CHECK("k_aras_prep");
  if (mode == 0) {
    hipLaunchKernelGGL(k_aras_prep, dim3(nm2 * amax), dim3(64), 0, st, B, P, E, status);
    FMLH_LAUNCH_CHECK("k_aras_prep");
  }
  hipLaunchKernelGGL(k_aras_elems, dim3(1), dim3(1), 0, st, E, status);
  FMLH_LAUNCH_CHECK("k_aras_elems");
  hipLaunchKernelGGL(k_aras_proj, dim3(nm1), dim3(64), 0, st, A, P, E);
  FMLH_LAUNCH_CHECK("k_aras_proj");
  hipLaunchKernelGGL(k_aras_self, dim3(nm1 * amax), dim3(64), 0, st, A, P, E);
  FMLH_LAUNCH_CHECK("k_aras_self");
  if (mode == 0) {
    hipLaunchKernelGGL(k_aras_proj, dim3(nm2), dim3(64), 0, st, B, P, E);
    FMLH_LAUNCH_CHECK("k_aras_proj");
    hipLaunchKernelGGL(k_aras_self, dim3(nm2 * amax), dim3(64), 0, st, B, P, E);
    FMLH_LAUNCH_CHECK("k_aras_self");
  }
  const int per_blk = max(1, 64 / amax);
  const int nblk = blk1 ? nblk1 : ceil_div(nm1, per_blk);
  const int64_t nw = (int64_t)nblk * ceil_div(nm2, kJB);
  if (nw > 0x7fffffff) return -5;
  const ArasUniform U = {B.n, B.nb, B.nt, B.tg, B.gq, B.dist, B.ksi, B.cs, B.self};
  timing_begin(st);
  hipLaunchKernelGGL(k_aras_tiles<4>, dim3((unsigned)nw), dim3(64), 0, st, A, U, nm2, P, blk1, ord1, nblk,
                     per_blk, mode, 0, E, out, status);
  hipLaunchKernelGGL(k_aras_tiles<5>, dim3((unsigned)nw), dim3(64), 0, st, A, U, nm2, P, blk1, ord1, nblk,
                     per_blk, mode, 4, E, out, status);
  hipLaunchKernelGGL(k_aras_tiles<6>, dim3((unsigned)nw), dim3(64), 0, st, A, U, nm2, P, blk1, ord1, nblk,
                     per_blk, mode, 5, E, out, status);
  hipLaunchKernelGGL(k_aras_tiles<8>, dim3((unsigned)nw), dim3(64), 0, st, A, U, nm2, P, blk1, ord1, nblk,
                     per_blk, mode, 6, E, out, status);
  hipLaunchKernelGGL(k_aras_tiles<16>, dim3((unsigned)nw), dim3(64), 0, st, A, U, nm2, P, blk1, ord1, nblk,
                     per_blk, mode, 8, E, out, status);
  timing_end(st);
  FMLH_LAUNCH_CHECK("k_aras_tiles");
  return 0;
}
