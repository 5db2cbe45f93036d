// (d) ARAD kernels on gfx950.
//
// Reference: fml/kernels/farad_kernels.f90 (module arad) and
// fml/math/farad.f90:452-600.
//
// The atom-pair overlap (atomic_distl2, farad_kernels.f90:29-90)
//   S = sum_{m1<c1} sum_{m2<c2} [(d2-d1)^2 < (8w)^2]
//         ((exp((d2-d1)^2 * -1/(4w^2)) * s1) * s2) * T(types) * ((1 + a1*a2) + b1*b2)
// is evaluated by one lane per atom pair in the reference's loop order
// (m1 outer, m2 inner, ascending) and operation order (FP contraction off).
// c = number of neighbours before the first distance > cut.
//
// Work decomposition
//   1. k_arad_prep: per atom slot, compact the (5, amax) descriptor block
//      into neighbour arrays (distance, sine term, the two angular channels,
//      a 9-bit element-pair type when the PTP row/column channels are small
//      integers), the cut-exit count and a "distances ascending" flag.
//   2. k_arad_self: self overlap S(a, a) per atom with the same device
//      function, so D(x, x) = 0 exactly.
//   3. k_arad_pairs: persistent workgroups walk molecule pairs; both
//      molecules' neighbour arrays are staged in LDS, each lane owns atom
//      pairs, the per-sigma exp sums are reduced with wave shuffles in a
//      fixed order.  The result of a molecule pair depends only on that
//      pair (no cross-pair reduction), so it is identical whatever else
//      shares the launch.
//   4. k_arad_atomic: 16x16 atom tiles for the per-atom forms.
//
// Band walk: when both neighbour lists are ascending in distance (always
// true for ARAD.describe output), the m2 that pass (d2-d1)^2 < (8w)^2 form
// one contiguous interval [lo, hi) whose ends only move forward as m1
// grows (fl(d2-d1) is monotone in both arguments and squaring is monotone
// in |x|).  Visiting only that interval adds exactly the same terms in the
// same order as the reference's full loop with its if-test, so the sum is
// unchanged.  Lists that are not ascending use the full loop.
#include <math.h>

#include "common.h"

#pragma clang fp contract(off)

namespace {

using namespace fmlh;

constexpr int kTypeRows = 8;     // PTP row differences 0..7
constexpr int kTypeCols = 64;    // PTP column differences 0..63
constexpr int kTypeTable = kTypeRows * kTypeCols;
constexpr uint16_t kTypeGeneric = 0xFFFF;
constexpr int kMaxSig = 8;
constexpr int kPairThreads = 256;

struct AradParams {
  double width, cut, rw, cw;
  double inv_w;   // -1/(4 w^2)          farad_kernels.f90:58
  double maxg;    // (8 w)^2             farad_kernels.f90:60
  double rw2, cw2;
  double inv_cut; // pi/(2 cut)          farad_kernels.f90:510
  double eps;     // real(5e-12) or real(1e-11), widened
};

// Per-set prepared neighbour data (workspace).
struct NbSet {
  double *d, *s, *c4, *c5, *row, *col;  // [slot*amax + m]
  uint16_t *ty;                         // [slot*amax + m]
  int32_t *cnt;                         // [slot]
  int32_t *asc;                         // [slot] 1 when d[0..cnt) ascending
  double *self;                         // [slot] self overlap (FP64 form)
  float *self32;                        // [slot] self overlap (FP32 form)
  double *zr, *zc;                      // [slot] PTP (row, col) of the atom
};

__device__ __forceinline__ bool small_int(double v, int limit) {
  return v >= 0.0 && v < (double)limit && v == (double)(int)v;
}

// One thread per atom slot.  per_atom_n: n is indexed by slot (atomic
// forms), else by molecule (slot / amax) and slots a >= n are empty.
__global__ __launch_bounds__(256) void k_arad_prep(const double *__restrict__ q,
                                                   const double *__restrict__ zt,
                                                   const int32_t *__restrict__ n, int nslots,
                                                   int amax, int per_atom_n, int sin_guard,
                                                   AradParams p, NbSet set,
                                                   int32_t *__restrict__ generic_flag) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  int N;
  bool valid;
  if (per_atom_n) {
    N = n[slot];
    valid = true;
  } else {
    N = n[slot / amax];
    valid = (slot % amax) < N;
  }
  if (N > amax) N = amax;
  if (N < 0) N = 0;
  const double *x = q + (int64_t)slot * 5 * amax;
  int cnt = 0;
  int asc = 1;
  bool generic = false;
  if (valid) {
    for (int m = 0; m < N; ++m) {
      const double d = x[m];
      if (d > p.cut) break;  // farad_kernels.f90:68,72
      const double rowv = x[amax + m], colv = x[2 * amax + m];
      double sv;
      if (!sin_guard || d < p.cut)
        sv = 1.0 - sin(d * p.inv_cut);
      else
        sv = 0.0;
      const int64_t o = (int64_t)slot * amax + m;
      set.d[o] = d;
      set.s[o] = sv;
      set.c4[o] = x[3 * amax + m];
      set.c5[o] = x[4 * amax + m];
      set.row[o] = rowv;
      set.col[o] = colv;
      uint16_t ty = kTypeGeneric;
      if (small_int(rowv, kTypeRows) && small_int(colv, kTypeCols))
        ty = (uint16_t)((int)rowv * kTypeCols + (int)colv);
      else
        generic = true;
      set.ty[o] = ty;
      if (cnt > 0 && !(set.d[o - 1] <= d)) asc = 0;
      ++cnt;
    }
    set.zr[slot] = zt[(int64_t)slot * 2 + 0];
    set.zc[slot] = zt[(int64_t)slot * 2 + 1];
  } else {
    set.zr[slot] = 0.0;
    set.zc[slot] = 0.0;
  }
  set.cnt[slot] = cnt;
  set.asc[slot] = asc;
  if (generic) atomicOr(generic_flag, 1);
}

__global__ void k_zero_flag(int32_t *f) { *f = 0; }

// ((rw2/(rw2+dr^2))*cw2)/(cw2+dc^2), farad_kernels.f90:80-81 and 590-592.
__device__ __forceinline__ double stoch(double dr, double dc, double rw2, double cw2) {
  return rw2 / (rw2 + dr * dr) * cw2 / (cw2 + dc * dc);
}

template <typename TC>
struct NbView {
  const double *d, *s, *c4, *c5, *row, *col;
  const uint16_t *ty;
};

template <typename TC>
__device__ __forceinline__ TC tcast(double v) { return (TC)v; }

// One inner term, farad_kernels.f90:74-83.
template <typename TC, bool kGeneric>
__device__ __forceinline__ void arad_term(TC &acc, TC r2, TC inv_w, TC s1, TC s2, TC a1, TC b1,
                                          TC a2, TC b2, int ty1, int ty2, double row1,
                                          double col1, double row2, double col2,
                                          const TC *__restrict__ ttab, double rw2, double cw2) {
  TC e;
  if (sizeof(TC) == 8)
    e = (TC)exp((double)(r2 * inv_w));
  else
    e = (TC)expf((float)(r2 * inv_w));
  TC v = e * s1;
  v = v * s2;
  TC t;
  if (kGeneric) {
    t = (TC)stoch(row1 - row2, col1 - col2, rw2, cw2);
  } else {
    const int r1 = ty1 / kTypeCols, c1 = ty1 % kTypeCols;
    const int rr = ty2 / kTypeCols, cc = ty2 % kTypeCols;
    const int dr = r1 > rr ? r1 - rr : rr - r1;
    const int dc = c1 > cc ? c1 - cc : cc - c1;
    t = ttab[dr * kTypeCols + dc];
  }
  v = v * t;
  TC ang = TC(1) + a1 * a2;
  ang = ang + b1 * b2;
  acc = acc + v * ang;
}

// atomic_distl2 for atoms x (list 1) and y (list 2).
template <typename TC, bool kGeneric>
__device__ TC arad_distl2(const NbView<TC> &x, int cnt1, bool asc1, const NbView<TC> &y, int cnt2,
                          bool asc2, TC inv_w, TC maxg, const TC *__restrict__ ttab, double rw2,
                          double cw2) {
  TC acc = TC(0);
  if (asc1 && asc2) {
    int lo = 0, hi = 0;
    for (int m1 = 0; m1 < cnt1; ++m1) {
      const TC d1 = tcast<TC>(x.d[m1]);
      while (lo < cnt2) {
        const TC d2 = tcast<TC>(y.d[lo]);
        const TC dr = d2 - d1;
        if (dr * dr < maxg || d2 > d1) break;
        ++lo;
      }
      if (hi < lo) hi = lo;
      while (hi < cnt2) {
        const TC dr = tcast<TC>(y.d[hi]) - d1;
        if (!(dr * dr < maxg)) break;
        ++hi;
      }
      if (lo == hi) continue;
      const TC s1 = tcast<TC>(x.s[m1]);
      const TC a1 = tcast<TC>(x.c4[m1]);
      const TC b1 = tcast<TC>(x.c5[m1]);
      const int ty1 = kGeneric ? 0 : x.ty[m1];
      const double row1 = kGeneric ? x.row[m1] : 0.0;
      const double col1 = kGeneric ? x.col[m1] : 0.0;
      for (int m2 = lo; m2 < hi; ++m2) {
        const TC dr = tcast<TC>(y.d[m2]) - d1;
        const TC r2 = dr * dr;
        arad_term<TC, kGeneric>(acc, r2, inv_w, s1, tcast<TC>(y.s[m2]), a1, b1,
                                tcast<TC>(y.c4[m2]), tcast<TC>(y.c5[m2]), ty1,
                                kGeneric ? 0 : y.ty[m2], row1, col1,
                                kGeneric ? y.row[m2] : 0.0, kGeneric ? y.col[m2] : 0.0, ttab,
                                rw2, cw2);
      }
    }
  } else {
    for (int m1 = 0; m1 < cnt1; ++m1) {
      const TC d1 = tcast<TC>(x.d[m1]);
      const TC s1 = tcast<TC>(x.s[m1]);
      const TC a1 = tcast<TC>(x.c4[m1]);
      const TC b1 = tcast<TC>(x.c5[m1]);
      const int ty1 = kGeneric ? 0 : x.ty[m1];
      const double row1 = kGeneric ? x.row[m1] : 0.0;
      const double col1 = kGeneric ? x.col[m1] : 0.0;
      for (int m2 = 0; m2 < cnt2; ++m2) {
        const TC dr = tcast<TC>(y.d[m2]) - d1;
        const TC r2 = dr * dr;
        if (r2 < maxg)
          arad_term<TC, kGeneric>(acc, r2, inv_w, s1, tcast<TC>(y.s[m2]), a1, b1,
                                  tcast<TC>(y.c4[m2]), tcast<TC>(y.c5[m2]), ty1,
                                  kGeneric ? 0 : y.ty[m2], row1, col1,
                                  kGeneric ? y.row[m2] : 0.0, kGeneric ? y.col[m2] : 0.0, ttab,
                                  rw2, cw2);
      }
    }
  }
  return acc;
}

template <typename TC>
__device__ __forceinline__ NbView<TC> global_view(const NbSet &s, int64_t slot, int amax) {
  const int64_t o = slot * amax;
  return NbView<TC>{s.d + o, s.s + o, s.c4 + o, s.c5 + o, s.row + o, s.col + o, s.ty + o};
}

template <typename TC>
__device__ void fill_ttab(TC *ttab, double rw2, double cw2) {
  for (int t = threadIdx.x; t < kTypeTable; t += blockDim.x)
    ttab[t] = (TC)stoch((double)(t / kTypeCols), (double)(t % kTypeCols), rw2, cw2);
}

// Self overlaps, one thread per slot (global-memory operands).
template <typename TC>
__global__ __launch_bounds__(256) void k_arad_self(NbSet set, int nslots, int amax, AradParams p,
                                                   const int32_t *__restrict__ generic_flag) {
  __shared__ TC ttab[kTypeTable];
  fill_ttab(ttab, p.rw2, p.cw2);
  __syncthreads();
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  const int c = set.cnt[slot];
  const bool asc = set.asc[slot] != 0;
  NbView<TC> v = global_view<TC>(set, slot, amax);
  TC r;
  if (*generic_flag)
    r = arad_distl2<TC, true>(v, c, asc, v, c, asc, (TC)p.inv_w, (TC)p.maxg, ttab, p.rw2, p.cw2);
  else
    r = arad_distl2<TC, false>(v, c, asc, v, c, asc, (TC)p.inv_w, (TC)p.maxg, ttab, p.rw2, p.cw2);
  if (sizeof(TC) == 8)
    set.self[slot] = (double)r;
  else
    set.self32[slot] = (float)r;
}

// Shared-memory staging of one molecule's neighbour arrays.
struct MolStage {
  double *d, *s, *c4, *c5;
  uint16_t *ty;
  int *cnt, *asc;
  double *self, *zr, *zc;
};

__device__ __forceinline__ MolStage carve_stage(char *&p, int amax) {
  MolStage m;
  const int nn = amax * amax;
  m.d = (double *)p; p += sizeof(double) * nn;
  m.s = (double *)p; p += sizeof(double) * nn;
  m.c4 = (double *)p; p += sizeof(double) * nn;
  m.c5 = (double *)p; p += sizeof(double) * nn;
  m.self = (double *)p; p += sizeof(double) * amax;
  m.zr = (double *)p; p += sizeof(double) * amax;
  m.zc = (double *)p; p += sizeof(double) * amax;
  m.ty = (uint16_t *)p; p += sizeof(uint16_t) * ((nn + 3) & ~3);
  m.cnt = (int *)p; p += sizeof(int) * amax;
  m.asc = (int *)p; p += sizeof(int) * amax;
  return m;
}

size_t stage_bytes(int amax) {
  const size_t nn = (size_t)amax * amax;
  return sizeof(double) * (4 * nn + 3 * amax) + sizeof(uint16_t) * ((nn + 3) & ~size_t(3)) +
         sizeof(int) * 2 * amax;
}

template <bool kF32Self>
__device__ void stage_molecule(MolStage &m, const NbSet &set, int mol, int natoms, int amax) {
  const int64_t slot0 = (int64_t)mol * amax;
  for (int a = threadIdx.x; a < natoms; a += blockDim.x) {
    m.cnt[a] = set.cnt[slot0 + a];
    m.asc[a] = set.asc[slot0 + a];
    m.self[a] = kF32Self ? (double)set.self32[slot0 + a] : set.self[slot0 + a];
    m.zr[a] = set.zr[slot0 + a];
    m.zc[a] = set.zc[slot0 + a];
  }
  const int total = natoms * amax;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int64_t o = slot0 * amax + t;
    m.d[t] = set.d[o];
    m.s[t] = set.s[o];
    m.c4[t] = set.c4[o];
    m.c5[t] = set.c5[o];
    m.ty[t] = set.ty[o];
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}


struct PairArgs {
  int mode;  // FMLH_ARAD_*
  int nm1, nm2, amax, nsig;
  int triangular;
  int64_t npairs;
  double inv_s2[kMaxSig];
  const int32_t *n1, *n2;
  double *out;
};

__device__ __forceinline__ int64_t dist_index(int i, int j, int a, int b, int nm1, int nm2,
                                              int amax) {
  // dmatrix(i, j, a, b) Fortran order (nm1, nm2, amax, amax), farad.f90:485
  return (int64_t)i + (int64_t)nm1 * ((int64_t)j + (int64_t)nm2 * ((int64_t)a + (int64_t)amax * b));
}

// Molecule-pair kernel.  kStage: stage both molecules' neighbour arrays in
// LDS (amax <= 29); otherwise operands are read from the prepared global
// arrays and only per-atom scalars are staged.
template <typename TC, bool kStage>
__global__ __launch_bounds__(kPairThreads) void k_arad_pairs(NbSet set1, NbSet set2, PairArgs pa,
                                                              AradParams p,
                                                              const int32_t *__restrict__ flag) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  char *ptr = smem;
  TC *ttab = (TC *)ptr;
  ptr += sizeof(double) * kTypeTable;
  double *red = (double *)ptr;
  ptr += sizeof(double) * kMaxSig * (kPairThreads / 64);
  MolStage mi, mj;
  if (kStage) {
    mi = carve_stage(ptr, pa.amax);
    mj = carve_stage(ptr, pa.amax);
  } else {
    // scalars only
    const int amax = pa.amax;
    mi.self = (double *)ptr; ptr += sizeof(double) * amax;
    mi.zr = (double *)ptr; ptr += sizeof(double) * amax;
    mi.zc = (double *)ptr; ptr += sizeof(double) * amax;
    mi.cnt = (int *)ptr; ptr += sizeof(int) * amax;
    mi.asc = (int *)ptr; ptr += sizeof(int) * amax;
    mj.self = (double *)ptr; ptr += sizeof(double) * amax;
    mj.zr = (double *)ptr; ptr += sizeof(double) * amax;
    mj.zc = (double *)ptr; ptr += sizeof(double) * amax;
    mj.cnt = (int *)ptr; ptr += sizeof(int) * amax;
    mj.asc = (int *)ptr; ptr += sizeof(int) * amax;
  }
  fill_ttab(ttab, p.rw2, p.cw2);

  const bool generic = *flag != 0;
  const int amax = pa.amax;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const TC inv_w = (TC)p.inv_w, maxg = (TC)p.maxg;
  int cur_i = -1, cur_j = -1;

  for (int64_t pidx = blockIdx.x; pidx < pa.npairs; pidx += gridDim.x) {
    int i, j;
    if (pa.triangular) {
      // pidx = j(j+1)/2 + i with i <= j
      int jj = (int)((sqrt(8.0 * (double)pidx + 1.0) - 1.0) * 0.5);
      while ((int64_t)(jj + 1) * (jj + 2) / 2 <= pidx) ++jj;
      while ((int64_t)jj * (jj + 1) / 2 > pidx) --jj;
      j = jj;
      i = (int)(pidx - (int64_t)jj * (jj + 1) / 2);
    } else {
      i = (int)(pidx % pa.nm1);
      j = (int)(pidx / pa.nm1);
    }
    const int ni = min(max(pa.n1[i], 0), amax);
    const int nj = min(max(pa.n2[j], 0), amax);
    __syncthreads();  // the previous pair's readers are done with the stages
    if (i != cur_i) {
      if (kStage) {
        stage_molecule<sizeof(TC) == 4>(mi, set1, i, ni, amax);
      } else {
        const int64_t s0 = (int64_t)i * amax;
        for (int a = threadIdx.x; a < ni; a += blockDim.x) {
          mi.cnt[a] = set1.cnt[s0 + a];
          mi.asc[a] = set1.asc[s0 + a];
          mi.self[a] = sizeof(TC) == 4 ? (double)set1.self32[s0 + a] : set1.self[s0 + a];
          mi.zr[a] = set1.zr[s0 + a];
          mi.zc[a] = set1.zc[s0 + a];
        }
      }
      cur_i = i;
    }
    if (j != cur_j) {
      if (kStage) {
        stage_molecule<sizeof(TC) == 4>(mj, set2, j, nj, amax);
      } else {
        const int64_t s0 = (int64_t)j * amax;
        for (int b = threadIdx.x; b < nj; b += blockDim.x) {
          mj.cnt[b] = set2.cnt[s0 + b];
          mj.asc[b] = set2.asc[s0 + b];
          mj.self[b] = sizeof(TC) == 4 ? (double)set2.self32[s0 + b] : set2.self[s0 + b];
          mj.zr[b] = set2.zr[s0 + b];
          mj.zc[b] = set2.zc[s0 + b];
        }
      }
      cur_j = j;
    }
    __syncthreads();

    double part[kMaxSig];
#pragma unroll
    for (int k = 0; k < kMaxSig; ++k) part[k] = 0.0;

    const bool dist_mode = (pa.mode == FMLH_ARAD_DISTANCE);
    const int arow = dist_mode ? amax : ni;
    const int npair_atoms = dist_mode ? amax * amax : ni * nj;
    for (int t = threadIdx.x; t < npair_atoms; t += kPairThreads) {
      const int a = t % arow, b = t / arow;
      if (a >= ni || b >= nj) {  // DISTANCE mode zero padding (farad.f90:565)
        pa.out[dist_index(i, j, a, b, pa.nm1, pa.nm2, amax)] = 0.0;
        continue;
      }
      NbView<TC> x, y;
      const int64_t gx = ((int64_t)i * amax + a) * amax, gy = ((int64_t)j * amax + b) * amax;
      if (kStage) {
        x = NbView<TC>{mi.d + a * amax, mi.s + a * amax, mi.c4 + a * amax, mi.c5 + a * amax,
                       set1.row + gx, set1.col + gx, mi.ty + a * amax};
        y = NbView<TC>{mj.d + b * amax, mj.s + b * amax, mj.c4 + b * amax, mj.c5 + b * amax,
                       set2.row + gy, set2.col + gy, mj.ty + b * amax};
      } else {
        x = NbView<TC>{set1.d + gx, set1.s + gx, set1.c4 + gx, set1.c5 + gx, set1.row + gx,
                       set1.col + gx, set1.ty + gx};
        y = NbView<TC>{set2.d + gy, set2.s + gy, set2.c4 + gy, set2.c5 + gy, set2.row + gy,
                       set2.col + gy, set2.ty + gy};
      }
      const TC S = generic ? arad_distl2<TC, true>(x, mi.cnt[a], mi.asc[a] != 0, y, mj.cnt[b],
                                                   mj.asc[b] != 0, inv_w, maxg, ttab, p.rw2, p.cw2)
                           : arad_distl2<TC, false>(x, mi.cnt[a], mi.asc[a] != 0, y, mj.cnt[b],
                                                    mj.asc[b] != 0, inv_w, maxg, ttab, p.rw2, p.cw2);
      // farad_kernels.f90:590-594
      const double T = stoch(mi.zr[a] - mj.zr[b], mi.zc[a] - mj.zc[b], p.rw2, p.cw2);
      double D = mi.self[a] + mj.self[b] - 2.0 * (double)S * T;
      if (fabs(D) < p.eps) D = 0.0;
      if (dist_mode) {
        pa.out[dist_index(i, j, a, b, pa.nm1, pa.nm2, amax)] = D;
      } else if (pa.mode == FMLH_ARAD_KERNEL) {
        part[0] += exp(pa.inv_s2[0] * D);  // farad_kernels.f90:211
      } else {
#pragma unroll
        for (int k = 0; k < kMaxSig; ++k)
          if (k < pa.nsig) part[k] += exp(D * pa.inv_s2[k]);  // farad_kernels.f90:603
      }
    }
    if (dist_mode) continue;

    const int nsig = (pa.mode == FMLH_ARAD_KERNEL) ? 1 : pa.nsig;
#pragma unroll
    for (int k = 0; k < kMaxSig; ++k) {
      if (k < nsig) {
        const double w = wave_sum(part[k]);
        if (lane == 0) red[k * (kPairThreads / 64) + wave] = w;
      }
    }
    __syncthreads();
    if ((int)threadIdx.x < nsig) {
      const int k = threadIdx.x;
      double v = 0.0;
      for (int w = 0; w < kPairThreads / 64; ++w) v += red[k * (kPairThreads / 64) + w];
      if (pa.mode == FMLH_ARAD_KERNEL) {
        pa.out[(int64_t)i + (int64_t)pa.nm1 * j] = v;
        pa.out[(int64_t)j + (int64_t)pa.nm1 * i] = v;
      } else {
        pa.out[k + (int64_t)pa.nsig * ((int64_t)i + (int64_t)pa.nm1 * j)] = v;
        if (pa.triangular) pa.out[k + (int64_t)pa.nsig * ((int64_t)j + (int64_t)pa.nm1 * i)] = v;
      }
    }
  }
}

// Per-atom forms (farad_kernels.f90:757-1030): one thread per atom pair on
// a 16 x 16 tile; operands come from the prepared global arrays.
template <typename TC>
__global__ __launch_bounds__(256) void k_arad_atomic(NbSet set1, NbSet set2, int na1, int na2,
                                                     int amax, int mode, PairArgs pa, AradParams p,
                                                     const int32_t *__restrict__ flag) {
  __shared__ TC ttab[kTypeTable];
  fill_ttab(ttab, p.rw2, p.cw2);
  __syncthreads();
  const bool generic = *flag != 0;
  const int i = blockIdx.x * 16 + (threadIdx.x & 15);
  const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i >= na1 || j >= na2) return;
  NbView<TC> x = global_view<TC>(set1, i, amax);
  NbView<TC> y = global_view<TC>(set2, j, amax);
  const TC inv_w = (TC)p.inv_w, maxg = (TC)p.maxg;
  const TC S = generic ? arad_distl2<TC, true>(x, set1.cnt[i], set1.asc[i] != 0, y, set2.cnt[j],
                                               set2.asc[j] != 0, inv_w, maxg, ttab, p.rw2, p.cw2)
                       : arad_distl2<TC, false>(x, set1.cnt[i], set1.asc[i] != 0, y, set2.cnt[j],
                                                set2.asc[j] != 0, inv_w, maxg, ttab, p.rw2, p.cw2);
  const double T = stoch(set1.zr[i] - set2.zr[j], set1.zc[i] - set2.zc[j], p.rw2, p.cw2);
  const double s1 = sizeof(TC) == 4 ? (double)set1.self32[i] : set1.self[i];
  const double s2 = sizeof(TC) == 4 ? (double)set2.self32[j] : set2.self[j];
  double D = s1 + s2 - 2.0 * (double)S * T;
  if (fabs(D) < p.eps) D = 0.0;
  if (mode == FMLH_ARAD_ATOMIC_DISTANCE) {
    pa.out[(int64_t)i + (int64_t)na1 * j] = D;
  } else {
#pragma unroll
    for (int k = 0; k < kMaxSig; ++k)
      if (k < pa.nsig)
        pa.out[k + (int64_t)pa.nsig * ((int64_t)i + (int64_t)na1 * j)] = exp(D * pa.inv_s2[k]);
  }
}

// ---- host side --------------------------------------------------------

size_t set_bytes(int64_t nslots, int amax) {
  const size_t nn = (size_t)nslots * amax;
  size_t b = 0;
  b += WsCarver::need(sizeof(double) * nn) * 6;
  b += WsCarver::need(sizeof(uint16_t) * nn);
  b += WsCarver::need(sizeof(int32_t) * nslots) * 2;
  b += WsCarver::need(sizeof(double) * nslots) * 3;
  b += WsCarver::need(sizeof(float) * nslots);
  return b;
}

bool carve_set(WsCarver &w, int64_t nslots, int amax, NbSet &s) {
  const size_t nn = (size_t)nslots * amax;
  s.d = w.take<double>(nn);
  s.s = w.take<double>(nn);
  s.c4 = w.take<double>(nn);
  s.c5 = w.take<double>(nn);
  s.row = w.take<double>(nn);
  s.col = w.take<double>(nn);
  s.ty = w.take<uint16_t>(nn);
  s.cnt = w.take<int32_t>(nslots);
  s.asc = w.take<int32_t>(nslots);
  s.self = w.take<double>(nslots);
  s.zr = w.take<double>(nslots);
  s.zc = w.take<double>(nslots);
  s.self32 = w.take<float>(nslots);
  return s.d && s.s && s.c4 && s.c5 && s.row && s.col && s.ty && s.cnt && s.asc && s.self &&
         s.zr && s.zc && s.self32;
}

AradParams make_params(double width, double cut, double rw, double cw, bool eps_1e11) {
  AradParams p;
  p.width = width;
  p.cut = cut;
  p.rw = rw;
  p.cw = cw;
  p.inv_w = -1.0 / (4.0 * width * width);   // farad_kernels.f90:58
  p.maxg = (8.0 * width) * (8.0 * width);   // :60
  p.rw2 = rw * rw;
  p.cw2 = cw * cw;
  const double pi = 4.0 * atan(1.0);
  p.inv_cut = pi / (2.0 * cut);             // :510
  p.eps = eps_1e11 ? (double)1.0e-11f : (double)5.0e-12f;  // :148, :505
  return p;
}

int device_cus() {
  static int cus = 0;
  if (cus == 0) cus = fmlh_device_cu_count();
  return cus > 0 ? cus : 256;
}

size_t pair_lds_bytes(int amax, bool stage) {
  size_t b = sizeof(double) * (kTypeTable + kMaxSig * (kPairThreads / 64));
  if (stage)
    b += 2 * stage_bytes(amax);
  else
    b += 2 * (sizeof(double) * 3 + sizeof(int) * 2) * (size_t)amax;
  return b;
}

constexpr size_t kLdsLimit = 65536;

// Shared preparation for both entry points.
int prepare_sets(const double *q1, const double *zt1, const int32_t *n1, int64_t ns1,
                 const double *q2, const double *zt2, const int32_t *n2, int64_t ns2, int amax,
                 bool one_set, int per_atom_n, int guard, int precision, const AradParams &p,
                 void *ws, size_t ws_bytes, NbSet &s1, NbSet &s2, int32_t *&flag,
                 hipStream_t st) {
  WsCarver w(ws, ws_bytes);
  flag = w.take<int32_t>(64);
  if (!flag || !carve_set(w, ns1, amax, s1)) return -1;
  if (!one_set) {
    if (!carve_set(w, ns2, amax, s2)) return -1;
  } else {
    s2 = s1;
  }
  hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(1), 0, st, flag);
  FMLH_LAUNCH_CHECK("k_zero_flag");
  hipLaunchKernelGGL(k_arad_prep, dim3((unsigned)ceil_div64(ns1, 256)), dim3(256), 0, st, q1, zt1,
                     n1, (int)ns1, amax, per_atom_n, guard, p, s1, flag);
  FMLH_LAUNCH_CHECK("k_arad_prep");
  if (!one_set) {
    hipLaunchKernelGGL(k_arad_prep, dim3((unsigned)ceil_div64(ns2, 256)), dim3(256), 0, st, q2,
                       zt2, n2, (int)ns2, amax, per_atom_n, guard, p, s2, flag);
    FMLH_LAUNCH_CHECK("k_arad_prep");
  }
  for (int k = 0; k < (one_set ? 1 : 2); ++k) {
    NbSet &s = k == 0 ? s1 : s2;
    const int64_t ns = k == 0 ? ns1 : ns2;
    if (precision)
      hipLaunchKernelGGL(k_arad_self<float>, dim3((unsigned)ceil_div64(ns, 256)), dim3(256), 0, st,
                         s, (int)ns, amax, p, flag);
    else
      hipLaunchKernelGGL(k_arad_self<double>, dim3((unsigned)ceil_div64(ns, 256)), dim3(256), 0,
                         st, s, (int)ns, amax, p, flag);
    FMLH_LAUNCH_CHECK("k_arad_self");
  }
  return 0;
}

}  // namespace

extern "C" {

size_t fmlh_arad_ws_bytes(int nm1, int nm2, int amax, int mode) {
  if (nm1 < 0 || nm2 < 0 || amax < 1) return 0;
  const bool one_set = (mode == FMLH_ARAD_SYMMETRIC || mode == FMLH_ARAD_KERNEL);
  size_t b = WsCarver::need(sizeof(int32_t) * 64) + set_bytes((int64_t)nm1 * amax, amax);
  if (!one_set) b += set_bytes((int64_t)nm2 * amax, amax);
  return b;
}

size_t fmlh_arad_atomic_ws_bytes(int na1, int na2, int amax) {
  if (na1 < 0 || na2 < 0 || amax < 1) return 0;
  return WsCarver::need(sizeof(int32_t) * 64) + set_bytes(na1, amax) + set_bytes(na2, amax);
}

int fmlh_arad_kernels(int mode, const double *q1, const double *zt1, const int32_t *n1, int nm1,
                      const double *q2, const double *zt2, const int32_t *n2, int nm2, int amax,
                      const double *sigmas_host, int nsig, double width, double cut,
                      double r_width, double c_width, int precision, double *out, void *ws,
                      size_t ws_bytes, void *stream) {
  FMLH_ARG_CHECK(mode == FMLH_ARAD_GLOBAL || mode == FMLH_ARAD_SYMMETRIC ||
                     mode == FMLH_ARAD_KERNEL || mode == FMLH_ARAD_DISTANCE,
                 1);
  const bool one_set = (mode == FMLH_ARAD_SYMMETRIC || mode == FMLH_ARAD_KERNEL);
  FMLH_ARG_CHECK(nm1 >= 0, 5);
  if (one_set) nm2 = nm1;
  FMLH_ARG_CHECK(nm2 >= 0, 9);
  FMLH_ARG_CHECK(amax >= 1 && amax <= 1024, 10);
  if (mode != FMLH_ARAD_DISTANCE) {
    FMLH_ARG_CHECK(sigmas_host != nullptr, 11);
    FMLH_ARG_CHECK(nsig >= 1 && nsig <= kMaxSig, 12);
    if (mode == FMLH_ARAD_KERNEL) FMLH_ARG_CHECK(nsig == 1, 12);
    for (int k = 0; k < nsig; ++k) FMLH_ARG_CHECK(sigmas_host[k] != 0.0, 11);
  }
  FMLH_ARG_CHECK(width > 0.0 || mode == FMLH_ARAD_KERNEL, 13);
  FMLH_ARG_CHECK(cut > 0.0 || mode == FMLH_ARAD_KERNEL, 14);
  FMLH_ARG_CHECK(precision == 0 || precision == 1, 17);
  if (nm1 == 0 || nm2 == 0) return 0;
  FMLH_ARG_CHECK(q1 && zt1 && n1, 2);
  if (!one_set) FMLH_ARG_CHECK(q2 && zt2 && n2, 6);
  FMLH_ARG_CHECK(out != nullptr, 18);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= fmlh_arad_ws_bytes(nm1, nm2, amax, mode), 19);

  const AradParams p = (mode == FMLH_ARAD_KERNEL) ? make_params(0.2, 6.0, 1.0, 0.5, true)
                                                  : make_params(width, cut, r_width, c_width, false);
  hipStream_t st = as_stream(stream);
  NbSet s1, s2;
  int32_t *flag = nullptr;
  const int guard = (mode == FMLH_ARAD_GLOBAL) ? 1 : 0;  // farad_kernels.f90:526,539
  int rc = prepare_sets(q1, zt1, n1, (int64_t)nm1 * amax, q2, zt2, n2, (int64_t)nm2 * amax, amax,
                        one_set, 0, guard, precision, p, ws, ws_bytes, s1, s2, flag, st);
  if (rc) return rc < 0 && rc > -100 ? -19 : rc;

  PairArgs pa;
  pa.mode = mode;
  pa.nm1 = nm1;
  pa.nm2 = nm2;
  pa.amax = amax;
  pa.nsig = (mode == FMLH_ARAD_DISTANCE) ? 0 : nsig;
  pa.triangular = one_set ? 1 : 0;
  pa.npairs = one_set ? (int64_t)nm1 * (nm1 + 1) / 2 : (int64_t)nm1 * nm2;
  for (int k = 0; k < kMaxSig; ++k) pa.inv_s2[k] = 0.0;
  if (mode != FMLH_ARAD_DISTANCE)
    for (int k = 0; k < nsig; ++k) pa.inv_s2[k] = -1.0 / (sigmas_host[k] * sigmas_host[k]);
  pa.n1 = n1;
  pa.n2 = one_set ? n1 : n2;
  pa.out = out;

  const bool stage = pair_lds_bytes(amax, true) <= kLdsLimit;
  const size_t lds = pair_lds_bytes(amax, stage);
  if (lds > kLdsLimit) return -10;
  const int64_t cap = (int64_t)device_cus() * 16;
  const int grid = (int)(pa.npairs < cap ? pa.npairs : cap);
  timing_begin(st);
  if (precision) {
    if (stage)
      hipLaunchKernelGGL((k_arad_pairs<float, true>), dim3(grid), dim3(kPairThreads), lds, st, s1,
                         s2, pa, p, flag);
    else
      hipLaunchKernelGGL((k_arad_pairs<float, false>), dim3(grid), dim3(kPairThreads), lds, st, s1,
                         s2, pa, p, flag);
  } else {
    if (stage)
      hipLaunchKernelGGL((k_arad_pairs<double, true>), dim3(grid), dim3(kPairThreads), lds, st, s1,
                         s2, pa, p, flag);
    else
      hipLaunchKernelGGL((k_arad_pairs<double, false>), dim3(grid), dim3(kPairThreads), lds, st,
                         s1, s2, pa, p, flag);
  }
  timing_end(st);
  FMLH_LAUNCH_CHECK("k_arad_pairs");
  return 0;
}

int fmlh_arad_atomic(int mode, const double *q1, const double *zt1, const int32_t *n1, int na1,
                     const double *q2, const double *zt2, const int32_t *n2, int na2, int amax,
                     const double *sigmas_host, int nsig, double width, double cut,
                     double r_width, double c_width, double *out, void *ws, size_t ws_bytes,
                     void *stream) {
  FMLH_ARG_CHECK(mode == FMLH_ARAD_ATOMIC_KERNELS || mode == FMLH_ARAD_ATOMIC_DISTANCE, 1);
  FMLH_ARG_CHECK(na1 >= 0, 5);
  FMLH_ARG_CHECK(na2 >= 0, 9);
  FMLH_ARG_CHECK(amax >= 1 && amax <= 1024, 10);
  if (mode == FMLH_ARAD_ATOMIC_KERNELS) {
    FMLH_ARG_CHECK(sigmas_host != nullptr, 11);
    FMLH_ARG_CHECK(nsig >= 1 && nsig <= kMaxSig, 12);
    for (int k = 0; k < nsig; ++k) FMLH_ARG_CHECK(sigmas_host[k] != 0.0, 11);
  }
  FMLH_ARG_CHECK(width > 0.0, 13);
  FMLH_ARG_CHECK(cut > 0.0, 14);
  if (na1 == 0 || na2 == 0) return 0;
  FMLH_ARG_CHECK(q1 && zt1 && n1, 2);
  FMLH_ARG_CHECK(q2 && zt2 && n2, 6);
  FMLH_ARG_CHECK(out != nullptr, 17);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= fmlh_arad_atomic_ws_bytes(na1, na2, amax), 18);
  const AradParams p = make_params(width, cut, r_width, c_width, false);
  hipStream_t st = as_stream(stream);
  NbSet s1, s2;
  int32_t *flag = nullptr;
  int rc = prepare_sets(q1, zt1, n1, na1, q2, zt2, n2, na2, amax, false, 1, 0, 0, p, ws, ws_bytes,
                        s1, s2, flag, st);
  if (rc) return rc < 0 && rc > -100 ? -18 : rc;
  PairArgs pa;
  pa.mode = mode;
  pa.nm1 = na1;
  pa.nm2 = na2;
  pa.amax = amax;
  pa.nsig = mode == FMLH_ARAD_ATOMIC_KERNELS ? nsig : 0;
  pa.triangular = 0;
  pa.npairs = (int64_t)na1 * na2;
  for (int k = 0; k < kMaxSig; ++k) pa.inv_s2[k] = 0.0;
  for (int k = 0; k < pa.nsig; ++k) pa.inv_s2[k] = -1.0 / (sigmas_host[k] * sigmas_host[k]);
  pa.n1 = n1;
  pa.n2 = n2;
  pa.out = out;
  dim3 grid(ceil_div(na1, 16), ceil_div(na2, 16));
  timing_begin(st);
  hipLaunchKernelGGL(k_arad_atomic<double>, grid, dim3(256), 0, st, s1, s2, na1, na2, amax, mode,
                     pa, p, flag);
  timing_end(st);
  FMLH_LAUNCH_CHECK("k_arad_atomic");
  return 0;
}

}  // extern "C"
