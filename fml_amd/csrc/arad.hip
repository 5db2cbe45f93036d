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
// c = number of neighbours before the first distance > cut; s = 1-sin(d pi/(2 cut));
// T(types) = ((rw^2/(rw^2+drow^2)) cw^2)/(cw^2+dcol^2).
//
// Work decomposition (one call = a handful of launches on one stream)
//   1. k_arad_prep: per atom slot, compact the (5, amax) descriptor block
//      into 32-byte neighbour records {d, s, a, b} plus the raw PTP
//      (row, col) pair, the cut-exit count and a "distances ascending"
//      flag; mark every (row, col) type seen in a 512-entry bitmap.
//   2. k_arad_types: one workgroup numbers the types seen (at most 16 get a
//      compact id) and builds the 16 x 16 type-pair table with the
//      reference's expression; k_arad_type_ids writes each neighbour's id.
//      Data with non-integer type channels or more than 16 types takes the
//      generic path, which evaluates T from the raw channels per term.
//   3. k_arad_self: self overlap S(a, a) per atom with the same device
//      function, so D(x, x) = 0 exactly.
//   4. k_arad_tiles (amax <= 29): persistent workgroups take size-sorted
//      molecule-pair tiles; k_arad_pairs (larger amax, distance form) walks
//      molecule pairs.  Each lane owns atom pairs; the per-sigma sums run
//      sequentially in the reference's order, so a molecule pair's result
//      depends only on that pair.
//   5. k_arad_atomic: 16 x 16 atom tiles for the per-atom forms.
//
// Band walk: when both neighbour lists are ascending in distance (always
// true for ARAD.describe output), the m2 that pass (d2-d1)^2 < (8w)^2 form
// one contiguous interval [lo, hi) whose ends only move forward as m1
// grows (fl(d2-d1) is monotone in both arguments and squaring is monotone
// in |x|).  Visiting only that interval adds exactly the same terms in the
// same order as the reference's full loop with its if-test, so the sum is
// unchanged.  Lists that are not ascending use the full loop.
//
// Inner exp: its argument (d2-d1)^2 * (-1/(4w^2)) lies in (-16, 0] whenever
// the term is taken, because (8w)^2 * (-1/(4w^2)) = -16.  exp_neg() and the
// sine terms' sin_libm() (exp_neg.h) round exactly as the host C library's
// exp and sin that the reference calls (the same table-driven algorithms;
// exp_neg's table is held in LDS).
//
// Every term is formed with the reference's operations in the reference's
// order, FP contraction off: ((exp * s1) * s2) * T, times (1 + a1 a2) + b1 b2
// (farad_kernels.f90:78-83), and the per-sigma kernel sums run sequentially
// over the atom pairs in the reference's column-major order (:602-604,
// :216-217), so overlaps, distances and kernel entries are the reference's
// bit for bit.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.h"
#include "exp_neg.h"

#pragma clang fp contract(off)

namespace {

using namespace fmlh;

constexpr int kMaxTypes = 16;            // compact type ids
constexpr int kTypeRows = 8;             // PTP rows 0..7 representable
constexpr int kTypeCols = 64;            // PTP columns 0..63 representable
constexpr int kTypeBitmap = kTypeRows * kTypeCols;
constexpr int32_t kTypeNone = 0xFFFF;
constexpr int kMaxSig = 8;
constexpr int kPairThreads = 256;
constexpr int kTileRounds = 4;  // atom-pair rounds per thread in a tile
constexpr int kMaxAmax = 29;    // largest amax the tile path takes
constexpr int kStageAtoms = 128;  // most set-2 atoms one tile stages

struct AradParams {
  double width, cut, rw, cw;
  double inv_w;    // -1/(4 w^2)          farad_kernels.f90:58
  double maxg;     // (8 w)^2             farad_kernels.f90:60
  double rw2, cw2;
  double inv_cut;  // pi/(2 cut)          farad_kernels.f90:510
  double eps;      // real(5e-12) or real(1e-11), widened
};

struct Rec {  // one neighbour: distance, sine term, two angular channels, compact type id
  double d, s, a, b;
  int32_t ty, pad;
};

// Records in memory (workspace and LDS stages) as two arrays: the four
// doubles {d, s, a, b} (32 bytes, read as two 16-byte loads) and the type id,
// 36 bytes a record, so that the stage of a tile with one 29-atom set-2
// molecule (841 records) leaves room for four workgroups in a CU's LDS.
struct alignas(16) Rec4 {
  double d, s, a, b;
};
struct RecArr {
  Rec4 *r;
  int32_t *ty;
  __device__ __forceinline__ Rec at(int64_t m) const {
    const Rec4 v = r[m];
    return Rec{v.d, v.s, v.a, v.b, ty[m], 0};
  }
  __device__ __forceinline__ double d(int64_t m) const { return r[m].d; }
  __device__ __forceinline__ void put(int64_t m, const Rec &x) const {
    r[m] = Rec4{x.d, x.s, x.a, x.b};
    ty[m] = x.ty;
  }
  __device__ __forceinline__ void copy(int64_t m, const RecArr &src, int64_t o) const {
    r[m] = src.r[o];
    ty[m] = src.ty[o];
  }
  __host__ __device__ __forceinline__ RecArr offset(int64_t o) const {
    return RecArr{r + o, ty + o};
  }
};

// bytes of n records as carved from LDS (each array 16-byte aligned)
__host__ __device__ constexpr size_t rec_bytes(size_t n) {
  return 32 * n + ((4 * n + 15) & ~size_t(15));
}

// carves n records from an LDS pointer
__device__ __forceinline__ RecArr carve_recs(char *&p, size_t n) {
  RecArr a;
  a.r = (Rec4 *)p;
  a.ty = (int32_t *)(p + 32 * n);
  p += rec_bytes(n);
  return a;
}

// Per-set prepared neighbour data (workspace).
struct NbSet {
  RecArr rec;       // [slot*amax + m]
  double2 *rc;      // [slot*amax + m] raw (row, col) type channels
  int32_t *cnt;     // [slot]
  int32_t *asc;     // [slot] 1 when d[0..cnt) is ascending
  double *self;     // [slot] self overlap (FP64 form)
  float *self32;    // [slot] self overlap (FP32 form)
  double *zr, *zc;  // [slot] PTP (row, col) of the atom itself
};

// Per-call control block in the workspace.
struct Ctl {
  int32_t generic;              // 1: use the raw-channel path
  int32_t ntypes;
  int32_t bitmap[kTypeBitmap];  // types seen
  int32_t id[kTypeBitmap];      // compact id per type, or -1
  double ttab[kMaxTypes * kMaxTypes];
};

__device__ __forceinline__ bool small_int(double v, int limit) {
  return v >= 0.0 && v < (double)limit && v == (double)(int)v;
}

// ((rw2/(rw2+dr^2))*cw2)/(cw2+dc^2), farad_kernels.f90:80-81 and 590-592.
__device__ __forceinline__ double stoch(double dr, double dc, double rw2, double cw2) {
  return rw2 / (rw2 + dr * dr) * cw2 / (cw2 + dc * dc);
}

// exp_neg's 2^(j/128) table (copied to LDS by the kernels) and sin_libm's
// sin/cos table.
__constant__ ExpPair c_exp_pairs[128] = {FMLH_EXP2_128_PAIRS};
__constant__ SinCosEntry c_sincos[116] = {FMLH_SINCOS_128};

// ---- preparation ------------------------------------------------------

// One thread per atom slot.  per_atom_n: n is indexed by slot (atomic
// forms), else by molecule (slot / amax) and slots a >= n are empty.
__global__ __launch_bounds__(256) void k_arad_prep(const double *__restrict__ q,
                                                   const double *__restrict__ zt,
                                                   const int32_t *__restrict__ n, int nslots,
                                                   int amax, int per_atom_n, int sin_guard,
                                                   AradParams p, NbSet set, Ctl *ctl) {
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
  N = min(max(N, 0), amax);
  const double *x = q + (int64_t)slot * 5 * amax;
  int cnt = 0, asc = 1;
  bool generic = false;
  double prev = 0.0;
  if (valid) {
    for (int m = 0; m < N; ++m) {
      const double d = x[m];
      if (d > p.cut) break;  // farad_kernels.f90:68,72
      const double rowv = x[amax + m], colv = x[2 * amax + m];
      // global kernel: sine only where d < cut (farad_kernels.f90:526,539)
      const double sv = (!sin_guard || d < p.cut) ? 1.0 - sin_libm(d * p.inv_cut, c_sincos) : 0.0;
      const int64_t o = (int64_t)slot * amax + m;
      set.rec.put(o, Rec{d, sv, x[3 * amax + m], x[4 * amax + m], kTypeNone, 0});
      set.rc[o] = make_double2(rowv, colv);
      if (small_int(rowv, kTypeRows) && small_int(colv, kTypeCols)) {
        // a handful of types cover every neighbour: set each bit once
        // instead of serialising thousands of atomics on the same words
        int32_t *bit = &ctl->bitmap[(int)rowv * kTypeCols + (int)colv];
        if (__atomic_load_n(bit, __ATOMIC_RELAXED) == 0) atomicOr(bit, 1);
      } else {
        generic = true;
      }
      if (cnt > 0 && !(prev <= d)) asc = 0;
      prev = d;
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
  if (generic) atomicOr(&ctl->generic, 1);
}

__global__ void k_ctl_clear(Ctl *ctl) {
  for (int i = threadIdx.x; i < kTypeBitmap; i += blockDim.x) ctl->bitmap[i] = 0;
  if (threadIdx.x == 0) {
    ctl->generic = 0;
    ctl->ntypes = 0;
  }
}

// One workgroup of 512 threads: number the types seen in bitmap order and
// fill the compact type-pair table.
__global__ __launch_bounds__(512) void k_arad_types(Ctl *ctl, double rw2, double cw2) {
  __shared__ int ids[kTypeBitmap];
  __shared__ int rows[kMaxTypes], cols[kMaxTypes];
  __shared__ int ntypes;
  const int t = threadIdx.x;
  if (t == 0) {
    int k = 0;
    for (int i = 0; i < kTypeBitmap; ++i) {
      if (ctl->bitmap[i]) {
        if (k < kMaxTypes) {
          rows[k] = i / kTypeCols;
          cols[k] = i % kTypeCols;
        }
        ids[i] = k < kMaxTypes ? k : -1;
        ++k;
      } else {
        ids[i] = -1;
      }
    }
    ntypes = k;
    ctl->ntypes = k;
    if (k > kMaxTypes) ctl->generic = 1;
  }
  __syncthreads();
  ctl->id[t] = ids[t];
  if (t < kMaxTypes * kMaxTypes) {
    const int i = t / kMaxTypes, j = t % kMaxTypes;
    const int k = min(ntypes, kMaxTypes);
    ctl->ttab[t] = (i < k && j < k)
                       ? stoch((double)(rows[i] - rows[j]), (double)(cols[i] - cols[j]), rw2, cw2)
                       : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_arad_type_ids(NbSet set, int nslots, int amax,
                                                       const Ctl *__restrict__ ctl) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= (int64_t)nslots * amax) return;
  const int slot = (int)(o / amax), m = (int)(o % amax);
  int32_t id = kTypeNone;
  if (m < set.cnt[slot]) {
    const double2 v = set.rc[o];
    if (small_int(v.x, kTypeRows) && small_int(v.y, kTypeCols)) {
      const int c = ctl->id[(int)v.x * kTypeCols + (int)v.y];
      if (c >= 0) id = c;
    }
  }
  if (m < set.cnt[slot]) set.rec.ty[o] = id;
}

// ---- the overlap ------------------------------------------------------

// One atom's neighbour list: the arrays' bases (the same for every lane of
// a launch or workgroup, so they stay in scalar registers) and the list's
// first index.
struct NbView {
  RecArr rec;
  const double2 *rc;  // raw channels, generic path only
  int off;            // first record
  int roff;           // first raw-channel pair
  __device__ __forceinline__ Rec at(int m) const { return rec.at(off + m); }
  __device__ __forceinline__ double d(int m) const { return rec.d(off + m); }
  __device__ __forceinline__ double2 raw(int m) const { return rc[roff + m]; }
};

template <typename TC>
__device__ __forceinline__ TC texp(TC x, const ExpPair *tab, double c2, double c4);
template <>
__device__ __forceinline__ double texp<double>(double x, const ExpPair *tab, double c2, double c4) {
  return exp_neg(x, tab, c2, c4);
}
template <>
__device__ __forceinline__ float texp<float>(float x, const ExpPair *, double, double) {
  return expf(x);
}

// atomic_distl2 for atoms x (list 1) and y (list 2).  kGeneric evaluates T
// from the raw channels; otherwise ttab is the compact type-pair table.
// etab: exp_neg's table.
template <typename TC, bool kGeneric>
__device__ TC arad_distl2(const NbView &x, int cnt1, bool asc1, const NbView &y, int cnt2,
                          bool asc2, TC inv_w, TC maxg, const TC *__restrict__ ttab,
                          const ExpPair *__restrict__ etab, double rw2, double cw2) {
  TC acc = TC(0);
  // exp_neg's constants C2 and C4 held in vector registers for the whole
  // call (the empty asm makes them opaque values), so its two FMAs with a
  // second constant need no register copy per term
  double c2 = FMLH_EXP_C2, c4 = FMLH_EXP_C4;
#if __HIP_DEVICE_COMPILE__
  asm volatile("" : "+v"(c2), "+v"(c4));
#endif
  // the quantity farad_kernels.f90:78-83 adds for record yr (index m2) of
  // list 2, in the reference's order: d = (exp(r2 inv_w) s1) s2; d = d T;
  // d ((1 + a1 a2) + b1 b2)
  auto rec_value = [&](const Rec &yr, int m2, TC d1, TC s1, TC a1, TC b1, int trow,
                       double2 rc1) -> TC {
    const TC dr = (TC)yr.d - d1;
    const TC r2 = dr * dr;
    TC v = texp<TC>(r2 * inv_w, etab, c2, c4) * s1;
    v = v * (TC)yr.s;
    TC t;
    if (kGeneric) {
      const double2 rc2 = y.raw(m2);
      t = (TC)stoch(rc1.x - rc2.x, rc1.y - rc2.y, rw2, cw2);
    } else {
      t = ttab[trow + yr.ty];
    }
    v = v * t;
    TC ang = TC(1) + a1 * (TC)yr.a;
    ang = ang + b1 * (TC)yr.b;
    return v * ang;
  };
  auto term = [&](int m2, TC d1, TC s1, TC a1, TC b1, int trow, double2 rc1) {
    acc = acc + rec_value(y.at(m2), m2, d1, s1, a1, b1, trow, rc1);
  };
  if (asc1 && asc2) {
    int lo = 0, hi = 0;
    for (int m1 = 0; m1 < cnt1; ++m1) {
      const Rec xr = x.at(m1);
      const TC d1 = (TC)xr.d;
      while (lo < cnt2) {
        const TC d2 = (TC)y.d(lo);
        const TC dr = d2 - d1;
        if (dr * dr < maxg || d2 > d1) break;
        ++lo;
      }
      if (hi < lo) hi = lo;
      while (hi < cnt2) {
        const TC dr = (TC)y.d(hi) - d1;
        if (!(dr * dr < maxg)) break;
        ++hi;
      }
      if (lo < hi) {
        const int trow = kGeneric ? 0 : xr.ty * kMaxTypes;
        const double2 rc1 = kGeneric ? x.raw(m1) : make_double2(0.0, 0.0);
        const TC s1 = (TC)xr.s, a1 = (TC)xr.a, b1 = (TC)xr.b;
        // two terms per trip, added to acc one after the other in m2 order;
        // their records are read at the top of the trip (the other waves of
        // the SIMD cover the LDS latency)
        int m2 = lo;
        for (int k = (hi - lo) >> 1; k > 0; --k, m2 += 2) {
          const Rec r0 = y.at(m2), r1 = y.at(m2 + 1);
          const TC v0 = rec_value(r0, m2, d1, s1, a1, b1, trow, rc1);
          const TC v1 = rec_value(r1, m2 + 1, d1, s1, a1, b1, trow, rc1);
          acc = acc + v0;
          acc = acc + v1;
        }
        if ((hi - lo) & 1) acc = acc + rec_value(y.at(m2), m2, d1, s1, a1, b1, trow, rc1);
      }
    }
  } else {
    for (int m1 = 0; m1 < cnt1; ++m1) {
      const Rec xr = x.at(m1);
      const TC d1 = (TC)xr.d;
      const int trow = kGeneric ? 0 : xr.ty * kMaxTypes;
      const double2 rc1 = kGeneric ? x.raw(m1) : make_double2(0.0, 0.0);
      const TC s1 = (TC)xr.s, a1 = (TC)xr.a, b1 = (TC)xr.b;
      for (int m2 = 0; m2 < cnt2; ++m2) {
        const TC dr = (TC)y.d(m2) - d1;
        if (dr * dr < maxg) term(m2, d1, s1, a1, b1, trow, rc1);
      }
    }
  }
  return acc;
}

__device__ __forceinline__ NbView global_view(const NbSet &s, int64_t slot, int amax) {
  return NbView{s.rec, s.rc, (int)(slot * amax), (int)(slot * amax)};
}

// LDS copies of the type-pair table and of exp_neg's table.
template <typename TC>
__device__ void load_tables(TC *ttab, ExpPair *etab, const Ctl *ctl) {
  for (int t = threadIdx.x; t < kMaxTypes * kMaxTypes; t += blockDim.x) ttab[t] = (TC)ctl->ttab[t];
  for (int t = threadIdx.x; t < 128; t += blockDim.x) etab[t] = c_exp_pairs[t];
}

// bytes of the two tables at the start of a kernel's dynamic LDS (the
// type-pair table, then exp_neg's, 16-byte aligned)
constexpr size_t kTabBytes = sizeof(double) * kMaxTypes * kMaxTypes + sizeof(ExpPair) * 128;

// Self overlaps, one thread per slot (global-memory operands).
template <typename TC>
__global__ __launch_bounds__(256) void k_arad_self(NbSet set, int nslots, int amax, AradParams p,
                                                   const Ctl *__restrict__ ctl) {
  __shared__ TC ttab[kMaxTypes * kMaxTypes];
  __shared__ ExpPair etab[128];
  load_tables(ttab, etab, ctl);
  __syncthreads();
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  const int c = set.cnt[slot];
  const bool asc = set.asc[slot] != 0;
  NbView v = global_view(set, slot, amax);
  TC r;
  if (ctl->generic)
    r = arad_distl2<TC, true>(v, c, asc, v, c, asc, (TC)p.inv_w, (TC)p.maxg, ttab, etab, p.rw2,
                              p.cw2);
  else
    r = arad_distl2<TC, false>(v, c, asc, v, c, asc, (TC)p.inv_w, (TC)p.maxg, ttab, etab, p.rw2,
                               p.cw2);
  if (sizeof(TC) == 8)
    set.self[slot] = (double)r;
  else
    set.self32[slot] = (float)r;
}

// ---- molecule-pair kernel --------------------------------------------

// LDS image of one molecule.
struct MolStage {
  RecArr rec;
  int *cnt, *asc;
  double *self, *zr, *zc;
};

__device__ __forceinline__ MolStage carve_stage(char *&p, int amax, bool records) {
  MolStage m;
  const int nn = amax * amax;
  m.rec = RecArr{nullptr, nullptr};
  if (records) m.rec = carve_recs(p, nn);
  m.self = (double *)p;
  p += sizeof(double) * amax;
  m.zr = (double *)p;
  p += sizeof(double) * amax;
  m.zc = (double *)p;
  p += sizeof(double) * amax;
  m.cnt = (int *)p;
  p += sizeof(int) * amax;
  m.asc = (int *)p;
  p += sizeof(int) * amax;
  return m;
}

size_t stage_bytes(int amax, bool records) {
  const size_t nn = (size_t)amax * amax;
  size_t b = (sizeof(double) * 3 + sizeof(int) * 2) * (size_t)amax;
  if (records) b += rec_bytes(nn);
  return b;
}

template <bool kF32Self>
__device__ void stage_molecule(MolStage &m, const NbSet &set, int mol, int natoms, int amax,
                               bool records) {
  const int64_t slot0 = (int64_t)mol * amax;
  for (int a = threadIdx.x; a < natoms; a += blockDim.x) {
    m.cnt[a] = set.cnt[slot0 + a];
    m.asc[a] = set.asc[slot0 + a];
    m.self[a] = kF32Self ? (double)set.self32[slot0 + a] : set.self[slot0 + a];
    m.zr[a] = set.zr[slot0 + a];
    m.zc[a] = set.zc[slot0 + a];
  }
  if (!records) return;
  const int total = natoms * amax;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int64_t o = slot0 * amax + t;
    m.rec.copy(t, set.rec, o);
  }
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

// kStage: both molecules' records live in LDS (amax <= 29); otherwise the
// records are read from the prepared global arrays.
template <typename TC, bool kStage>
__global__ __launch_bounds__(kPairThreads) void k_arad_pairs(NbSet set1, NbSet set2, PairArgs pa,
                                                              AradParams p,
                                                              const Ctl *__restrict__ ctl) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  char *ptr = smem;
  TC *ttab = (TC *)ptr;
  ExpPair *etab = (ExpPair *)(ptr + sizeof(double) * kMaxTypes * kMaxTypes);
  ptr += kTabBytes;
  double *chunk = (double *)ptr;  // kMaxSig x kPairThreads exponentials
  ptr += sizeof(double) * kMaxSig * kPairThreads;
  MolStage mi = carve_stage(ptr, pa.amax, kStage);
  MolStage mj = carve_stage(ptr, pa.amax, kStage);
  load_tables(ttab, etab, ctl);

  const bool generic = ctl->generic != 0;
  const int amax = pa.amax;
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
      stage_molecule<sizeof(TC) == 4>(mi, set1, i, ni, amax, kStage);
      cur_i = i;
    }
    if (j != cur_j) {
      stage_molecule<sizeof(TC) == 4>(mj, set2, j, nj, amax, kStage);
      cur_j = j;
    }
    __syncthreads();

    // atom pairs t = a + arow b in chunks of kPairThreads (one per thread);
    // each chunk's exponentials go to LDS and thread k adds them to its
    // running sum for sigma k in t order: the reference's column-major sum
    // (farad_kernels.f90:602-604, :216-217)
    const bool dist_mode = (pa.mode == FMLH_ARAD_DISTANCE);
    const int nsig = (pa.mode == FMLH_ARAD_KERNEL) ? 1 : pa.nsig;
    const int arow = dist_mode ? amax : ni;
    const int npair_atoms = dist_mode ? amax * amax : ni * nj;
    double acc = 0.0;  // thread k < nsig: the sum for sigma k
    for (int t0 = 0; t0 < npair_atoms; t0 += kPairThreads) {
      const int t = t0 + (int)threadIdx.x;
      if (t < npair_atoms) {
        const int a = t % arow, b = t / arow;
        if (a >= ni || b >= nj) {  // DISTANCE mode zero padding (farad.f90:565)
          pa.out[dist_index(i, j, a, b, pa.nm1, pa.nm2, amax)] = 0.0;
        } else {
          const int64_t gx = ((int64_t)i * amax + a) * amax, gy = ((int64_t)j * amax + b) * amax;
          NbView x, y;
          if (kStage) {
            x = NbView{mi.rec, set1.rc, a * amax, (int)gx};
            y = NbView{mj.rec, set2.rc, b * amax, (int)gy};
          } else {
            x = NbView{set1.rec, set1.rc, (int)gx, (int)gx};
            y = NbView{set2.rec, set2.rc, (int)gy, (int)gy};
          }
          const TC S = generic ? arad_distl2<TC, true>(x, mi.cnt[a], mi.asc[a] != 0, y, mj.cnt[b],
                                                       mj.asc[b] != 0, inv_w, maxg, ttab, etab,
                                                       p.rw2, p.cw2)
                               : arad_distl2<TC, false>(x, mi.cnt[a], mi.asc[a] != 0, y, mj.cnt[b],
                                                        mj.asc[b] != 0, inv_w, maxg, ttab, etab,
                                                        p.rw2, p.cw2);
          // farad_kernels.f90:590-594
          const double T = stoch(mi.zr[a] - mj.zr[b], mi.zc[a] - mj.zc[b], p.rw2, p.cw2);
          double D = mi.self[a] + mj.self[b] - 2.0 * (double)S * T;
          if (fabs(D) < p.eps) D = 0.0;
          if (dist_mode) {
            pa.out[dist_index(i, j, a, b, pa.nm1, pa.nm2, amax)] = D;
          } else {
            for (int k = 0; k < nsig; ++k)  // :211, :603
              chunk[k * kPairThreads + threadIdx.x] = exp_neg_wide(D * pa.inv_s2[k], etab);
          }
        }
      }
      if (dist_mode) continue;
      __syncthreads();
      if ((int)threadIdx.x < nsig) {
        const double *ck = chunk + threadIdx.x * kPairThreads;
        const int m = min(kPairThreads, npair_atoms - t0);
        for (int u = 0; u < m; ++u) acc = acc + ck[u];
      }
      __syncthreads();
    }
    if (dist_mode) continue;
    if ((int)threadIdx.x < nsig) {
      const int k = threadIdx.x;
      if (pa.mode == FMLH_ARAD_KERNEL) {
        pa.out[(int64_t)i + (int64_t)pa.nm1 * j] = acc;
        pa.out[(int64_t)j + (int64_t)pa.nm1 * i] = acc;
      } else {
        pa.out[k + (int64_t)pa.nsig * ((int64_t)i + (int64_t)pa.nm1 * j)] = acc;
        if (pa.triangular) pa.out[k + (int64_t)pa.nsig * ((int64_t)j + (int64_t)pa.nm1 * i)] = acc;
      }
    }
  }
}

// ---- size-sorted molecule-pair tiles ---------------------------------
//
// A tile is one set-1 molecule i and J set-2 molecules that all have the
// same atom count s (a run of the size class's member list, ascending
// index), J chosen on the host so J * n_i * s atom pairs fill the workgroup.
// Workgroups take tiles from an atomic counter until the list is exhausted
// (every workgroup reaches that exit).  Both sides are staged in LDS with
// stride = atom count; a pair's per-sigma sum is reduced over its own lanes
// in a fixed order, so it depends only on that pair.

struct Tile {
  int32_t i;    // set-1 molecule
  int32_t off;  // first member in the class list
  int32_t cnt;  // J
  int32_t s;    // atom count of the J set-2 molecules
};

struct TileArgs {
  int mode;
  int nm1, amax, nsig, triangular;
  const int64_t *tile_start;  // tile_start[nm1] = number of tiles (written by the planner)
  int stage_j_recs;  // LDS capacity (records) of the J-molecule stage
  int stage_j_atoms;
  double inv_s2[kMaxSig];
  const int32_t *n1;
  const Tile *tiles;
  const int32_t *members;
  int32_t *counter;
  double *out;
  int bfast;  // atom-pair order of a tile's threads: 0 a fastest, 1 set-2 atom fastest
};

// J (set-2 molecules per tile) for a set-1 molecule of ni atoms against the
// size class s: fill `lanes` atom pairs, fit the J stage (j_recs_cap
// records) and the reduction (nsig sums per molecule pair).
__host__ __device__ inline int tile_j(int ni, int s, int nsig, int lanes, int j_recs_cap) {
  const int jmax_red = kPairThreads / (nsig > 1 ? nsig : 1);
  const int pairs = ni * s;
  int J = pairs > 0 ? (lanes / pairs > 1 ? lanes / pairs : 1) : jmax_red;
  if (s > 0) {
    const int jc = j_recs_cap / (s * s);
    J = J < (jc > 1 ? jc : 1) ? J : (jc > 1 ? jc : 1);
    const int ja = kStageAtoms / s;  // per-atom stage (32 bytes an atom) stays small
    J = J < (ja > 1 ? ja : 1) ? J : (ja > 1 ? ja : 1);
  }
  return J < jmax_red ? J : jmax_red;
}

__device__ __forceinline__ int size_class(int n, int amax) { return min(max(n, 0), amax); }

// Size classes of set 2: cls_off[s] = offset of atom count s in members
// (s = 0..amax, cls_off[amax + 1] = nm2), members = set-2 indices grouped by
// atom count, ascending within a class.  One workgroup of 1024 threads:
// counts, a scan, then one wave per class appends its members in index
// order (a ballot prefix per 64 indices), so the grouping is stable.
__global__ __launch_bounds__(1024) void k_plan_classes(const int32_t *__restrict__ n2, int nm2,
                                                       int amax, int32_t *__restrict__ cls_off,
                                                       int32_t *__restrict__ members) {
  __shared__ int cnt[kMaxAmax + 2];
  __shared__ int off[kMaxAmax + 2];
  const int tid = threadIdx.x;
  if (tid <= amax + 1) cnt[tid] = 0;
  __syncthreads();
  for (int j = tid; j < nm2; j += blockDim.x) atomicAdd(&cnt[size_class(n2[j], amax)], 1);
  __syncthreads();
  if (tid == 0) {
    int a = 0;
    for (int s = 0; s <= amax; ++s) {
      off[s] = a;
      cls_off[s] = a;
      a += cnt[s];
    }
    cls_off[amax + 1] = a;
  }
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int s = wave; s <= amax; s += nwaves) {
    int base = off[s];
    for (int j0 = 0; j0 < nm2; j0 += 64) {
      const int j = j0 + lane;
      const bool hit = j < nm2 && size_class(n2[j], amax) == s;
      const uint64_t m = __ballot(hit);
      if (hit) members[base + __popcll(m & below)] = j;
      base += __popcll(m);
    }
  }
}

// Number of tiles of each set-1 molecule i (one thread per i): per size
// class, the class's member run (for the triangular form only members
// j >= i, found by bisection; members ascend within a class) in groups of J.
__global__ __launch_bounds__(256) void k_plan_counts(const int32_t *__restrict__ n1, int nm1,
                                                     int amax, int tri, int nsig, int lanes,
                                                     int j_recs_cap,
                                                     const int32_t *__restrict__ cls_off,
                                                     const int32_t *__restrict__ members,
                                                     int32_t *__restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm1) return;
  const int ni = size_class(n1[i], amax);
  int t = 0;
  for (int s = amax; s >= 0; --s) {
    int lo = cls_off[s];
    const int hi = cls_off[s + 1];
    if (tri) {
      int a = lo, b = hi;
      while (a < b) {
        const int m = (a + b) >> 1;
        if (members[m] < i) a = m + 1; else b = m;
      }
      lo = a;
    }
    if (lo < hi) {
      const int J = tile_j(ni, s, nsig, lanes, j_recs_cap);
      t += (hi - lo + J - 1) / J;
    }
  }
  counts[i] = t;
}

// tile_start[i] = sum of counts[0..i) for i = 0..nm1 (tile_start[nm1] = the
// number of tiles).  One workgroup of 1024 threads; a wave-shuffle scan per
// 1024-entry chunk, chunks chained through a 64-bit carry.
__global__ __launch_bounds__(1024) void k_plan_scan(const int32_t *__restrict__ counts, int nm1,
                                                    int64_t *__restrict__ tile_start) {
  __shared__ int wsum[16];
  __shared__ int wpre[16];
  __shared__ int64_t carry;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nm1; base += 1024) {
    const int i = base + tid;
    const int v = i < nm1 ? counts[i] : 0;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (tid == 0) {
      int a = 0;
      for (int w = 0; w < 16; ++w) {
        wpre[w] = a;
        a += wsum[w];
      }
      wsum[0] = a;  // chunk total, read below before the carry moves
    }
    __syncthreads();
    if (i < nm1) tile_start[i] = carry + (int64_t)(wpre[wave] + x - v);
    __syncthreads();
    if (tid == 0) carry += wsum[0];
    __syncthreads();
  }
  if (tid == 0) tile_start[nm1] = carry;
}

// Tile list on the device, one wave per set-1 molecule i: its tiles start at
// tile_start[i]; size classes largest first, each class's member run (for
// the triangular form only members j >= i) chunked into groups of J.
__global__ __launch_bounds__(64) void k_plan_emit(const int32_t *__restrict__ n1, int nm1, int amax,
                                                  int tri, int nsig, int lanes, int j_recs_cap,
                                                  const int32_t *__restrict__ cls_off,
                                                  const int32_t *__restrict__ members,
                                                  const int64_t *__restrict__ tile_start,
                                                  Tile *__restrict__ tiles) {
  __shared__ int cnt_s[kMaxAmax + 1], lo_s[kMaxAmax + 1], j_s[kMaxAmax + 1];
  __shared__ int64_t base_s[kMaxAmax + 1];
  const int i = blockIdx.x;
  if (i >= nm1) return;
  const int ni = min(max(n1[i], 0), amax);
  const int s = threadIdx.x;
  if (s <= amax) {
    int lo = cls_off[s];
    const int hi = cls_off[s + 1];
    if (tri) {  // first member >= i (members ascending within a class)
      int a = lo, b = hi;
      while (a < b) {
        const int m = (a + b) >> 1;
        if (members[m] < i) a = m + 1; else b = m;
      }
      lo = a;
    }
    const int J = tile_j(ni, s, nsig, lanes, j_recs_cap);
    lo_s[s] = lo;
    j_s[s] = J;
    cnt_s[s] = lo < hi ? (hi - lo + J - 1) / J : 0;
  }
  __syncthreads();
  if (s <= amax) {
    int64_t b = tile_start[i];
    for (int t = amax; t > s; --t) b += cnt_s[t];
    base_s[s] = b;
  }
  __syncthreads();
  for (int c = amax; c >= 0; --c) {
    const int lo = lo_s[c], J = j_s[c], hi = cls_off[c + 1];
    for (int k = threadIdx.x; k < cnt_s[c]; k += blockDim.x) {
      const int o = lo + k * J;
      tiles[base_s[c] + k] = Tile{i, o, min(J, hi - o), c};
    }
  }
}

template <bool kF32Self, bool kRecs>
__device__ void stage_compact(const RecArr &rec, int *cnt, int *asc, double *self,
                              double *zr, double *zc, const NbSet &set, int mol, int natoms,
                              int amax, int atom0) {
  const int64_t slot0 = (int64_t)mol * amax;
  for (int a = threadIdx.x; a < natoms; a += blockDim.x) {
    const int64_t sl = slot0 + a;
    cnt[atom0 + a] = set.cnt[sl];
    asc[atom0 + a] = set.asc[sl];
    self[atom0 + a] = kF32Self ? (double)set.self32[sl] : set.self[sl];
    zr[atom0 + a] = set.zr[sl];
    zc[atom0 + a] = set.zc[sl];
  }
  if (!kRecs) return;
  const int total = natoms * natoms;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int a = t / natoms, m = t % natoms;
    const int64_t o = (slot0 + a) * amax + m;
    rec.copy((atom0 + a) * natoms + m, set.rec, o);
  }
}

// Instantiated per type path; the host launches both and the one that does
// not match the set's path (decided on the device) exits at once.  kWaves:
// waves per SIMD the register budget is cut for (5 where the LDS footprint
// lets five workgroups share a CU, else 4).
template <typename TC, bool kGeneric, int kWaves>
__global__ __launch_bounds__(kPairThreads, kWaves) void k_arad_tiles(NbSet set1, NbSet set2,
                                                                   TileArgs ta, AradParams p,
                                                                   const Ctl *__restrict__ ctl) {
  if ((ctl->generic != 0) != kGeneric) return;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  char *ptr = smem;
  TC *ttab = (TC *)ptr;
  ExpPair *etab = (ExpPair *)(ptr + sizeof(double) * kMaxTypes * kMaxTypes);
  ptr += kTabBytes;
  int *tile_slot = (int *)ptr;
  ptr += 16;
  // molecule-i stage: per-atom values only; its neighbour records are read
  // from the prepared global arrays (one record per m1, read ahead)
  double *selfi = (double *)ptr;
  ptr += sizeof(double) * ta.amax;
  double *zri = (double *)ptr;
  ptr += sizeof(double) * ta.amax;
  double *zci = (double *)ptr;
  ptr += sizeof(double) * ta.amax;
  int *cnti = (int *)ptr;
  ptr += sizeof(int) * ta.amax;
  int *asci = (int *)ptr;
  ptr += sizeof(int) * ta.amax;
  // J-molecule stage (reused as the reduction buffer after compute)
  char *jbase = ptr;
  const RecArr recj = carve_recs(ptr, ta.stage_j_recs);
  double *selfj = (double *)ptr;
  ptr += sizeof(double) * ta.stage_j_atoms;
  double *zrj = (double *)ptr;
  ptr += sizeof(double) * ta.stage_j_atoms;
  double *zcj = (double *)ptr;
  ptr += sizeof(double) * ta.stage_j_atoms;
  int *cntj = (int *)ptr;
  ptr += sizeof(int) * ta.stage_j_atoms;
  int *ascj = (int *)ptr;
  ptr += sizeof(int) * ta.stage_j_atoms;
  double *red = (double *)jbase;
  load_tables(ttab, etab, ctl);

  const int amax = ta.amax;
  const TC inv_w = (TC)p.inv_w, maxg = (TC)p.maxg;
  const int nsig = (ta.mode == FMLH_ARAD_KERNEL) ? 1 : ta.nsig;
  const int64_t ntiles = ta.tile_start[ta.nm1];
  int cur_i = -1;
  for (;;) {
    __syncthreads();  // previous tile's readers are done with stages/red
    if (threadIdx.x == 0) tile_slot[0] = atomicAdd(ta.counter, 1);
    __syncthreads();
    const int tix = tile_slot[0];
    if (tix >= ntiles) break;
    const Tile tl = ta.tiles[tix];
    const int i = tl.i, J = tl.cnt, s = tl.s;
    const int ni = min(max(ta.n1[i], 0), amax);
    if (i != cur_i) {
      stage_compact<sizeof(TC) == 4, false>(RecArr{nullptr, nullptr}, cnti, asci, selfi, zri, zci, set1, i,
                                            ni, amax, 0);
      cur_i = i;
    }
    for (int jj = 0; jj < J; ++jj)
      stage_compact<sizeof(TC) == 4, true>(recj, cntj, ascj, selfj, zrj, zcj, set2,
                                           ta.members[tl.off + jj], s, amax, jj * s);
    __syncthreads();

    // a thread's atom pairs (rounds of kPairThreads; at most kTileRounds,
    // the host uses tiles only where amax^2 fits) keep their D values in
    // registers until the per-sigma exponentials below.  Pair u of the tile
    // is (jj, a, b): with bfast = 0, u = jj P + a + n_i b (a fastest); with
    // bfast = 1, u = a (J s) + jj s + b, so a wave's lanes share the set-1
    // atom a (one neighbour list, one m1 loop) and differ in the set-2 atom.
    double D0 = 0.0, D1 = 0.0, D2 = 0.0, D3 = 0.0;
    const int P = ni * s;
    const int total = J * P;
    const int Js = J * s;
    auto decode = [&](int u, int &jj, int &a, int &b) {
      if (ta.bfast) {
        a = u / Js;
        const int bj = u - a * Js;
        jj = bj / s;
        b = bj - jj * s;
      } else {
        jj = u / P;
        const int rem = u - jj * P;
        a = rem % ni;
        b = rem / ni;
      }
    };
    int nr = 0;
    for (int t = threadIdx.x; t < total; t += kPairThreads, ++nr) {
      int jj, a, b;
      decode(t, jj, a, b);
      const int bj = jj * s + b;  // staged atom index of b
      const int j = ta.members[tl.off + jj];
      const int64_t xs = ((int64_t)i * amax + a) * amax;
      NbView x{set1.rec, set1.rc, (int)xs, (int)xs};
      NbView y{recj, set2.rc, bj * s, (j * amax + b) * amax};
      const TC S = arad_distl2<TC, kGeneric>(x, cnti[a], asci[a] != 0, y, cntj[bj],
                                                   ascj[bj] != 0, inv_w, maxg, ttab, etab, p.rw2,
                                                   p.cw2);
      const double T = stoch(zri[a] - zrj[bj], zci[a] - zcj[bj], p.rw2, p.cw2);
      double D = selfi[a] + selfj[bj] - 2.0 * (double)S * T;  // farad_kernels.f90:590-592
      if (fabs(D) < p.eps) D = 0.0;                            // :594
      D0 = nr == 0 ? D : D0;
      D1 = nr == 1 ? D : D1;
      D2 = nr == 2 ? D : D2;
      D3 = nr == 3 ? D : D3;
    }
    __syncthreads();  // the J stage becomes the reduction buffer
    // per sigma k, the exponential of atom pair (jj, a, b) goes to
    // red[(a + n_i b) JK + jj nsig + k], JK = J nsig: column c = jj nsig + k
    // of the buffer lists one molecule pair's terms in the reference's
    // column-major order (a fastest, farad_kernels.f90:602-604, :216-217)
    const int JK = J * nsig;
    for (int r = 0; r < nr; ++r) {
      const double D = r == 0 ? D0 : r == 1 ? D1 : r == 2 ? D2 : D3;
      int jj, a, b;
      decode(threadIdx.x + r * kPairThreads, jj, a, b);
      double *rk = red + (int64_t)(a + ni * b) * JK + jj * nsig;
      for (int k = 0; k < nsig; ++k) rk[k] = exp_neg_wide(D * ta.inv_s2[k], etab);
    }
    __syncthreads();
    // one thread per (molecule pair, sigma) adds its column sequentially: the
    // reference's sum, in its order, whatever the tile shape
    for (int c = threadIdx.x; c < JK; c += kPairThreads) {
      const int jj = c / nsig, k = c - jj * nsig;
      double v = 0.0;
      for (int u = 0; u < P; ++u) v = v + red[(int64_t)u * JK + c];
      const int j = ta.members[tl.off + jj];
      if (ta.mode == FMLH_ARAD_KERNEL) {
        ta.out[(int64_t)i + (int64_t)ta.nm1 * j] = v;
        ta.out[(int64_t)j + (int64_t)ta.nm1 * i] = v;
      } else {
        ta.out[k + (int64_t)ta.nsig * ((int64_t)i + (int64_t)ta.nm1 * j)] = v;
        if (ta.triangular) ta.out[k + (int64_t)ta.nsig * ((int64_t)j + (int64_t)ta.nm1 * i)] = v;
      }
    }
  }
}

__global__ void k_zero_counter(int32_t *c) { *c = 0; }

// Per-atom forms (farad_kernels.f90:757-1030): one thread per atom pair on
// a 16 x 16 tile; operands come from the prepared global arrays.
template <typename TC>
__global__ __launch_bounds__(256) void k_arad_atomic(NbSet set1, NbSet set2, int na1, int na2,
                                                     int amax, int mode, PairArgs pa, AradParams p,
                                                     const Ctl *__restrict__ ctl) {
  __shared__ TC ttab[kMaxTypes * kMaxTypes];
  __shared__ ExpPair etab[128];
  load_tables(ttab, etab, ctl);
  __syncthreads();
  const bool generic = ctl->generic != 0;
  const int i = blockIdx.x * 16 + (threadIdx.x & 15);
  const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i >= na1 || j >= na2) return;
  NbView x = global_view(set1, i, amax);
  NbView y = global_view(set2, j, amax);
  const TC inv_w = (TC)p.inv_w, maxg = (TC)p.maxg;
  const TC S = generic ? arad_distl2<TC, true>(x, set1.cnt[i], set1.asc[i] != 0, y, set2.cnt[j],
                                               set2.asc[j] != 0, inv_w, maxg, ttab, etab, p.rw2,
                                               p.cw2)
                       : arad_distl2<TC, false>(x, set1.cnt[i], set1.asc[i] != 0, y, set2.cnt[j],
                                                set2.asc[j] != 0, inv_w, maxg, ttab, etab, p.rw2,
                                                p.cw2);
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
        pa.out[k + (int64_t)pa.nsig * ((int64_t)i + (int64_t)na1 * j)] =
            exp_neg_wide(D * pa.inv_s2[k], etab);  // farad_kernels.f90:887
  }
}

// ---- host side --------------------------------------------------------

size_t set_bytes(int64_t nslots, int amax) {
  const size_t nn = (size_t)nslots * amax;
  size_t b = 0;
  b += WsCarver::need(sizeof(Rec4) * nn) + WsCarver::need(sizeof(int32_t) * nn);
  b += WsCarver::need(sizeof(double2) * nn);
  b += WsCarver::need(sizeof(int32_t) * nslots) * 2;
  b += WsCarver::need(sizeof(double) * nslots) * 3;
  b += WsCarver::need(sizeof(float) * nslots);
  return b;
}

bool carve_set(WsCarver &w, int64_t nslots, int amax, NbSet &s) {
  const size_t nn = (size_t)nslots * amax;
  s.rec.r = w.take<Rec4>(nn);
  s.rec.ty = w.take<int32_t>(nn);
  s.rc = w.take<double2>(nn);
  s.cnt = w.take<int32_t>(nslots);
  s.asc = w.take<int32_t>(nslots);
  s.self = w.take<double>(nslots);
  s.zr = w.take<double>(nslots);
  s.zc = w.take<double>(nslots);
  s.self32 = w.take<float>(nslots);
  return s.rec.r && s.rec.ty && s.rc && s.cnt && s.asc && s.self && s.zr && s.zc && s.self32;
}

AradParams make_params(double width, double cut, double rw, double cw, bool eps_1e11) {
  AradParams p;
  p.width = width;
  p.cut = cut;
  p.rw = rw;
  p.cw = cw;
  p.inv_w = -1.0 / (4.0 * width * width);  // farad_kernels.f90:58
  p.maxg = (8.0 * width) * (8.0 * width);  // :60
  p.rw2 = rw * rw;
  p.cw2 = cw * cw;
  const double pi = 4.0 * atan(1.0);
  p.inv_cut = pi / (2.0 * cut);  // :510
  p.eps = eps_1e11 ? (double)1.0e-11f : (double)5.0e-12f;  // :148, :505
  return p;
}

int device_cus() {
  static int cus = 0;
  if (cus == 0) cus = fmlh_device_cu_count();
  return cus > 0 ? cus : 256;
}

size_t pair_lds_bytes(int amax, bool stage) {
  return kTabBytes + sizeof(double) * kMaxSig * kPairThreads + 2 * stage_bytes(amax, stage);
}

constexpr size_t kLdsLimit = 65536;
constexpr size_t kLdsPerCu = 160 * 1024;

template <typename TC, int kWaves>
void launch_tiles(int grid, size_t lds, hipStream_t st, const NbSet &s1, const NbSet &s2,
                  const TileArgs &ta, const AradParams &p, const Ctl *ctl) {
  hipLaunchKernelGGL((k_arad_tiles<TC, false, kWaves>), dim3(grid), dim3(kPairThreads), lds, st, s1,
                     s2, ta, p, ctl);
  hipLaunchKernelGGL((k_arad_tiles<TC, true, kWaves>), dim3(grid), dim3(kPairThreads), lds, st, s1,
                     s2, ta, p, ctl);
}

// ---- tile shapes ------------------------------------------------------

struct TileLds {
  int j_recs, j_atoms;
  size_t bytes;
};

// LDS of the tile kernel for the largest J stage any tile can need: the
// planner runs on the device, so the host bounds J * s^2 records and J * s
// atoms over every (n_i, s) the tile shape allows.
TileLds tile_lds(int amax, int nsig, int lanes) {
  int j_recs = 1, j_atoms = 1;
  for (int ni = 0; ni <= amax; ++ni)
    for (int s = 1; s <= amax; ++s) {
      const int J = tile_j(ni, s, nsig, lanes, amax * amax);
      j_recs = std::max(j_recs, J * s * s);
      j_atoms = std::max(j_atoms, J * s);
    }
  // largest atom-pair count J n_i s of a tile (at most kTileRounds x 256)
  int pairs = 1;
  for (int ni = 1; ni <= amax; ++ni)
    for (int s = 1; s <= amax; ++s) pairs = std::max(pairs, tile_j(ni, s, nsig, lanes, amax * amax) * ni * s);
  TileLds L;
  L.j_atoms = j_atoms;
  // the J stage doubles as the reduction buffer (nsig x pairs doubles)
  const size_t red = sizeof(double) * (size_t)pairs * std::max(nsig, 1);
  int jr = j_recs;
  auto stage = [&](int r) {
    return rec_bytes((size_t)r) + 32 * (size_t)L.j_atoms;
  };
  while (stage(jr) < red) jr += 64;  // grow the record area until the stage covers the buffer
  L.j_recs = jr;
  L.bytes = kTabBytes + 16 + 32 * (size_t)amax + stage(jr);
  return L;
}

// Process-wide tiling table: (mode kind, amax, precision, nsig) -> tiling,
// filled by the autotuner or by fmlh_arad_tiling_set.
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, fmlh_arad_tiling> g_tune;

std::tuple<int, int, int, int> tune_key(int mode, int amax, int precision, int nsig) {
  const int kind = (mode == FMLH_ARAD_GLOBAL) ? 0 : 1;  // rectangular / triangular tile list
  return std::make_tuple(kind, amax, precision, mode == FMLH_ARAD_KERNEL ? 1 : nsig);
}

fmlh_arad_tiling default_tiling() { return fmlh_arad_tiling{kPairThreads, 0, 16, 0}; }

fmlh_arad_tiling lookup_tiling(int mode, int amax, int precision, int nsig) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(tune_key(mode, amax, precision, nsig));
  return it == g_tune.end() ? default_tiling() : it->second;
}

bool tiling_ok(const fmlh_arad_tiling &t) {
  return t.pairs_per_tile >= 64 && t.pairs_per_tile <= kTileRounds * kPairThreads &&
         (t.waves_per_simd == 0 || t.waves_per_simd == 4 || t.waves_per_simd == 5) &&
         t.groups_per_cu >= 1 && t.groups_per_cu <= 64 && (t.pair_order == 0 || t.pair_order == 1);
}

int64_t max_tiles(int nm1, int nm2, bool tri) {
  return tri ? (int64_t)nm1 * (nm1 + 1) / 2 + nm1 : (int64_t)nm1 * nm2 + nm1;
}

// Preparation shared by both entry points: control block, neighbour
// records, type ids, self overlaps.
int prepare_sets(const double *q1, const double *zt1, const int32_t *n1, int64_t ns1,
                 const double *q2, const double *zt2, const int32_t *n2, int64_t ns2, int amax,
                 bool one_set, int per_atom_n, int guard, int precision, const AradParams &p,
                 void *ws, size_t ws_bytes, NbSet &s1, NbSet &s2, Ctl *&ctl, hipStream_t st) {
  WsCarver w(ws, ws_bytes);
  ctl = w.take<Ctl>(1);
  if (!ctl || !carve_set(w, ns1, amax, s1)) return -1;
  if (!one_set) {
    if (!carve_set(w, ns2, amax, s2)) return -1;
  } else {
    s2 = s1;
  }
  hipLaunchKernelGGL(k_ctl_clear, dim3(1), dim3(256), 0, st, ctl);
  FMLH_LAUNCH_CHECK("k_ctl_clear");
  hipLaunchKernelGGL(k_arad_prep, dim3((unsigned)ceil_div64(ns1, 256)), dim3(256), 0, st, q1, zt1,
                     n1, (int)ns1, amax, per_atom_n, guard, p, s1, ctl);
  FMLH_LAUNCH_CHECK("k_arad_prep");
  if (!one_set) {
    hipLaunchKernelGGL(k_arad_prep, dim3((unsigned)ceil_div64(ns2, 256)), dim3(256), 0, st, q2,
                       zt2, n2, (int)ns2, amax, per_atom_n, guard, p, s2, ctl);
    FMLH_LAUNCH_CHECK("k_arad_prep");
  }
  hipLaunchKernelGGL(k_arad_types, dim3(1), dim3(kTypeBitmap), 0, st, ctl, p.rw2, p.cw2);
  FMLH_LAUNCH_CHECK("k_arad_types");
  for (int k = 0; k < (one_set ? 1 : 2); ++k) {
    NbSet &s = k == 0 ? s1 : s2;
    const int64_t ns = k == 0 ? ns1 : ns2;
    hipLaunchKernelGGL(k_arad_type_ids, dim3((unsigned)ceil_div64(ns * amax, 256)), dim3(256), 0,
                       st, s, (int)ns, amax, ctl);
    FMLH_LAUNCH_CHECK("k_arad_type_ids");
  }
  for (int k = 0; k < (one_set ? 1 : 2); ++k) {
    NbSet &s = k == 0 ? s1 : s2;
    const int64_t ns = k == 0 ? ns1 : ns2;
    if (precision)
      hipLaunchKernelGGL(k_arad_self<float>, dim3((unsigned)ceil_div64(ns, 256)), dim3(256), 0, st,
                         s, (int)ns, amax, p, ctl);
    else
      hipLaunchKernelGGL(k_arad_self<double>, dim3((unsigned)ceil_div64(ns, 256)), dim3(256), 0,
                         st, s, (int)ns, amax, p, ctl);
    FMLH_LAUNCH_CHECK("k_arad_self");
  }
  return 0;
}

}  // namespace

extern "C" {

size_t fmlh_arad_ws_bytes(int nm1, int nm2, int amax, int mode) {
  if (nm1 < 0 || nm2 < 0 || amax < 1) return 0;
  const bool one_set = (mode == FMLH_ARAD_SYMMETRIC || mode == FMLH_ARAD_KERNEL);
  size_t b = WsCarver::need(sizeof(Ctl)) + set_bytes((int64_t)nm1 * amax, amax);
  if (!one_set) b += set_bytes((int64_t)nm2 * amax, amax);
  if (mode != FMLH_ARAD_DISTANCE)  // tiles, members, classes, starts, counts, counter
    b += WsCarver::need(sizeof(Tile) * (size_t)max_tiles(nm1, one_set ? nm1 : nm2, one_set)) +
         WsCarver::need(sizeof(int32_t) * (size_t)(one_set ? nm1 : nm2)) +
         WsCarver::need(sizeof(int32_t) * (size_t)(amax + 2)) +
         WsCarver::need(sizeof(int64_t) * ((size_t)nm1 + 1)) +
         WsCarver::need(sizeof(int32_t) * (size_t)nm1) + WsCarver::need(64);
  return b;
}

size_t fmlh_arad_atomic_ws_bytes(int na1, int na2, int amax) {
  if (na1 < 0 || na2 < 0 || amax < 1) return 0;
  return WsCarver::need(sizeof(Ctl)) + set_bytes(na1, amax) + set_bytes(na2, amax);
}

}  // extern "C"

namespace {

int arad_check(int mode, int nm1, int &nm2, int amax, const double *sigmas_host, int nsig,
               double width, double cut, int precision) {
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
  return 0;
}

bool uses_tiles(int mode, int amax) {
  return mode != FMLH_ARAD_DISTANCE && amax <= kMaxAmax && amax * amax <= kTileRounds * kPairThreads;
}

// Body of fmlh_arad_kernels after argument checks; t is a valid tiling.
int arad_run(int mode, const double *q1, const double *zt1, const int32_t *n1, int nm1,
             const double *q2, const double *zt2, const int32_t *n2, int nm2, int amax,
             const double *sigmas_host, int nsig, double width, double cut, double r_width,
             double c_width, int precision, double *out, void *ws, size_t ws_bytes,
             const fmlh_arad_tiling &t, hipStream_t st) {
  const bool one_set = (mode == FMLH_ARAD_SYMMETRIC || mode == FMLH_ARAD_KERNEL);
  const AradParams p = (mode == FMLH_ARAD_KERNEL) ? make_params(0.2, 6.0, 1.0, 0.5, true)
                                                  : make_params(width, cut, r_width, c_width, false);
  NbSet s1, s2;
  Ctl *ctl = nullptr;
  const int guard = (mode == FMLH_ARAD_GLOBAL) ? 1 : 0;  // farad_kernels.f90:526,539
  int rc = prepare_sets(q1, zt1, n1, (int64_t)nm1 * amax, q2, zt2, n2, (int64_t)nm2 * amax, amax,
                        one_set, 0, guard, precision, p, ws, ws_bytes, s1, s2, ctl, st);
  if (rc) return rc == -1 ? -19 : rc;

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

  if (uses_tiles(mode, amax)) {
    const int nsig_eff = (mode == FMLH_ARAD_KERNEL) ? 1 : nsig;
    const int lanes = t.pairs_per_tile, jcap = amax * amax;
    const TileLds L = tile_lds(amax, nsig_eff, lanes);
    if (L.bytes > kLdsLimit) return -10;
    WsCarver w(ws, ws_bytes);
    w.take<Ctl>(1);
    NbSet d1, d2;
    carve_set(w, (int64_t)nm1 * amax, amax, d1);
    if (!one_set) carve_set(w, (int64_t)nm2 * amax, amax, d2);
    const int nm2e = one_set ? nm1 : nm2;
    Tile *dt = w.take<Tile>((size_t)max_tiles(nm1, nm2e, one_set));
    int32_t *dm = w.take<int32_t>((size_t)nm2e);
    int32_t *dcls = w.take<int32_t>((size_t)(amax + 2));
    int64_t *dstart = w.take<int64_t>((size_t)nm1 + 1);
    int32_t *dcount = w.take<int32_t>((size_t)nm1);
    int32_t *counter = w.take<int32_t>(16);
    if (!dt || !dm || !dcls || !dstart || !dcount || !counter) return -19;
    const int32_t *n2e = one_set ? n1 : n2;
    // the plan: size classes, tiles per set-1 molecule, their prefix, the list
    hipLaunchKernelGGL(k_plan_classes, dim3(1), dim3(1024), 0, st, n2e, nm2e, amax, dcls, dm);
    FMLH_LAUNCH_CHECK("k_plan_classes");
    hipLaunchKernelGGL(k_plan_counts, dim3(ceil_div(nm1, 256)), dim3(256), 0, st, n1, nm1, amax,
                       one_set ? 1 : 0, nsig_eff, lanes, jcap, dcls, dm, dcount);
    FMLH_LAUNCH_CHECK("k_plan_counts");
    hipLaunchKernelGGL(k_plan_scan, dim3(1), dim3(1024), 0, st, dcount, nm1, dstart);
    FMLH_LAUNCH_CHECK("k_plan_scan");
    hipLaunchKernelGGL(k_plan_emit, dim3(nm1), dim3(64), 0, st, n1, nm1, amax, one_set ? 1 : 0,
                       nsig_eff, lanes, jcap, dcls, dm, dstart, dt);
    FMLH_LAUNCH_CHECK("k_plan_emit");
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, counter);
    FMLH_LAUNCH_CHECK("k_zero_counter");
    TileArgs ta;
    ta.mode = mode;
    ta.nm1 = nm1;
    ta.amax = amax;
    ta.nsig = pa.nsig;
    ta.triangular = pa.triangular;
    ta.tile_start = dstart;
    ta.stage_j_recs = L.j_recs;
    ta.stage_j_atoms = L.j_atoms;
    for (int k = 0; k < kMaxSig; ++k) ta.inv_s2[k] = pa.inv_s2[k];
    ta.n1 = n1;
    ta.tiles = dt;
    ta.members = dm;
    ta.counter = counter;
    ta.out = out;
    ta.bfast = t.pair_order;
    const int64_t tmax = std::max<int64_t>(max_tiles(nm1, nm2e, one_set), 1);
    const int grid = (int)std::min<int64_t>(tmax, (int64_t)device_cus() * t.groups_per_cu);
    const bool five = t.waves_per_simd == 5 || (t.waves_per_simd == 0 && L.bytes * 5 <= kLdsPerCu);
    timing_begin(st);
    if (precision) {
      if (five)
        launch_tiles<float, 5>(grid, L.bytes, st, s1, s2, ta, p, ctl);
      else
        launch_tiles<float, 4>(grid, L.bytes, st, s1, s2, ta, p, ctl);
    } else {
      if (five)
        launch_tiles<double, 5>(grid, L.bytes, st, s1, s2, ta, p, ctl);
      else
        launch_tiles<double, 4>(grid, L.bytes, st, s1, s2, ta, p, ctl);
    }
    timing_end(st);
    FMLH_LAUNCH_CHECK("k_arad_tiles");
    return 0;
  }

  const int64_t cap = (int64_t)device_cus() * 16;
  const bool stage = pair_lds_bytes(amax, true) <= kLdsLimit;
  const size_t lds = pair_lds_bytes(amax, stage);
  if (lds > kLdsLimit) return -10;
  const int grid = (int)(pa.npairs < cap ? pa.npairs : cap);
  timing_begin(st);
  if (precision) {
    if (stage)
      hipLaunchKernelGGL((k_arad_pairs<float, true>), dim3(grid), dim3(kPairThreads), lds, st, s1,
                         s2, pa, p, ctl);
    else
      hipLaunchKernelGGL((k_arad_pairs<float, false>), dim3(grid), dim3(kPairThreads), lds, st, s1,
                         s2, pa, p, ctl);
  } else {
    if (stage)
      hipLaunchKernelGGL((k_arad_pairs<double, true>), dim3(grid), dim3(kPairThreads), lds, st, s1,
                         s2, pa, p, ctl);
    else
      hipLaunchKernelGGL((k_arad_pairs<double, false>), dim3(grid), dim3(kPairThreads), lds, st,
                         s1, s2, pa, p, ctl);
  }
  timing_end(st);
  FMLH_LAUNCH_CHECK("k_arad_pairs");
  return 0;
}

}  // namespace

extern "C" {

int fmlh_arad_kernels(int mode, const double *q1, const double *zt1, const int32_t *n1, int nm1,
                      const double *q2, const double *zt2, const int32_t *n2, int nm2, int amax,
                      const double *sigmas_host, int nsig, double width, double cut,
                      double r_width, double c_width, int precision, double *out, void *ws,
                      size_t ws_bytes, const fmlh_arad_tiling *tiling, void *stream) {
  int rc = arad_check(mode, nm1, nm2, amax, sigmas_host, nsig, width, cut, precision);
  if (rc) return rc;
  const bool one_set = (mode == FMLH_ARAD_SYMMETRIC || mode == FMLH_ARAD_KERNEL);
  if (tiling) FMLH_ARG_CHECK(tiling_ok(*tiling), 21);
  if (nm1 == 0 || nm2 == 0) return 0;
  FMLH_ARG_CHECK(q1 && zt1 && n1, 2);
  if (!one_set) FMLH_ARG_CHECK(q2 && zt2 && n2, 6);
  FMLH_ARG_CHECK(out != nullptr, 18);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= fmlh_arad_ws_bytes(nm1, nm2, amax, mode), 19);
  const fmlh_arad_tiling t = tiling ? *tiling : lookup_tiling(mode, amax, precision, nsig);
  return arad_run(mode, q1, zt1, n1, nm1, q2, zt2, n2, nm2, amax, sigmas_host, nsig, width, cut,
                  r_width, c_width, precision, out, ws, ws_bytes, t, as_stream(stream));
}

int fmlh_arad_autotune(int mode, const double *q1, const double *zt1, const int32_t *n1, int nm1,
                       const double *q2, const double *zt2, const int32_t *n2, int nm2, int amax,
                       const double *sigmas_host, int nsig, double width, double cut,
                       double r_width, double c_width, int precision, double *out, void *ws,
                       size_t ws_bytes, int reps, fmlh_arad_tiling *best, double *best_ms,
                       void *stream) {
  int rc = arad_check(mode, nm1, nm2, amax, sigmas_host, nsig, width, cut, precision);
  if (rc) return rc;
  const bool one_set = (mode == FMLH_ARAD_SYMMETRIC || mode == FMLH_ARAD_KERNEL);
  FMLH_ARG_CHECK(reps >= 1 && reps <= 1000, 21);
  FMLH_ARG_CHECK(best != nullptr, 22);
  FMLH_ARG_CHECK(uses_tiles(mode, amax), 1);
  FMLH_ARG_CHECK(nm1 > 0 && nm2 > 0, 5);
  FMLH_ARG_CHECK(q1 && zt1 && n1, 2);
  if (!one_set) FMLH_ARG_CHECK(q2 && zt2 && n2, 6);
  FMLH_ARG_CHECK(out != nullptr, 18);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= fmlh_arad_ws_bytes(nm1, nm2, amax, mode), 19);
  hipStream_t st = as_stream(stream);
  const int nsig_eff = (mode == FMLH_ARAD_KERNEL) ? 1 : nsig;
  std::vector<fmlh_arad_tiling> cand;
  for (int lanes : {128, 256, 384, 512, 768, 1024}) {
    const TileLds L = tile_lds(amax, nsig_eff, lanes);
    if (L.bytes > kLdsLimit) continue;
    for (int order = 0; order < 2; ++order) {
      cand.push_back(fmlh_arad_tiling{lanes, 4, 16, order});
      if (L.bytes * 5 <= kLdsPerCu) cand.push_back(fmlh_arad_tiling{lanes, 5, 16, order});
    }
  }
  if (cand.empty()) return -10;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) return record_error(e, "hipEventCreate(autotune)");
  float best_t = 0.f;
  int best_k = -1;
  for (size_t k = 0; k < cand.size() && rc == 0; ++k) {
    rc = arad_run(mode, q1, zt1, n1, nm1, q2, zt2, n2, nm2, amax, sigmas_host, nsig, width, cut,
                  r_width, c_width, precision, out, ws, ws_bytes, cand[k], st);
    if (rc) break;
    e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && rc == 0 && e == hipSuccess; ++r)
      rc = arad_run(mode, q1, zt1, n1, nm1, q2, zt2, n2, nm2, amax, sigmas_host, nsig, width, cut,
                    r_width, c_width, precision, out, ws, ws_bytes, cand[k], st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) {
      rc = record_error(e, "autotune timing");
      break;
    }
    if (best_k < 0 || ms < best_t) {
      best_t = ms;
      best_k = (int)k;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  // leave `out` computed with the chosen tiling
  rc = arad_run(mode, q1, zt1, n1, nm1, q2, zt2, n2, nm2, amax, sigmas_host, nsig, width, cut,
                r_width, c_width, precision, out, ws, ws_bytes, cand[best_k], st);
  if (rc) return rc;
  *best = cand[best_k];
  if (best_ms) *best_ms = best_t / reps;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[tune_key(mode, amax, precision, nsig)] = cand[best_k];
  }
  return 0;
}

int fmlh_arad_tiling_get(int mode, int amax, int precision, int nsig, fmlh_arad_tiling *out) {
  FMLH_ARG_CHECK(out != nullptr, 5);
  *out = lookup_tiling(mode, amax, precision, nsig);
  return 0;
}

int fmlh_arad_tiling_set(int mode, int amax, int precision, int nsig, const fmlh_arad_tiling *t) {
  FMLH_ARG_CHECK(mode >= FMLH_ARAD_GLOBAL && mode <= FMLH_ARAD_KERNEL, 1);
  FMLH_ARG_CHECK(amax >= 1 && amax <= kMaxAmax, 2);
  FMLH_ARG_CHECK(precision == 0 || precision == 1, 3);
  FMLH_ARG_CHECK(nsig >= 1 && nsig <= kMaxSig, 4);
  FMLH_ARG_CHECK(t != nullptr && tiling_ok(*t), 5);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[tune_key(mode, amax, precision, nsig)] = *t;
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
  Ctl *ctl = nullptr;
  int rc = prepare_sets(q1, zt1, n1, na1, q2, zt2, n2, na2, amax, false, 1, 0, 0, p, ws, ws_bytes,
                        s1, s2, ctl, st);
  if (rc) return rc == -1 ? -18 : rc;
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
                     pa, p, ctl);
  timing_end(st);
  FMLH_LAUNCH_CHECK("k_arad_atomic");
  return 0;
}

}  // extern "C"
