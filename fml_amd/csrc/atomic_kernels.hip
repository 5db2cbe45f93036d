// Atomic KRR path: molecule-pair sums over atom-pair kernels.
//
//   fget_vector_kernels_gaussian / _laplacian (fml/kernels/fkernels.f90:23-162):
//     K_s(i, j) = sum over atoms a of molecule i (set 1), b of molecule j
//                 (set 2) of exp(inv_s * dist(a, b)),
//     dist = sum_t (x_a,t - x_b,t)^2 (inv = -0.5/sigma^2) or
//            sum_t |x_a,t - x_b,t| (inv = -1/sigma);
//   fget_alpha's molecular kernel (fkernels.f90:380-435): the Laplacian form
//     with set 2 = set 1;
//   fget_alpha_from_distance's kernel (fkernels.f90:541-599): the same sums
//     of exp(inv * d) over a given atom distance matrix;
//   fget_prediction / _from_distance (fkernels.f90:480-539, 645-702):
//     y(m) = sum over test atoms a of m, training atoms b of
//            exp(inv * dist(a, b)) * alpha(b).
//
// One workgroup owns a tile of up to 128 x 128 atoms.  Row and column
// blocks are molecule aligned: the caller's block list (molecules packed
// greedily into blocks of at most 128 atoms), or by default block b holds
// the molecules whose first atom index lies in [b S, (b+1) S) with
// S = 129 - amax, so no molecule crosses a block and a block spans at most
// S - 1 + amax = 128 atoms.  (Prediction
// columns are plain 128-atom blocks: they are summed over, not split by
// molecule.)  The tile's distances come from a VALU loop (each thread an
// 8 x 8 register block of rows ty + 16 i and columns tx + 16 j, so a wave's
// LDS reads are broadcasts or 16 consecutive doubles; the sum over t in the
// reference's order, so every distance is the reference's bit for bit),
// from the matrix cores for the fast squared-distance form (|a|^2 + |b|^2 -
// 2 a.b, 16x16x4 FP64 MFMA tiles, as the dense Gaussian kernel), or from the
// distance matrix itself.
// Exponentials use the table-driven exp_neg_wide (Tang's method, rounding like a correctly rounded exp).  The epilogue passes over the tile in
// two 64-row halves through LDS: per row, the exponentials of each column
// molecule are summed in column order and stored over that molecule's first
// column (step A); per molecule pair the row sums are added in row order
// (step B), the molecule crossing row 64 carrying its partial from the first
// half to the second.  Every sum has a fixed order, so results do not depend
// on the launch or on the other molecules of the batch.
#include <math.h>

#include "common.h"
#include "exp_neg.h"
#include "gemm_nt.h"

#pragma clang fp contract(off)

namespace {

using namespace fmlh;

constexpr int kT = 128;             // atoms per tile side
constexpr int kBK = 8;              // descriptor elements per staging step
constexpr int kStage = kT + 2;      // staging stride (doubles)
constexpr int kHalf = 64;           // tile rows per epilogue pass
constexpr int kDS = kT + 1;         // distance tile stride (doubles)
constexpr int kMaxSig = 8;          // sigmas per launch
constexpr int kMaxAmax = 64;        // largest molecule

// kL2: squared distances summed over t in the reference's order (difference,
// product, sum; no fused multiply-add), bitwise the reference's; kL1: the
// L1 distance in the same order; kL2M: squared distances in the norm form
// on the matrix cores; kMatrix: distances given.
enum Source { kL2 = 0, kL1 = 1, kL2M = 2, kMatrix = 3 };

__constant__ ExpPair c_exp_pairs[128] = {FMLH_EXP2_128_PAIRS};

struct Side {
  const double *x;        // atom rows (kL1/kL2), ld doubles apart
  int64_t ld;
  const int32_t *off;     // molecule atom offsets (nm + 1); null: plain blocks
  const int32_t *blk;     // optional block starts (nb + 1 molecule indices)
  int nm, na, span, nb;   // molecules, atoms, block start interval, blocks
};

struct TileArgs {
  Side r, c;              // rows: set 1 / test atoms; columns: set 2 / training
  const double *dmat;     // kMatrix: row-major (r.na, c.na), leading dim ldd
  int64_t ldd;
  int d;                  // descriptor length (kL1/kL2)
  int nsig;
  double inv[kMaxSig];
  double *out;            // pairs: out[(j * r.nm + i) * ostride + s]
  int ostride;
  const double *w;        // weighted: column weights (c.na)
  double *part;           // weighted: part[i * c.nb + column block]
  int32_t *status;        // status[0] = 1: a block wider than 128 atoms or
                          // a block list outside the molecules
};

// First m in [0, nm) with off[m] >= v, nm if none.
__device__ int first_at_least(const int32_t *off, int nm, int v) {
  int lo = 0, hi = nm;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] >= v)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

struct Block {
  int m0, m1;  // molecules [m0, m1)
  int a0, a1;  // atoms [a0, a1)
};

__device__ Block side_block(const Side &s, int b) {
  Block g;
  if (s.off == nullptr) {
    g.m0 = g.m1 = 0;
    g.a0 = b * kT;
    g.a1 = min(s.na, g.a0 + kT);
    return g;
  }
  if (s.blk) {
    g.m0 = s.blk[b];
    g.m1 = s.blk[b + 1];
  } else {
    g.m0 = b == 0 ? 0 : first_at_least(s.off, s.nm, b * s.span);
    g.m1 = b == s.nb - 1 ? s.nm : first_at_least(s.off, s.nm, (b + 1) * s.span);
  }
  g.a0 = g.m0 < s.nm ? s.off[g.m0] : s.na;
  g.a1 = g.m1 < s.nm ? s.off[g.m1] : s.na;
  return g;
}

template <int kSrc, bool kWeighted>
__global__ __launch_bounds__(256, 2) void k_atom_tiles(TileArgs a) {
  __shared__ union {
    struct {
      double A[2][kBK * kStage];
      double B[2][kBK * kStage];
    } st;
    TileSmem<double, kT> mf;
    double D[kHalf * kDS];
  } sm;
  __shared__ double carry[kMaxSig * kT];
  __shared__ ExpPair etab[128];

  int bx, by;  // column block, row block
  if (!xcd_tile(a.c.nb, a.r.nb, bx, by)) return;
  const Block rb = side_block(a.r, by);
  const Block cb = side_block(a.c, bx);
  const int nrows = rb.a1 - rb.a0, ncols = cb.a1 - cb.a0;
  const int tid = threadIdx.x;
  const bool mols_ok = rb.m0 >= 0 && rb.m0 <= rb.m1 && rb.m1 <= a.r.nm &&
                       cb.m0 >= 0 && cb.m0 <= cb.m1 && cb.m1 <= a.c.nm;
  if (!mols_ok || nrows > kT || ncols > kT || nrows < 0 || ncols < 0) {
    if (tid == 0 && a.status) a.status[0] = 1;
    return;
  }
  const int tx = tid & 15, ty = tid >> 4;
  if (tid < 128) etab[tid] = c_exp_pairs[tid];
  double acc[8][8];
  TileAcc<double, kT> mac;

  if constexpr (kSrc == kL2M) {
    // squared distances in the norm form |a|^2 + |b|^2 - 2 a.b: the dot
    // products on the matrix cores, the norms sequential fused sums
    tile_nt<double, true, kT>(a.r.x + (int64_t)rb.a0 * a.r.ld, a.r.ld, nrows,
                              a.c.x + (int64_t)cb.a0 * a.c.ld, a.c.ld, ncols, a.d, sm.mf, mac);
    __syncthreads();  // (tile_nt's last LDS reads)
    constexpr int NB = kT / 32;
#pragma unroll
    for (int im = 0; im < NB; ++im)
#pragma unroll
      for (int in = 0; in < NB; ++in)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = tile_row<double, kT>(im, r), n = tile_col<kT>(in);
          mac.v[im][in][r] = fma(-2.0, mac.v[im][in][r], sm.mf.pn[m] + sm.mf.qn[n]);
        }
    __syncthreads();  // the norms are read before the epilogue reuses the space
  } else if constexpr (kSrc == kL1 || kSrc == kL2) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i][j] = 0.0;
    const int lrow = tid >> 1, lk = (tid & 1) * 4;
    const bool rok = lrow < nrows, cok = lrow < ncols;
    const double *xr = a.r.x + (int64_t)(rb.a0 + (rok ? lrow : 0)) * a.r.ld;
    const double *xc = a.c.x + (int64_t)(cb.a0 + (cok ? lrow : 0)) * a.c.ld;
    double pa[4], pb[4];
    auto gload = [&](int k0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = k0 + lk + q;
        pa[q] = (rok && k < a.d) ? xr[k] : 0.0;
        pb[q] = (cok && k < a.d) ? xc[k] : 0.0;
      }
    };
    auto lstore = [&](int buf) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        sm.st.A[buf][(lk + q) * kStage + lrow] = pa[q];
        sm.st.B[buf][(lk + q) * kStage + lrow] = pb[q];
      }
    };
    // Staged elements past d are zero on both sides and add exactly 0 to
    // every distance, so each step runs all kBK elements.
    const int nk = (a.d + kBK - 1) / kBK;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int buf = kt & 1;
      if (kt + 1 < nk) gload((kt + 1) * kBK);
#pragma unroll 2
      for (int k = 0; k < kBK; ++k) {
        double ra[8], rbv[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) ra[i] = sm.st.A[buf][k * kStage + ty + 16 * i];
#pragma unroll
        for (int j = 0; j < 8; ++j) rbv[j] = sm.st.B[buf][k * kStage + tx + 16 * j];
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
          for (int j = 0; j < 8; ++j) {
          {
            const double t = ra[i] - rbv[j];
            if constexpr (kSrc == kL2)
              acc[i][j] = acc[i][j] + t * t;
            else
              acc[i][j] = acc[i][j] + fabs(t);
          }
          }
      }
      if (kt + 1 < nk) lstore(buf ^ 1);
      __syncthreads();
    }
  }

  // ---- epilogue -----------------------------------------------------------
  const int32_t *roff = a.r.off + rb.m0;  // row molecule il: rows roff[il]-a0 ..
  const int Jr = rb.m1 - rb.m0;
  const int32_t *coff = kWeighted ? nullptr : a.c.off + cb.m0;
  const int Jc = kWeighted ? 1 : cb.m1 - cb.m0;
  const int nsig = kWeighted ? 1 : a.nsig;

  for (int h = 0; h < 2; ++h) {
    if (h == 1 && nrows <= kHalf) break;  // every row molecule ended in the first half
    const int hr0 = h * kHalf;
    const int hrows = min(kHalf, nrows - hr0);  // may be <= 0
    for (int s = 0; s < nsig; ++s) {
      // stage this half's distances
      if constexpr (kSrc == kMatrix) {
        for (int t = tid; t < kHalf * kT; t += 256) {
          const int p = t / kT, q = t % kT;
          sm.D[p * kDS + q] = (p < hrows && q < ncols)
                                  ? a.dmat[(int64_t)(rb.a0 + hr0 + p) * a.ldd + cb.a0 + q]
                                  : 0.0;
        }
      } else if constexpr (kSrc == kL2M) {
        // wave (wm, wn) holds rows wm*64 .. wm*64+63: all in half wm
        if (tile_wm() == h) {
          constexpr int NB = kT / 32;
#pragma unroll
          for (int im = 0; im < NB; ++im)
#pragma unroll
            for (int in = 0; in < NB; ++in)
#pragma unroll
              for (int r = 0; r < 4; ++r)
                sm.D[(tile_row<double, kT>(im, r) - hr0) * kDS + tile_col<kT>(in)] = mac.v[im][in][r];
        }
      } else {
        // thread (tx, ty) holds rows ty + 16 i, columns tx + 16 j
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          if ((i >> 2) != h) continue;
#pragma unroll
          for (int j = 0; j < 8; ++j) sm.D[(ty + 16 * i - hr0) * kDS + tx + 16 * j] = acc[i][j];
        }
      }
      __syncthreads();
      // step A: per row, column-molecule sums over each molecule's first column
      const double inv = a.inv[s];
      {
        const int p = tid >> 2, qtr = tid & 3;
        if (p < hrows) {
          double *Dp = sm.D + p * kDS;
          if constexpr (kWeighted) {
            const int q0 = qtr * 32, q1 = min(ncols, q0 + 32);
            double v = 0.0;
            for (int q = q0; q < q1; ++q) v = v + exp_neg_wide(inv * Dp[q], etab) * a.w[cb.a0 + q];
            Dp[q0] = v;
          } else {
            const int lo = cb.a0 + qtr * 32;
            int jl = first_at_least(coff, Jc, lo);
            for (; jl < Jc; ++jl) {
              const int cs = coff[jl] - cb.a0;
              if (cs >= qtr * 32 + 32) break;
              const int ce = (jl + 1 < Jc ? coff[jl + 1] : cb.a1) - cb.a0;
              if (ce == cs) continue;
              double v = 0.0;
              for (int q = cs; q < ce; ++q) v = v + exp_neg_wide(inv * Dp[q], etab);
              Dp[cs] = v;
            }
          }
        }
      }
      __syncthreads();
      // step B: per molecule pair, row sums in row order
      for (int t = tid; t < Jr * Jc; t += 256) {
        const int il = t / Jc, jl = t % Jc;
        const int rs = roff[il] - rb.a0;
        const int re = (il + 1 < Jr ? roff[il + 1] : rb.a1) - rb.a0;
        const int lo = max(rs, hr0), hi = min(re, hr0 + kHalf);
        const bool finished = re <= hr0 + kHalf && (re > hr0 || h == 0);
        if (lo >= hi && !finished) continue;
        int cs, ce;
        if constexpr (kWeighted) {
          cs = 0;
          ce = ncols;
        } else {
          cs = coff[jl] - cb.a0;
          ce = (jl + 1 < Jc ? coff[jl + 1] : cb.a1) - cb.a0;
        }
        const bool cnon = ce > cs;
        double v = 0.0;
        if (h == 1 && cnon && rs < kHalf && re > kHalf) v = carry[s * kT + cs];
        if (cnon) {
          for (int p = lo; p < hi; ++p) {
            const double *Dp = sm.D + (p - hr0) * kDS;
            if constexpr (kWeighted)
              v = v + (((Dp[0] + Dp[32]) + Dp[64]) + Dp[96]);
            else
              v = v + Dp[cs];
          }
        }
        if (finished) {
          const int64_t i = rb.m0 + il;
          if constexpr (kWeighted)
            a.part[i * a.c.nb + bx] = v;
          else
            a.out[((int64_t)(cb.m0 + jl) * a.r.nm + i) * a.ostride + s] = v;
        } else if (cnon) {
          carry[s * kT + cs] = v;
        }
      }
      __syncthreads();
    }
  }
}

// y[i] = sum over column blocks of part[i * nb + b], in block order.
__global__ void k_part_rows(const double *__restrict__ part, int nm, int nb, double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  double v = 0.0;
  for (int b = 0; b < nb; ++b) v = v + part[(int64_t)i * nb + b];
  y[i] = v;
}

__global__ void k_zero_status(int32_t *s) { s[0] = 0; }

// Molecule-aligned blocks: the caller's block starts (blk, nb blocks), or
// the molecules whose first atom lies in [b S, (b+1) S), S = 129 - amax.
Side mol_side(const double *x, int64_t ld, const int32_t *off, int nm, int na, int amax,
              const int32_t *blk, int nb) {
  Side s;
  s.x = x;
  s.ld = ld;
  s.off = off;
  s.blk = blk;
  s.nm = nm;
  s.na = na;
  s.span = kT + 1 - amax;
  s.nb = blk ? nb : (na > 0 ? (na + s.span - 1) / s.span : 1);
  return s;
}

Side plain_side(const double *x, int64_t ld, int na) {
  Side s;
  s.x = x;
  s.ld = ld;
  s.off = nullptr;
  s.blk = nullptr;
  s.nm = 0;
  s.na = na;
  s.span = kT;
  s.nb = (na + kT - 1) / kT;
  return s;
}

template <int kSrc, bool kWeighted>
int launch_tiles(const TileArgs &a, hipStream_t st, const char *name) {
  const int64_t grid = xcd_grid(a.c.nb, a.r.nb);
  if (grid > 0x7fffffff) return -1;
  hipLaunchKernelGGL((k_atom_tiles<kSrc, kWeighted>), dim3((unsigned)grid), dim3(256), 0, st, a);
  FMLH_LAUNCH_CHECK(name);
  return 0;
}

int zero_status(int32_t *status, hipStream_t st) {
  if (!status) return 0;
  hipLaunchKernelGGL(k_zero_status, dim3(1), dim3(1), 0, st, status);
  FMLH_LAUNCH_CHECK("k_zero_status");
  return 0;
}

size_t predict_ws(int nm_test, int na_train) {
  const int64_t nb = na_train > 0 ? ((int64_t)na_train + kT - 1) / kT : 0;
  return WsCarver::need(sizeof(double) * (size_t)nm_test * (size_t)nb);
}

}  // namespace

extern "C" int fmlh_vector_kernels(int kind, const double *x1, int64_t ldx1, const int32_t *off1,
                                   int nm1, int na1, const double *x2, int64_t ldx2,
                                   const int32_t *off2, int nm2, int na2, int d, int amax,
                                   const double *sigmas, int nsig, const int32_t *blk1, int nb1,
                                   const int32_t *blk2, int nb2, double *out, int32_t *status,
                                   void *stream) {
  FMLH_ARG_CHECK(kind >= 0 && kind <= 2, 1);
  FMLH_ARG_CHECK(nm1 >= 0, 5);
  FMLH_ARG_CHECK(na1 >= 0, 6);
  FMLH_ARG_CHECK(nm2 >= 0, 10);
  FMLH_ARG_CHECK(na2 >= 0, 11);
  FMLH_ARG_CHECK(d >= 1, 12);
  FMLH_ARG_CHECK(amax >= 1 && amax <= kMaxAmax, 13);
  FMLH_ARG_CHECK(nsig >= 1 && nsig <= kMaxSig && sigmas != nullptr, 15);
  for (int s = 0; s < nsig; ++s) FMLH_ARG_CHECK(sigmas[s] > 0.0, 14);
  FMLH_ARG_CHECK(ldx1 >= d && ldx2 >= d, 2);
  if (nm1 == 0 || nm2 == 0) return 0;
  FMLH_ARG_CHECK(off1 != nullptr && off2 != nullptr, 4);
  FMLH_ARG_CHECK((x1 != nullptr || na1 == 0) && (x2 != nullptr || na2 == 0), 2);
  FMLH_ARG_CHECK(blk1 == nullptr || nb1 >= 1, 17);
  FMLH_ARG_CHECK(blk2 == nullptr || nb2 >= 1, 19);
  FMLH_ARG_CHECK(out != nullptr, 20);
  const hipStream_t st = as_stream(stream);
  int rc = zero_status(status, st);
  if (rc) return rc;
  TileArgs a = {};
  a.r = mol_side(x1, ldx1, off1, nm1, na1, amax, blk1, nb1);
  a.c = mol_side(x2, ldx2, off2, nm2, na2, amax, blk2, nb2);
  a.d = d;
  a.nsig = nsig;
  for (int s = 0; s < nsig; ++s)
    a.inv[s] = kind == 1 ? -1.0 / sigmas[s] : -0.5 / (sigmas[s] * sigmas[s]);
  a.out = out;
  a.ostride = nsig;
  a.status = status;
  timing_begin(st);
  rc = kind == 1   ? launch_tiles<kL1, false>(a, st, "k_atom_tiles<L1>")
       : kind == 2 ? launch_tiles<kL2M, false>(a, st, "k_atom_tiles<L2 MFMA>")
                   : launch_tiles<kL2, false>(a, st, "k_atom_tiles<L2>");
  timing_end(st);
  return rc;
}

extern "C" int fmlh_distance_mol_kernel(const double *dmat, int64_t ldd, const int32_t *off, int nm,
                                        int na, int amax, double gamma, const int32_t *blk, int nb,
                                        double *k, int32_t *status, void *stream) {
  FMLH_ARG_CHECK(nm >= 0, 4);
  FMLH_ARG_CHECK(na >= 0, 5);
  FMLH_ARG_CHECK(ldd >= na, 2);
  FMLH_ARG_CHECK(amax >= 1 && amax <= kMaxAmax, 6);
  FMLH_ARG_CHECK(gamma > 0.0, 7);
  if (nm == 0) return 0;
  FMLH_ARG_CHECK(off != nullptr, 3);
  FMLH_ARG_CHECK(dmat != nullptr || na == 0, 1);
  FMLH_ARG_CHECK(blk == nullptr || nb >= 1, 8);
  FMLH_ARG_CHECK(k != nullptr, 10);
  const hipStream_t st = as_stream(stream);
  int rc = zero_status(status, st);
  if (rc) return rc;
  TileArgs a = {};
  a.r = mol_side(nullptr, 0, off, nm, na, amax, blk, nb);
  a.c = a.r;
  a.dmat = dmat;
  a.ldd = ldd;
  a.nsig = 1;
  a.inv[0] = -1.0 / gamma;
  a.out = k;
  a.ostride = 1;
  a.status = status;
  return launch_tiles<kMatrix, false>(a, st, "k_atom_tiles<matrix>");
}

extern "C" size_t fmlh_atomic_predict_ws_bytes(int nm_test, int na_train) {
  return nm_test >= 0 && na_train >= 0 ? predict_ws(nm_test, na_train) : 0;
}

// Shared by the two prediction forms: the weighted tiles, then the
// per-molecule sum over column blocks.
static int predict_common(TileArgs &a, int src, double *y, void *ws, size_t ws_bytes,
                          hipStream_t st) {
  WsCarver wc(ws, ws_bytes);
  a.part = wc.take<double>((size_t)a.r.nm * (size_t)a.c.nb);
  if (a.part == nullptr) return -99;
  int rc = src == kMatrix ? launch_tiles<kMatrix, true>(a, st, "k_atom_tiles<matrix,w>")
                          : launch_tiles<kL1, true>(a, st, "k_atom_tiles<L1,w>");
  if (rc) return rc;
  hipLaunchKernelGGL(k_part_rows, dim3(ceil_div(a.r.nm, 256)), dim3(256), 0, st, a.part, a.r.nm,
                     a.c.nb, y);
  FMLH_LAUNCH_CHECK("k_part_rows");
  return 0;
}

extern "C" int fmlh_atomic_predict(const double *xt, int64_t ldt, const int32_t *off_t, int nm_t,
                                   int na_t, const double *x, int64_t ldx, int na, int d, int amax,
                                   const double *alpha, double sigma, const int32_t *blk_t,
                                   int nb_t, double *y, int32_t *status, void *ws,
                                   size_t ws_bytes, void *stream) {
  FMLH_ARG_CHECK(nm_t >= 0, 4);
  FMLH_ARG_CHECK(na_t >= 0, 5);
  FMLH_ARG_CHECK(na >= 0, 8);
  FMLH_ARG_CHECK(d >= 1, 9);
  FMLH_ARG_CHECK(ldt >= d && ldx >= d, 2);
  FMLH_ARG_CHECK(amax >= 1 && amax <= kMaxAmax, 10);
  FMLH_ARG_CHECK(sigma > 0.0, 12);
  if (nm_t == 0) return 0;
  FMLH_ARG_CHECK(off_t != nullptr, 3);
  FMLH_ARG_CHECK(blk_t == nullptr || nb_t >= 1, 13);
  FMLH_ARG_CHECK(y != nullptr, 15);
  FMLH_ARG_CHECK((xt != nullptr || na_t == 0) && (x != nullptr || na == 0), 1);
  FMLH_ARG_CHECK(alpha != nullptr || na == 0, 11);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= predict_ws(nm_t, na), 17);
  const hipStream_t st = as_stream(stream);
  int rc = zero_status(status, st);
  if (rc) return rc;
  TileArgs a = {};
  a.r = mol_side(xt, ldt, off_t, nm_t, na_t, amax, blk_t, nb_t);
  a.c = plain_side(x, ldx, na);
  a.d = d;
  a.nsig = 1;
  a.inv[0] = -1.0 / sigma;
  a.w = alpha;
  a.status = status;
  if (na == 0) {  // no training atoms: every prediction is an empty sum
    a.c.nb = 0;
    hipLaunchKernelGGL(k_part_rows, dim3(ceil_div(nm_t, 256)), dim3(256), 0, st, nullptr, nm_t, 0,
                       y);
    FMLH_LAUNCH_CHECK("k_part_rows");
    return 0;
  }
  timing_begin(st);
  rc = predict_common(a, kL1, y, ws, ws_bytes, st);
  timing_end(st);
  return rc == -99 ? -17 : rc;
}

extern "C" int fmlh_distance_predict(const double *d2, int64_t ldd, const int32_t *off_t, int nm_t,
                                     int na_t, int na, int amax, const double *alpha, double gamma,
                                     const int32_t *blk_t, int nb_t, double *y, int32_t *status,
                                     void *ws, size_t ws_bytes, void *stream) {
  FMLH_ARG_CHECK(nm_t >= 0, 4);
  FMLH_ARG_CHECK(na_t >= 0, 5);
  FMLH_ARG_CHECK(na >= 0, 6);
  FMLH_ARG_CHECK(ldd >= na, 2);
  FMLH_ARG_CHECK(amax >= 1 && amax <= kMaxAmax, 7);
  FMLH_ARG_CHECK(gamma > 0.0, 9);
  if (nm_t == 0) return 0;
  FMLH_ARG_CHECK(off_t != nullptr, 3);
  FMLH_ARG_CHECK(blk_t == nullptr || nb_t >= 1, 10);
  FMLH_ARG_CHECK(y != nullptr, 12);
  FMLH_ARG_CHECK(d2 != nullptr || na_t == 0 || na == 0, 1);
  FMLH_ARG_CHECK(alpha != nullptr || na == 0, 8);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= predict_ws(nm_t, na), 14);
  const hipStream_t st = as_stream(stream);
  int rc = zero_status(status, st);
  if (rc) return rc;
  TileArgs a = {};
  a.r = mol_side(nullptr, 0, off_t, nm_t, na_t, amax, blk_t, nb_t);
  a.c = plain_side(nullptr, 0, na);
  a.dmat = d2;
  a.ldd = ldd;
  a.nsig = 1;
  a.inv[0] = -1.0 / gamma;
  a.w = alpha;
  a.status = status;
  if (na == 0) {
    hipLaunchKernelGGL(k_part_rows, dim3(ceil_div(nm_t, 256)), dim3(256), 0, st, nullptr, nm_t, 0,
                       y);
    FMLH_LAUNCH_CHECK("k_part_rows");
    return 0;
  }
  rc = predict_common(a, kMatrix, y, ws, ws_bytes, st);
  return rc == -99 ? -14 : rc;
}
