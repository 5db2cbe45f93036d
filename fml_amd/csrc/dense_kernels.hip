// (b) Gaussian and (c) Laplacian kernels, plus the small FP64 helpers the
// KRR composition uses (diagonal shift, symmetric residual, K^T alpha).
//
// Gaussian  K(j,i) = exp(-0.5/sigma^2 * ||a_j - b_i||^2)
//   fml/kernels/fkernels.f90:317-349.  ||a||^2 + ||b||^2 - 2 a.b with the dot
//   products on MFMA (128x128 or 64x64 tiles, see gemm_nt.h) and exp fused in the
//   epilogue; an entry whose GEMM-form distance is within the cancellation
//   band (d2 <= tau*(||a||^2+||b||^2)) is recomputed from direct differences
//   after the tile's other entries are stored (recompute_near_deferred,
//   near_pairs.h).
// Laplacian K(j,i) = exp(-1/sigma * sum_t |a_jt - b_it|)
//   fml/kernels/fkernels.f90:351-377.  64x64 output tile per 256 threads,
//   4x4 entries per thread, operands staged k-major in LDS; every entry's
//   sum runs over t ascending, so the sums equal the reference's sequential
//   sum() bit for bit.
#include <math.h>

#include "common.h"
#include "gemm_nt.h"
#include "near_pairs.h"

namespace {

using namespace fmlh;



__device__ __forceinline__ double kexp(double x) { return exp(x); }
__device__ __forceinline__ float kexp(float x) { return expf(x); }

template <typename O>
__device__ __forceinline__ void kstore(O *p, double v) {
  __builtin_nontemporal_store((O)v, p);
}

// exp(inv d2), formed in double for double output
template <typename T, typename OutT>
__device__ __forceinline__ double gauss_value(T d2, double inv) {
  if (sizeof(OutT) == 8) return exp(inv * (double)d2);
  return (double)kexp((T)((T)inv * d2));
}

// Gaussian on MFMA, MT x MT output tiles (MT = 128 for large kernels, 64
// for small ones).  A 1-D grid in XCD-contiguous tile order (gemm_nt.h):
// gx tiles over A (na), gy over B (nb).  OutT is float or double; the
// exponent is formed in double for double output.  The distance of an
// entry does not depend on MT: the dot product runs over k in the same
// 4-wide MFMA steps and the row norms are sequential fused sums.
// kSym (A and B the same rows): only the lower-triangle tiles run, in
// tri_tile order, and an off-diagonal tile stores each value at (m, n) and
// (n, m).  An entry is bitwise the same either way round -- the MFMA sums
// the same products over k in the same order, the norms' sum commutes and
// a near pair's sum of squared differences is symmetric -- so the kernel
// equals the one the full grid writes, at half the dot products.
template <typename T, typename OutT, int MT, bool kSym = false>
__global__ __launch_bounds__(MT == 128 ? 512 : 256, (MT == 128 ? 2 : 1)) void k_gaussian(const T *__restrict__ A, int64_t lda, int na,
                                                  const T *__restrict__ B, int64_t ldb, int nb,
                                                  int d, double inv, OutT *__restrict__ K,
                                                  int64_t ldk, T tau, int gx, int gy) {
  __shared__ TileSmem<T, MT> s;
  int tx, ty;
  if (kSym) {
    const int64_t t = (int64_t)gy * (gy + 1) / 2;
    const int64_t per = (t + kXcds - 1) / kXcds;
    const int64_t L = blockIdx.x;
    if (!tri_tile(gy, (L % kXcds) * per + L / kXcds, tx, ty)) return;
  } else if (!xcd_tile(gx, gy, tx, ty)) {
    return;
  }
  const bool mirror = kSym && tx != ty;
  const int n0 = tx * MT;
  const int m0 = ty * MT;
  constexpr int WR = MT == 128 ? 4 : 2;
  TileAcc<T, MT, WR> acc;
  tile_nt<T, true, MT, WR>(B + (int64_t)m0 * ldb, ldb, nb - m0, A + (int64_t)n0 * lda, lda, na - n0,
                           d, s, acc);
  constexpr int NBM = TileAcc<T, MT, WR>::NBM, NBN = TileAcc<T, MT, WR>::NBN;
  static_assert(NBM * NBN * 4 <= 64, "one ballot bit per epilogue step");
  // entries in the cancellation band are stored after the tile's others, from
  // direct differences: step b = 4 (NBN im + in) + r records its ballot
  __shared__ unsigned long long s_near[2 * WR][NBM * NBN * 4];
  unsigned long long *masks = s_near[threadIdx.x >> 6];
  unsigned long long steps = 0;
  // FP32 tiles: a lane's four r are four consecutive rows gm, so the mirror
  // (row gn, columns gm .. gm + 3) is one vector store when all four are
  // plain entries of a whole tile (scattered 4-byte stores cost more than
  // the FP32 tile's arithmetic)
  const bool mvec = mirror && sizeof(T) == 4 && n0 + MT <= na && m0 + MT <= nb && (ldk & 3) == 0 &&
                    (reinterpret_cast<uintptr_t>(K) & (4 * sizeof(OutT) - 1)) == 0;
#pragma unroll
  for (int im = 0; im < NBM; ++im)
#pragma unroll
    for (int in = 0; in < NBN; ++in) {
      OutT kvs[4];
      bool plain4 = true;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = tile_row<T, MT, WR>(im, r);
        const int n = tile_col<MT>(in);
        const int gm = m0 + m, gn = n0 + n;
        const bool valid = gm < nb && gn < na;
        const T pn = s.pn[m], qn = s.qn[n];
        T d2 = fma(T(-2), acc.v[im][in][r], pn + qn);
        const bool near = valid && d2 <= tau * (pn + qn);
        const unsigned long long mk = __ballot(near);
        if (mk) {
          masks[4 * (NBN * im + in) + r] = mk;
          steps |= 1ull << (4 * (NBN * im + in) + r);
        }
        plain4 = plain4 && valid && !near;
        if (d2 < T(0)) d2 = T(0);
        const double kv = gauss_value<T, OutT>(d2, inv);
        kvs[r] = (OutT)kv;
        if (valid && !near) {
          kstore(K + (int64_t)gn + (int64_t)gm * ldk, kv);
          if (mirror && !mvec) kstore(K + (int64_t)gm + (int64_t)gn * ldk, kv);
        }
      }
      if (mvec) {
        const int gm0 = m0 + tile_row<T, MT, WR>(im, 0), gn = n0 + tile_col<MT>(in);
        OutT *dst = K + (int64_t)gm0 + (int64_t)gn * ldk;
        if (plain4) {
          typedef OutT o4v __attribute__((ext_vector_type(4)));
          __builtin_nontemporal_store(o4v{kvs[0], kvs[1], kvs[2], kvs[3]}, reinterpret_cast<o4v *>(dst));
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int gm = m0 + tile_row<T, MT, WR>(im, r);
            const T d2r = fma(T(-2), acc.v[im][in][r], s.pn[gm - m0] + s.qn[gn - n0]);
            if (!(d2r <= tau * (s.pn[gm - m0] + s.qn[gn - n0])))
              __builtin_nontemporal_store(kvs[r], dst + r);
          }
        }
      }
    }
  if (steps) {
    auto idx = [&](int b, int L, int &gn, int &gm) {
      const int im = b / (4 * NBN), in = (b >> 2) % NBN, r = b & 3, q = L >> 4;
      gm = m0 + tile_wm() * (MT / WR) + im * 16 + (sizeof(T) == 8 ? q + 4 * r : 4 * q + r);
      gn = n0 + tile_wn() * (MT / 2) + in * 16 + (L & 15);
    };
    auto put = [&](int gn, int gm, T d2) {
      const double kv = gauss_value<T, OutT>(d2, inv);
      kstore(K + (int64_t)gn + (int64_t)gm * ldk, kv);
      if (mirror) kstore(K + (int64_t)gm + (int64_t)gn * ldk, kv);
    };
    recompute_near_deferred<T>(steps, masks, idx, A, lda, B, ldb, d, put);
  }
}

// Tile size for an na x nb kernel: the 128 tile once both sides span a few
// of them and the grid still covers the chip several times.
inline int gauss_tile(int na, int nb) {
  return (na >= 512 && nb >= 512 && (int64_t)ceil_div(na, 128) * ceil_div(nb, 128) >= 1024) ? 128
                                                                                          : kTile;
}

template <typename T, typename OutT>
void launch_gaussian(const T *a, int64_t lda, int na, const T *b, int64_t ldb, int nb, int d,
                     double inv, OutT *k, int64_t ldk, T tau, hipStream_t st) {
  if (a == b && na == nb && lda == ldb && gauss_tile(na, nb) == 128) {
    // a kernel of a set with itself: lower-triangle tiles, mirrored stores
    const int g = ceil_div(na, 128);
    const int64_t t = (int64_t)g * (g + 1) / 2;
    hipLaunchKernelGGL((k_gaussian<T, OutT, 128, true>), dim3((unsigned)xcd_grid((int)t, 1)),
                       dim3(512), 0, st, a, lda, na, b, ldb, nb, d, inv, k, ldk, tau, g, g);
  } else if (gauss_tile(na, nb) == 128) {
    const int gx = ceil_div(na, 128), gy = ceil_div(nb, 128);
    hipLaunchKernelGGL((k_gaussian<T, OutT, 128>), dim3((unsigned)xcd_grid(gx, gy)), dim3(512), 0,
                       st, a, lda, na, b, ldb, nb, d, inv, k, ldk, tau, gx, gy);
  } else {
    const int gx = ceil_div(na, kTile), gy = ceil_div(nb, kTile);
    hipLaunchKernelGGL((k_gaussian<T, OutT, kTile>), dim3((unsigned)xcd_grid(gx, gy)), dim3(256),
                       0, st, a, lda, na, b, ldb, nb, d, inv, k, ldk, tau, gx, gy);
  }
}

// Laplacian on VALU over MT x MT tiles (64 or 128), 256 threads: thread
// (tx, ty) owns A samples n = tx+16j and B samples m = ty+16i, R = MT/16
// (strided, so the 16 lanes of a row group read 16 consecutive LDS words),
// and sums |a - b| over k in ascending order (the same order for either
// tile).  Operands are staged k-major through two LDS buffers, kLapBK
// columns at a time.  The launcher uses the 64 tile: the 128 tile (8 x 8
// outputs per thread) needs 210 registers, runs two waves per SIMD and is
// no faster.  kSym (A and B the same rows): lower-triangle tiles only, each
// off-diagonal value stored at (m, n) and (n, m) -- |a - b| and |b - a| are
// the same number and the sum runs over k in the same order, so the kernel
// is bitwise the one the full grid writes.
constexpr int kLapBK = 16;

template <typename T, typename OutT, int MT, bool kSym = false>
__global__ __launch_bounds__(256, (MT == 128 ? 2 : 1)) void k_laplacian(
    const T *__restrict__ A, int64_t lda, int na, const T *__restrict__ B, int64_t ldb, int nb,
    int d, double inv, OutT *__restrict__ K, int64_t ldk, int gx, int gy) {
#pragma clang fp contract(off)
  constexpr int R = MT / 16;
  constexpr int kStride = MT + 2;
  constexpr int EPT = MT * kLapBK / 256;  // staged elements per thread and operand
  constexpr int TPR = kLapBK / EPT;       // threads per staged row
  // the staging buffers, then (mirrored tiles) the tile's values for the
  // transposed store
  constexpr int kStage = 2 * kLapBK * kStride;
  constexpr size_t kStageBytes = 2 * kStage * sizeof(T);
  constexpr size_t kMirrorBytes = kSym ? (size_t)MT * (MT + 1) * sizeof(OutT) : 0;
  __shared__ __attribute__((aligned(16))) unsigned char smem[kStageBytes > kMirrorBytes ? kStageBytes : kMirrorBytes];
  T (*As)[kLapBK * kStride] = reinterpret_cast<T (*)[kLapBK * kStride]>(smem);
  T (*Bs)[kLapBK * kStride] = reinterpret_cast<T (*)[kLapBK * kStride]>(smem + kStage * sizeof(T));
  int bx, by;
  if (kSym) {
    const int64_t t = (int64_t)gy * (gy + 1) / 2;
    const int64_t per = (t + kXcds - 1) / kXcds;
    const int64_t L = blockIdx.x;
    if (!tri_tile(gy, (L % kXcds) * per + L / kXcds, bx, by)) return;
  } else if (!xcd_tile(gx, gy, bx, by)) {
    return;
  }
  const bool mirror = kSym && bx != by;
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int n0 = bx * MT, m0 = by * MT;
  const int lrow = tid / TPR, lk = (tid % TPR) * EPT;
  const bool arow_ok = n0 + lrow < na, brow_ok = m0 + lrow < nb;
  const T *arow = A + (int64_t)(n0 + lrow) * lda;
  const T *brow = B + (int64_t)(m0 + lrow) * ldb;
  T ra[EPT], rb[EPT];
  T acc[R][R];
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int j = 0; j < R; ++j) acc[i][j] = T(0);

  auto gload = [&](int k0) {
#pragma unroll
    for (int c = 0; c < EPT; ++c) {
      const int k = k0 + lk + c;
      ra[c] = (arow_ok && k < d) ? arow[k] : T(0);
      rb[c] = (brow_ok && k < d) ? brow[k] : T(0);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int c = 0; c < EPT; ++c) {
      As[buf][(lk + c) * kStride + lrow] = ra[c];
      Bs[buf][(lk + c) * kStride + lrow] = rb[c];
    }
  };
  const int nk = (d + kLapBK - 1) / kLapBK;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * kLapBK);
    const int kmax = min(kLapBK, d - kt * kLapBK);
    for (int k = 0; k < kmax; ++k) {
      T a[R], b[R];
#pragma unroll
      for (int j = 0; j < R; ++j) a[j] = As[buf][k * kStride + tx + 16 * j];
#pragma unroll
      for (int i = 0; i < R; ++i) b[i] = Bs[buf][k * kStride + ty + 16 * i];
#pragma unroll
      for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) acc[i][j] = acc[i][j] + fabs(a[j] - b[i]);
    }
    if (kt + 1 < nk) lstore(buf ^ 1);
    __syncthreads();
  }
  // (the loop's last barrier: every read of the staging buffers is done)
  OutT *Cs = reinterpret_cast<OutT *>(smem);  // Cs[n][m], row stride MT + 1
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int gm = m0 + ty + 16 * i;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int gn = n0 + tx + 16 * j;
      double v;
      if (sizeof(OutT) == 8)
        v = exp(inv * (double)acc[i][j]);
      else
        v = (double)kexp((T)((T)inv * acc[i][j]));
      if (gm < nb && gn < na) kstore(K + (int64_t)gn + (int64_t)gm * ldk, v);
      if (mirror) Cs[(tx + 16 * j) * (MT + 1) + ty + 16 * i] = (OutT)v;
    }
  }
  if (mirror) {
    // the mirror, row n0 + c over columns m0 .. m0 + MT, from the tile
    // transposed through LDS (single-value stores down the columns scatter)
    __syncthreads();
    // a wave writes whole rows, its lanes over 64 consecutive columns
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int cc = 0; cc < MT / 4; ++cc) {
      const int c = wave * (MT / 4) + cc, gn = n0 + c;
#pragma unroll
      for (int h = 0; h < MT / 64; ++h) {
        const int r = 64 * h + lane;
        if (gn < na && m0 + r < nb)
          __builtin_nontemporal_store(Cs[c * (MT + 1) + r], K + (int64_t)gn * ldk + m0 + r);
      }
    }
  }
}

template <typename T, typename OutT>
void launch_laplacian(const T *a, int64_t lda, int na, const T *b, int64_t ldb, int nb, int d,
                      double inv, OutT *k, int64_t ldk, hipStream_t st) {
  const int gx = ceil_div(na, 64), gy = ceil_div(nb, 64);
  if (a == b && na == nb && lda == ldb) {  // a set with itself: mirrored lower triangle
    const int64_t t = (int64_t)gy * (gy + 1) / 2;
    hipLaunchKernelGGL((k_laplacian<T, OutT, 64, true>), dim3((unsigned)xcd_grid((int)t, 1)),
                       dim3(256), 0, st, a, lda, na, b, ldb, nb, d, inv, k, ldk, gy, gy);
    return;
  }
  hipLaunchKernelGGL((k_laplacian<T, OutT, 64>), dim3((unsigned)xcd_grid(gx, gy)), dim3(256), 0, st,
                     a, lda, na, b, ldb, nb, d, inv, k, ldk, gx, gy);
}

__global__ void k_fill(double *p, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_add_diagonal(double *a, int64_t lda, int n, double lambda) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[(int64_t)i + (int64_t)i * lda] += lambda;
}

// Upper-triangle symmetric product, part 1: t_i = sum_{j>i} A(i,j) x_j.
// Thread i walks j; consecutive threads read consecutive addresses.
__global__ __launch_bounds__(256) void k_symv_rows(const double *__restrict__ a, int64_t lda, int n,
                                                   const double *__restrict__ x,
                                                   double *__restrict__ t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = i + 1; j < n; ++j) s += a[(int64_t)i + (int64_t)j * lda] * x[j];
  t[i] = s;
}

// Part 2 and the residual: one workgroup per column j,
// r_j = y_j - (sum_{i<=j} A(i,j) x_i + t_j + lambda x_j).
__global__ __launch_bounds__(256) void k_symv_cols_residual(
    const double *__restrict__ a, int64_t lda, int n, double lambda,
    const double *__restrict__ x, const double *__restrict__ y, const double *t, double *r) {
  __shared__ double part[256];
  const int j = blockIdx.x;
  const double *col = a + (int64_t)j * lda;
  double s = 0.0;
  for (int i = threadIdx.x; i <= j; i += 256) s += col[i] * x[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) r[j] = y[j] - (part[0] + t[j] + lambda * x[j]);
}

// y_i = sum_j K(j,i) alpha_j for K (na, nb) Fortran order; one workgroup
// per output, fixed reduction order.
__global__ __launch_bounds__(256) void k_gemv_t(const double *__restrict__ k, int64_t ldk, int na,
                                                const double *__restrict__ alpha,
                                                double *__restrict__ y) {
  __shared__ double part[256];
  const int i = blockIdx.x;
  const double *col = k + (int64_t)i * ldk;
  double s = 0.0;
  for (int j = threadIdx.x; j < na; j += 256) s += col[j] * alpha[j];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[i] = part[0];
}

// Predictions y_i = sum_j K(j,i) alpha_j in two passes, so that a batch of
// a few molecules still spreads over the chip: k_gemv_part sums a fixed
// chunk of kGemvChunk training rows of one column (eight loads in flight
// per thread), k_gemv_sum adds a column's chunk sums in chunk order.  The
// chunking depends on the training-set size only, so every prediction is
// the same sum in the same order whatever batch its molecule came in.
constexpr int kGemvChunk = 2048;

__global__ __launch_bounds__(256) void k_gemv_part(const double *__restrict__ k, int64_t ldk,
                                                   int na, int nchunk,
                                                   const double *__restrict__ alpha,
                                                   double *__restrict__ part) {
  __shared__ double wsum[4];
  const int chunk = blockIdx.x % nchunk;
  const int64_t col = blockIdx.x / nchunk;
  const double *c = k + col * ldk;
  const int j0 = chunk * kGemvChunk + (int)threadIdx.x;
  double kv[8], av[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int j = j0 + 256 * u;
    kv[u] = j < na ? c[j] : 0.0;
    av[u] = j < na ? alpha[j] : 0.0;
  }
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < 8; ++u) s += kv[u] * av[u];
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) s += __shfl_xor(s, w, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[col * nchunk + chunk] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

__global__ __launch_bounds__(256) void k_gemv_sum(const double *__restrict__ part, int nb,
                                                  int nchunk, double *__restrict__ y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nb) return;
  const double *p = part + (int64_t)i * nchunk;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += p[c];
  y[i] = s;
}

// r_i = y_i - (sum_j K(j,i) x_j + lambda x_i) for K (n, n) Fortran order
// (column i contiguous), one workgroup per output, every product and sum
// error-free: the product by TwoProduct (p = k x, e = fma(k, x, -p)) and
// the running sum as an unevaluated pair (hi, lo) through TwoSum, the
// accumulation of Ogita, Rump and Oishi ("Accurate sum and dot product",
// SIAM J. Sci. Comput. 26 (2005), Dot2), so the sum is as accurate as if
// computed in twice the working precision and then rounded once.  Partial
// pairs are combined in a fixed tree order.  Contraction is off in all
// three: fused into an FMA, hi + a * b would not be the rounded sum the
// error terms assume.
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
#pragma clang fp contract(off)
  s = a + b;
  const double z = s - a;
  e = (a - (s - z)) + (b - z);
}

__device__ __forceinline__ void dd_add(double &hi, double &lo, double bh, double bl) {
#pragma clang fp contract(off)
  double s, e;
  two_sum(hi, bh, s, e);
  e += lo + bl;
  hi = s + e;
  lo = e - (hi - s);
}

__global__ __launch_bounds__(256) void k_residual_dd(const double *__restrict__ k, int64_t ldk,
                                                     int n, double lambda,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ y,
                                                     double *__restrict__ r) {
#pragma clang fp contract(off)
  __shared__ double ph[256], pl[256];
  const int i = blockIdx.x;
  const double *col = k + (int64_t)i * ldk;
  double hi = 0.0, lo = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) {
    const double a = col[j], b = x[j];
    const double p = a * b;
    const double e = fma(a, b, -p);
    double s, t;
    two_sum(hi, p, s, t);
    hi = s;
    lo += t + e;
  }
  ph[threadIdx.x] = hi;
  pl[threadIdx.x] = lo;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      double h = ph[threadIdx.x], l = pl[threadIdx.x];
      dd_add(h, l, ph[threadIdx.x + w], pl[threadIdx.x + w]);
      ph[threadIdx.x] = h;
      pl[threadIdx.x] = l;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double h = ph[0], l = pl[0];
    const double p = lambda * x[i];
    dd_add(h, l, p, fma(lambda, x[i], -p));  // + lambda x_i, exactly
    double s, e;
    two_sum(y[i], -h, s, e);  // y_i - (h + l)
    r[i] = s + (e - l);
  }
}

// The same residual for a matrix that shares its storage with its Cholesky
// factor: K's strictly lower triangle (Fortran order; column j, contiguous,
// holds K(i, j) for i > j) and its diagonal in a separate vector, the upper
// triangle not read.  Pass one (k_residual_dd_tiles) reads every 64 x 64
// tile of the strictly lower triangle once and uses it twice, for the
// outputs of its columns (sums down a column, from registers) and of its
// rows (sums along a row, through LDS), each sum over the tile's 64 entries
// error-free and double length; output i receives exactly one partial per
// tile index t of the other side, written to part[t][i].  Pass two
// (k_residual_dd_finish) adds the partials of output i in t order, then
// K(i,i) x_i and lambda x_i, and rounds y_i minus the sum once.  Every
// order is fixed by n, so the result is deterministic.
constexpr int kRT = 64;

__device__ __forceinline__ void dot2_acc(double a, double b, double &hi, double &lo) {
#pragma clang fp contract(off)
  const double p = a * b;
  const double e = fma(a, b, -p);
  double s, t;
  two_sum(hi, p, s, t);
  hi = s;
  lo += t + e;
}

// the four lanes of a quad (lanes 4q .. 4q + 3) combined in a fixed order;
// lane 4q ends with ((p0 + p1) + (p2 + p3))
__device__ __forceinline__ void dd_quad_sum(double &hi, double &lo) {
#pragma unroll
  for (int m = 1; m <= 2; m <<= 1) {
    const double oh = __shfl_xor(hi, m, 64), ol = __shfl_xor(lo, m, 64);
    dd_add(hi, lo, oh, ol);
  }
}

// This thread's 16 entries of tile (tc, tr): column j = 64 tc + q, rows
// 64 tr + e0 + e.  Edge tiles load every entry in bounds (clamped) and zero
// those outside the strictly lower triangle afterwards.
__device__ __forceinline__ void residual_tile_load(const double *__restrict__ a, int64_t lda, int n,
                                                   int tc, int tr, int q, int e0, double *v) {
  const int j = tc * kRT + q, i0 = tr * kRT + e0;
  if (tr > tc && (tr + 1) * kRT <= n) {
    const double *src = a + (int64_t)j * lda + i0;
    if (((lda & 1) == 0) && ((reinterpret_cast<uintptr_t>(a) & 15) == 0)) {
#pragma unroll
      for (int e = 0; e < 16; e += 2) {
        const double2 w = *reinterpret_cast<const double2 *>(src + e);
        v[e] = w.x;
        v[e + 1] = w.y;
      }
    } else {
#pragma unroll
      for (int e = 0; e < 16; ++e) v[e] = src[e];
    }
  } else {
    const double *col = a + (int64_t)min(j, n - 1) * lda;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int i = i0 + e;
      const double w = col[min(i, n - 1)];
      v[e] = (j < n && i < n && i > j) ? w : 0.0;
    }
  }
}

// Persistent: workgroup g takes tiles g, g + G, g + 2G, ... of the
// triangle (tri_tile order) and loads the next tile's entries while it works
// on the current one; x's 64 values at the tile's rows and at its columns
// go through LDS (per-thread loads of them from global memory took twice as
// long as the whole pass).
__global__ __launch_bounds__(256) void k_residual_dd_tiles(const double *__restrict__ a, int64_t lda,
                                                           int n, int nt,
                                                           const double *__restrict__ x,
                                                           double2 *__restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double tile[kRT][kRT + 1];
  __shared__ double xs[2][kRT];  // x at the tile's rows, at its columns
  const int64_t ntiles = (int64_t)nt * (nt + 1) / 2;
  const int64_t G = gridDim.x;
  int64_t cur = blockIdx.x;
  if (cur >= ntiles) return;
  const int tid = threadIdx.x, q = tid >> 2, e0 = (tid & 3) * 16;
  int tc, tr;  // column tile tc <= row tile tr
  tri_tile(nt, cur, tc, tr);
  double v[16];
  residual_tile_load(a, lda, n, tc, tr, q, e0, v);
  while (true) {
    const int64_t nxt = cur + G;
    const bool more = nxt < ntiles;
    int tc2 = 0, tr2 = 0;
    double w[16];
    if (more) {
      tri_tile(nt, nxt, tc2, tr2);
      residual_tile_load(a, lda, n, tc2, tr2, q, e0, w);
    }
    if (tid < 2 * kRT) {
      const int k = tid < kRT ? tr * kRT + tid : tc * kRT + tid - kRT;
      xs[tid >> 6][tid & 63] = k < n ? x[k] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) tile[q][e0 + e] = v[e];
    __syncthreads();
    // sums down column j = 64 tc + q over this thread's rows (output j of
    // K x), and along row i = 64 tr + q over columns 64 tc + e0 + e (output i)
    double ch = 0.0, cl = 0.0, rh = 0.0, rl = 0.0;
#pragma unroll
    for (int e = 0; e < 16; ++e) dot2_acc(v[e], xs[0][e0 + e], ch, cl);
#pragma unroll
    for (int e = 0; e < 16; ++e) dot2_acc(tile[e0 + e][q], xs[1][e0 + e], rh, rl);
    dd_quad_sum(ch, cl);
    dd_quad_sum(rh, rl);
    if ((tid & 3) == 0) {
      const int j = tc * kRT + q, i = tr * kRT + q;
      if (tr == tc) {  // a diagonal tile: both sums belong to output i = j
        if (i < n) {
          dd_add(ch, cl, rh, rl);
          part[(int64_t)tc * n + i] = double2{ch, cl};
        }
      } else {
        if (j < n) part[(int64_t)tr * n + j] = double2{ch, cl};
        if (i < n) part[(int64_t)tc * n + i] = double2{rh, rl};
      }
    }
    __syncthreads();  // the LDS tile and x slices are rewritten next
    if (!more) break;
    cur = nxt;
    tc = tc2;
    tr = tr2;
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = w[e];
  }
}

__global__ __launch_bounds__(256) void k_residual_dd_finish(const double2 *__restrict__ part, int n,
                                                            int nt, const double *__restrict__ diag,
                                                            double lambda,
                                                            const double *__restrict__ x,
                                                            const double *__restrict__ y,
                                                            double *__restrict__ r) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double2 *p = part + i;
  double hi = 0.0, lo = 0.0;
  int t = 0;
  for (; t + 8 <= nt; t += 8) {
    double2 w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) w[u] = p[(int64_t)(t + u) * n];
#pragma unroll
    for (int u = 0; u < 8; ++u) dd_add(hi, lo, w[u].x, w[u].y);
  }
  for (; t < nt; ++t) {
    const double2 w = p[(int64_t)t * n];
    dd_add(hi, lo, w.x, w.y);
  }
  const double xi = x[i];
  double pr = diag[i] * xi;
  dd_add(hi, lo, pr, fma(diag[i], xi, -pr));  // + K(i,i) x_i, exactly
  pr = lambda * xi;
  dd_add(hi, lo, pr, fma(lambda, xi, -pr));  // + lambda x_i, exactly
  double s, e;
  two_sum(y[i], -hi, s, e);
  r[i] = s + (e - lo);
}

template <typename T>
int check_dense_args(const T *a, int64_t lda, int na, const T *b, int64_t ldb, int nb, int d,
                     double sigma, const void *k, int64_t ldk) {
  FMLH_ARG_CHECK(na >= 0, 3);
  FMLH_ARG_CHECK(nb >= 0, 6);
  FMLH_ARG_CHECK(d >= 0, 7);
  FMLH_ARG_CHECK(lda >= d, 2);
  FMLH_ARG_CHECK(ldb >= d, 5);
  FMLH_ARG_CHECK(sigma != 0.0, 8);
  FMLH_ARG_CHECK(ldk >= na, 10);
  if (na > 0 && nb > 0) {
    FMLH_ARG_CHECK(a != nullptr || d == 0, 1);
    FMLH_ARG_CHECK(b != nullptr || d == 0, 4);
    FMLH_ARG_CHECK(k != nullptr, 9);
  }
  return 0;
}

}  // namespace

extern "C" {

int fmlh_gaussian_kernel(const double *a, int64_t lda, int na, const double *b, int64_t ldb,
                         int nb, int d, double sigma, double *k, int64_t ldk, void *stream) {
  int rc = check_dense_args(a, lda, na, b, ldb, nb, d, sigma, k, ldk);
  if (rc) return rc;
  if (na == 0 || nb == 0) return 0;
  const double inv = -0.5 / (sigma * sigma);
  timing_begin(as_stream(stream));
  launch_gaussian<double, double>(a, lda, na, b, ldb, nb, d, inv, k, ldk, 1e-3, as_stream(stream));
  timing_end(as_stream(stream));
  FMLH_LAUNCH_CHECK("k_gaussian<f64>");
  return 0;
}

}  // extern "C"

namespace fmlh {  // precision.hip
size_t gaussian_bf16x3_ws_bytes(const float *a, int64_t lda, int na, const float *b, int64_t ldb,
                                int nb, int d);
int gaussian_bf16x3(const float *a, int64_t lda, int na, const float *b, int64_t ldb, int nb,
                    int d, double sigma, float *k32, double *k64, int64_t ldk, void *ws,
                    size_t ws_bytes, unsigned long long *recomputed, hipStream_t st);
}  // namespace fmlh

extern "C" {

size_t fmlh_gaussian_kernel_f32_ws_bytes(int na, int nb, int d, int mode, int same) {
  if (na < 0 || nb < 0 || d < 0 || mode != 1) return 0;
  const float *a = reinterpret_cast<const float *>(16);
  const float *b = same ? a : reinterpret_cast<const float *>(32);
  return gaussian_bf16x3_ws_bytes(a, d, na, b, d, same ? na : nb, d);
}

int fmlh_gaussian_kernel_f32(const float *a, int64_t lda, int na, const float *b, int64_t ldb,
                             int nb, int d, double sigma, int mode, float *k32, double *k64,
                             int64_t ldk, void *ws, size_t ws_bytes, unsigned long long *recomputed,
                             void *stream) {
  int rc = check_dense_args(a, lda, na, b, ldb, nb, d, sigma, k32 ? (void *)k32 : (void *)k64, ldk);
  if (rc) return rc;
  FMLH_ARG_CHECK(mode == 0 || mode == 1, 9);
  FMLH_ARG_CHECK((k32 == nullptr) != (k64 == nullptr), 10);
  if (na == 0 || nb == 0) return 0;
  if (mode == 1) {
    FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= gaussian_bf16x3_ws_bytes(a, lda, na, b, ldb, nb, d), 13);
    return gaussian_bf16x3(a, lda, na, b, ldb, nb, d, sigma, k32, k64, ldk, ws, ws_bytes, recomputed,
                           as_stream(stream));
  }
  const double inv = -0.5 / (sigma * sigma);
  timing_begin(as_stream(stream));
  if (k32)
    launch_gaussian<float, float>(a, lda, na, b, ldb, nb, d, inv, k32, ldk, 1e-2f, as_stream(stream));
  else
    launch_gaussian<float, double>(a, lda, na, b, ldb, nb, d, inv, k64, ldk, 1e-2f, as_stream(stream));
  timing_end(as_stream(stream));
  FMLH_LAUNCH_CHECK("k_gaussian<f32>");
  return 0;
}

int fmlh_laplacian_kernel(const double *a, int64_t lda, int na, const double *b, int64_t ldb,
                          int nb, int d, double sigma, double *k, int64_t ldk, void *stream) {
  int rc = check_dense_args(a, lda, na, b, ldb, nb, d, sigma, k, ldk);
  if (rc) return rc;
  if (na == 0 || nb == 0) return 0;
  const double inv = -1.0 / sigma;
  timing_begin(as_stream(stream));
  launch_laplacian<double, double>(a, lda, na, b, ldb, nb, d, inv, k, ldk, as_stream(stream));
  timing_end(as_stream(stream));
  FMLH_LAUNCH_CHECK("k_laplacian<f64>");
  return 0;
}

int fmlh_laplacian_kernel_f32(const float *a, int64_t lda, int na, const float *b, int64_t ldb,
                              int nb, int d, double sigma, float *k32, double *k64, int64_t ldk,
                              void *stream) {
  int rc = check_dense_args(a, lda, na, b, ldb, nb, d, sigma, k32 ? (void *)k32 : (void *)k64, ldk);
  if (rc) return rc;
  FMLH_ARG_CHECK((k32 == nullptr) != (k64 == nullptr), 9);
  if (na == 0 || nb == 0) return 0;
  const double inv = -1.0 / sigma;
  timing_begin(as_stream(stream));
  if (k32)
    launch_laplacian<float, float>(a, lda, na, b, ldb, nb, d, inv, k32, ldk, as_stream(stream));
  else
    launch_laplacian<float, double>(a, lda, na, b, ldb, nb, d, inv, k64, ldk, as_stream(stream));
  timing_end(as_stream(stream));
  FMLH_LAUNCH_CHECK("k_laplacian<f32>");
  return 0;
}

int fmlh_add_diagonal(double *a, int64_t lda, int n, double lambda, void *stream) {
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(lda >= n, 2);
  if (n == 0) return 0;
  FMLH_ARG_CHECK(a != nullptr, 1);
  hipLaunchKernelGGL(k_add_diagonal, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), a,
                     lda, n, lambda);
  FMLH_LAUNCH_CHECK("k_add_diagonal");
  return 0;
}

int fmlh_sym_residual(const double *a, int64_t lda, int n, double lambda, const double *x,
                      const double *y, double *r, void *stream) {
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(lda >= n, 2);
  if (n == 0) return 0;
  FMLH_ARG_CHECK(a && x && y && r, 1);
  // r doubles as the scratch vector t: each column's workgroup reads t_j
  // and then writes r_j = ... for the same j only.
  hipLaunchKernelGGL(k_symv_rows, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), a, lda,
                     n, x, r);
  FMLH_LAUNCH_CHECK("k_symv_rows");
  hipLaunchKernelGGL(k_symv_cols_residual, dim3(n), dim3(256), 0, as_stream(stream), a, lda, n,
                     lambda, x, y, r, r);
  FMLH_LAUNCH_CHECK("k_symv_cols_residual");
  return 0;
}

int fmlh_gemv_t(const double *k, int64_t ldk, int na, int nb, const double *alpha, double *y,
                void *stream) {
  FMLH_ARG_CHECK(na >= 0, 3);
  FMLH_ARG_CHECK(nb >= 0, 4);
  FMLH_ARG_CHECK(ldk >= na, 2);
  if (nb == 0) return 0;
  FMLH_ARG_CHECK(k && alpha && y, 1);
  hipLaunchKernelGGL(k_gemv_t, dim3(nb), dim3(256), 0, as_stream(stream), k, ldk, na, alpha, y);
  FMLH_LAUNCH_CHECK("k_gemv_t");
  return 0;
}

int fmlh_residual_dd(const double *k, int64_t ldk, int n, double lambda, const double *x,
                     const double *y, double *r, void *stream) {
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(ldk >= n, 2);
  if (n == 0) return 0;
  FMLH_ARG_CHECK(k != nullptr, 1);
  FMLH_ARG_CHECK(x != nullptr, 5);
  FMLH_ARG_CHECK(y != nullptr, 6);
  FMLH_ARG_CHECK(r != nullptr && r != x && r != y, 7);
  hipLaunchKernelGGL(k_residual_dd, dim3(n), dim3(256), 0, as_stream(stream), k, ldk, n, lambda, x,
                     y, r);
  FMLH_LAUNCH_CHECK("k_residual_dd");
  return 0;
}

size_t fmlh_residual_dd_lower_ws_bytes(int n) {
  if (n <= 0) return 0;
  const size_t nt = (size_t)ceil_div(n, kRT);
  return WsCarver::need(sizeof(double2) * nt * (size_t)n);
}

int fmlh_residual_dd_lower(const double *a, int64_t lda, int n, const double *diag, double lambda,
                           const double *x, const double *y, double *r, void *ws, size_t ws_bytes,
                           void *stream) {
  FMLH_ARG_CHECK(n >= 0, 3);
  FMLH_ARG_CHECK(lda >= (n > 1 ? n : 1), 2);
  if (n == 0) return 0;
  FMLH_ARG_CHECK(a != nullptr, 1);
  FMLH_ARG_CHECK(diag != nullptr, 4);
  FMLH_ARG_CHECK(x != nullptr, 6);
  FMLH_ARG_CHECK(y != nullptr, 7);
  FMLH_ARG_CHECK(r != nullptr && r != x && r != y && r != diag, 8);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= fmlh_residual_dd_lower_ws_bytes(n), 9);
  const int nt = ceil_div(n, kRT);
  const int64_t ntiles = (int64_t)nt * (nt + 1) / 2;
  hipStream_t st = as_stream(stream);
  double2 *part = static_cast<double2 *>(ws);
  // persistent workgroups, eight a CU (2048 measured best at 36,000 rows)
  const unsigned grid = (unsigned)(ntiles < 2048 ? ntiles : 2048);
  hipLaunchKernelGGL(k_residual_dd_tiles, dim3(grid), dim3(256), 0, st, a, lda, n, nt, x, part);
  FMLH_LAUNCH_CHECK("k_residual_dd_tiles");
  hipLaunchKernelGGL(k_residual_dd_finish, dim3(ceil_div(n, 256)), dim3(256), 0, st, part, n, nt, diag,
                     lambda, x, y, r);
  FMLH_LAUNCH_CHECK("k_residual_dd_finish");
  return 0;
}

size_t fmlh_krr_predict_ws_bytes(int nmol, int ntrain, int nmax) {
  if (nmol < 0 || ntrain < 0 || nmax < 1) return 0;
  const size_t packed = (size_t)nmax * (nmax + 1) / 2;
  // the kernel block is taken with at least one column, also for an empty
  // training set (whose predictions are zeros)
  const size_t nchunk = (size_t)ceil_div(ntrain > 0 ? ntrain : 1, kGemvChunk);
  return WsCarver::need(sizeof(double) * (size_t)nmol * packed) +
         WsCarver::need(sizeof(double) * (size_t)nmol * (size_t)(ntrain > 0 ? ntrain : 1)) +
         WsCarver::need(sizeof(double) * (size_t)nmol * nchunk);
}

int fmlh_coulomb_sorted(const double *z, const double *r, const int32_t *offsets, int nmol,
                        int nmax, double *cm, float *cm32, int32_t *perm, int32_t *status,
                        void *stream);

int fmlh_krr_predict(const double *z, const double *r, const int32_t *offsets, int nmol, int nmax,
                     const double *xtrain, int64_t ldx, int ntrain, const double *alpha,
                     double sigma, double *y, int32_t *status, void *ws, size_t ws_bytes,
                     void *stream) {
  FMLH_ARG_CHECK(nmol >= 0, 4);
  FMLH_ARG_CHECK(nmax >= 1, 5);
  FMLH_ARG_CHECK(ntrain >= 0, 8);
  const int packed = nmax * (nmax + 1) / 2;
  FMLH_ARG_CHECK(ldx >= packed, 7);
  FMLH_ARG_CHECK(sigma != 0.0, 10);
  if (nmol == 0) return 0;
  FMLH_ARG_CHECK(z && r && offsets, 1);
  FMLH_ARG_CHECK(y != nullptr, 11);
  if (ntrain > 0) FMLH_ARG_CHECK(xtrain && alpha, 6);
  FMLH_ARG_CHECK(ws != nullptr && ws_bytes >= fmlh_krr_predict_ws_bytes(nmol, ntrain, nmax), 13);
  WsCarver w(ws, ws_bytes);
  double *cm = w.take<double>((size_t)nmol * packed);
  double *k = w.take<double>((size_t)nmol * (size_t)(ntrain > 0 ? ntrain : 1));
  const int nchunk = ceil_div(ntrain > 0 ? ntrain : 1, kGemvChunk);
  double *part = w.take<double>((size_t)nmol * nchunk);
  if (!cm || !k || !part) return -13;
  FMLH_ARG_CHECK((int64_t)nmol * nchunk <= INT32_MAX, 4);
  int rc = fmlh_coulomb_sorted(z, r, offsets, nmol, nmax, cm, nullptr, nullptr, status, stream);
  if (rc) return rc;
  hipStream_t st = as_stream(stream);
  if (ntrain == 0) {
    hipLaunchKernelGGL(k_fill, dim3(ceil_div(nmol, 256)), dim3(256), 0, st, y, nmol, 0.0);
    FMLH_LAUNCH_CHECK("k_fill");
    return 0;
  }
  const double inv = -0.5 / (sigma * sigma);
  launch_gaussian<double, double>(xtrain, ldx, ntrain, cm, packed, nmol, packed, inv, k, ntrain,
                                  1e-3, st);
  FMLH_LAUNCH_CHECK("k_gaussian<f64>(predict)");
  hipLaunchKernelGGL(k_gemv_part, dim3(nmol * nchunk), dim3(256), 0, st, k, (int64_t)ntrain, ntrain,
                     nchunk, alpha, part);
  FMLH_LAUNCH_CHECK("k_gemv_part(predict)");
  hipLaunchKernelGGL(k_gemv_sum, dim3(ceil_div(nmol, 256)), dim3(256), 0, st, part, nmol, nchunk, y);
  FMLH_LAUNCH_CHECK("k_gemv_sum(predict)");
  return 0;
}

}  // extern "C"

