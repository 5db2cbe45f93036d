// (a) Sorted Coulomb matrices, one wavefront (workgroup of 64) per molecule.
//
// Follows fml/representations/frepresentations.f90:23-110:
//   M_ii = 0.5 Z_i^2.4, M_ij = Z_i Z_j / sqrt(sum_c (R_jc - R_ic)^2)
//   row norm r_k = diag_k + (sum over partners in index order, from 0)
//   order: descending r, equal norms with the higher original index first
//   output: packed row-major lower triangle of the sorted matrix.
// The arithmetic is evaluated in the reference's operation order with FP
// contraction off, so every value, and therefore the permutation, is the
// IEEE result of the Fortran expressions (sqrt and division are correctly
// rounded on gfx950).  0.5*Z^2.4 for integral Z comes from a table filled on
// the host with the C library's pow(), the function the reference's `**`
// lowers to.
#include <math.h>

#include <mutex>

#include "common.h"

#pragma clang fp contract(off)

namespace {

constexpr int kMaxAtoms = 256;  // LDS capacity per molecule
constexpr int kDiagTable = 129;
__constant__ double c_cm_diag[kDiagTable];

std::mutex g_init_mu;
bool g_init_done[64] = {};

int ensure_tables() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fmlh::record_error(e, "hipGetDevice");
  if (dev < 0 || dev >= 64) return FMLH_EHIP;
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_init_done[dev]) return 0;
  double host[kDiagTable];
  for (int z = 0; z < kDiagTable; ++z) host[z] = 0.5 * pow((double)z, 2.4);
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_cm_diag), host, sizeof(host), 0, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fmlh::record_error(e, "hipMemcpyToSymbol(c_cm_diag)");
  g_init_done[dev] = true;
  return 0;
}

__device__ __forceinline__ double cm_diag(double z) {
  int zi = (int)z;
  if ((double)zi == z && zi >= 0 && zi < kDiagTable) return c_cm_diag[zi];
  return 0.5 * pow(z, 2.4);
}

// Pair value for atoms i < j, written as frepresentations.f90:75-76.
__device__ __forceinline__ double cm_pair(const double *zs, const double *rs, int i, int j) {
  double dx = rs[3 * j + 0] - rs[3 * i + 0];
  double dy = rs[3 * j + 1] - rs[3 * i + 1];
  double dz = rs[3 * j + 2] - rs[3 * i + 2];
  double s = dx * dx;
  s = s + dy * dy;
  s = s + dz * dz;
  return zs[i] * zs[j] / sqrt(s);
}

__device__ __forceinline__ double cm_entry(const double *zs, const double *rs, int i, int j) {
  if (i == j) return cm_diag(zs[i]);
  return i < j ? cm_pair(zs, rs, i, j) : cm_pair(zs, rs, j, i);
}

__global__ __launch_bounds__(64) void k_coulomb_sorted(
    const double *__restrict__ z, const double *__restrict__ r,
    const int32_t *__restrict__ offsets, int nmol, int nmax, double *__restrict__ cm,
    float *__restrict__ cm32, int32_t *__restrict__ perm, int32_t *__restrict__ status) {
  __shared__ double zs[kMaxAtoms];
  __shared__ double rs[3 * kMaxAtoms];
  __shared__ double norms[kMaxAtoms];
  __shared__ int sorted_atoms[kMaxAtoms];
  __shared__ int all_finite;

  const int mol = blockIdx.x;
  const int tid = threadIdx.x;
  const int a0 = offsets[mol];
  const int n = offsets[mol + 1] - a0;
  const int64_t packed = (int64_t)nmax * (nmax + 1) / 2;
  double *out = cm + (int64_t)mol * packed;
  float *out32 = cm32 ? cm32 + (int64_t)mol * packed : nullptr;

  const bool bad = (n < 0 || n > nmax || n > kMaxAtoms);
  if (status && tid == 0) status[mol] = bad ? 1 : 0;
  if (bad) {
    for (int64_t t = tid; t < packed; t += 64) {
      out[t] = 0.0;
      if (out32) out32[t] = 0.0f;
    }
    if (perm)
      for (int m = tid; m < nmax; m += 64) perm[(int64_t)mol * nmax + m] = -1;
    return;
  }

  for (int k = tid; k < n; k += 64) {
    zs[k] = z[a0 + k];
    rs[3 * k + 0] = r[3 * (int64_t)(a0 + k) + 0];
    rs[3 * k + 1] = r[3 * (int64_t)(a0 + k) + 1];
    rs[3 * k + 2] = r[3 * (int64_t)(a0 + k) + 2];
  }
  if (tid == 0) all_finite = 1;
  __syncthreads();

  // Row norms: r_k = diag_k + (partners summed in index order from zero).
  for (int k = tid; k < n; k += 64) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      s = s + (i < k ? cm_pair(zs, rs, i, k) : cm_pair(zs, rs, k, i));
    }
    double rk = cm_diag(zs[k]) + s;
    norms[k] = rk;
    if (!(fabs(rk) < 1.7976931348623157e308)) atomicAnd(&all_finite, 0);
  }
  __syncthreads();

  if (all_finite) {
    // rank = #{j : r_j > r_k} + #{j : r_j == r_k and j > k}; identical to the
    // reference's repeated-minloc selection sort when every norm is finite.
    for (int k = tid; k < n; k += 64) {
      const double rk = norms[k];
      int rank = 0;
      for (int j = 0; j < n; ++j) {
        const double rj = norms[j];
        rank += (rj > rk || (rj == rk && j > k)) ? 1 : 0;
      }
      sorted_atoms[rank] = k;
    }
  } else if (tid == 0) {
    // Literal selection sort (frepresentations.f90:87-91) for inf/NaN norms.
    for (int i = 0; i < n; ++i) {
      int jmin = 0;
      for (int j = 1; j < n; ++j)
        if (norms[j] < norms[jmin]) jmin = j;
      sorted_atoms[n - 1 - i] = jmin;
      norms[jmin] = 1.7976931348623157e308;
    }
  }
  __syncthreads();

  // Packed lower triangle: t = m(m+1)/2 + q, q <= m.
  for (int64_t t = tid; t < packed; t += 64) {
    int m = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((int64_t)(m + 1) * (m + 2) / 2 <= t) ++m;
    while ((int64_t)m * (m + 1) / 2 > t) --m;
    const int q = (int)(t - (int64_t)m * (m + 1) / 2);
    double v = 0.0;
    if (m < n) v = cm_entry(zs, rs, sorted_atoms[m], sorted_atoms[q]);
    out[t] = v;
    if (out32) out32[t] = (float)v;
  }
  if (perm)
    for (int m = tid; m < nmax; m += 64)
      perm[(int64_t)mol * nmax + m] = m < n ? sorted_atoms[m] : -1;
}

// Unsorted, local and atomic Coulomb matrices (frepresentations.f90:112-390),
// one wavefront per molecule.  Mode 0 writes the molecule's packed lower
// triangle in input atom order (one row per molecule); modes 1 and 2 write
// one packed row per atom k (row offsets[m] + k), the atoms taken in the
// order:
//   mode 1 (local): atom k, then every other atom by descending row norm
//     (the sorted Coulomb matrix's order, equal norms higher index first);
//   mode 2 (atomic): every atom by ascending distance from k, equal
//     distances lowest index first (k itself has distance 0).
// Entries are the pair values Z_i Z_j / |R_j - R_i| (i < j) and 0.5 Z^2.4,
// evaluated in the reference's operation order, so each row is bitwise the
// reference's.
__global__ __launch_bounds__(64) void k_coulomb_forms(
    int mode, const double *__restrict__ z, const double *__restrict__ r,
    const int32_t *__restrict__ offsets, int nmol, int nmax, double *__restrict__ cm,
    int32_t *__restrict__ status) {
  __shared__ double zs[kMaxAtoms];
  __shared__ double rs[3 * kMaxAtoms];
  __shared__ double key[kMaxAtoms];
  __shared__ int sorted_atoms[kMaxAtoms];
  __shared__ int pos[kMaxAtoms];
  __shared__ int ord[kMaxAtoms];
  __shared__ int all_finite;

  const int mol = blockIdx.x;
  const int tid = threadIdx.x;
  const int a0 = offsets[mol];
  const int n = offsets[mol + 1] - a0;
  const int64_t packed = (int64_t)nmax * (nmax + 1) / 2;
  const bool bad = (n < 0 || n > nmax || n > kMaxAtoms);
  if (status && tid == 0) status[mol] = bad ? 1 : 0;
  if (mode == 0) {
    double *out = cm + (int64_t)mol * packed;
    if (bad) {
      for (int64_t t = tid; t < packed; t += 64) out[t] = 0.0;
      return;
    }
  } else if (bad) {
    if (n > 0)
      for (int64_t t = tid; t < (int64_t)n * packed; t += 64) cm[(int64_t)a0 * packed + t] = 0.0;
    return;
  }
  for (int k = tid; k < n; k += 64) {
    zs[k] = z[a0 + k];
    rs[3 * k + 0] = r[3 * (int64_t)(a0 + k) + 0];
    rs[3 * k + 1] = r[3 * (int64_t)(a0 + k) + 1];
    rs[3 * k + 2] = r[3 * (int64_t)(a0 + k) + 2];
  }
  if (tid == 0) all_finite = 1;
  __syncthreads();

  // writes row `out` from the order ord[0..n)
  auto write_row = [&](double *out) {
    for (int64_t t = tid; t < packed; t += 64) {
      int m = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
      while ((int64_t)(m + 1) * (m + 2) / 2 <= t) ++m;
      while ((int64_t)m * (m + 1) / 2 > t) --m;
      const int q = (int)(t - (int64_t)m * (m + 1) / 2);
      out[t] = m < n ? cm_entry(zs, rs, ord[m], ord[q]) : 0.0;
    }
  };

  if (mode == 0) {
    for (int k = tid; k < n; k += 64) ord[k] = k;
    __syncthreads();
    write_row(cm + (int64_t)mol * packed);
    return;
  }

  if (mode == 1) {
    // row norms and the descending order, as k_coulomb_sorted
    for (int k = tid; k < n; k += 64) {
      double s = 0.0;
      for (int i = 0; i < n; ++i) {
        if (i == k) continue;
        s = s + (i < k ? cm_pair(zs, rs, i, k) : cm_pair(zs, rs, k, i));
      }
      const double rk = cm_diag(zs[k]) + s;
      key[k] = rk;
      if (!(fabs(rk) < 1.7976931348623157e308)) atomicAnd(&all_finite, 0);
    }
    __syncthreads();
    if (all_finite) {
      for (int k = tid; k < n; k += 64) {
        const double rk = key[k];
        int rank = 0;
        for (int j = 0; j < n; ++j) {
          const double rj = key[j];
          rank += (rj > rk || (rj == rk && j > k)) ? 1 : 0;
        }
        sorted_atoms[rank] = k;
        pos[k] = rank;
      }
    } else if (tid == 0) {
      for (int i = 0; i < n; ++i) {
        int jmin = 0;
        for (int j = 1; j < n; ++j)
          if (key[j] < key[jmin]) jmin = j;
        sorted_atoms[n - 1 - i] = jmin;
        key[jmin] = 1.7976931348623157e308;
      }
      for (int i = 0; i < n; ++i) pos[sorted_atoms[i]] = i;
    }
    __syncthreads();
  }

  for (int k = 0; k < n; ++k) {
    if (mode == 1) {
      // atom k first, then the sorted list without k
      const int pk = pos[k];
      for (int e = tid; e < n; e += 64)
        ord[e] = e == 0 ? k : sorted_atoms[(e - 1) < pk ? e - 1 : e];
    } else {
      // ascending distance from k, ties lowest index first
      for (int j = tid; j < n; j += 64) {
        double dj = 0.0;
        if (j != k) {
          const int i0 = j < k ? j : k, i1 = j < k ? k : j;
          const double dx = rs[3 * i1 + 0] - rs[3 * i0 + 0];
          const double dy = rs[3 * i1 + 1] - rs[3 * i0 + 1];
          const double dz = rs[3 * i1 + 2] - rs[3 * i0 + 2];
          double s = dx * dx;
          s = s + dy * dy;
          s = s + dz * dz;
          dj = sqrt(s);
        }
        key[j] = dj;
        if (!(dj < 1.7976931348623157e308)) atomicAnd(&all_finite, 0);
      }
      __syncthreads();
      if (all_finite) {
        for (int j = tid; j < n; j += 64) {
          const double dj = key[j];
          int rank = 0;
          for (int l = 0; l < n; ++l) {
            const double dl = key[l];
            rank += (dl < dj || (dl == dj && l < j)) ? 1 : 0;
          }
          ord[rank] = j;
        }
      } else if (tid == 0) {
        // literal repeated minloc (frepresentations.f90:366-372)
        for (int i = 0; i < n; ++i) {
          int jmin = 0;
          for (int j = 1; j < n; ++j)
            if (key[j] < key[jmin]) jmin = j;
          ord[i] = jmin;
          key[jmin] = 1.7976931348623157e308;
        }
      }
    }
    __syncthreads();
    write_row(cm + (int64_t)(a0 + k) * packed);
    __syncthreads();  // the next atom overwrites ord / key
  }
}

}  // namespace

namespace fmlh {
int potrs_scratch_init();
}

extern "C" int fmlh_init(void) {
  int rc = ensure_tables();
  if (rc) return rc;
  return fmlh::potrs_scratch_init();
}

extern "C" int fmlh_coulomb_form(int mode, const double *z, const double *r,
                                 const int32_t *offsets, int nmol, int nmax, double *cm,
                                 int32_t *status, void *stream) {
  FMLH_ARG_CHECK(mode >= 0 && mode <= 2, 1);
  FMLH_ARG_CHECK(nmol >= 0, 5);
  FMLH_ARG_CHECK(nmax >= 1, 6);
  if (nmol == 0) return 0;
  FMLH_ARG_CHECK(z != nullptr, 2);
  FMLH_ARG_CHECK(r != nullptr, 3);
  FMLH_ARG_CHECK(offsets != nullptr, 4);
  FMLH_ARG_CHECK(cm != nullptr, 7);
  int rc = ensure_tables();
  if (rc) return rc;
  hipLaunchKernelGGL(k_coulomb_forms, dim3(nmol), dim3(64), 0, fmlh::as_stream(stream), mode, z, r,
                     offsets, nmol, nmax, cm, status);
  FMLH_LAUNCH_CHECK("k_coulomb_forms");
  return 0;
}

extern "C" int fmlh_coulomb_sorted(const double *z, const double *r, const int32_t *offsets,
                                   int nmol, int nmax, double *cm, float *cm32, int32_t *perm,
                                   int32_t *status, void *stream) {
  FMLH_ARG_CHECK(nmol >= 0, 4);
  FMLH_ARG_CHECK(nmax >= 1, 5);
  FMLH_ARG_CHECK(cm != nullptr, 6);
  if (nmol == 0) return 0;
  FMLH_ARG_CHECK(z != nullptr, 1);
  FMLH_ARG_CHECK(r != nullptr, 2);
  FMLH_ARG_CHECK(offsets != nullptr, 3);
  int rc = ensure_tables();
  if (rc) return rc;
  hipLaunchKernelGGL(k_coulomb_sorted, dim3(nmol), dim3(64), 0, fmlh::as_stream(stream), z, r,
                     offsets, nmol, nmax, cm, cm32, perm, status);
  FMLH_LAUNCH_CHECK("k_coulomb_sorted");
  return 0;
}
