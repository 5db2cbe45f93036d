/*
 * fml_oracle.c -- CPU restatement of the fml KRR hot path, used ONLY as a
 * test oracle and as the "port" CPU baseline.  Nothing in the product
 * (fml_amd/) links or calls this file; tests, the smoke check and the CPU
 * baseline leg of bench.py are its only users.
 *
 * Every routine follows the loop nest and the expression order of the
 * reference Fortran it cites, evaluated in IEEE double precision with FP
 * contraction disabled (built with -ffp-contract=off), so that a sum is the
 * sequential left-to-right sum the Fortran source writes.
 *
 * Layout conventions (all row-major C arrays unless stated):
 *   coordinates        R[natoms][3]
 *   descriptor sets    A[n][d]   (one sample per row; this is the memory
 *                      layout of the reference's Fortran (d, n) arrays)
 *   dense kernels      K stored Fortran order: K(j,i) at K[j + i*na]
 *   ARAD descriptors   q[nm][amax][5][amax]  (numpy C order of the
 *                      reference's (nm, amax, 5, amax) arrays)
 *   ARAD types         z[nm][amax][2]
 *   ARAD atomic descr. q[na][5][amax], z[na][2]
 *
 * Parity status of each routine is stated where it is defined.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* Sorted Coulomb matrix                                               */
/* ------------------------------------------------------------------ */

/*
 * fml/representations/frepresentations.f90:23-110 (fgenerate_coulomb_matrix)
 *
 * diag: 0.5*Z**2.4 (line 66); pair: Zi*Zj/sqrt(sum((Rj-Ri)**2)) (75-76).
 * Row norms follow the single-thread order of the two OpenMP REDUCTION
 * loops (64, 72): the first loop leaves r_k = diag_k, the second
 * accumulates a private sum starting at zero over partners in index order
 * (i<k as "j" first, then i=k over j>k) and adds it to r_k at the end.
 * Selection sort (87-91): minloc picks the FIRST minimum and places it last,
 * so equal norms are ordered with the higher original index first.
 * Packed output (94-103): cm(m(m-1)/2 + n) = M(sorted(m), sorted(n)), n<=m,
 * zero padded to nmax(nmax+1)/2.
 *
 * Returns 0, or -1 when natoms > nmax (the reference writes out of bounds
 * there, frepresentations.f90:96-103) or natoms < 0.
 * perm (optional, may be NULL) receives sorted_atoms - 1.
 * Parity: pinned by the flang build of the reference (oracle/_ref), by hand
 * computed molecules and by invariants (tests/test_oracle.py).
 */
EXPORT int oracle_coulomb_matrix(const double *z, const double *r, int natoms,
                                 int nmax, double *cm, int *perm) {
  if (natoms < 0 || natoms > nmax) return -1;
  const int n = natoms;
  const int packed = nmax * (nmax + 1) / 2;
  double *pdm = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n * n : 1));
  double *norms = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double *priv = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  int *sorted = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));

  for (int i = 0; i < n; ++i) {
    double p = 0.5 * pow(z[i], 2.4);
    norms[i] = 0.0 + p;
    pdm[i * n + i] = p;
  }
  for (int i = 0; i < n; ++i) priv[i] = 0.0;
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      double dx = r[j * 3 + 0] - r[i * 3 + 0];
      double dy = r[j * 3 + 1] - r[i * 3 + 1];
      double dz = r[j * 3 + 2] - r[i * 3 + 2];
      double s = dx * dx;
      s = s + dy * dy;
      s = s + dz * dz;
      double p = z[i] * z[j] / sqrt(s);
      priv[j] = priv[j] + p;
      priv[i] = priv[i] + p;
      pdm[i * n + j] = p;
      pdm[j * n + i] = p;
    }
  }
  for (int i = 0; i < n; ++i) norms[i] = norms[i] + priv[i];

  for (int i = 0; i < n; ++i) {
    int jmin = 0;
    for (int j = 1; j < n; ++j)
      if (norms[j] < norms[jmin]) jmin = j;
    sorted[n - 1 - i] = jmin;
    norms[jmin] = 1.7976931348623157e308; /* huge(1d0), line 59 */
  }

  for (int k = 0; k < packed; ++k) cm[k] = 0.0;
  for (int m = 0; m < n; ++m) {
    int i = sorted[m];
    int idx = m * (m + 1) / 2;
    for (int q = 0; q <= m; ++q) cm[idx + q] = pdm[i * n + sorted[q]];
  }
  if (perm)
    for (int m = 0; m < n; ++m) perm[m] = sorted[m];
  free(pdm);
  free(norms);
  free(priv);
  free(sorted);
  return 0;
}

/* Batched form over concatenated molecules: offsets has nmol+1 entries. */
EXPORT int oracle_coulomb_matrices(const double *z, const double *r,
                                   const int *offsets, int nmol, int nmax,
                                   double *cm, int *perm) {
  const int packed = nmax * (nmax + 1) / 2;
  int rc = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : rc)
  for (int m = 0; m < nmol; ++m) {
    int a0 = offsets[m];
    int na = offsets[m + 1] - a0;
    rc |= oracle_coulomb_matrix(z + a0, r + 3 * (size_t)a0, na, nmax,
                                cm + (size_t)m * packed,
                                perm ? perm + (size_t)m * nmax : NULL);
  }
  return rc ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* Gaussian and Laplacian kernels                                      */
/* ------------------------------------------------------------------ */

/*
 * fml/kernels/fkernels.f90:317-349 (fgaussian_kernel)
 * inv = -0.5/(sigma*sigma) (334); K(j,i) = exp(inv * sum((a_j-b_i)**2))
 * (341-342), the sum sequential over d.  A: na x d rows, B: nb x d rows,
 * K Fortran order (na, nb).
 * Parity: closed form, brute force numpy, flang build (oracle/_ref).
 */
EXPORT void oracle_gaussian_kernel(const double *a, int na, const double *b,
                                   int nb, int d, double sigma, double *k) {
  const double inv = -0.5 / (sigma * sigma);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nb; ++i) {
    const double *bi = b + (size_t)i * d;
    for (int j = 0; j < na; ++j) {
      const double *aj = a + (size_t)j * d;
      double s = 0.0;
      for (int t = 0; t < d; ++t) {
        double diff = aj[t] - bi[t];
        s = s + diff * diff;
      }
      k[(size_t)i * na + j] = exp(inv * s);
    }
  }
}

/*
 * fml/kernels/fkernels.f90:351-377 (flaplacian_kernel)
 * inv = -1/sigma (367); K(j,i) = exp(inv * sum(abs(a_j-b_i))) (372).
 */
EXPORT void oracle_laplacian_kernel(const double *a, int na, const double *b,
                                    int nb, int d, double sigma, double *k) {
  const double inv = -1.0 / sigma;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nb; ++i) {
    const double *bi = b + (size_t)i * d;
    for (int j = 0; j < na; ++j) {
      const double *aj = a + (size_t)j * d;
      double s = 0.0;
      for (int t = 0; t < d; ++t) s = s + fabs(aj[t] - bi[t]);
      k[(size_t)i * na + j] = exp(inv * s);
    }
  }
}

/* ------------------------------------------------------------------ */
/* ARAD                                                                */
/* ------------------------------------------------------------------ */

typedef struct {
  double width, cut, rw, cw;
} arad_par;

/*
 * fml/kernels/farad_kernels.f90:29-90 (atomic_distl2); identical copy at
 * fml/math/farad.f90:62-123.
 * x1, x2 point at one atom's (5, amax) block stored channel-major:
 * channel c of neighbour m at x[c*amax + m].  s1, s2 are the precomputed
 * 1-sin(d*inv_cut) vectors of length >= n.
 * Expression order: d=(exp(r2*inv_w)*s1)*s2 (78); d = d*(((rw2/(rw2+dr^2))
 * *cw2)/(cw2+dc^2)) (80-81); acc = acc + d*((1+x4*x4')+x5*x5') (83).
 */
static double arad_distl2(const double *x1, const double *x2, int amax, int n1,
                          int n2, const double *s1, const double *s2,
                          const arad_par *p) {
  const double inv_w = -1.0 / (4.0 * p->width * p->width);
  const double maxg = (8.0 * p->width) * (8.0 * p->width);
  const double rw2 = p->rw * p->rw;
  const double cw2 = p->cw * p->cw;
  double acc = 0.0;
  for (int m1 = 0; m1 < n1; ++m1) {
    if (x1[m1] > p->cut) break;
    for (int m2 = 0; m2 < n2; ++m2) {
      if (x2[m2] > p->cut) break;
      double dr = x2[m2] - x1[m1];
      double r2 = dr * dr;
      if (r2 < maxg) {
        double d = exp(r2 * inv_w) * s1[m1] * s2[m2];
        double e1 = x1[amax + m1] - x2[amax + m2];
        double e2 = x1[2 * amax + m1] - x2[2 * amax + m2];
        d = d * (rw2 / (rw2 + e1 * e1) * cw2 / (cw2 + e2 * e2));
        double ang = 1.0 + x1[3 * amax + m1] * x2[3 * amax + m2];
        ang = ang + x1[4 * amax + m1] * x2[4 * amax + m2];
        acc = acc + d * ang;
      }
    }
  }
  return acc;
}

/* (rw2/(rw2+dz1^2) * cw2/(cw2+dz2^2)) as written at farad_kernels.f90:590-592 */
static inline double arad_stoch(double za0, double za1, double zb0, double zb1,
                                const arad_par *p) {
  const double rw2 = p->rw * p->rw;
  const double cw2 = p->cw * p->cw;
  double e1 = za0 - zb0;
  double e2 = za1 - zb1;
  return rw2 / (rw2 + e1 * e1) * cw2 / (cw2 + e2 * e2);
}

/* single-precision eps literals of the reference, widened to double */
static const double EPS_5E12 = (double)5.0e-12f;  /* farad_kernels.f90:505 */
static const double EPS_1E11 = (double)1.0e-11f;  /* farad_kernels.f90:148 */

/* sin terms for a (nm, amax, 5, amax) descriptor set.
 * guard=1: only where d < cut, else 0 (farad_kernels.f90:521-545);
 * guard=0: unconditional (farad_kernels.f90:689-698, 163-171). */
static double *arad_sin_mol(const double *q, const int *n, int nm, int amax,
                            double cut, int guard) {
  const double pi = 4.0 * atan(1.0);
  const double inv_cut = pi / (2.0 * cut);
  double *s = (double *)calloc((size_t)nm * amax * amax, sizeof(double));
  for (int i = 0; i < nm; ++i)
    for (int a = 0; a < n[i]; ++a)
      for (int m = 0; m < n[i]; ++m) {
        double d = q[(((size_t)i * amax + a) * 5 + 0) * amax + m];
        if (!guard || d < cut)
          s[((size_t)i * amax + a) * amax + m] = 1.0 - sin(d * inv_cut);
      }
  return s;
}

static double *arad_self_mol(const double *q, const int *n, int nm, int amax,
                             const double *s, const arad_par *p) {
  double *self = (double *)calloc((size_t)nm * amax, sizeof(double));
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < nm; ++i)
    for (int a = 0; a < n[i]; ++a) {
      const double *x = q + ((size_t)i * amax + a) * 5 * amax;
      const double *sa = s + ((size_t)i * amax + a) * amax;
      self[(size_t)i * amax + a] = arad_distl2(x, x, amax, n[i], n[i], sa, sa, p);
    }
  return self;
}

/* D = self1+self2 - 2*S*T, clamped by eps (farad_kernels.f90:590-594) */
static inline double arad_pair_distance(double self1, double self2, double S,
                                        double T, double eps) {
  double l2 = self1 + self2 - 2.0 * S * T;
  if (fabs(l2) < eps) l2 = 0.0;
  return l2;
}

/*
 * fml/kernels/farad_kernels.f90:439-616 (fget_kernels_arad)
 * Output kernels(nsig, nm1, nm2) in Fortran order:
 *   K(k,i,j) at out[k + nsig*(i + nm1*j)].
 * The per-sigma sum runs over atomic_distance(:ni,:nj) in column-major
 * order (i1 fastest), exp(D*inv_sigma2(k)) (602-604).
 */
EXPORT void oracle_kernels_arad(const double *q1, const double *q2,
                                const double *z1, const double *z2,
                                const int *n1, const int *n2,
                                const double *sigmas, int nm1, int nm2,
                                int nsig, int amax, double width, double cut,
                                double rw, double cw, double *out) {
  arad_par p = {width, cut, rw, cw};
  double *inv_s2 = (double *)malloc(sizeof(double) * (size_t)nsig);
  for (int k = 0; k < nsig; ++k) inv_s2[k] = -1.0 / (sigmas[k] * sigmas[k]);
  double *s1 = arad_sin_mol(q1, n1, nm1, amax, cut, 1);
  double *s2 = arad_sin_mol(q2, n2, nm2, amax, cut, 1);
  double *self1 = arad_self_mol(q1, n1, nm1, amax, s1, &p);
  double *self2 = arad_self_mol(q2, n2, nm2, amax, s2, &p);
#pragma omp parallel
  {
    double *ad = (double *)malloc(sizeof(double) * (size_t)amax * amax);
#pragma omp for schedule(dynamic, 1)
    for (int j = 0; j < nm2; ++j) {
      for (int i = 0; i < nm1; ++i) {
        int ni = n1[i], nj = n2[j];
        for (int a = 0; a < ni; ++a)
          for (int b = 0; b < nj; ++b) {
            const double *x1 = q1 + ((size_t)i * amax + a) * 5 * amax;
            const double *x2 = q2 + ((size_t)j * amax + b) * 5 * amax;
            double S = arad_distl2(x1, x2, amax, ni, nj,
                                   s1 + ((size_t)i * amax + a) * amax,
                                   s2 + ((size_t)j * amax + b) * amax, &p);
            const double *za = z1 + ((size_t)i * amax + a) * 2;
            const double *zb = z2 + ((size_t)j * amax + b) * 2;
            double T = arad_stoch(za[0], za[1], zb[0], zb[1], &p);
            ad[a + (size_t)amax * b] =
                arad_pair_distance(self1[(size_t)i * amax + a],
                                   self2[(size_t)j * amax + b], S, T, EPS_5E12);
          }
        for (int k = 0; k < nsig; ++k) {
          double acc = 0.0;
          for (int b = 0; b < nj; ++b)
            for (int a = 0; a < ni; ++a)
              acc = acc + exp(ad[a + (size_t)amax * b] * inv_s2[k]);
          out[k + (size_t)nsig * (i + (size_t)nm1 * j)] = acc;
        }
      }
    }
    free(ad);
  }
  free(inv_s2);
  free(s1);
  free(s2);
  free(self1);
  free(self2);
}

/*
 * fml/kernels/farad_kernels.f90:619-755 (fget_symmetric_kernels_arad)
 * As fget_kernels_arad for X vs X, sin terms without the <cut guard
 * (689-698), i<=j computed with set-1 as i, then mirrored (743-744).
 */
EXPORT void oracle_symmetric_kernels_arad(const double *q1, const double *z1,
                                          const int *n1, const double *sigmas,
                                          int nm1, int nsig, int amax,
                                          double width, double cut, double rw,
                                          double cw, double *out) {
  arad_par p = {width, cut, rw, cw};
  double *inv_s2 = (double *)malloc(sizeof(double) * (size_t)nsig);
  for (int k = 0; k < nsig; ++k) inv_s2[k] = -1.0 / (sigmas[k] * sigmas[k]);
  double *s1 = arad_sin_mol(q1, n1, nm1, amax, cut, 0);
  double *self1 = arad_self_mol(q1, n1, nm1, amax, s1, &p);
#pragma omp parallel
  {
    double *ad = (double *)malloc(sizeof(double) * (size_t)amax * amax);
#pragma omp for schedule(dynamic, 1)
    for (int j = 0; j < nm1; ++j) {
      for (int i = 0; i <= j; ++i) {
        int ni = n1[i], nj = n1[j];
        for (int a = 0; a < ni; ++a)
          for (int b = 0; b < nj; ++b) {
            const double *x1 = q1 + ((size_t)i * amax + a) * 5 * amax;
            const double *x2 = q1 + ((size_t)j * amax + b) * 5 * amax;
            double S = arad_distl2(x1, x2, amax, ni, nj,
                                   s1 + ((size_t)i * amax + a) * amax,
                                   s1 + ((size_t)j * amax + b) * amax, &p);
            const double *za = z1 + ((size_t)i * amax + a) * 2;
            const double *zb = z1 + ((size_t)j * amax + b) * 2;
            double T = arad_stoch(za[0], za[1], zb[0], zb[1], &p);
            ad[a + (size_t)amax * b] =
                arad_pair_distance(self1[(size_t)i * amax + a],
                                   self1[(size_t)j * amax + b], S, T, EPS_5E12);
          }
        for (int k = 0; k < nsig; ++k) {
          double acc = 0.0;
          for (int b = 0; b < nj; ++b)
            for (int a = 0; a < ni; ++a)
              acc = acc + exp(ad[a + (size_t)amax * b] * inv_s2[k]);
          out[k + (size_t)nsig * (i + (size_t)nm1 * j)] = acc;
          out[k + (size_t)nsig * (j + (size_t)nm1 * i)] = acc;
        }
      }
    }
    free(ad);
  }
  free(inv_s2);
  free(s1);
  free(self1);
}

/*
 * fml/kernels/farad_kernels.f90:95-227 (fget_kernel_arad)
 * One sigma; width .2, cut 6.0, rw 1.0, cw .5 hard-coded (101-104);
 * eps = real(1.0e-11) (148); no <cut guard on sin (163-171);
 * K(i,j) = K(j,i) = sum of the full atomic_kernel work array (216-217),
 * whose untouched entries are zero, so the sum equals the sum over
 * (:ni,:nj) in column-major order.  Output (nm, nm) Fortran order.
 */
EXPORT void oracle_kernel_arad(const double *q, const double *z, const int *n,
                               double sigma, int nm, int amax, double *out) {
  arad_par p = {0.2, 6.0, 1.0, 0.5};
  const double inv_s2 = -1.0 / (sigma * sigma);
  double *s1 = arad_sin_mol(q, n, nm, amax, p.cut, 0);
  double *self1 = arad_self_mol(q, n, nm, amax, s1, &p);
#pragma omp parallel
  {
    double *ad = (double *)malloc(sizeof(double) * (size_t)amax * amax);
#pragma omp for schedule(dynamic, 1)
    for (int j = 0; j < nm; ++j) {
      for (int i = 0; i <= j; ++i) {
        int ni = n[i], nj = n[j];
        for (int a = 0; a < ni; ++a)
          for (int b = 0; b < nj; ++b) {
            const double *x1 = q + ((size_t)i * amax + a) * 5 * amax;
            const double *x2 = q + ((size_t)j * amax + b) * 5 * amax;
            double S = arad_distl2(x1, x2, amax, ni, nj,
                                   s1 + ((size_t)i * amax + a) * amax,
                                   s1 + ((size_t)j * amax + b) * amax, &p);
            const double *za = z + ((size_t)i * amax + a) * 2;
            const double *zb = z + ((size_t)j * amax + b) * 2;
            double T = arad_stoch(za[0], za[1], zb[0], zb[1], &p);
            double l2 = arad_pair_distance(self1[(size_t)i * amax + a],
                                           self1[(size_t)j * amax + b], S, T,
                                           EPS_1E11);
            ad[a + (size_t)amax * b] = exp(inv_s2 * l2);
          }
        double acc = 0.0;
        for (int b = 0; b < nj; ++b)
          for (int a = 0; a < ni; ++a) acc = acc + ad[a + (size_t)amax * b];
        out[i + (size_t)nm * j] = acc;
        out[j + (size_t)nm * i] = acc;
      }
    }
    free(ad);
  }
  free(s1);
  free(self1);
}

/* sin terms for per-atom descriptors q[na][5][amax] (no guard;
 * farad_kernels.f90:836-852) */
static double *arad_sin_atom(const double *q, const int *n, int na, int amax,
                             double cut) {
  const double pi = 4.0 * atan(1.0);
  const double inv_cut = pi / (2.0 * cut);
  double *s = (double *)calloc((size_t)na * amax, sizeof(double));
  for (int i = 0; i < na; ++i)
    for (int m = 0; m < n[i]; ++m)
      s[(size_t)i * amax + m] = 1.0 - sin(q[(size_t)i * 5 * amax + m] * inv_cut);
  return s;
}

/* shared body of fget_atomic_kernels_arad / fget_atomic_distance_arad:
 * D(i,j) for atom i of set 1 and atom j of set 2 (farad_kernels.f90:873-889,
 * 1006-1019). */
static void arad_atomic_distances(const double *q1, const double *q2,
                                  const double *z1, const double *z2,
                                  const int *n1, const int *n2, int na1,
                                  int na2, int amax, const arad_par *p,
                                  double *dmat /* D(i,j) at i + na1*j */) {
  double *s1 = arad_sin_atom(q1, n1, na1, amax, p->cut);
  double *s2 = arad_sin_atom(q2, n2, na2, amax, p->cut);
  double *self1 = (double *)malloc(sizeof(double) * (size_t)(na1 > 0 ? na1 : 1));
  double *self2 = (double *)malloc(sizeof(double) * (size_t)(na2 > 0 ? na2 : 1));
  for (int i = 0; i < na1; ++i) {
    const double *x = q1 + (size_t)i * 5 * amax;
    self1[i] = arad_distl2(x, x, amax, n1[i], n1[i], s1 + (size_t)i * amax,
                           s1 + (size_t)i * amax, p);
  }
  for (int i = 0; i < na2; ++i) {
    const double *x = q2 + (size_t)i * 5 * amax;
    self2[i] = arad_distl2(x, x, amax, n2[i], n2[i], s2 + (size_t)i * amax,
                           s2 + (size_t)i * amax, p);
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 0; j < na2; ++j)
    for (int i = 0; i < na1; ++i) {
      double S = arad_distl2(q1 + (size_t)i * 5 * amax, q2 + (size_t)j * 5 * amax,
                             amax, n1[i], n2[j], s1 + (size_t)i * amax,
                             s2 + (size_t)j * amax, p);
      double T = arad_stoch(z1[2 * (size_t)i], z1[2 * (size_t)i + 1],
                            z2[2 * (size_t)j], z2[2 * (size_t)j + 1], p);
      dmat[i + (size_t)na1 * j] =
          arad_pair_distance(self1[i], self2[j], S, T, EPS_5E12);
    }
  free(s1);
  free(s2);
  free(self1);
  free(self2);
}

/*
 * fml/kernels/farad_kernels.f90:757-899 (fget_atomic_kernels_arad)
 * kernels(k,i,j) = exp(D(i,j)*inv_sigma2(k)) (887), Fortran order
 * (nsig, na1, na2).
 */
EXPORT void oracle_atomic_kernels_arad(const double *q1, const double *q2,
                                       const double *z1, const double *z2,
                                       const int *n1, const int *n2,
                                       const double *sigmas, int na1, int na2,
                                       int nsig, int amax, double width,
                                       double cut, double rw, double cw,
                                       double *out) {
  arad_par p = {width, cut, rw, cw};
  double *dm = (double *)malloc(sizeof(double) * (size_t)na1 * na2 + 8);
  arad_atomic_distances(q1, q2, z1, z2, n1, n2, na1, na2, amax, &p, dm);
  for (int k = 0; k < nsig; ++k) {
    double inv = -1.0 / (sigmas[k] * sigmas[k]);
    for (size_t t = 0; t < (size_t)na1 * na2; ++t)
      out[k + (size_t)nsig * t] = exp(dm[t] * inv);
  }
  free(dm);
}

/*
 * fml/kernels/farad_kernels.f90:902-1030 (fget_atomic_distance_arad)
 * The reference stores D(i,j) as kernel(j,i) into a (na1, na2) array (932,
 * 1019): correct only for na1 == na2.  This restatement writes the
 * well-defined (na1, na2) Fortran-order matrix D(i,j) at out[i + na1*j];
 * the transposed reference layout equals it for na1 == na2 after a
 * transpose.
 */
EXPORT void oracle_atomic_distance_arad(const double *q1, const double *q2,
                                        const double *z1, const double *z2,
                                        const int *n1, const int *n2, int na1,
                                        int na2, int amax, double width,
                                        double cut, double rw, double cw,
                                        double *out) {
  arad_par p = {width, cut, rw, cw};
  arad_atomic_distances(q1, q2, z1, z2, n1, n2, na1, na2, amax, &p, out);
}

/*
 * fml/math/farad.f90:452-600 (atomic_arad_l2_distance_all)
 * dmatrix(i,j,i1,j1) Fortran order (nm1, nm2, amax, amax), zero where
 * i1 >= n1(i) or j1 >= n2(j) (565); sin terms without the guard (520-540).
 */
EXPORT void oracle_arad_l2_distance_all(const double *q1, const double *q2,
                                        const double *z1, const double *z2,
                                        const int *n1, const int *n2, int nm1,
                                        int nm2, int amax, double width,
                                        double cut, double rw, double cw,
                                        double *out) {
  arad_par p = {width, cut, rw, cw};
  double *s1 = arad_sin_mol(q1, n1, nm1, amax, cut, 0);
  double *s2 = arad_sin_mol(q2, n2, nm2, amax, cut, 0);
  double *self1 = arad_self_mol(q1, n1, nm1, amax, s1, &p);
  double *self2 = arad_self_mol(q2, n2, nm2, amax, s2, &p);
  const size_t total = (size_t)nm1 * nm2 * amax * amax;
  for (size_t t = 0; t < total; ++t) out[t] = 0.0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int j = 0; j < nm2; ++j)
    for (int i = 0; i < nm1; ++i) {
      int ni = n1[i], nj = n2[j];
      for (int a = 0; a < ni; ++a)
        for (int b = 0; b < nj; ++b) {
          const double *x1 = q1 + ((size_t)i * amax + a) * 5 * amax;
          const double *x2 = q2 + ((size_t)j * amax + b) * 5 * amax;
          double S = arad_distl2(x1, x2, amax, ni, nj,
                                 s1 + ((size_t)i * amax + a) * amax,
                                 s2 + ((size_t)j * amax + b) * amax, &p);
          const double *za = z1 + ((size_t)i * amax + a) * 2;
          const double *zb = z2 + ((size_t)j * amax + b) * 2;
          double T = arad_stoch(za[0], za[1], zb[0], zb[1], &p);
          out[i + (size_t)nm1 * (j + (size_t)nm2 * (a + (size_t)amax * b))] =
              arad_pair_distance(self1[(size_t)i * amax + a],
                                 self2[(size_t)j * amax + b], S, T, EPS_5E12);
        }
    }
  free(s1);
  free(s2);
  free(self1);
  free(self2);
}

/* ------------------------------------------------------------------ */
/* Unblocked Cholesky (upper), used only to cross-check LAPACK on tiny  */
/* closed-form cases.  fml/math/fcho_solve.f90:23-47 calls dpotrf('U')  */
/* + dpotrs('U'); A is Fortran order and only its upper triangle is     */
/* read.  Returns the LAPACK-style info (0 ok, k>0 leading minor k).    */
/* ------------------------------------------------------------------ */
EXPORT int oracle_cho_solve(const double *a_in, int n, const double *y,
                            double *x) {
  double *u = (double *)malloc(sizeof(double) * (size_t)n * n + 8);
  memcpy(u, a_in, sizeof(double) * (size_t)n * n);
  /* U(i,j) at u[i + n*j], i <= j */
  for (int j = 0; j < n; ++j) {
    double s = u[j + (size_t)n * j];
    for (int k = 0; k < j; ++k) s -= u[k + (size_t)n * j] * u[k + (size_t)n * j];
    if (!(s > 0.0)) {
      free(u);
      return j + 1;
    }
    double ujj = sqrt(s);
    u[j + (size_t)n * j] = ujj;
    for (int i = j + 1; i < n; ++i) {
      double t = u[j + (size_t)n * i];
      for (int k = 0; k < j; ++k) t -= u[k + (size_t)n * j] * u[k + (size_t)n * i];
      u[j + (size_t)n * i] = t / ujj;
    }
  }
  for (int i = 0; i < n; ++i) x[i] = y[i];
  for (int i = 0; i < n; ++i) { /* U^T w = y */
    double t = x[i];
    for (int k = 0; k < i; ++k) t -= u[k + (size_t)n * i] * x[k];
    x[i] = t / u[i + (size_t)n * i];
  }
  for (int i = n - 1; i >= 0; --i) { /* U x = w */
    double t = x[i];
    for (int k = i + 1; k < n; ++k) t -= u[i + (size_t)n * k] * x[k];
    x[i] = t / u[i + (size_t)n * i];
  }
  free(u);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Atomic KRR path: per-atom Coulomb matrices, vector kernels,          */
/* get_alpha / get_prediction                                          */
/* ------------------------------------------------------------------ */

/* Pair matrix and row norms as frepresentations.f90:152-172 (identical in
 * :238-258): diag 0.5*Z**2.4, pair Zi*Zj/sqrt(sum((Rj-Ri)**2)) for i<j;
 * norms in the single-thread order of the two REDUCTION loops. */
static void cm_pairs(const double *z, const double *r, int n, double *pdm,
                     double *norms) {
  double *priv = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    double p = 0.5 * pow(z[i], 2.4);
    norms[i] = 0.0 + p;
    pdm[i * n + i] = p;
    priv[i] = 0.0;
  }
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      double dx = r[j * 3 + 0] - r[i * 3 + 0];
      double dy = r[j * 3 + 1] - r[i * 3 + 1];
      double dz = r[j * 3 + 2] - r[i * 3 + 2];
      double s = dx * dx;
      s = s + dy * dy;
      s = s + dz * dz;
      double p = z[i] * z[j] / sqrt(s);
      priv[j] = priv[j] + p;
      priv[i] = priv[i] + p;
      pdm[i * n + j] = p;
      pdm[j * n + i] = p;
    }
  }
  for (int i = 0; i < n; ++i) norms[i] = norms[i] + priv[i];
  free(priv);
}

/* Writes row `row` of a packed per-atom matrix: entries (m, q), q <= m,
 * of the pair matrix taken in the order ord[] (frepresentations.f90:280-291,
 * :376-387), zero padded. */
static void cm_pack_row(const double *pdm, int n, const int *ord, int packed,
                        double *out) {
  for (int t = 0; t < packed; ++t) out[t] = 0.0;
  for (int m = 0; m < n; ++m) {
    int i = ord[m];
    int idx = m * (m + 1) / 2;
    for (int q = 0; q <= m; ++q) out[idx + q] = pdm[i * n + ord[q]];
  }
}

/*
 * Per-molecule Coulomb matrix forms; each returns 0, or -1 when
 * natoms > nmax (the reference writes out of bounds there) or natoms < 0.
 *
 * mode 0: fgenerate_unsorted_coulomb_matrix (frepresentations.f90:112-201):
 *   the packed lower triangle in input atom order, cm[packed].
 * mode 1: fgenerate_local_coulomb_matrix (:204-294): atoms sorted once by
 *   descending row norm (repeated minloc placed last, :262-266: equal norms
 *   with the higher index first); row k puts atom k first and then the
 *   sorted list without k (:268-277); cm[natoms][packed].
 * mode 2: fgenerate_atomic_coulomb_matrix (:297-390): row k orders the atoms
 *   by ascending distance from atom k (repeated minloc over distance
 *   column k, :366-372: equal distances lowest index first); distances
 *   sqrt(sum((Rj-Ri)**2)) for i<j, pair values Zi*Zj/distance (:339-357);
 *   cm[natoms][packed].
 * Parity: pinned bitwise by the flang build of the reference (oracle/_ref)
 * and by hand-computed molecules (tests/test_oracle.py).
 */
EXPORT int oracle_coulomb_form(int mode, const double *z, const double *r,
                               int natoms, int nmax, double *cm) {
  if (natoms < 0 || natoms > nmax) return -1;
  const int n = natoms;
  const int packed = nmax * (nmax + 1) / 2;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  double *pdm = (double *)malloc(sizeof(double) * nn * nn);
  double *norms = (double *)malloc(sizeof(double) * nn);
  int *sorted = (int *)malloc(sizeof(int) * nn);
  int *ord = (int *)malloc(sizeof(int) * nn);
  cm_pairs(z, r, n, pdm, norms);
  if (mode == 0) {
    for (int i = 0; i < n; ++i) ord[i] = i;
    cm_pack_row(pdm, n, ord, packed, cm);
  } else if (mode == 1) {
    for (int i = 0; i < n; ++i) {
      int jmin = 0;
      for (int j = 1; j < n; ++j)
        if (norms[j] < norms[jmin]) jmin = j;
      sorted[n - 1 - i] = jmin;
      norms[jmin] = 1.7976931348623157e308;
    }
    for (int k = 0; k < n; ++k) {
      ord[0] = k;
      int m = 0;
      for (int i = 0; i < n - 1; ++i) {
        if (sorted[i] == k) m = 1;
        ord[i + 1] = sorted[i + m];
      }
      cm_pack_row(pdm, n, ord, packed, cm + (size_t)k * packed);
    }
  } else {
    double *dist = (double *)malloc(sizeof(double) * nn * nn);
    for (int i = 0; i < n; ++i) dist[i * n + i] = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        double dx = r[j * 3 + 0] - r[i * 3 + 0];
        double dy = r[j * 3 + 1] - r[i * 3 + 1];
        double dz = r[j * 3 + 2] - r[i * 3 + 2];
        double s = dx * dx;
        s = s + dy * dy;
        s = s + dz * dz;
        double d = sqrt(s);
        dist[i * n + j] = d;
        dist[j * n + i] = d;
      }
    /* pair values from the distance: Zi*Zj/distance_matrix(j,i) (:355-356) */
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        double p = z[i] * z[j] / dist[j * n + i];
        pdm[i * n + j] = p;
        pdm[j * n + i] = p;
      }
    double *col = (double *)malloc(sizeof(double) * nn);
    for (int k = 0; k < n; ++k) {
      for (int i = 0; i < n; ++i) col[i] = dist[i * n + k];
      for (int i = 0; i < n; ++i) {
        int jmin = 0;
        for (int j = 1; j < n; ++j)
          if (col[j] < col[jmin]) jmin = j;
        ord[i] = jmin;
        col[jmin] = 1.7976931348623157e308;
      }
      cm_pack_row(pdm, n, ord, packed, cm + (size_t)k * packed);
    }
    free(col);
    free(dist);
  }
  free(pdm);
  free(norms);
  free(sorted);
  free(ord);
  return 0;
}

/* Batched form over concatenated molecules (offsets: nmol+1 entries).
 * Mode 0 writes one packed row per molecule, modes 1-2 one per atom (row
 * offsets[m] + k for atom k of molecule m). */
EXPORT int oracle_coulomb_forms(int mode, const double *z, const double *r,
                                const int *offsets, int nmol, int nmax,
                                double *cm) {
  const int packed = nmax * (nmax + 1) / 2;
  int rc = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : rc)
  for (int m = 0; m < nmol; ++m) {
    int a0 = offsets[m];
    int na = offsets[m + 1] - a0;
    size_t row = mode == 0 ? (size_t)m : (size_t)a0;
    rc |= oracle_coulomb_form(mode, z + a0, r + 3 * (size_t)a0, na, nmax,
                              cm + row * packed);
  }
  return rc ? -1 : 0;
}

/*
 * fml/kernels/fkernels.f90:23-162 (fget_vector_kernels_laplacian, kind 1;
 * fget_vector_kernels_gaussian, kind 0, at :93).
 * q1[nm1][amax1][d], q2[nm2][amax2][d]: the D-vector of atom ia of molecule
 * i is q1(:, ia, i) in the reference.  Atom-pair distance
 * sum(abs(q1-q2)) (:76) or sum((q1-q2)**2) (:147), sequential over d;
 * inv = -1/sigma (:58) or -0.5/sigma**2 (:128);
 * kernels(k, i, j) = sum(exp(dist(:ni, :nj) * inv(k))) (:82, :153): one
 * accumulator over the section in array element order (ia fastest).
 * out: Fortran (nsig, nm1, nm2) -> out[k + nsig*(i + nm1*j)].
 * Parity: pinned by the flang build of the reference (oracle/_ref) and by
 * closed forms (tests/test_oracle.py).
 */
EXPORT void oracle_vector_kernels(int kind, const double *q1, const double *q2,
                                  const int *n1, const int *n2,
                                  const double *sigmas, int nm1, int nm2,
                                  int nsig, int amax1, int amax2, int d,
                                  double *out) {
  double *inv = (double *)malloc(sizeof(double) * (size_t)(nsig > 0 ? nsig : 1));
  for (int k = 0; k < nsig; ++k)
    inv[k] = kind == 1 ? -1.0 / sigmas[k] : -0.5 / (sigmas[k] * sigmas[k]);
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 0; j < nm2; ++j) {
    double *dist = (double *)malloc(sizeof(double) * (size_t)(amax1 * amax2 > 0 ? amax1 * amax2 : 1));
    const int nj = n2[j];
    for (int i = 0; i < nm1; ++i) {
      const int ni = n1[i];
      for (int ja = 0; ja < nj; ++ja) {
        const double *b = q2 + ((size_t)j * amax2 + ja) * d;
        for (int ia = 0; ia < ni; ++ia) {
          const double *a = q1 + ((size_t)i * amax1 + ia) * d;
          double s = 0.0;
          if (kind == 1)
            for (int t = 0; t < d; ++t) s = s + fabs(a[t] - b[t]);
          else
            for (int t = 0; t < d; ++t) {
              double df = a[t] - b[t];
              s = s + df * df;
            }
          dist[ja * ni + ia] = s;
        }
      }
      for (int k = 0; k < nsig; ++k) {
        double acc = 0.0;
        for (int t = 0; t < ni * nj; ++t) acc = acc + exp(dist[t] * inv[k]);
        out[k + (size_t)nsig * (i + (size_t)nm1 * j)] = acc;
      }
    }
    free(dist);
  }
  free(inv);
}

/*
 * fml/kernels/fkernels.f90:380-478 (fget_alpha), the molecular kernel part:
 * q[na][d] atom rows (q(:, k) in the reference), n[nm] atoms per molecule
 * in order.  For each molecule j: ltc(i) = sum over atoms k of j, in order,
 * of exp(inv * sum(abs(q(:,k) - q(:,i)))) with inv = -1/sigma (:407, :430);
 * ltcl(i, j) = sum(ltc(atoms of i)) for i <= j (:433-435).  Writes the
 * upper triangle of ltcl, Fortran (nm, nm): ltcl[i + nm*j].
 * Parity: flang build of the reference (oracle/_ref) through the full
 * fget_alpha, and closed forms.
 */
EXPORT void oracle_alpha_kernel(const double *q, const int *n, int nm, int d,
                                double sigma, double *ltcl) {
  const double inv = -1.0 / sigma;
  int *start = (int *)malloc(sizeof(int) * (size_t)(nm + 1));
  start[0] = 0;
  for (int i = 0; i < nm; ++i) start[i + 1] = start[i] + n[i];
  const int na = start[nm];
#pragma omp parallel for schedule(dynamic, 1)
  for (int j = 0; j < nm; ++j) {
    double *ltc = (double *)malloc(sizeof(double) * (size_t)(na > 0 ? na : 1));
    for (int i = 0; i < na; ++i) {
      double acc = 0.0;
      const double *qi = q + (size_t)i * d;
      for (int k = start[j]; k < start[j + 1]; ++k) {
        const double *qk = q + (size_t)k * d;
        double s = 0.0;
        for (int t = 0; t < d; ++t) s = s + fabs(qk[t] - qi[t]);
        acc = acc + exp(inv * s);
      }
      ltc[i] = acc;
    }
    for (int i = 0; i <= j; ++i) {
      double s = 0.0;
      for (int a = start[i]; a < start[i + 1]; ++a) s = s + ltc[a];
      ltcl[i + (size_t)nm * j] = s;
    }
    free(ltc);
  }
  free(start);
}

/*
 * fml/kernels/fkernels.f90:480-539 (fget_prediction): for each test atom i,
 * csa(i) = sum over training atoms j, in order, of
 * exp(inv * sum(abs(q(:,j) - q2(:,i)))) * alpha(j), inv = -1/sigma (:504,
 * :522); Y(m) = sum(csa(atoms of test molecule m)) (:530).
 * q[na][d], q2[na2][d] atom rows; n2[nm2].
 * Parity: flang build of the reference (oracle/_ref) and closed forms.
 */
EXPORT void oracle_prediction(const double *q, int na, const double *q2, int na2,
                              const int *n2, int nm2, int d, const double *alpha,
                              double sigma, double *y) {
  const double inv = -1.0 / sigma;
  double *csa = (double *)malloc(sizeof(double) * (size_t)(na2 > 0 ? na2 : 1));
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < na2; ++i) {
    const double *b = q2 + (size_t)i * d;
    double acc = 0.0;
    for (int j = 0; j < na; ++j) {
      const double *a = q + (size_t)j * d;
      double s = 0.0;
      for (int t = 0; t < d; ++t) s = s + fabs(a[t] - b[t]);
      acc = acc + exp(inv * s) * alpha[j];
    }
    csa[i] = acc;
  }
  int a0 = 0;
  for (int m = 0; m < nm2; ++m) {
    double s = 0.0;
    for (int a = a0; a < a0 + n2[m]; ++a) s = s + csa[a];
    y[m] = s;
    a0 += n2[m];
  }
  free(csa);
}

/*
 * fml/kernels/fkernels.f90:541-641 (fget_alpha_from_distance), the molecular
 * kernel part: dmat Fortran (na, na), d(k, i) at dmat[k + na*i];
 * ltc(j, i) = sum over atoms k of j of exp(inv * d(k, i)), inv = -1/dgamma
 * (:568, :588-590); ltcl(i, j) = sum(ltc(j, atoms of i)) for all i, j
 * (:597-599).  Writes ltcl Fortran (nm, nm).
 */
EXPORT void oracle_alpha_kernel_from_distance(const double *dmat, const int *n, int nm,
                                              double dgamma, double *ltcl) {
  const double inv = -1.0 / dgamma;
  int *start = (int *)malloc(sizeof(int) * (size_t)(nm + 1));
  start[0] = 0;
  for (int i = 0; i < nm; ++i) start[i + 1] = start[i] + n[i];
  const int na = start[nm];
  double *ltc = (double *)malloc(sizeof(double) * (size_t)(nm > 0 ? nm : 1) * (size_t)(na > 0 ? na : 1));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nm; ++j) {
      double acc = 0.0;
      for (int k = start[j]; k < start[j + 1]; ++k) acc = acc + exp(inv * dmat[k + (size_t)na * i]);
      ltc[j + (size_t)nm * i] = acc;
    }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < nm; ++j)
    for (int i = 0; i < nm; ++i) {
      double s = 0.0;
      for (int a = start[i]; a < start[i + 1]; ++a) s = s + ltc[j + (size_t)nm * a];
      ltcl[i + (size_t)nm * j] = s;
    }
  free(ltc);
  free(start);
}

/*
 * fml/kernels/fkernels.f90:645-702 (fget_prediction_from_distance):
 * d2 Fortran (na, na2), D2(j, i) at d2[j + na*i];
 * csa(i) = sum_j exp(inv * D2(j, i)) * alpha(j), inv = -1/dgamma (:668,
 * :685); Y(m) = sum(csa(atoms of test molecule m)) (:693).
 * The reference never zeroes csa before accumulating (allocated at :680,
 * unlike fget_prediction's :517); it is taken as starting from zero here.
 */
EXPORT void oracle_prediction_from_distance(const double *d2, int na, int na2,
                                            const int *n2, int nm2, const double *alpha,
                                            double dgamma, double *y) {
  const double inv = -1.0 / dgamma;
  double *csa = (double *)malloc(sizeof(double) * (size_t)(na2 > 0 ? na2 : 1));
  for (int i = 0; i < na2; ++i) {
    double acc = 0.0;
    for (int j = 0; j < na; ++j) acc = acc + exp(inv * d2[j + (size_t)na * i]) * alpha[j];
    csa[i] = acc;
  }
  int a0 = 0;
  for (int m = 0; m < nm2; ++m) {
    double s = 0.0;
    for (int a = a0; a < a0 + n2[m]; ++a) s = s + csa[a];
    y[m] = s;
    a0 += n2[m];
  }
  free(csa);
}

/* ------------------------------------------------------------------ */
/* ARAS kernels (fml/kernels/faras_kernels.f90)                        */
/* ------------------------------------------------------------------ */

typedef struct {
  double t_width, r_width, c_width, d_width, cut, angular_scale;
  int order;
} aras_par;

/* x[nm][amax][3 + amax][amax]: channel 0 distances, 1 nuclear charges,
 * 3 + m angles to neighbour m (aras.py:123-129). */
#define ARAS_X(x, amax, a, i, c, m) \
  (x)[((((size_t)(a) * (amax) + (i)) * (size_t)(3 + (amax)) + (c)) * (amax)) + (m)]

/* Per-atom tables of one set (faras_kernels.f90:258-361 / :540-620):
 * ksi[a][i][m] (radial, neighbours m >= 1 inside the cut: 1/d^6 q0 qm),
 * cosp/sinp[a][i][o][j][p] (p = nuclear charge 0..pmax): sum over
 * neighbours k != j (both >= 1, < nneigh) with charge p of
 * (cos(o' th) - cos((th + pi) o')) ksi3(j, k), o' = o + 1, th the angle
 * between j and k, and ksi3 the three-body factor
 * q0 qj qk (1 + cos(th) cos_i cos_j) / (di dj dij)^3 with dij, cos_i, cos_j
 * from the law of cosines (:288-306). */
typedef struct {
  int nm, amax, order, pmax;
  double *ksi;
  double *cosp, *sinp;
} aras_tables;

static double powi6(double x) { return x * x * x * x * x * x; }

static void aras_build(const double *x, const int *n, const int *nn, int nm, int amax,
                       const aras_par *pr, int pmax, aras_tables *t) {
  const double pi = 3.141592653589793;
  t->nm = nm;
  t->amax = amax;
  t->order = pr->order;
  t->pmax = pmax;
  const size_t na = (size_t)nm * amax;
  t->ksi = (double *)calloc(na * amax + 1, sizeof(double));
  const size_t per = (size_t)pr->order * amax * (pmax + 1);
  t->cosp = (double *)calloc(na * per + 1, sizeof(double));
  t->sinp = (double *)calloc(na * per + 1, sizeof(double));
  double *k3 = (double *)calloc((size_t)amax * amax + 1, sizeof(double));
  for (int a = 0; a < nm; ++a)
    for (int i = 0; i < n[a]; ++i) {
      const int nb = nn[a * amax + i];
      const double q0 = ARAS_X(x, amax, a, i, 1, 0);
      double *ks = t->ksi + ((size_t)a * amax + i) * amax;
      for (int m = 1; m < nb; ++m) {
        const double dm = ARAS_X(x, amax, a, i, 0, m);
        if (dm < pr->cut) ks[m] = 1.0 / powi6(dm) * q0 * ARAS_X(x, amax, a, i, 1, m);
      }
      for (int m1 = 1; m1 < nb; ++m1)
        for (int m2 = 1; m2 < nb; ++m2) {
          if (m1 == m2) continue;
          const double di = ARAS_X(x, amax, a, i, 0, m1);
          const double dj = ARAS_X(x, amax, a, i, 0, m2);
          const double ang = ARAS_X(x, amax, a, i, 3 + m1, m2);
          const double dij = sqrt(di * di + dj * dj - 2.0 * di * dj * cos(ang));
          const double cosi = (dj * dj + dij * dij - di * di) / (2.0 * dj * dij);
          const double cosj = (di * di + dij * dij - dj * dj) / (2.0 * di * dij);
          const double p3 = di * dj * dij;
          k3[m1 * amax + m2] = q0 * ARAS_X(x, amax, a, i, 1, m1) * ARAS_X(x, amax, a, i, 1, m2) *
                               (1.0 + cos(ang) * cosi * cosj) / (p3 * p3 * p3);
        }
      double *cp = t->cosp + ((size_t)a * amax + i) * per;
      double *sp = t->sinp + ((size_t)a * amax + i) * per;
      for (int o = 0; o < pr->order; ++o) {
        const double mo = (double)(o + 1);
        for (int j = 1; j < nb; ++j)
          for (int k = 1; k < nb; ++k) {
            if (j == k) continue;
            const int p = (int)ARAS_X(x, amax, a, i, 1, k);
            const double th = ARAS_X(x, amax, a, i, 3 + j, k);
            const size_t at = ((size_t)o * amax + j) * (pmax + 1) + p;
            cp[at] = cp[at] + (cos(mo * th) - cos((th + pi) * mo)) * k3[j * amax + k];
            sp[at] = sp[at] + (sin(mo * th) - sin((th + pi) * mo)) * k3[j * amax + k];
          }
      }
    }
  free(k3);
}

static void aras_free(aras_tables *t) {
  free(t->ksi);
  free(t->cosp);
  free(t->sinp);
}

/* faras_kernels.f90:198-207: the angular normalisation. */
static double aras_ang_norm2(double t_width) {
  const double pi = 3.141592653589793;
  double s = 0.0;
  for (int n = -10000; n <= 10000; ++n) {
    const double tn = t_width * n;
    s = s + exp(-(tn * tn)) * (2.0 - 2.0 * cos(n * pi));
  }
  return sqrt(s * pi) * 2.0;
}

/* faras_kernels.f90:7-147 (aras_utils::atomic_distl2) for atom i1 of
 * molecule a1 (set 1) against atom i2 of molecule a2 (set 2). */
static double aras_distl2(const double *x1, const aras_tables *t1, int a1, int i1, int nb1,
                          const double *x2, const aras_tables *t2, int a2, int i2, int nb2,
                          const aras_par *pr, const double *pd, int emax, const double *s) {
  const int amax1 = t1->amax, amax2 = t2->amax;
  int pmax1 = 0, pmax2 = 0;
  for (int m = 0; m < nb1; ++m) {
    const int q = (int)ARAS_X(x1, amax1, a1, i1, 1, m);
    if (q > pmax1) pmax1 = q;
  }
  for (int m = 0; m < nb2; ++m) {
    const int q = (int)ARAS_X(x2, amax2, a2, i2, 1, m);
    if (q > pmax2) pmax2 = q;
  }
  char *use1 = (char *)calloc((size_t)pmax1 + 1, 1);
  char *use2 = (char *)calloc((size_t)pmax2 + 1, 1);
  for (int m = 0; m < nb1; ++m) use1[(int)ARAS_X(x1, amax1, a1, i1, 1, m)] = 1;
  for (int m = 0; m < nb2; ++m) use2[(int)ARAS_X(x2, amax2, a2, i2, 1, m)] = 1;
  const double inv_width = -1.0 / (4.0 * (pr->d_width * pr->d_width));
  const double g8 = 8.0 * pr->d_width;
  const double maxg2 = g8 * g8;
#define PD(p, q) pd[((size_t)(q) - 1) * emax + ((p) - 1)] /* Fortran pd(p, q) */
  const int q01 = (int)ARAS_X(x1, amax1, a1, i1, 1, 0), q02 = (int)ARAS_X(x2, amax2, a2, i2, 1, 0);
  double aad = PD(q01, q02);
  const double *ks1 = t1->ksi + ((size_t)a1 * amax1 + i1) * amax1;
  const double *ks2 = t2->ksi + ((size_t)a2 * amax2 + i2) * amax2;
  const size_t per1 = (size_t)t1->order * amax1 * (t1->pmax + 1);
  const size_t per2 = (size_t)t2->order * amax2 * (t2->pmax + 1);
  const double *c1 = t1->cosp + ((size_t)a1 * amax1 + i1) * per1;
  const double *s1 = t1->sinp + ((size_t)a1 * amax1 + i1) * per1;
  const double *c2 = t2->cosp + ((size_t)a2 * amax2 + i2) * per2;
  const double *s2 = t2->sinp + ((size_t)a2 * amax2 + i2) * per2;
  for (int m1 = 1; m1 < nb1; ++m1) {
    const double d1 = ARAS_X(x1, amax1, a1, i1, 0, m1);
    if (d1 > pr->cut) break;
    for (int m2 = 1; m2 < nb2; ++m2) {
      const double d2 = ARAS_X(x2, amax2, a2, i2, 0, m2);
      if (d2 > pr->cut) break;
      const double dd = d2 - d1;
      const double r2 = dd * dd;
      if (!(r2 < maxg2)) continue;
      const double d = exp(r2 * inv_width) *
                       PD((int)ARAS_X(x1, amax1, a1, i1, 1, m1), (int)ARAS_X(x2, amax2, a2, i2, 1, m2));
      double angular = 0.0 * 0.0;
      for (int o = 0; o < pr->order; ++o) {
        double temp = 0.0;
        for (int p1 = 1; p1 <= pmax1; ++p1) {
          if (!use1[p1]) continue;
          const double cc = c1[((size_t)o * amax1 + m1) * (t1->pmax + 1) + p1];
          const double ss = s1[((size_t)o * amax1 + m1) * (t1->pmax + 1) + p1];
          for (int p2 = 1; p2 <= pmax2; ++p2) {
            if (!use2[p2]) continue;
            temp = temp + (cc * c2[((size_t)o * amax2 + m2) * (t2->pmax + 1) + p2] +
                           ss * s2[((size_t)o * amax2 + m2) * (t2->pmax + 1) + p2]) * PD(p2, p1);
          }
        }
        angular = angular + temp * s[o];
      }
      aad = aad + d * (ks1[m1] * ks2[m2] + angular * pr->angular_scale);
    }
  }
#undef PD
  free(use1);
  free(use2);
  return aad;
}

static int aras_pmax(const double *x, const int *n, const int *nn, int nm, int amax) {
  int p = 0;
  for (int a = 0; a < nm; ++a)
    for (int i = 0; i < n[a]; ++i)
      for (int m = 0; m < nn[a * amax + i]; ++m) {
        const int q = (int)ARAS_X(x, amax, a, i, 1, m);
        if (q > p) p = q;
      }
  return p;
}

/*
 * fml/kernels/faras_kernels.f90:150-483 (fget_kernels_aras; sym = 0) and
 * :486-723 (fget_symmetric_kernels_aras; sym = 1, x2 = x1, each pair i >= j
 * computed with molecule i in the first role and written to both halves).
 * nn[a][i]: neighbours of atom i inside the cut, self included (the
 * Python wrapper's neighbors arrays, kernels.py:460-472).  pd: Fortran
 * (emax, emax) element similarity (kernels.py gen_pd).  The element axis of
 * the projection tables is sized by the largest charge any atom sees (the
 * reference sizes it from each molecule's first atom and writes past it
 * when another atom sees a heavier element, :210-215).  Expression order
 * as the reference; l2 = self1 + self2 - 2 l2 pd(q1, q2), zeroed below
 * 5e-12 (:452-455); kernels(s, i, j) = sum over the (ni, nj) atom block,
 * column-major, of exp(l2 * (-1/sigma^2)).
 * out: Fortran (nsig, nm1, nm2).
 * Parity: flang build of the reference (oracle/_ref) and closed forms.
 */
EXPORT void oracle_kernels_aras(int sym, const double *x1, const double *x2, const int *n1,
                                const int *n2, const int *nn1, const int *nn2,
                                const double *sigmas, int nm1, int nm2, int nsig, int amax,
                                double t_width, double r_width, double c_width, double d_width,
                                double cut, int order, const double *pd, int emax,
                                double angular_scale, double *out) {
  aras_par pr = {t_width, r_width, c_width, d_width, cut, angular_scale, order};
  if (sym) {
    x2 = x1;
    n2 = n1;
    nn2 = nn1;
    nm2 = nm1;
  }
  const double pi = 3.141592653589793;
  const double ang_norm2 = aras_ang_norm2(t_width);
  const double g1 = sqrt(2.0 * pi) / ang_norm2;
  double *s = (double *)malloc(sizeof(double) * (size_t)(order > 0 ? order : 1));
  for (int o = 0; o < order; ++o) {
    const double tm = t_width * (o + 1);
    s[o] = g1 * exp(-(tm * tm) / 2.0);
  }
  int pmax = aras_pmax(x1, n1, nn1, nm1, amax);
  const int pm2 = aras_pmax(x2, n2, nn2, nm2, amax);
  if (pm2 > pmax) pmax = pm2;
  aras_tables t1, t2;
  aras_build(x1, n1, nn1, nm1, amax, &pr, pmax, &t1);
  if (!sym) aras_build(x2, n2, nn2, nm2, amax, &pr, pmax, &t2);
  const aras_tables *tb2 = sym ? &t1 : &t2;
  double *self1 = (double *)calloc((size_t)nm1 * amax + 1, sizeof(double));
  double *self2 = sym ? self1 : (double *)calloc((size_t)nm2 * amax + 1, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
  for (int a = 0; a < nm1; ++a)
    for (int i = 0; i < n1[a]; ++i)
      self1[a * amax + i] = aras_distl2(x1, &t1, a, i, nn1[a * amax + i], x1, &t1, a, i,
                                        nn1[a * amax + i], &pr, pd, emax, s);
  if (!sym) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int a = 0; a < nm2; ++a)
      for (int i = 0; i < n2[a]; ++i)
        self2[a * amax + i] = aras_distl2(x2, tb2, a, i, nn2[a * amax + i], x2, tb2, a, i,
                                          nn2[a * amax + i], &pr, pd, emax, s);
  }
  double *inv = (double *)malloc(sizeof(double) * (size_t)(nsig > 0 ? nsig : 1));
  for (int k = 0; k < nsig; ++k) inv[k] = -1.0 / (sigmas[k] * sigmas[k]);
#define PD(p, q) pd[((size_t)(q) - 1) * emax + ((p) - 1)]
#pragma omp parallel for schedule(dynamic, 1)
  for (int j = 0; j < nm2; ++j) {
    double *ad = (double *)calloc((size_t)amax * amax + 1, sizeof(double));
    const int nj = n2[j];
    for (int i = sym ? j : 0; i < nm1; ++i) {
      const int ni = n1[i];
      for (int i1 = 0; i1 < ni; ++i1)
        for (int j1 = 0; j1 < nj; ++j1) {
          double l2 = aras_distl2(x1, &t1, i, i1, nn1[i * amax + i1], x2, tb2, j, j1,
                                  nn2[j * amax + j1], &pr, pd, emax, s);
          l2 = self1[i * amax + i1] + self2[j * amax + j1] -
               2.0 * l2 * PD((int)ARAS_X(x1, amax, i, i1, 1, 0), (int)ARAS_X(x2, amax, j, j1, 1, 0));
          if (fabs(l2) < 5.0e-12) l2 = 0.0;
          ad[j1 * ni + i1] = l2;
        }
      for (int k = 0; k < nsig; ++k) {
        double acc = 0.0;
        for (int t = 0; t < ni * nj; ++t) acc = acc + exp(ad[t] * inv[k]);
        out[k + (size_t)nsig * (i + (size_t)nm1 * j)] = acc;
        if (sym) out[k + (size_t)nsig * (j + (size_t)nm1 * i)] = acc;
      }
    }
    free(ad);
  }
#undef PD
  free(inv);
  if (!sym) free(self2);
  free(self1);
  aras_free(&t1);
  if (!sym) aras_free(&t2);
  free(s);
}
