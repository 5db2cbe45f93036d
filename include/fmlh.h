/*
 * fmlh.h -- C ABI of libfmlhip.so, the MI355X (gfx950) implementation of the
 * fml kernel-ridge-regression hot path.
 *
 * Conventions
 *   - Every pointer argument is DEVICE memory unless named *_host.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Every call is asynchronous on that stream unless documented otherwise.
 *   - Return value: 0 on success; -k when argument k (1-based) is invalid;
 *     FMLH_EHIP when a HIP runtime call failed (see fmlh_last_error()).
 *   - Matrices are described by (pointer, leading dimension in elements).
 *     "Sample-major" descriptor sets hold one sample per row of `ld`
 *     elements, which is the memory layout of the reference's Fortran
 *     (d, n) column-per-sample arrays.
 *   - Dense kernel outputs are Fortran ordered like the reference's:
 *     K(j, i) for sample j of A and sample i of B lives at k[j + i*ldk].
 *   - No entry point allocates device memory.  Scratch space is passed in by
 *     the caller (`ws`, `ws_bytes`) and sized by the matching *_ws_bytes().
 */
#ifndef FMLH_H
#define FMLH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMLH_EHIP (-1000)

/* ABI version (bumped on any signature change or added entry point). */
int fmlh_abi_version(void);
/* Text of the last HIP error seen by this thread ("" if none). */
const char *fmlh_last_error(void);
/* Number of compute units of the current device (0 on error). */
int fmlh_device_cu_count(void);

/* Instrumentation: while enabled, every entry point records a pair of HIP
 * events on its stream around its dominant kernel (the ARAD pair kernel,
 * the Gaussian / Laplacian tile kernel, the whole factorisation for
 * potrf).  fmlh_timing_collect waits for the recorded events and returns
 * the summed milliseconds and the number of pairs since the last collect. */
void fmlh_timing_enable(int on);
int fmlh_timing_collect(double *total_ms, int *count);

/* One-time per-device set-up (the Coulomb diagonal table, the address of
 * fmlh_potrs_noalloc's scratch array).  Entry points call it themselves;
 * call it before capturing a graph on a device the library has not used
 * yet, since the set-up copies synchronously. */
int fmlh_init(void);

/* A stream (returned as hipStream_t in *stream) whose kernels run only on
 * the first cu_count compute units of the current device (all of them when
 * cu_count <= 0 or >= the device's count), with a hardware queue of its
 * own.  No reference counterpart: the serving loop runs fits on such a
 * stream so that predictions always find free CUs.  Default priority.
 * Destroy a stream only when nothing can still record work or events on it:
 * a stream handed to torch (ExternalStream) is recorded on by torch's
 * pinned-memory allocator whenever a pinned block used on it is freed, so
 * the serving loop keeps its streams for the life of the process.
 * fmlh_potrf called on such a stream runs without its panel look-ahead
 * (which would fork work onto a stream of the library's own, outside the
 * mask and on a further hardware queue). */
int fmlh_stream_create(int cu_count, void **stream);
int fmlh_stream_destroy(void *stream);

/* HIP graphs (no reference counterpart).  fmlh_graph_begin starts capturing
 * the calling thread's work on `stream`; every entry point called on that
 * stream until fmlh_graph_end is recorded instead of run (Coulomb matrices,
 * dense kernels, the ARAD kernel modes, potrf/potrs and the helpers are all
 * capture-safe; fmlh_arad_autotune is not).  fmlh_graph_end instantiates
 * the recording; fmlh_graph_launch replays it on any stream, reading the
 * same device buffers the captured calls were given (their contents may
 * change between replays, their sizes may not). */
typedef struct fmlh_graph fmlh_graph;
int fmlh_graph_begin(void *stream);
int fmlh_graph_end(void *stream, fmlh_graph **graph);
int fmlh_graph_launch(fmlh_graph *graph, void *stream);
int fmlh_graph_destroy(fmlh_graph *graph);

/* ------------------------------------------------------------------------
 * (a) Sorted Coulomb matrices, one workgroup per molecule.
 * Replaces fgenerate_coulomb_matrix
 *   (fml/representations/frepresentations.f90:23-110; Python entry
 *    Molecule.generate_coulomb_matrix, fml/fml.py:64-66).
 * z:       concatenated nuclear charges (total atoms) f64
 * r:       concatenated coordinates (total atoms, 3) f64 row-major
 * offsets: nmol+1 int32 prefix offsets into z / r
 * cm:      (nmol, nmax*(nmax+1)/2) f64 row-major; each row is the packed
 *          row-major lower triangle of the row-norm-sorted matrix
 * cm32:    optional (may be NULL) float copy of cm, same layout
 * perm:    optional (may be NULL) (nmol, nmax) int32: sorted position ->
 *          original atom index (0-based), -1 padded
 * status:  optional (may be NULL) (nmol) int32: 0 ok, 1 natoms > nmax
 *          (that molecule's row is left zero; the reference writes out of
 *          bounds there, frepresentations.f90:96-103)
 * ---------------------------------------------------------------------- */
int fmlh_coulomb_sorted(const double *z, const double *r, const int32_t *offsets,
                        int nmol, int nmax, double *cm, float *cm32,
                        int32_t *perm, int32_t *status, void *stream);

/* Unsorted, local and atomic Coulomb matrices:
 *   mode 0: fgenerate_unsorted_coulomb_matrix (frepresentations.f90:112-201;
 *           Molecule.generate_unsorted_coulomb_matrix, fml/fml.py:68-70):
 *           cm (nmol, nmax(nmax+1)/2), the packed lower triangle in input
 *           atom order;
 *   mode 1: fgenerate_local_coulomb_matrix (:204-294; fml/fml.py:72-74);
 *   mode 2: fgenerate_atomic_coulomb_matrix (:297-390; fml/fml.py:76-78):
 *           cm (offsets[nmol], nmax(nmax+1)/2), one row per atom k of each
 *           molecule (row offsets[m] + k): atom k first, then the others by
 *           descending row norm (mode 1), or all atoms by ascending distance
 *           from k (mode 2).
 * z, r, offsets as for fmlh_coulomb_sorted; status as there (a molecule
 * with more than nmax atoms gets zero rows; the reference writes out of
 * bounds).  Bitwise equal to the reference for finite coordinates. */
int fmlh_coulomb_form(int mode, const double *z, const double *r,
                      const int32_t *offsets, int nmol, int nmax, double *cm,
                      int32_t *status, void *stream);

/* ------------------------------------------------------------------------
 * (b) Gaussian kernel K(j,i) = exp(-0.5/sigma^2 * ||a_j - b_i||^2)
 * Replaces fgaussian_kernel (fml/kernels/fkernels.f90:317-349; Python entry
 * gaussian_kernel, fml/kernels/kernels.py:167-196).
 * FP64: row norms + FP64 MFMA GEMM + fused exp epilogue; pairs whose
 * GEMM-form distance is within the cancellation band are recomputed by
 * direct differences.
 * ---------------------------------------------------------------------- */
int fmlh_gaussian_kernel(const double *a, int64_t lda, int na,
                         const double *b, int64_t ldb, int nb, int d,
                         double sigma, double *k, int64_t ldk, void *stream);

/* Precision forms on float descriptors.  mode: 0 = FP32 MFMA products with
 * FP32 accumulation; 1 = split-BF16x3 MFMA (hi*hi + hi*lo + lo*hi) with FP32
 * accumulation, on hi/lo bf16 planes split once per operand into ws (no
 * reference counterpart; fmlh_gaussian_kernel_f32_ws_bytes(na, nb, d, mode,
 * same) bytes, same = 1 when a and b are the same rows; 0 bytes for mode 0).
 * Both recompute near pairs (d2 <= tau (|a|^2 + |b|^2)) by direct FP32
 * differences; for mode 1, recomputed (device counter, may be NULL) is
 * incremented by their number.  Output is float (k32) or double (k64);
 * exactly one must be non-NULL. */
size_t fmlh_gaussian_kernel_f32_ws_bytes(int na, int nb, int d, int mode, int same);
int fmlh_gaussian_kernel_f32(const float *a, int64_t lda, int na,
                             const float *b, int64_t ldb, int nb, int d,
                             double sigma, int mode, float *k32, double *k64,
                             int64_t ldk, void *ws, size_t ws_bytes,
                             unsigned long long *recomputed, void *stream);

/* ------------------------------------------------------------------------
 * (c) Laplacian kernel K(j,i) = exp(-1/sigma * sum_t |a_jt - b_it|)
 * Replaces flaplacian_kernel (fml/kernels/fkernels.f90:351-377; Python entry
 * laplacian_kernel, fml/kernels/kernels.py:136-164).  LDS-tiled VALU L1
 * distance; each entry's sum runs over t in ascending order exactly as the
 * reference's sum() does.
 * ---------------------------------------------------------------------- */
int fmlh_laplacian_kernel(const double *a, int64_t lda, int na,
                          const double *b, int64_t ldb, int nb, int d,
                          double sigma, double *k, int64_t ldk, void *stream);
int fmlh_laplacian_kernel_f32(const float *a, int64_t lda, int na,
                              const float *b, int64_t ldb, int nb, int d,
                              double sigma, float *k32, double *k64,
                              int64_t ldk, void *stream);

/* ------------------------------------------------------------------------
 * (d) ARAD kernels.
 * Molecular descriptor sets: q (nm, amax, 5, amax) f64 C order, i.e. atom
 * (i, a) owns a contiguous (5, amax) channel-major block; zt (nm, amax, 2)
 * f64 PTP (row, col) types; n (nm) int32 atom counts.
 * mode selects the reference routine reproduced:
 *   FMLH_ARAD_GLOBAL     fget_kernels_arad           farad_kernels.f90:439-616
 *                        out (nsig, nm1, nm2) Fortran order
 *   FMLH_ARAD_SYMMETRIC  fget_symmetric_kernels_arad farad_kernels.f90:619-755
 *                        (set 2 ignored) out (nsig, nm1, nm1) Fortran order
 *   FMLH_ARAD_KERNEL     fget_kernel_arad            farad_kernels.f90:95-227
 *                        (set 2 ignored; one sigma; width .2, cut 6, rw 1,
 *                        cw .5 forced) out (nm1, nm1) Fortran order
 *   FMLH_ARAD_DISTANCE   atomic_arad_l2_distance_all farad.f90:452-600
 *                        out (nm1, nm2, amax, amax) Fortran order
 *                        (sigmas unused)
 * Python entries: get_atomic_kernels_arad / get_atomic_symmetric_kernels_arad
 * (fml/kernels/kernels.py:229-349), get_l2_distance_arad
 * (fml/math/distance.py:120-189).
 * ---------------------------------------------------------------------- */
#define FMLH_ARAD_GLOBAL 0
#define FMLH_ARAD_SYMMETRIC 1
#define FMLH_ARAD_KERNEL 2
#define FMLH_ARAD_DISTANCE 3

/* ------------------------------------------------------------------------
 * Atomic KRR path (molecule sums over atom-pair kernels)
 *
 * Atoms are rows: x (na, d) f64 row-major with leading dimension ld; off
 * (nm+1) int32 prefix offsets, molecule m owns rows off[m]..off[m+1]-1
 * (off[0] = 0, off[nm] = na); amax (1..64) bounds every molecule's atom
 * count.  Molecules are processed in blocks of at most 128 atoms: blk
 * (optional, device, nb + 1 int32 molecule indices, blk[0] = 0,
 * blk[nb] = nm, non-decreasing) lists the caller's blocks (e.g. molecules
 * packed greedily); NULL takes the molecules whose first atom lies in
 * [b S, (b+1) S), S = 129 - amax.  status (may be NULL) (1) int32: set to 1
 * when a block spans more than 128 atoms (a molecule larger than amax, or a
 * bad block list; the block is skipped).  Sums have a fixed order: a
 * molecule pair's value does not depend on the blocks or on the rest of
 * the batch.
 * ---------------------------------------------------------------------- */

/* fget_vector_kernels_gaussian (kinds 0 and 2, fkernels.f90:93-162) and
 * fget_vector_kernels_laplacian (kind 1, :23-91); Python entries
 * get_atomic_kernels_gaussian / get_atomic_kernels_laplacian
 * (fml/kernels/kernels.py:351-362, 380-406).
 * K_s(i, j) = sum over atoms a of molecule i (set 1), b of molecule j
 * (set 2) of exp(inv_s dist(a, b)); dist = sum_t (x1_a,t - x2_b,t)^2 with
 * inv = -0.5/sigma^2 (kind 0: summed over t in the reference's order, the
 * reference's distances bit for bit; kind 2: |a|^2 + |b|^2 - 2 a.b with the
 * products on the matrix cores, faster, distances to about 1e-13 of
 * |a|^2 + |b|^2), or dist = sum_t |x1_a,t - x2_b,t| with inv = -1/sigma
 * (kind 1, in the reference's order).  sigmas: host array of nsig (1..8) values.
 * out (nm2, nm1, nsig) C order = the reference's Fortran (nsig, nm1, nm2). */
int fmlh_vector_kernels(int kind, const double *x1, int64_t ldx1,
                        const int32_t *off1, int nm1, int na1,
                        const double *x2, int64_t ldx2, const int32_t *off2,
                        int nm2, int na2, int d, int amax,
                        const double *sigmas, int nsig, const int32_t *blk1,
                        int nb1, const int32_t *blk2, int nb2, double *out,
                        int32_t *status, void *stream);

/* The molecular kernel of fget_alpha_from_distance (fkernels.f90:541-599;
 * Python entry get_alpha_from_distance, fml/kernels/kernels.py:127-134):
 * dmat: the reference's Fortran (na, na) distance array d(k, a), i.e.
 * row-major rows a with leading dimension ldd;
 * k(i, j) = sum over atoms a of i, k of j of exp(-d(k, a) / gamma), written
 * Fortran (nm, nm): k[i + nm j].  (The same sums for d = the L1 distances
 * of atom rows give fget_alpha's kernel, fkernels.f90:420-435, which is
 * fmlh_vector_kernels kind 1 with set 2 = set 1.) */
int fmlh_distance_mol_kernel(const double *dmat, int64_t ldd,
                             const int32_t *off, int nm, int na, int amax,
                             double gamma, const int32_t *blk, int nb,
                             double *k, int32_t *status, void *stream);

/* fget_prediction (fkernels.f90:480-539; get_prediction,
 * fml/kernels/kernels.py:101-108): test atoms xt (na_t, d) in molecules
 * off_t, training atoms x (na, d), alpha (na):
 * y(m) = sum over test atoms a of m, training atoms b of
 * exp(-sum_t |x_b,t - xt_a,t| / sigma) alpha(b).
 * ws: fmlh_atomic_predict_ws_bytes(nm_t, na) bytes. */
size_t fmlh_atomic_predict_ws_bytes(int nm_t, int na);
int fmlh_atomic_predict(const double *xt, int64_t ldt, const int32_t *off_t,
                        int nm_t, int na_t, const double *x, int64_t ldx,
                        int na, int d, int amax, const double *alpha,
                        double sigma, const int32_t *blk_t, int nb_t,
                        double *y, int32_t *status, void *ws, size_t ws_bytes,
                        void *stream);

/* fget_prediction_from_distance (fkernels.f90:645-702;
 * get_prediction_from_distance, fml/kernels/kernels.py:110-117): d2 the
 * reference's Fortran (na, na_t) array D2(b, a), i.e. row-major rows a
 * (test atoms) with leading dimension ldd;
 * y(m) = sum over test atoms a of m, training atoms b of
 * exp(-D2(b, a) / gamma) alpha(b).  The reference accumulates into an
 * uninitialised array (fkernels.f90:680-686); here the sums start at 0.
 * ws: fmlh_atomic_predict_ws_bytes(nm_t, na) bytes. */
int fmlh_distance_predict(const double *d2, int64_t ldd, const int32_t *off_t,
                          int nm_t, int na_t, int na, int amax,
                          const double *alpha, double gamma,
                          const int32_t *blk_t, int nb_t, double *y,
                          int32_t *status, void *ws, size_t ws_bytes,
                          void *stream);

/* ARAD descriptors on the device: ARAD.describe (fml/arad.py:70-145, the
 * non-periodic form; Python entry Molecule.generate_arad_descriptor,
 * fml/fml.py:80-86), one wavefront per molecule.
 * z, r, offsets: molecules as for fmlh_coulomb_sorted (at most 64 atoms
 * each); ptp: (ptp_rows, 2) f64 PTP (row, column) per nuclear charge, a row
 * with a non-positive period marks a charge without an entry.
 * Outputs, in the layouts the kernel modes read: q (nmol, amax, 5, amax),
 * zt (nmol, amax, 2), n (nmol) atom counts; status (may be NULL) (nmol):
 * 1 where a molecule has more than amax atoms or a charge that is not an
 * integer with a PTP entry (its block is all padding and n = 0).
 * Distances, types and the neighbour order equal the host computation's
 * bit for bit; the angular channels differ only by the device sin/cos/acos
 * rounding. */
int fmlh_arad_describe(const double *z, const double *r, const int32_t *offsets,
                       int nmol, int amax, double cut, const double *ptp,
                       int ptp_rows, double *q, double *zt, int32_t *n,
                       int32_t *status, void *stream);

/* ------------------------------------------------------------------------
 * ARAS kernels: fget_kernels_aras (mode 0, fml/kernels/faras_kernels.f90:
 * 150-483) and fget_symmetric_kernels_aras (mode 1, :486-723; set 2 = set 1,
 * each pair i >= j computed with molecule i first and written to both
 * halves); Python entries get_atomic_kernels_aras /
 * get_atomic_symmetric_kernels_aras (fml/kernels/kernels.py:408-560).
 * x1, x2: ARAS descriptors (nm, amax, 3 + amax, amax) f64 as ARAS.describe
 * makes them (fml/aras.py:70-131); n1, n2 (nm) int32 atom counts; nn1, nn2
 * (optional, (nm, amax) int32): neighbour counts per atom, self included
 * (NULL: the number of distances < cut, as the reference's wrapper counts
 * them, kernels.py:460-472).  sigmas: host array of nsig (1..8) values;
 * pd: device Fortran (emax, emax) element similarity (gen_pd,
 * kernels.py:68-98, built with r_width and c_width); order 1..8; amax <= 64.
 * out (nm2, nm1, nsig) C order = the reference's Fortran (nsig, nm1, nm2).
 * blk1 (optional, device int32, nblk1 + 1 ascending positions from 0 to
 * nm1) and ord1 (optional, device int32 permutation of the nm1 set-1
 * molecules, NULL: the identity): block b holds molecules ord1[blk1[b]] ..
 * ord1[blk1[b + 1] - 1], at most 64 atoms, one workgroup row per block
 * (blk1 NULL: blocks of 64 / amax consecutive positions).  Blocks of
 * molecules with the same elements and similar sizes run fastest.
 * status (may be NULL) (1) int32: 1 when a charge is not an integer in
 * 1..emax (or above 127), the two sets hold more than 16 distinct elements,
 * a block more than 64 atoms, or a molecule more than amax atoms.
 * ws: fmlh_aras_ws_bytes(nm1, nm2, amax, order, mode).
 * Every expression of the atom distance in the reference's order; entries
 * differ from the reference by the device exp/sin/cos rounding and the
 * order of the final sum over atom pairs (a fixed order: values do not
 * depend on the batch or the blocks).
 * ---------------------------------------------------------------------- */
size_t fmlh_aras_ws_bytes(int nm1, int nm2, int amax, int order, int mode);
int fmlh_aras_kernels(int mode, const double *x1, const int32_t *n1,
                      const int32_t *nn1, int nm1, const double *x2,
                      const int32_t *n2, const int32_t *nn2, int nm2, int amax,
                      const double *sigmas, int nsig, double t_width,
                      double d_width, double cut, int order, const double *pd,
                      int emax, double angular_scale, const int32_t *blk1,
                      const int32_t *ord1, int nblk1, double *out,
                      int32_t *status, void *ws, size_t ws_bytes, void *stream);

/* precision: 0 = FP64 throughout (the plain form); 1 = FP32 inner overlap
 * (exp, products and accumulation in float), FP64 self terms, distance
 * combination and kernel sums.
 *
 * The kernel modes (GLOBAL, SYMMETRIC, KERNEL; amax <= 29) run over
 * size-sorted molecule-pair tiles: one set-1 molecule i against J set-2
 * molecules of one atom count s.  The tile list is planned on the device
 * from the device atom counts, so a call never waits on the host and can
 * be captured into a HIP graph (fmlh_graph_begin / fmlh_graph_end).
 * tiling: tile shape and launch shape, or NULL for the library's current
 * choice for (mode, amax, precision, nsig): an autotuned entry when
 * fmlh_arad_autotune / fmlh_arad_tiling_set provided one, else the
 * built-in default {256, 0, 16, 0}. */
typedef struct fmlh_arad_tiling {
  int pairs_per_tile; /* target atom pairs J * n_i * s per tile, 64..1024 */
  int waves_per_simd; /* register budget of the tile kernel: 4 or 5 waves
                         per SIMD; 0 = 5 where five workgroups fit a CU's
                         LDS, else 4 */
  int groups_per_cu;  /* persistent workgroups launched per CU, 1..64 */
  int pair_order;     /* order of a tile's atom pairs over its threads:
                         0 = set-1 atom fastest, 1 = set-2 atom fastest
                         (a wave's lanes share one set-1 atom); results are
                         bitwise the same for every tiling */
} fmlh_arad_tiling;

size_t fmlh_arad_ws_bytes(int nm1, int nm2, int amax, int mode);
int fmlh_arad_kernels(int mode, const double *q1, const double *zt1,
                      const int32_t *n1, int nm1, const double *q2,
                      const double *zt2, const int32_t *n2, int nm2, int amax,
                      const double *sigmas_host, int nsig, double width,
                      double cut, double r_width, double c_width,
                      int precision, double *out, void *ws, size_t ws_bytes,
                      const fmlh_arad_tiling *tiling, void *stream);

/* Autotuner (no reference counterpart): runs fmlh_arad_kernels on the given
 * inputs once per candidate tiling (pairs_per_tile in {128, 256, 384, 512,
 * 768, 1024} x waves_per_simd in {4, 5} where five fit x pair_order in
 * {0, 1}), `reps` timed calls
 * each after one warm-up, timed with HIP events on `stream`; stores the
 * fastest as the library's choice for (mode, amax, precision, nsig) and
 * returns it with its mean time per call.  Synchronises the stream; out is
 * left holding a valid result.  Not capturable. */
int fmlh_arad_autotune(int mode, const double *q1, const double *zt1,
                       const int32_t *n1, int nm1, const double *q2,
                       const double *zt2, const int32_t *n2, int nm2, int amax,
                       const double *sigmas_host, int nsig, double width,
                       double cut, double r_width, double c_width,
                       int precision, double *out, void *ws, size_t ws_bytes,
                       int reps, fmlh_arad_tiling *best, double *best_ms,
                       void *stream);
/* The library's current tiling for (mode, amax, precision, nsig) (default
 * or tuned), and a setter for loading persisted tuning results. */
int fmlh_arad_tiling_get(int mode, int amax, int precision, int nsig,
                         fmlh_arad_tiling *out);
int fmlh_arad_tiling_set(int mode, int amax, int precision, int nsig,
                         const fmlh_arad_tiling *t);

/* Atomic (per-atom) forms.  q (na, 5, amax), zt (na, 2), n (na) int32 =
 * atom count of the molecule each atom belongs to.
 *   FMLH_ARAD_ATOMIC_KERNELS  fget_atomic_kernels_arad
 *                             farad_kernels.f90:757-899
 *                             out (nsig, na1, na2) Fortran order
 *   FMLH_ARAD_ATOMIC_DISTANCE fget_atomic_distance_arad
 *                             farad_kernels.f90:902-1030, written as the
 *                             well-defined (na1, na2) Fortran-order D(i,j)
 *                             (the reference stores kernel(j,i) into an
 *                             (na1,na2) array, farad_kernels.f90:932,1019) */
#define FMLH_ARAD_ATOMIC_KERNELS 4
#define FMLH_ARAD_ATOMIC_DISTANCE 5
size_t fmlh_arad_atomic_ws_bytes(int na1, int na2, int amax);
int fmlh_arad_atomic(int mode, const double *q1, const double *zt1,
                     const int32_t *n1, int na1, const double *q2,
                     const double *zt2, const int32_t *n2, int na2, int amax,
                     const double *sigmas_host, int nsig, double width,
                     double cut, double r_width, double c_width, double *out,
                     void *ws, size_t ws_bytes, void *stream);

/* ------------------------------------------------------------------------
 * (e) Cholesky factorisation and solve of a symmetric positive definite
 * matrix held in Fortran order, reading and writing only its UPPER
 * triangle: A = U^T U, as LAPACK dpotrf('U') / dpotrs('U').
 * Replaces fcho_solve (fml/math/fcho_solve.f90:23-47 -> dpotrf line 35,
 * dpotrs line 42; Python entry cho_solve, fml/math/cho_solve.py:33-58).
 * a: (n, n) Fortran order, lda >= n; overwritten by U in its upper triangle
 * (the strictly lower part is not touched).
 * info (device int32): 0, or k > 0 when the leading minor of order k is not
 * positive definite (as LAPACK).  The factor is undefined after a failure.
 * ---------------------------------------------------------------------- */
size_t fmlh_potrf_ws_bytes(int n);
int fmlh_potrf(double *a, int64_t lda, int n, int32_t *info, void *ws,
               size_t ws_bytes, void *stream);
/* Precision form of fmlh_potrf (no reference counterpart): the same
 * factorisation, the panels in FP64, with the next-panel and trailing
 * updates computed on INT8 matrix cores from an exact split of each
 * finished 1024-column panel: every row scaled by a power of two, truncated
 * to a 40-bit integer and written in five signed base-256 digits (the
 * error-free splitting of Ozaki, Ogita, Oishi and Rump, with integer slices
 * as Mukunoki et al. use on tensor cores), the 15 digit products with
 * t + u <= 6 summed exactly in int32 by v_mfma_i32_32x32x32_i8 and combined
 * in FP64.  An update's error is below about 2^-38 |row| |column| summed
 * over the panel (dropped digit products and the truncation; zero-mean),
 * so the factor serves as the solver of FP64 iterative refinement
 * (KRR(precision="ozaki")), not as a stand-in for the FP64 factor.
 * ws: fmlh_potrf_ozaki_ws_bytes(n) bytes (slice planes and row exponents). */
size_t fmlh_potrf_ozaki_ws_bytes(int n);
int fmlh_potrf_ozaki(double *a, int64_t lda, int n, int32_t *info, void *ws,
                     size_t ws_bytes, void *stream);
/* Solve U^T U x = b in place for nrhs right-hand sides, x (n, nrhs) Fortran
 * order with leading dimension ldx. */
size_t fmlh_potrs_ws_bytes(int n, int nrhs);
int fmlh_potrs(const double *a, int64_t lda, int n, double *x, int64_t ldx,
               int nrhs, void *ws, size_t ws_bytes, void *stream);

/* The correction solve of iterative refinement (no reference counterpart):
 * U^T U x = b in place for one right-hand side x (n), as fmlh_potrs with
 * nrhs = 1, except that every full 64 x 64 diagonal block is applied as its
 * inverse, formed while the block waits for the blocks before it, instead
 * of by 64 dependent substitution steps (the solve is bound by that chain:
 * 36,000 rows in about half the time).  The rounding of a block's values
 * then grows with its diagonal block's condition number instead of meeting
 * a triangular solve's backward-error bound, which a refinement correction
 * (accurate to a few digits is enough) can afford and fmlh_potrs keeps.
 * ws: fmlh_potrs_ws_bytes(n, 1) bytes. */
int fmlh_potrs_correction(const double *a, int64_t lda, int n, double *x, void *ws,
                          size_t ws_bytes, void *stream);

/* fmlh_potrs without a workspace argument: no reference counterpart.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same kernels,
 * with their scratch (ready flags and tickets of the one-launch solve, or
 * the blocked solve's partial products; potrs_ws(n, nrhs) bytes) in a
 * file-scope device array of fmlh_potrs_noalloc_max_bytes() = 1 MiB that
 * each device holds until the process ends.  A call whose scratch would
 * exceed it returns -3 (n) with the reason in fmlh_last_error(); nothing is
 * allocated or freed.  Every call on a device shares that device's array,
 * which keeps the last call's contents (the solve clears what must start at
 * zero itself): calls that overlap on one device -- two streams, or two
 * host threads -- race and must not be issued. */
size_t fmlh_potrs_noalloc_max_bytes(void);
int fmlh_potrs_noalloc(const double *a, int64_t lda, int n, double *x, int64_t ldx,
                       int nrhs, void *stream);

/* Inverse from the factor, as LAPACK dpotri('U'): a holds U from fmlh_potrf
 * (upper triangle, Fortran order) and receives the upper triangle of
 * inv(A) = inv(U) inv(U)^T; the strictly lower part is not touched.
 * info (device int32): 0, or k > 0 when U(k,k) is zero.  Replaces the
 * dpotri call of fcho_invert (fml/math/fcho_solve.f90:49-68, Python entry
 * cho_invert, fml/math/cho_solve.py:27-30) and of fget_alpha_arad
 * (fml/kernels/farad_kernels.f90:414).  ws: fmlh_potri_ws_bytes(n) bytes
 * (an n x n scratch matrix). */
size_t fmlh_potri_ws_bytes(int n);
int fmlh_potri(double *a, int64_t lda, int n, int32_t *info, void *ws,
               size_t ws_bytes, void *stream);

/* Single-precision factor for the mixed-precision form (same layout and
 * semantics on float data). */
size_t fmlh_spotrf_ws_bytes(int n);
int fmlh_spotrf(float *a, int64_t lda, int n, int32_t *info, void *ws,
                size_t ws_bytes, void *stream);
size_t fmlh_spotrs_ws_bytes(int n, int nrhs);
int fmlh_spotrs(const float *a, int64_t lda, int n, float *x, int64_t ldx,
                int nrhs, void *ws, size_t ws_bytes, void *stream);

/* ------------------------------------------------------------------------
 * Helpers for the KRR composition (no single reference counterpart; the
 * README's user pattern, README.md:33-69).
 * ---------------------------------------------------------------------- */
/* a(i,i) += lambda for i < n (Fortran order, lda). */
int fmlh_add_diagonal(double *a, int64_t lda, int n, double lambda,
                      void *stream);
/* r = y - (A + lambda I) x using only A's upper triangle (symmetric), all
 * in FP64.  A is the UNfactored matrix (n, n) Fortran order. */
int fmlh_sym_residual(const double *a, int64_t lda, int n, double lambda,
                      const double *x, const double *y, double *r,
                      void *stream);
/* r = y - (K + lambda I) x for the full symmetric K (n, n) Fortran order,
 * each r(i) from error-free products and a double-length running sum
 * (Ogita-Rump-Oishi Dot2), rounded once: the residual of iterative
 * refinement in extra precision, so the refined x converges to the solution
 * of the FP64 system to working accuracy instead of to its condition
 * number times the unit roundoff.  r must not alias x or y. */
int fmlh_residual_dd(const double *k, int64_t ldk, int n, double lambda,
                     const double *x, const double *y, double *r, void *stream);
/* fmlh_residual_dd for a matrix that shares its storage with its own
 * Cholesky factor: K's strictly lower triangle in a (n, n) Fortran order
 * (K(i,j), i > j, at a[i + j lda]) and its diagonal in diag (n); a's upper
 * triangle, where fmlh_potrf / fmlh_potrf_ozaki leave the factor of
 * K + lambda I, is not read.  Each tile of the triangle is read once and
 * serves both its rows' and its columns' sums, so the residual reads half
 * the bytes of the full-matrix form, and the caller keeps one n x n matrix
 * instead of K and a factored copy.  Same accuracy as fmlh_residual_dd
 * (error-free products, double-length sums, one rounding; the order of
 * the sums differs).  r must not alias x, y or diag.
 * ws: fmlh_residual_dd_lower_ws_bytes(n) bytes (double-length partial
 * sums, 16 ceil(n/64) n bytes). */
size_t fmlh_residual_dd_lower_ws_bytes(int n);
int fmlh_residual_dd_lower(const double *a, int64_t lda, int n, const double *diag,
                           double lambda, const double *x, const double *y, double *r,
                           void *ws, size_t ws_bytes, void *stream);
/* y(i) = sum_j K(j, i) * alpha(j) for K (na, nb) Fortran order: the
 * prediction K_test^T alpha. */
int fmlh_gemv_t(const double *k, int64_t ldk, int na, int nb,
                const double *alpha, double *y, void *stream);

/* KRR prediction of a molecule batch against a resident Gaussian model in
 * one call (the serving loop's batch; no single reference counterpart: the
 * README's predict pattern, README.md:33-69): sorted Coulomb matrices of the
 * batch (as fmlh_coulomb_sorted, z / r / offsets / status alike), the
 * Gaussian kernel block K(j, i) between training row j of xtrain (ntrain
 * rows of nmax(nmax+1)/2 features, leading dimension ldx) and batch
 * molecule i, and y(i) = sum_j K(j, i) alpha(j).  Every y(i) depends only on
 * molecule i (fixed per-entry and per-output reduction orders).
 * ws: fmlh_krr_predict_ws_bytes(nmol, ntrain, nmax) bytes. */
size_t fmlh_krr_predict_ws_bytes(int nmol, int ntrain, int nmax);
int fmlh_krr_predict(const double *z, const double *r, const int32_t *offsets,
                     int nmol, int nmax, const double *xtrain, int64_t ldx,
                     int ntrain, const double *alpha, double sigma, double *y,
                     int32_t *status, void *ws, size_t ws_bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FMLH_H */
