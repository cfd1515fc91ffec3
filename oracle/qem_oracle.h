/*
 * qem_oracle.h -- CPU FP64 restatement of the QEM4Molly quasi-Ewald force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded only by tests/, by __graft_entry__.smoke() (as the comparison side)
 * and by bench.py's cpu_baseline leg.  The product library (libqem_hip.so)
 * never calls it.
 *
 * Each function restates one reference function (cited file:line, paths
 * relative to the QEM4Molly repository root).  The reference is Julia
 * (Julia 1.8, packages pinned in Manifest.toml: CellListMap 0.8.10,
 * SpecialFunctions 2.2.0, GaussQuadrature 0.5.8, StatsBase 0.33.21); Julia is
 * not available, so nothing here is produced by running it.  Third-party
 * numerics are replaced by published equivalents:
 *   - GaussQuadrature.legendre(n)  -> Newton iteration on P_n in long double
 *   - SpecialFunctions.besselj0/1  -> C library j0/j1
 *   - CellListMap.neighborlist     -> O(N^2) minimum-image enumeration
 *   - sortperm                     -> stable sort by z, ties by index
 *   - StatsBase.sample             -> not restated: k indices are inputs
 *
 * Every function is pinned on the CPU against checks that do not reuse its
 * formulas (tests/test_oracle_pins.py, tests/test_oracle_energy_pins.py):
 * explicit image-charge sums (point and in-plane-gaussian-smeared, by 2-D
 * quadrature), the image-series per-k energy and its gradient, O(N^2)
 * direct sums, closed-form K-set enumeration, 40-digit Gauss-Legendre rules
 * and finite differences.  The energy functions (qo_pair_potential,
 * qo_self_potential, qo_energy) have no reference counterpart: they are
 * the functional whose gradient is the corrected-semantics force.
 */
#ifndef QEM_ORACLE_H
#define QEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Semantic switches.  0 = the reference as written (the parity target). */
#define QO_SEM_SWAP_LIMITS   1u  /* gauss integrals on k_f2, point on k_f1   */
#define QO_SEM_MINIMAGE_DIR  2u  /* xy direction from minimum-image dx/rho   */
#define QO_SEM_NEWTON_IMAGES 4u  /* j-side z force uses d/dz_j (images 2,3)  */
#define QO_SEM_POLE_SPLIT    8u  /* gamma_1 gamma_2 > 1: principal-value rule
                                    of src/python_lib/Force_short_norm_njit.py */

typedef struct qo_params {
    double Lx, Ly, Lz;
    double r_cutoff;
    double alpha;
    double gamma_1;   /* lower interface (z = 0)  */
    double gamma_2;   /* upper interface (z = Lz) */
    double eps_0;
    double accuracy;
    int32_t N_t;
    uint32_t semantics;
    double pole_eps;  /* half-width of the middle segment (QO_SEM_POLE_SPLIT) */
} qo_params;

/* greens_element (src/greens_function.jl:4-15) */
typedef struct qo_elem {
    double gamma_1, gamma_2, rho;
    double a[4], b[4], s[4];
    double L_z, alpha;
    double k_f1[4], k_f2[4];
} qo_elem;

/* Gauss-Legendre nodes/weights on [-1,1], ascending nodes. */
int qo_legendre(int n, double *x, double *w);

void qo_elem_ij(double g1, double g2, double zi, double zj, double rho,
                double Lz, double alpha, double acc, qo_elem *e);
void qo_elem_i(double g1, double g2, double zi, double Lz, double alpha,
               double acc, qo_elem *e);
void qo_elem_long(double g1, double g2, double Lz, double alpha, qo_elem *e);

/* l = 0: sum over the four images; l = 1..4: one image (1-based as in Julia) */
double qo_Gamma_1(double k, const qo_elem *e, int l);
double qo_Gamma_2(double k, const qo_elem *e, int l);
double qo_dz_Gamma_1(double k, const qo_elem *e, int l);
double qo_dz_Gamma_2(double k, const qo_elem *e, int l);

/* F_short_ij: force on i from j.  fz_j receives the z force on j
 * (as written: -out[2]; with QO_SEM_NEWTON_IMAGES: from d/dz_j). */
void qo_F_short_ij(const qo_params *p, double qi, double qj, const double *ri,
                   const double *rj, int single, double *out, double *fz_j);
void qo_F_short_i(const qo_params *p, double qi, const double *ri, int single,
                  double *out);

/* Pole splitting for gamma_1 gamma_2 > 1 (src/python_lib/Force_short_norm_njit.py:
 * 11-414 with green_norm_njit.py:13-32): the integrand N(k)/D(k) of one short-range
 * integral, N summed over the images in mask (bit l-1 = image l), integrated over
 * [0, k0-eps], [k0-eps, k0+eps], [k0+eps, 2k0], [2k0, kf] with N1, 10, N1,
 * N_t - 2 N1 Gauss-Legendre nodes, k0 = ln(g1 g2)/(2 Lz), kf = max(kf_base, 2.1 k0),
 * N1 = floor(N_t k0 / kf); N(k0)/(2 Lz (k - k0)) is added on [0, 2k0] and the
 * middle segment uses N(k) (1/D + 1/x) - N'(k0)/(2 Lz).  kind: 0 rho gauss,
 * 1 rho point, 2 z gauss, 3 z point.  kf_base: -ln(acc)/(2 Lz) for the point
 * integrals, sqrt(-4 alpha ln acc) for the gaussian ones.  Returns NaN when
 * the rule does not exist (N1 < 1, N_t - 2 N1 < 1, or eps outside (0, k0)). */
int qo_pole_rule(double g12, double Lz, double kf_base, double eps, int N_t, double *k0,
                 double *kf, int *n1, int *n3);
double qo_pole_integral(int kind, const qo_elem *e, unsigned mask, double kf_base,
                        double eps, int N_t);

/* O(N^2) minimum-image 2-D pair enumeration, i < j, row-major order.
 * Returns the number of pairs; writes at most cap pairs (2 int32 each). */
int64_t qo_neighbor_pairs(int64_t n, const double *xyz, double Lx, double Ly,
                          double rc, int32_t *pairs, int64_t cap);

/* Molly.forces(::QEM_short) with an explicit pair list (src/QEM_short.jl:27-59).
 * nthreads > 1 evaluates pairs with OpenMP; accumulation order is fixed. */
void qo_short_forces(const qo_params *p, int64_t n, const double *xyz,
                     const double *q, int64_t npairs, const int32_t *pairs,
                     double *f, int nthreads);
/* The same forces for a subset of rows (f: nrows x 3), each row's pairs
 * found by an O(N) scan, so that large systems can be checked on a sample
 * of atoms without the O(N^2) pair list. */
void qo_short_forces_rows(const qo_params *p, int64_t n, const double *xyz, const double *q,
                          int64_t nrows, const int64_t *rows, double *f, int nthreads);
/* pair forces only (4 doubles per pair: fx, fy, fz_i, fz_j) for timing */
void qo_short_pair_forces(const qo_params *p, const double *xyz,
                          const double *q, int64_t npairs,
                          const int32_t *pairs, double *fpair, int nthreads);

/* K_set_generator (src/K_set_generator.jl:5-45) */
int64_t qo_kset_size(double Lx, double Ly, double alpha, double acc,
                     int32_t *mx_max, int32_t *my_max);
void qo_kset(double Lx, double Ly, double alpha, double acc, double *kxkyk,
             float *prob, double *sum_K);

/* sortperm of z (src/QEM_long.jl:39-40), 0-based */
void qo_sortperm_z(int64_t n, const double *xyz, int64_t *zlist);

/* Long-range pieces for one k = (kx, ky, k); outputs N x 3 (overwritten) */
void qo_Fl_non_sort(const double *k3, int64_t n, const double *xyz,
                    const double *q, double g1, double g2, double Lz,
                    double *f);
void qo_Fl_sort(const double *k3, int64_t n, const int64_t *zlist,
                const double *xyz, const double *q, double *f);
/* the reference's subtractive suffix sums (src/Force_long.jl:83-84,97-98) */
void qo_Fl_sort_subtractive(const double *k3, int64_t n, const int64_t *zlist,
                            const double *xyz, const double *q, double *f);
void qo_Flz_sort_k0(int64_t n, const int64_t *zlist, const double *q,
                    double *f);
void qo_Flz_self(const double *k3, int64_t n, const double *xyz,
                 const double *q, double g1, double g2, double Lz, double *f);
void qo_F_l_total(const double *k3, int64_t n, const double *xyz,
                  const double *q, const int64_t *zlist, const qo_params *p,
                  double *f);

/* F_long restricted to its random-batch branch (src/Force_long.jl:201-218,
 * 234-242) with the sampled k vectors given explicitly: kvecs is P x 3,
 * weight = sum_K / rbe_p.  full_sum != 0 instead weights each k by
 * exp(-k^2/(4 alpha)) (the expectation of the batch estimator). */
void qo_long_forces(const qo_params *p, int64_t n, const double *xyz,
                    const double *q, const double *kvecs, int64_t P,
                    double weight, int full_sum, double *f, int nthreads);

/* Per-atom batch variance of the random-batch long-range force (no
 * counterpart in the reference): X_ip = weight * F_l_total(k_p),
 * var_i = (P sum_p |X_ip|^2 - |sum_p X_ip|^2) / (P - 1), P >= 2. */
void qo_long_variance(const qo_params *p, int64_t n, const double *xyz, const double *q,
                      const double *kvecs, int64_t P, double weight, double *var,
                      int nthreads);

/* Energy (no counterpart in the reference; see qem_oracle.c for the
 * definition), phi = dE/dq, dE/dgamma_1, dE/dgamma_2. */
void qo_pair_potential(const qo_params *p, const double *ri, const double *rj,
                       double *u, double *du);
void qo_self_potential(const qo_params *p, const double *ri, double *s, double *ds);
void qo_energy(const qo_params *p, int64_t n, const double *xyz, const double *q,
               int64_t npairs, const int32_t *pairs, const double *kvecs, int64_t P,
               double weight, int full_sum, int with_short, int with_long,
               double *E, double *phi, double *dEdg, int nthreads);

/* Substrate_LJ wall (src/Substrate_LJ.jl:7-46): writes (accumulate == 0)
 * or adds the z forces of one wall at z = sub_pos.  Returns 0, or 1 (and
 * writes nothing) when direction * (z_i - sub_pos) <= 0 for some atom (the
 * reference's @assert, src/Substrate_LJ.jl:37).  sem bit 1: cutoff on |dz|;
 * bit 2: the true force -dU/dz instead of the as-written -dU/dz / |dz|. */
int qo_substrate_lj(double cutoff, double sigma, double eps, double sub_pos, int direction,
                    unsigned sem, int64_t n, const double *xyz, double *f, int accumulate);

#ifdef __cplusplus
}
#endif
#endif
