/*
 * qem_oracle.c -- CPU FP64 restatement of the QEM4Molly quasi-Ewald force path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 *
 * Reference files restated here (QEM4Molly repository, Julia):
 *   src/Gaussian_int.jl        Guass_int / Gauss_Legendre
 *   src/greens_function.jl     greens_element*, Gamma_1/2, dz_Gamma_1/2
 *   src/Force_short.jl         integrands, F_short_ij, F_short_i
 *   src/QEM_short.jl           Molly.forces(::QEM_short) accumulation
 *   src/K_set_generator.jl     K_set_generator
 *   src/Force_long.jl          Fl_non_sort, Fl_sort, Flz_sort_k0, Flz_self,
 *                              F_l_total, F_long (random-batch branch)
 *   src/QEM_long.jl            sortperm of z
 * Operation order follows the Julia expressions where it is cheap to do so;
 * the only intentional numerical departure is documented at qo_Fl_sort.
 *
 * Build: see oracle/Makefile (gcc -O3 -ffp-contract=off -fopenmp).
 */
#define _DEFAULT_SOURCE 1 /* j0/j1 are XSI, not ISO C */
#include "qem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define QO_TWO_PI 6.283185307179586 /* Float64(2pi), as Julia's 2π */

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre nodes: stands in for GaussQuadrature.legendre(n)
 * (GaussQuadrature 0.5.8, used at src/Gaussian_int.jl:11).  Newton iteration
 * on the three-term recurrence of P_n in long double; nodes ascending. */
int qo_legendre(int n, double *x, double *w)
{
    if (n < 1) return -1;
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int i = 0; i < n; ++i) {
        long double z = cosl(pi * ((long double)i + 0.75L) / ((long double)n + 0.5L));
        long double dp = 0.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (int k = 2; k <= n; ++k) {
                long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / (long double)k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p0 = 1.0L; p1 = z; }
            dp = (long double)n * (z * p1 - p0) / (z * z - 1.0L);
            long double dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        /* recompute derivative at the converged root */
        {
            long double p0 = 1.0L, p1 = z;
            for (int k = 2; k <= n; ++k) {
                long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / (long double)k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) p0 = 1.0L;
            dp = (long double)n * (z * p1 - p0) / (z * z - 1.0L);
        }
        x[n - 1 - i] = (double)z;
        w[n - 1 - i] = (double)(2.0L / ((1.0L - z * z) * dp * dp));
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* greens_element_ij_init -- src/greens_function.jl:17-30 */
static double jl_sign(double v) { return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0); }

static void elem_limits(qo_elem *e, double acc)
{
    /* k_f1 = -log(accuracy) ./ (2L_z .+ a)                      (line 26)
     * k_f2 = sqrt.(4α² a² - 4α log(acc)) .- 2α a                 (line 27) */
    double la = log(acc);
    for (int l = 0; l < 4; ++l) {
        e->k_f1[l] = -la / (2.0 * e->L_z + e->a[l]);
        e->k_f2[l] = sqrt((4.0 * (e->alpha * e->alpha)) * (e->a[l] * e->a[l])
                          - (4.0 * e->alpha) * la)
                     - (2.0 * e->alpha) * e->a[l];
    }
}

void qo_elem_ij(double g1, double g2, double zi, double zj, double rho,
                double Lz, double alpha, double acc, qo_elem *e)
{
    double z_n = fabs(zi - zj), z_p = zi + zj;
    e->gamma_1 = g1; e->gamma_2 = g2; e->rho = rho; e->L_z = Lz; e->alpha = alpha;
    e->a[0] = z_n; e->a[1] = z_p; e->a[2] = 2.0 * Lz - z_p; e->a[3] = 2.0 * Lz - z_n;
    e->b[0] = 1.0; e->b[1] = g1; e->b[2] = g2; e->b[3] = g1 * g2;
    double sg = jl_sign(zi - zj);
    e->s[0] = -sg; e->s[1] = -1.0; e->s[2] = 1.0; e->s[3] = sg;
    elem_limits(e, acc);
}

/* greens_element_i_init -- src/greens_function.jl:33-48 */
void qo_elem_i(double g1, double g2, double zi, double Lz, double alpha,
               double acc, qo_elem *e)
{
    double z_p = 2.0 * zi;
    e->gamma_1 = g1; e->gamma_2 = g2; e->rho = 0.0; e->L_z = Lz; e->alpha = alpha;
    e->a[0] = 0.0; e->a[1] = z_p; e->a[2] = 2.0 * Lz - z_p; e->a[3] = 2.0 * Lz;
    e->b[0] = 1.0; e->b[1] = g1; e->b[2] = g2; e->b[3] = g1 * g2;
    e->s[0] = 0.0; e->s[1] = -1.0; e->s[2] = 1.0; e->s[3] = 0.0;
    elem_limits(e, acc);
}

/* greens_element_init -- src/greens_function.jl:51-65 (limits unused) */
void qo_elem_long(double g1, double g2, double Lz, double alpha, qo_elem *e)
{
    e->gamma_1 = g1; e->gamma_2 = g2; e->rho = 0.0; e->L_z = Lz; e->alpha = alpha;
    e->a[0] = 0.0; e->a[1] = 0.0; e->a[2] = 2.0 * Lz; e->a[3] = 2.0 * Lz;
    e->b[0] = 1.0; e->b[1] = g1; e->b[2] = g2; e->b[3] = g1 * g2;
    e->s[0] = 0.0; e->s[1] = -1.0; e->s[2] = 1.0; e->s[3] = 0.0;
    for (int l = 0; l < 4; ++l) { e->k_f1[l] = 0.0; e->k_f2[l] = 0.0; }
}

/* ------------------------------------------------------------------------ */
/* Gamma_1 / Gamma_2 / dz_Gamma_1 / dz_Gamma_2 -- src/greens_function.jl:68-116 */
static double green_d(double k, const qo_elem *e)
{
    return e->gamma_1 * e->gamma_2 * exp(-2.0 * k * e->L_z) - 1.0;
}

static double green_u(double k, const qo_elem *e, int l)
{
    if (l == 0) {
        double s = 0.0;
        for (int i = 0; i < 4; ++i) s += e->b[i] * exp(-k * e->a[i]);
        return 0.5 * s;
    }
    return 0.5 * e->b[l - 1] * exp(-k * e->a[l - 1]);
}

static double dz_green_u(double k, const qo_elem *e, int l)
{
    if (l == 0) {
        double s = 0.0;
        for (int i = 0; i < 4; ++i) s += k * e->s[i] * e->b[i] * exp(-k * e->a[i]);
        return 0.5 * s;
    }
    return 0.5 * k * e->s[l - 1] * e->b[l - 1] * exp(-k * e->a[l - 1]);
}

double qo_Gamma_1(double k, const qo_elem *e, int l)
{
    return green_u(k, e, l) / green_d(k, e);
}

double qo_Gamma_2(double k, const qo_elem *e, int l)
{
    return e->gamma_1 * e->gamma_2 * green_u(k, e, l) * exp(-2.0 * k * e->L_z) / green_d(k, e);
}

double qo_dz_Gamma_1(double k, const qo_elem *e, int l)
{
    return dz_green_u(k, e, l) / green_d(k, e);
}

double qo_dz_Gamma_2(double k, const qo_elem *e, int l)
{
    return e->gamma_1 * e->gamma_2 * dz_green_u(k, e, l) * exp(-2.0 * k * e->L_z) / green_d(k, e);
}

/* ------------------------------------------------------------------------ */
/* Integrands -- src/Force_short.jl:4-22 */
enum { F_SR_GAUSS, F_SR_POINT, F_SZ_GAUSS, F_SZ_POINT };

static double integrand(int kind, double k, const qo_elem *e, int l)
{
    switch (kind) {
    case F_SR_GAUSS: return k * qo_Gamma_1(k, e, l) * exp(-(k * k) / (4.0 * e->alpha)) * j1(k * e->rho);
    case F_SR_POINT: return k * qo_Gamma_2(k, e, l) * j1(k * e->rho);
    case F_SZ_GAUSS: return qo_dz_Gamma_1(k, e, l) * exp(-(k * k) / (4.0 * e->alpha)) * j0(k * e->rho);
    default:         return qo_dz_Gamma_2(k, e, l) * j0(k * e->rho);
    }
}

/* Guass_int / Gauss_Legendre on region (0, K) -- src/Gaussian_int.jl:9-28.
 * The nodes are generated once per N_t instead of once per call (line 11);
 * the values are the same. */
typedef struct { int n; double x[256]; double w[256]; } gl_rule;

static const gl_rule *get_rule(int n)
{
    static gl_rule rules[257];
    static int ready[257];
    if (n < 1 || n > 256) return NULL;
    if (!ready[n]) {
#ifdef _OPENMP
#pragma omp critical(qo_rule)
#endif
        {
            if (!ready[n]) {
                rules[n].n = n;
                qo_legendre(n, rules[n].x, rules[n].w);
                ready[n] = 1;
            }
        }
    }
    return &rules[n];
}

static double gauss_legendre(int kind, const qo_elem *e, double K, int l, const gl_rule *r)
{
    double s = 0.0;
    for (int i = 0; i < r->n; ++i) {
        double xi = K / 2.0 + (K * r->x[i]) / 2.0;
        double wi = (K * r->w[i]) / 2.0;
        s += integrand(kind, xi, e, l) * wi;
    }
    return s;
}

/* ------------------------------------------------------------------------ */
/* Pole splitting for gamma_1 gamma_2 > 1 -- src/python_lib/Force_short_norm_njit.py
 * (gauss: 11-21, the four intcore functions: 23-73, 138-184, 298-344, segment
 * sums: 75-93, 186-229; the node counts of Force_short_range_njit.py:258-286) and
 * the regular part of 1/D, green_norm_njit.py:13-32.  The rule is applied to the
 * Julia integrands (src/Force_short.jl:4-22) with the images summed. */

/* the integrand times D(k): 1/D taken out */
static double pole_num(int kind, double k, const qo_elem *e, unsigned mask)
{
    double gu = 0.0, gz = 0.0;
    for (int i = 0; i < 4; ++i) {
        if (!(mask & (1u << i))) continue;
        const double t = e->b[i] * exp(-k * e->a[i]);
        gu += t;
        gz += k * e->s[i] * t;
    }
    gu *= 0.5;
    gz *= 0.5;
    const double pt = e->gamma_1 * e->gamma_2 * exp(-2.0 * k * e->L_z);
    const double ga = exp(-(k * k) / (4.0 * e->alpha));
    switch (kind) {
    case F_SR_GAUSS: return k * ga * gu * j1(k * e->rho);
    case F_SR_POINT: return k * pt * gu * j1(k * e->rho);
    case F_SZ_GAUSS: return ga * gz * j0(k * e->rho);
    default:         return pt * gz * j0(k * e->rho);
    }
}

/* 1/D + 1/x at x = 2 Lz (k - k0), where D = e^{-x} - 1: with
 * u = sum_{m=0}^{10} (-x)^m / (m+2)!  (the series of (e^{-x} - 1 + x)/x^2)
 * and d = -1 + x u (that of D/x), the value is u/d, truncated as the
 * reference truncates it */
static double pole_regular(double x)
{
    double c[11], f = 2.0;   /* c[m] = 1/(m+2)! */
    for (int m = 0; m <= 10; ++m) {
        c[m] = 1.0 / f;
        f *= (double)(m + 3);
    }
    double u = 0.0;
    for (int m = 10; m >= 0; --m) u = u * (-x) + c[m];
    return u / (-1.0 + x * u);
}

int qo_pole_rule(double g12, double Lz, double kf_base, double eps, int N_t, double *k0,
                 double *kf, int *n1, int *n3)
{
    const double kk0 = log(g12) / (2.0 * Lz);
    const double kkf = fmax(kf_base, 2.1 * kk0);
    const int m1 = (int)((double)N_t * kk0 / kkf);
    if (k0) *k0 = kk0;
    if (kf) *kf = kkf;
    if (n1) *n1 = m1;
    if (n3) *n3 = N_t - 2 * m1;
    if (!(g12 > 1.0) || m1 < 1 || N_t - 2 * m1 < 1 || N_t > 256 || !(eps > 0.0 && eps < kk0))
        return -1;
    return 0;
}

/* the integrand at one node (the three branches of each intcore) */
static double pole_core(int kind, double k, const qo_elem *e, unsigned mask, double k0,
                        double eps)
{
    const double two_lz = 2.0 * e->L_z;
    if (fabs(k - k0) <= eps) {
        const double k1 = k0 - 0.01 * eps, k2 = k0 + 0.01 * eps;
        const double df = (pole_num(kind, k2, e, mask) - pole_num(kind, k1, e, mask)) / (k2 - k1);
        return pole_num(kind, k, e, mask) * pole_regular(two_lz * (k - k0)) - df / two_lz;
    }
    double f = pole_num(kind, k, e, mask) / green_d(k, e);
    if (k <= 2.0 * k0) f += pole_num(kind, k0, e, mask) / (two_lz * (k - k0));
    return f;
}

static double pole_segment(int kind, const qo_elem *e, unsigned mask, double k0, double eps,
                           double a, double b, int n)
{
    const gl_rule *r = get_rule(n);
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        const double k = 0.5 * (b - a) * r->x[i] + 0.5 * (b + a);
        s += 0.5 * (b - a) * (r->w[i] * pole_core(kind, k, e, mask, k0, eps));
    }
    return s;
}

double qo_pole_integral(int kind, const qo_elem *e, unsigned mask, double kf_base,
                        double eps, int N_t)
{
    double k0, kf;
    int n1, n3;
    if (qo_pole_rule(e->gamma_1 * e->gamma_2, e->L_z, kf_base, eps, N_t, &k0, &kf, &n1, &n3))
        return NAN;
    return pole_segment(kind, e, mask, k0, eps, 0.0, k0 - eps, n1) +
           pole_segment(kind, e, mask, k0, eps, k0 - eps, k0 + eps, 10) +
           pole_segment(kind, e, mask, k0, eps, k0 + eps, 2.0 * k0, n1) +
           pole_segment(kind, e, mask, k0, eps, 2.0 * k0, kf, n3);
}

static int pole_mode(const qo_params *p)
{
    return (p->semantics & QO_SEM_POLE_SPLIT) && p->gamma_1 * p->gamma_2 > 1.0;
}

/* ------------------------------------------------------------------------ */
static void min_image(double dx, double dy, double Lx, double Ly, double *mx, double *my)
{
    *mx = dx - Lx * rint(dx / Lx);
    *my = dy - Ly * rint(dy / Ly);
}

/* F_short_ij -- src/Force_short.jl:25-78.  The per-pair 2-point
 * neighborlist (line 36) reduces to: minimum-image distance, empty if
 * beyond r_cutoff. */
void qo_F_short_ij(const qo_params *p, double qi, double qj, const double *ri,
                   const double *rj, int single, double *out, double *fz_j)
{
    out[0] = out[1] = out[2] = 0.0;
    if (fz_j) *fz_j = 0.0;
    double mx, my;
    min_image(ri[0] - rj[0], ri[1] - rj[1], p->Lx, p->Ly, &mx, &my);
    double d2 = mx * mx + my * my;
    if (d2 > p->r_cutoff * p->r_cutoff) return;
    double rho = sqrt(d2);

    const gl_rule *r = get_rule(p->N_t);
    qo_elem e;
    qo_elem_ij(p->gamma_1, p->gamma_2, ri[2], rj[2], rho, p->Lz, p->alpha, p->accuracy, &e);
    const int swap = (p->semantics & QO_SEM_SWAP_LIMITS) != 0;
    const double *Kp = swap ? e.k_f1 : e.k_f2; /* point-charge integrals */
    const double *Kg = swap ? e.k_f2 : e.k_f1; /* gaussian integrals     */
    /* pole splitting: common limits, the images summed (odd: 1 and 4, even:
     * 2 and 3, kept apart for the j-side option) */
    const int pole = pole_mode(p);
    const double kfp = -log(p->accuracy) / (2.0 * p->Lz);
    const double kfg = sqrt(-4.0 * p->alpha * log(p->accuracy));

    double Fsx = 0.0, Fsy = 0.0;
    if (rho != 0.0) {
        double p1 = 0.0, p2 = 0.0, g = 0.0;
        if (pole) p1 = qo_pole_integral(F_SR_POINT, &e, 15u, kfp, p->pole_eps, p->N_t);
        else
            for (int l = 1; l <= 4; ++l) p1 += gauss_legendre(F_SR_POINT, &e, Kp[l - 1], l, r);
        for (int l = 0; l < 4; ++l) p2 += e.b[l] * rho / pow(e.a[l] * e.a[l] + rho * rho, 1.5);
        p2 *= 0.5;
        if (!single && pole) g = qo_pole_integral(F_SR_GAUSS, &e, 15u, kfg, p->pole_eps, p->N_t);
        else if (!single)
            for (int l = 1; l <= 4; ++l) g += gauss_legendre(F_SR_GAUSS, &e, Kg[l - 1], l, r);
        double Fsr = -p1 + p2 + g;
        double dirx, diry;
        if (p->semantics & QO_SEM_MINIMAGE_DIR) { dirx = mx; diry = my; }
        else { dirx = ri[0] - rj[0]; diry = ri[1] - rj[1]; }
        Fsx = Fsr * dirx / rho;
        Fsy = Fsr * diry / rho;
    }
    /* z: keep the per-image parts for the j-side option */
    double zp1[4], zp2[4], zg[4];
    double P1 = 0.0, P2 = 0.0, G = 0.0;
    if (pole) {
        /* the odd and even image groups' integrals, carried on images 1 and 2 */
        zp1[0] = qo_pole_integral(F_SZ_POINT, &e, 9u, kfp, p->pole_eps, p->N_t);
        zp1[1] = qo_pole_integral(F_SZ_POINT, &e, 6u, kfp, p->pole_eps, p->N_t);
        zp1[2] = zp1[3] = 0.0;
        P1 = zp1[0] + zp1[1];
    } else {
        for (int l = 1; l <= 4; ++l) {
            zp1[l - 1] = gauss_legendre(F_SZ_POINT, &e, Kp[l - 1], l, r);
            P1 += zp1[l - 1];
        }
    }
    for (int l = 0; l < 4; ++l) {
        zp2[l] = e.b[l] * e.a[l] * e.s[l] / pow(e.a[l] * e.a[l] + rho * rho, 1.5);
        P2 += zp2[l];
    }
    P2 *= 0.5;
    if (pole) {
        zg[0] = single ? 0.0 : qo_pole_integral(F_SZ_GAUSS, &e, 9u, kfg, p->pole_eps, p->N_t);
        zg[1] = single ? 0.0 : qo_pole_integral(F_SZ_GAUSS, &e, 6u, kfg, p->pole_eps, p->N_t);
        zg[2] = zg[3] = 0.0;
        G = zg[0] + zg[1];
    } else {
        for (int l = 1; l <= 4; ++l) {
            zg[l - 1] = single ? 0.0 : gauss_legendre(F_SZ_GAUSS, &e, Kg[l - 1], l, r);
            G += zg[l - 1];
        }
    }
    double Fsz = P1 - P2 - G;
    double pref = qi * qj / (QO_TWO_PI * p->eps_0);
    out[0] = pref * Fsx;
    out[1] = pref * Fsy;
    out[2] = pref * Fsz;
    if (fz_j) {
        if ((p->semantics & QO_SEM_NEWTON_IMAGES) && pole) {
            /* odd images change sign under z_i <-> z_j */
            const double odd = zp1[0] - 0.5 * (zp2[0] + zp2[3]) - zg[0];
            const double even = zp1[1] - 0.5 * (zp2[1] + zp2[2]) - zg[1];
            *fz_j = pref * (-odd + even);
        } else if (p->semantics & QO_SEM_NEWTON_IMAGES) {
            double part[4];
            for (int l = 0; l < 4; ++l) part[l] = zp1[l] - 0.5 * zp2[l] - zg[l];
            *fz_j = pref * (-part[0] + part[1] + part[2] - part[3]);
        } else {
            *fz_j = -out[2];
        }
    }
}

/* F_short_i -- src/Force_short.jl:82-107 */
void qo_F_short_i(const qo_params *p, double qi, const double *ri, int single, double *out)
{
    const gl_rule *r = get_rule(p->N_t);
    qo_elem e;
    qo_elem_i(p->gamma_1, p->gamma_2, ri[2], p->Lz, p->alpha, p->accuracy, &e);
    const int swap = (p->semantics & QO_SEM_SWAP_LIMITS) != 0;
    const double *Kp = swap ? e.k_f1 : e.k_f2;
    const double *Kg = swap ? e.k_f2 : e.k_f1;
    double P1 = 0.0, P2 = 0.0, G = 0.0;
    const int pole = pole_mode(p);
    if (pole)
        P1 = qo_pole_integral(F_SZ_POINT, &e, 15u, -log(p->accuracy) / (2.0 * p->Lz),
                              p->pole_eps, p->N_t);
    else
        for (int l = 1; l <= 4; ++l) P1 += gauss_legendre(F_SZ_POINT, &e, Kp[l - 1], l, r);
    for (int l = 1; l <= 2; ++l) { /* images 2 and 3 (0-based 1, 2) */
        double a = e.a[l];
        P2 += e.b[l] * a * e.s[l] / (a * a * a);
    }
    P2 *= 0.5;
    if (!single && pole)
        G = qo_pole_integral(F_SZ_GAUSS, &e, 15u, sqrt(-4.0 * p->alpha * log(p->accuracy)),
                             p->pole_eps, p->N_t);
    else if (!single)
        for (int l = 1; l <= 4; ++l) G += gauss_legendre(F_SZ_GAUSS, &e, Kg[l - 1], l, r);
    double Fsz = -P1 + P2 + G;
    out[0] = 0.0;
    out[1] = 0.0;
    out[2] = -(qi * qi) / (QO_TWO_PI * p->eps_0) * Fsz;
}

/* ------------------------------------------------------------------------ */
/* 2-D neighbour pairs: stands in for CellListMap.neighborlist(xy, r_c;
 * unitcell=[Lx,Ly]) (src/QEM_short.jl:34-35): all unordered pairs with
 * minimum-image xy distance <= r_c, reported as i < j. */
int64_t qo_neighbor_pairs(int64_t n, const double *xyz, double Lx, double Ly,
                          double rc, int32_t *pairs, int64_t cap)
{
    int64_t np = 0;
    double rc2 = rc * rc;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t j = i + 1; j < n; ++j) {
            double mx, my;
            min_image(xyz[3 * i] - xyz[3 * j], xyz[3 * i + 1] - xyz[3 * j + 1], Lx, Ly, &mx, &my);
            double d2 = mx * mx + my * my;
            if (d2 <= rc2) {
                if (np < cap && pairs) {
                    pairs[2 * np] = (int32_t)i;
                    pairs[2 * np + 1] = (int32_t)j;
                }
                ++np;
            }
        }
    }
    return np;
}

/* ------------------------------------------------------------------------ */
void qo_short_pair_forces(const qo_params *p, const double *xyz, const double *q,
                          int64_t npairs, const int32_t *pairs, double *fpair,
                          int nthreads)
{
    (void)get_rule(p->N_t);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t t = 0; t < npairs; ++t) {
        int32_t i = pairs[2 * t], j = pairs[2 * t + 1];
        double out[3], fzj;
        qo_F_short_ij(p, q[i], q[j], xyz + 3 * i, xyz + 3 * j, 0, out, &fzj);
        fpair[4 * t] = out[0];
        fpair[4 * t + 1] = out[1];
        fpair[4 * t + 2] = out[2];
        fpair[4 * t + 3] = fzj;
    }
    (void)nthreads;
}

/* Molly.forces(::QEM_short) -- src/QEM_short.jl:27-59, with the list given:
 * F[i] += F_ij; F[j] -= F_ij (lines 41-48), then F[i] += F_short_i (51-56). */
void qo_short_forces(const qo_params *p, int64_t n, const double *xyz,
                     const double *q, int64_t npairs, const int32_t *pairs,
                     double *f, int nthreads)
{
    double *fp = (double *)malloc(sizeof(double) * 4 * (size_t)(npairs > 0 ? npairs : 1));
    double *fs = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    qo_short_pair_forces(p, xyz, q, npairs, pairs, fp, nthreads);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < n; ++i) qo_F_short_i(p, q[i], xyz + 3 * i, 0, fs + 3 * i);
    memset(f, 0, sizeof(double) * 3 * (size_t)n);
    for (int64_t t = 0; t < npairs; ++t) {
        int32_t i = pairs[2 * t], j = pairs[2 * t + 1];
        f[3 * i] += fp[4 * t];
        f[3 * i + 1] += fp[4 * t + 1];
        f[3 * i + 2] += fp[4 * t + 2];
        f[3 * j] -= fp[4 * t];
        f[3 * j + 1] -= fp[4 * t + 1];
        f[3 * j + 2] += fp[4 * t + 3];
    }
    for (int64_t i = 0; i < n; ++i) {
        f[3 * i + 2] += fs[3 * i + 2];
    }
    free(fp);
    free(fs);
}

/* Molly.forces(::QEM_short) -- src/QEM_short.jl:27-59 -- for the rows
 * `rows` only (f is nrows x 3).  Each row's pairs are found by an O(N) scan
 * over every other atom (minimum-image xy distance <= r_cutoff, the test of
 * qo_neighbor_pairs, evaluated with the lower index first), and added in the
 * order qo_short_forces adds them for that atom: the pairs (j, a), j < a, as
 * F[j] -= F_ja (its z part the j-side force), then (a, j), j > a, as
 * F[a] += F_aj, then the self term F_short_i.  The rows' forces are those of
 * the full list without building it, so they can be checked at any N. */
void qo_short_forces_rows(const qo_params *p, int64_t n, const double *xyz, const double *q,
                          int64_t nrows, const int64_t *rows, double *f, int nthreads)
{
    (void)get_rule(p->N_t);
    const double rc2 = p->r_cutoff * p->r_cutoff;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t r = 0; r < nrows; ++r) {
        const int64_t a = rows[r];
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == a) continue;
            const int64_t i0 = j < a ? j : a, j0 = j < a ? a : j;
            double mx, my;
            min_image(xyz[3 * i0] - xyz[3 * j0], xyz[3 * i0 + 1] - xyz[3 * j0 + 1], p->Lx,
                      p->Ly, &mx, &my);
            if (mx * mx + my * my > rc2) continue;
            double out[3], fzj;
            qo_F_short_ij(p, q[i0], q[j0], xyz + 3 * i0, xyz + 3 * j0, 0, out, &fzj);
            if (a == i0) {
                fx += out[0];
                fy += out[1];
                fz += out[2];
            } else {
                fx -= out[0];
                fy -= out[1];
                fz += fzj;
            }
        }
        double fs[3];
        qo_F_short_i(p, q[a], xyz + 3 * a, 0, fs);
        f[3 * r] = fx;
        f[3 * r + 1] = fy;
        f[3 * r + 2] = fz + fs[2];
    }
    (void)nthreads;
}

/* ------------------------------------------------------------------------ */
/* K_set_generator -- src/K_set_generator.jl:5-45 */
int64_t qo_kset_size(double Lx, double Ly, double alpha, double acc,
                     int32_t *mx_max, int32_t *my_max)
{
    double k_c = 2.0 * sqrt(-4.0 * alpha * log(acc));
    int32_t mx = (int32_t)ceil(k_c * Lx / QO_TWO_PI) + 1;
    int32_t my = (int32_t)ceil(k_c * Ly / QO_TWO_PI) + 1;
    if (mx_max) *mx_max = mx;
    if (my_max) *my_max = my;
    return (int64_t)(2 * mx + 1) * (int64_t)(2 * my + 1) - 1;
}

void qo_kset(double Lx, double Ly, double alpha, double acc, double *kxkyk,
             float *prob, double *sum_K)
{
    int32_t mx_max, my_max;
    qo_kset_size(Lx, Ly, alpha, acc, &mx_max, &my_max);
    double S_x = 0.0, S_y = 0.0;
    for (int32_t m = -mx_max; m <= mx_max; ++m)
        if (m != 0) {
            double k_x = (double)m * QO_TWO_PI / Lx;
            S_x += exp(-(k_x * k_x) / (4.0 * alpha));
        }
    for (int32_t m = -my_max; m <= my_max; ++m)
        if (m != 0) {
            double k_y = (double)m * QO_TWO_PI / Ly;
            S_y += exp(-(k_y * k_y) / (4.0 * alpha));
        }
    double sK = S_x + S_y + S_x * S_y;
    int64_t t = 0;
    for (int32_t mx = -mx_max; mx <= mx_max; ++mx)
        for (int32_t my = -my_max; my <= my_max; ++my) {
            double k_x = (double)mx * QO_TWO_PI / Lx;
            double k_y = (double)my * QO_TWO_PI / Ly;
            double k = sqrt(k_x * k_x + k_y * k_y);
            if (k != 0.0) {
                kxkyk[3 * t] = k_x;
                kxkyk[3 * t + 1] = k_y;
                kxkyk[3 * t + 2] = k;
                prob[t] = (float)(exp(-(k * k) / (4.0 * alpha)) / sK);
                ++t;
            }
        }
    *sum_K = sK;
}

/* ------------------------------------------------------------------------ */
/* sortperm(z) -- src/QEM_long.jl:39-40; ties in index order (stable). */
static const double *g_sort_xyz;
static int cmp_z(const void *a, const void *b)
{
    int64_t i = *(const int64_t *)a, j = *(const int64_t *)b;
    double zi = g_sort_xyz[3 * i + 2], zj = g_sort_xyz[3 * j + 2];
    if (zi < zj) return -1;
    if (zi > zj) return 1;
    return (i > j) - (i < j);
}

void qo_sortperm_z(int64_t n, const double *xyz, int64_t *zlist)
{
    for (int64_t i = 0; i < n; ++i) zlist[i] = i;
#ifdef _OPENMP
#pragma omp critical(qo_sort)
#endif
    {
        g_sort_xyz = xyz;
        qsort(zlist, (size_t)n, sizeof(int64_t), cmp_z);
    }
}

/* ------------------------------------------------------------------------ */
/* Fl_non_sort -- src/Force_long.jl:5-70 */
void qo_Fl_non_sort(const double *k3, int64_t n, const double *xyz,
                    const double *q, double g1, double g2, double Lz, double *f)
{
    double k_x = k3[0], k_y = k3[1], k = k3[2];
    double f0[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = 0; i < n; ++i) {
        double qi = q[i], xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
        double th = k_x * xi + k_y * yi;
        f0[0] += qi * cos(th) * exp(+k * zi);
        f0[1] += qi * sin(th) * exp(+k * zi);
        f0[2] += qi * cos(th) * exp(-k * zi);
        f0[3] += qi * sin(th) * exp(-k * zi);
    }
    double e2 = exp(-2.0 * k * Lz);
    double gamma_val = 0.5 / (g1 * g2 * e2 - 1.0);
    double GP[4] = {g2, 1.0, g1, g1 * g2};
    for (int64_t i = 0; i < n; ++i) {
        double qi = q[i], xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
        double th = k_x * xi + k_y * yi;
        double c = cos(th), s = sin(th), Ep = exp(k * zi), Em = exp(-k * zi);
        double ff[4];
        ff[0] = f0[0] - qi * c * exp(+k * zi);
        ff[1] = f0[1] - qi * s * exp(+k * zi);
        ff[2] = f0[2] - qi * c * exp(-k * zi);
        ff[3] = f0[3] - qi * s * exp(-k * zi);
        double sum_xy = qi * gamma_val *
            (c * (GP[0] * e2 * Ep * ff[1] + GP[2] * Em * ff[3] + GP[3] * e2 * Em * ff[1] + GP[3] * e2 * Ep * ff[3])
             - s * (GP[0] * e2 * Ep * ff[0] + GP[2] * Em * ff[2] + GP[3] * e2 * Em * ff[0] + GP[3] * e2 * Ep * ff[2]));
        double sum_z = qi * gamma_val *
            (c * (GP[0] * e2 * Ep * ff[0] - GP[2] * Em * ff[2] - GP[3] * e2 * Em * ff[0] + GP[3] * e2 * Ep * ff[2])
             + s * (GP[0] * e2 * Ep * ff[1] - GP[2] * Em * ff[3] - GP[3] * e2 * Em * ff[1] + GP[3] * e2 * Ep * ff[3]));
        f[3 * i] = (k_x / k) * sum_xy;
        f[3 * i + 1] = (k_y / k) * sum_xy;
        f[3 * i + 2] = sum_z;
    }
}

/* Fl_sort -- src/Force_long.jl:72-127.
 * A, B are the exclusive prefix sums of q cos/sin e^{+kz} in z order, as in
 * lines 95-96.  C, D are the exclusive suffix sums of q cos/sin e^{-kz}.
 * The reference forms C, D by subtracting from the total (lines 83-84,
 * 97-98); in FP64 that loses every digit of C_i e^{+kz_i} once
 * k (z_i - z_min) exceeds about 36.  qo_Fl_sort accumulates the suffix
 * from the top (same sums, no cancellation); qo_Fl_sort_subtractive keeps
 * the reference's rounding for comparison. */
static void fl_sort_impl(const double *k3, int64_t n, const int64_t *zlist,
                         const double *xyz, const double *q, double *f, int subtractive)
{
    double k_x = k3[0], k_y = k3[1], k = k3[2];
    double *A = (double *)calloc((size_t)(4 * (n > 0 ? n : 1)), sizeof(double));
    double *B = A + n, *C = B + n, *D = C + n;
    if (n == 0) { free(A); return; }
    if (subtractive) {
        for (int64_t i = 1; i < n; ++i) {
            int64_t l = zlist[i];
            double th = k_x * xyz[3 * l] + k_y * xyz[3 * l + 1];
            C[0] += q[l] * exp(-k * xyz[3 * l + 2]) * cos(th);
            D[0] += q[l] * exp(-k * xyz[3 * l + 2]) * sin(th);
        }
    }
    for (int64_t i = 0; i < n - 1; ++i) {
        int64_t l = zlist[i], lp1 = zlist[i + 1];
        double th = k_x * xyz[3 * l] + k_y * xyz[3 * l + 1];
        double th1 = k_x * xyz[3 * lp1] + k_y * xyz[3 * lp1 + 1];
        A[i + 1] = A[i] + q[l] * exp(+k * xyz[3 * l + 2]) * cos(th);
        B[i + 1] = B[i] + q[l] * exp(+k * xyz[3 * l + 2]) * sin(th);
        if (subtractive) {
            C[i + 1] = C[i] - q[lp1] * exp(-k * xyz[3 * lp1 + 2]) * cos(th1);
            D[i + 1] = D[i] - q[lp1] * exp(-k * xyz[3 * lp1 + 2]) * sin(th1);
        }
    }
    if (!subtractive) {
        C[n - 1] = 0.0;
        D[n - 1] = 0.0;
        for (int64_t i = n - 2; i >= 0; --i) {
            int64_t lp1 = zlist[i + 1];
            double th1 = k_x * xyz[3 * lp1] + k_y * xyz[3 * lp1 + 1];
            C[i] = C[i + 1] + q[lp1] * exp(-k * xyz[3 * lp1 + 2]) * cos(th1);
            D[i] = D[i + 1] + q[lp1] * exp(-k * xyz[3 * lp1 + 2]) * sin(th1);
        }
    }
    for (int64_t i = 0; i < n; ++i) {
        int64_t l = zlist[i];
        double qi = q[l], zi = xyz[3 * l + 2];
        double th = k_x * xyz[3 * l] + k_y * xyz[3 * l + 1];
        double c = cos(th), s = sin(th);
        double sum_xy = s * exp(-k * zi) * A[i] - c * exp(-k * zi) * B[i]
                      + s * exp(k * zi) * C[i] - c * exp(k * zi) * D[i];
        f[3 * l] = qi * k_x / (2.0 * k) * sum_xy;
        f[3 * l + 1] = qi * k_y / (2.0 * k) * sum_xy;
        f[3 * l + 2] = (qi / 2.0) * (c * exp(-k * zi) * A[i] + s * exp(-k * zi) * B[i]
                                     - c * exp(k * zi) * C[i] - s * exp(k * zi) * D[i]);
    }
    free(A);
}

void qo_Fl_sort(const double *k3, int64_t n, const int64_t *zlist,
                const double *xyz, const double *q, double *f)
{
    fl_sort_impl(k3, n, zlist, xyz, q, f, 0);
}

void qo_Fl_sort_subtractive(const double *k3, int64_t n, const int64_t *zlist,
                            const double *xyz, const double *q, double *f)
{
    fl_sort_impl(k3, n, zlist, xyz, q, f, 1);
}

/* Flz_sort_k0 -- src/Force_long.jl:130-161 (z component only; x, y = 0) */
void qo_Flz_sort_k0(int64_t n, const int64_t *zlist, const double *q, double *f)
{
    if (n == 0) return;
    double *Q1 = (double *)calloc((size_t)(2 * n), sizeof(double));
    double *Q2 = Q1 + n;
    for (int64_t i = 1; i < n; ++i) Q2[0] += q[zlist[i]];
    for (int64_t i = 0; i < n - 1; ++i) Q1[i + 1] = Q1[i] + q[zlist[i]];
    for (int64_t i = 1; i < n; ++i) Q2[i] = Q2[i - 1] - q[zlist[i]];
    for (int64_t j = 0; j < n; ++j) {
        int64_t l = zlist[j];
        f[3 * l] = 0.0;
        f[3 * l + 1] = 0.0;
        f[3 * l + 2] = q[l] * (Q1[j] - Q2[j]);
    }
    free(Q1);
}

/* Flz_self -- src/Force_long.jl:164-184 */
void qo_Flz_self(const double *k3, int64_t n, const double *xyz,
                 const double *q, double g1, double g2, double Lz, double *f)
{
    double k = k3[2];
    double gamma_val = g1 * g2 * exp(-2.0 * k * Lz) - 1.0;
    for (int64_t i = 0; i < n; ++i) {
        double zi = xyz[3 * i + 2];
        f[3 * i] = 0.0;
        f[3 * i + 1] = 0.0;
        f[3 * i + 2] = q[i] * q[i] * (g2 * exp(-2.0 * k * (Lz - zi)) - g1 * exp(-2.0 * k * zi))
                       / (2.0 * gamma_val);
    }
}

/* F_l_total -- src/Force_long.jl:186-199 */
void qo_F_l_total(const double *k3, int64_t n, const double *xyz,
                  const double *q, const int64_t *zlist, const qo_params *p,
                  double *f)
{
    double *a = (double *)malloc(sizeof(double) * 9 * (size_t)(n > 0 ? n : 1));
    double *b = a + 3 * n, *c = b + 3 * n;
    qo_Fl_non_sort(k3, n, xyz, q, p->gamma_1, p->gamma_2, p->Lz, a);
    qo_Fl_sort(k3, n, zlist, xyz, q, b);
    qo_Flz_self(k3, n, xyz, q, p->gamma_1, p->gamma_2, p->Lz, c);
    double den = p->Lx * p->Ly * p->eps_0;
    for (int64_t t = 0; t < 3 * n; ++t) f[t] = (a[t] + b[t] + c[t]) / den;
    free(a);
}

/* F_long, random-batch branch -- src/Force_long.jl:201-218, 234-242 */
void qo_long_forces(const qo_params *p, int64_t n, const double *xyz,
                    const double *q, const double *kvecs, int64_t P,
                    double weight, int full_sum, double *f, int nthreads)
{
    int64_t *zlist = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    qo_sortperm_z(n, xyz, zlist);
    memset(f, 0, sizeof(double) * 3 * (size_t)n);
    qo_Flz_sort_k0(n, zlist, q, f);
    double den0 = 2.0 * p->Lx * p->Ly * p->eps_0;
    for (int64_t t = 0; t < 3 * n; ++t) f[t] /= den0;
    int nt = nthreads > 0 ? nthreads : 1;
    double *part = (double *)calloc((size_t)nt * 3 * (size_t)n, sizeof(double));
    /* per-thread partial sums over contiguous k blocks, added in thread order */
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        int64_t lo = P * tid / nth, hi = P * (tid + 1) / nth;
        double *fk = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
        double *acc = part + (size_t)tid * 3 * (size_t)n;
        for (int64_t t = lo; t < hi; ++t) {
            const double *k3 = kvecs + 3 * t;
            qo_F_l_total(k3, n, xyz, q, zlist, p, fk);
            double wk = full_sum ? exp(-(k3[2] * k3[2]) / (4.0 * p->alpha)) : weight;
            for (int64_t u = 0; u < 3 * n; ++u) acc[u] += fk[u] * wk;
        }
        free(fk);
    }
    for (int th = 0; th < nt; ++th)
        for (int64_t u = 0; u < 3 * n; ++u) f[u] += part[(size_t)th * 3 * (size_t)n + u];
    free(part);
    free(zlist);
}

/* Batch variance of the random-batch long-range force (no counterpart in
 * the reference; the estimator is src/Force_long.jl:201-218,234-242):
 * X_ip = weight * F_l_total(k_p) on atom i, var_i = (P sum_p |X_ip|^2 -
 * |sum_p X_ip|^2) / (P - 1), the k = 0 term left out. */
void qo_long_variance(const qo_params *p, int64_t n, const double *xyz, const double *q,
                      const double *kvecs, int64_t P, double weight, double *var,
                      int nthreads)
{
    int64_t *zlist = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    qo_sortperm_z(n, xyz, zlist);
    int nt = nthreads > 0 ? nthreads : 1;
    /* per thread: 3 n sums of X, n sums of |X|^2 */
    double *part = (double *)calloc((size_t)nt * 4 * (size_t)n, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        int64_t lo = P * tid / nth, hi = P * (tid + 1) / nth;
        double *fk = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
        double *s1 = part + (size_t)tid * 4 * (size_t)n, *s2 = s1 + 3 * (size_t)n;
        for (int64_t t = lo; t < hi; ++t) {
            qo_F_l_total(kvecs + 3 * t, n, xyz, q, zlist, p, fk);
            for (int64_t i = 0; i < n; ++i) {
                const double x = weight * fk[3 * i], y = weight * fk[3 * i + 1],
                             z = weight * fk[3 * i + 2];
                s1[3 * i] += x;
                s1[3 * i + 1] += y;
                s1[3 * i + 2] += z;
                s2[i] += x * x + y * y + z * z;
            }
        }
        free(fk);
    }
    for (int64_t i = 0; i < n; ++i) {
        double x = 0.0, y = 0.0, z = 0.0, ss = 0.0;
        for (int th = 0; th < nt; ++th) {
            const double *s1 = part + (size_t)th * 4 * (size_t)n, *s2 = s1 + 3 * (size_t)n;
            x += s1[3 * i];
            y += s1[3 * i + 1];
            z += s1[3 * i + 2];
            ss += s2[i];
        }
        var[i] = ((double)P * ss - (x * x + y * y + z * z)) / ((double)P - 1.0);
    }
    free(part);
    free(zlist);
}

/* ------------------------------------------------------------------------ */
/* Energy, dE/dq and dE/dgamma.  The reference has no energy; this is the
 * functional whose negative gradient at frozen quadrature limits is the
 * force above (with QO_SEM_CORRECTED it is the physical slab energy split
 * by exp(-k^2/(4 alpha))).  Definitions (P = 1/(2 pi eps_0)):
 *   pair   q_i q_j P [ 1/2 sum_l b_l / sqrt(a_l^2 + rho^2)
 *                      - sum_l GL[Gamma_2,l J0; K_point,l]
 *                      + sum_l GL[Gamma_1,l e^{-k^2/4a} J0; K_gauss,l] ]
 *   self   q_i^2 P/2 [ 1/2 sum_{l=2..4} b_l / a_l - sum_{l=1..4} GL[Gamma_2,l]
 *                      + GL[Gamma_2,1 e] + sum_{l=2..4} GL[Gamma_1,l e] ]
 *          (images only: the charge's own 1/r is excluded, so image 1
 *           keeps only its multiple reflections Gamma_2,1)
 *   k = 0  -1/(2 Lx Ly eps_0) sum_{i<j} q_i q_j |z_i - z_j|
 *   k      w_k/(Lx Ly eps_0 2k) [ sum_{i<j} q_i q_j cos(k.rho_ij) e^{-k|z_ij|}
 *                       + 1/2 sum_{i,j} q_i q_j cos(k.rho_ij) R_ij/(1 - x) ]
 *          R_ij = g1 e^{-k z+} + g2 e^{-k(2Lz - z+)} + g1 g2 (e^{-k(2Lz+dz)} + e^{-k(2Lz-dz)}),
 *          x = g1 g2 e^{-2kLz}
 * phi_i = dE/dq_i; E = 1/2 sum_i q_i phi_i (E is quadratic in q). */

/* d Gamma / d gamma_w at fixed k (w = 0: gamma_1, w = 1: gamma_2) */
static double dGamma_dg(double k, const qo_elem *e, int l, int which2, int w)
{
    const double E = exp(-2.0 * k * e->L_z);
    const double g12 = e->gamma_1 * e->gamma_2;
    const double D = g12 * E - 1.0;
    const double dg12 = w == 0 ? e->gamma_2 : e->gamma_1;
    double db[4] = {0.0, w == 0 ? 1.0 : 0.0, w == 1 ? 1.0 : 0.0, dg12};
    const double ea = exp(-k * e->a[l - 1]);
    const double b = e->b[l - 1];
    if (!which2) return 0.5 * ea * (db[l - 1] / D - b * dg12 * E / (D * D));
    return 0.5 * ea * (db[l - 1] * g12 * E / D - b * dg12 * E / (D * D));
}

/* integrand kinds for the energy */
enum { U_POINT, U_GAUSS, U_GAUSS2, DU_POINT, DU_GAUSS, DU_GAUSS2 };

static double integrand_e(int kind, double k, const qo_elem *e, int l, int w)
{
    const double g = exp(-(k * k) / (4.0 * e->alpha));
    const double J = j0(k * e->rho);
    switch (kind) {
    case U_POINT:   return qo_Gamma_2(k, e, l) * J;
    case U_GAUSS:   return qo_Gamma_1(k, e, l) * g * J;
    case U_GAUSS2:  return qo_Gamma_2(k, e, l) * g * J;
    case DU_POINT:  return dGamma_dg(k, e, l, 1, w) * J;
    case DU_GAUSS:  return dGamma_dg(k, e, l, 0, w) * g * J;
    default:        return dGamma_dg(k, e, l, 1, w) * g * J;
    }
}

static double gl_e(int kind, const qo_elem *e, double K, int l, int w, const gl_rule *r)
{
    double s = 0.0;
    for (int i = 0; i < r->n; ++i) {
        double xi = K / 2.0 + (K * r->x[i]) / 2.0;
        double wi = (K * r->w[i]) / 2.0;
        s += integrand_e(kind, xi, e, l, w) * wi;
    }
    return s;
}

/* pair potential u_ij (without q_i q_j) and its gamma derivatives */
void qo_pair_potential(const qo_params *p, const double *ri, const double *rj,
                       double *u, double *du)
{
    *u = 0.0;
    du[0] = du[1] = 0.0;
    double mx, my;
    min_image(ri[0] - rj[0], ri[1] - rj[1], p->Lx, p->Ly, &mx, &my);
    double d2 = mx * mx + my * my;
    if (d2 > p->r_cutoff * p->r_cutoff) return;
    double rho = sqrt(d2);
    const gl_rule *r = get_rule(p->N_t);
    qo_elem e;
    qo_elem_ij(p->gamma_1, p->gamma_2, ri[2], rj[2], rho, p->Lz, p->alpha, p->accuracy, &e);
    const int swap = (p->semantics & QO_SEM_SWAP_LIMITS) != 0;
    const double *Kp = swap ? e.k_f1 : e.k_f2;
    const double *Kg = swap ? e.k_f2 : e.k_f1;
    const double P = 1.0 / (QO_TWO_PI * p->eps_0);
    double s = 0.0, ds[2] = {0.0, 0.0};
    const double dbs[2][4] = {{0.0, 1.0, 0.0, p->gamma_2}, {0.0, 0.0, 1.0, p->gamma_1}};
    for (int l = 1; l <= 4; ++l) {
        double inv = 1.0 / sqrt(e.a[l - 1] * e.a[l - 1] + rho * rho);
        s += 0.5 * e.b[l - 1] * inv - gl_e(U_POINT, &e, Kp[l - 1], l, 0, r)
             + gl_e(U_GAUSS, &e, Kg[l - 1], l, 0, r);
        for (int w = 0; w < 2; ++w)
            ds[w] += 0.5 * dbs[w][l - 1] * inv - gl_e(DU_POINT, &e, Kp[l - 1], l, w, r)
                     + gl_e(DU_GAUSS, &e, Kg[l - 1], l, w, r);
    }
    *u = P * s;
    du[0] = P * ds[0];
    du[1] = P * ds[1];
}

/* self energy factor s_i (E_self = q_i^2 s_i) and gamma derivatives */
void qo_self_potential(const qo_params *p, const double *ri, double *s_out, double *ds_out)
{
    const gl_rule *r = get_rule(p->N_t);
    qo_elem e;
    qo_elem_i(p->gamma_1, p->gamma_2, ri[2], p->Lz, p->alpha, p->accuracy, &e);
    const int swap = (p->semantics & QO_SEM_SWAP_LIMITS) != 0;
    const double *Kp = swap ? e.k_f1 : e.k_f2;
    const double *Kg = swap ? e.k_f2 : e.k_f1;
    const double P = 1.0 / (QO_TWO_PI * p->eps_0);
    const double dbs[2][4] = {{0.0, 1.0, 0.0, p->gamma_2}, {0.0, 0.0, 1.0, p->gamma_1}};
    double s = 0.0, ds[2] = {0.0, 0.0};
    for (int l = 2; l <= 4; ++l) {
        s += 0.5 * e.b[l - 1] / e.a[l - 1];
        for (int w = 0; w < 2; ++w) ds[w] += 0.5 * dbs[w][l - 1] / e.a[l - 1];
    }
    for (int l = 1; l <= 4; ++l) {
        s -= gl_e(U_POINT, &e, Kp[l - 1], l, 0, r);
        s += gl_e(l == 1 ? U_GAUSS2 : U_GAUSS, &e, Kg[l - 1], l, 0, r);
        for (int w = 0; w < 2; ++w) {
            ds[w] -= gl_e(DU_POINT, &e, Kp[l - 1], l, w, r);
            ds[w] += gl_e(l == 1 ? DU_GAUSS2 : DU_GAUSS, &e, Kg[l - 1], l, w, r);
        }
    }
    *s_out = 0.5 * P * s;
    ds_out[0] = 0.5 * P * ds[0];
    ds_out[1] = 0.5 * P * ds[1];
}

/* E, phi = dE/dq (may be NULL) and dE/dgamma_1,2 for the short range over
 * a pair list plus self terms, and the long range over explicit k vectors
 * (weight as in qo_long_forces).  O(N^2) per k: test sizes only. */
void qo_energy(const qo_params *p, int64_t n, const double *xyz, const double *q,
               int64_t npairs, const int32_t *pairs, const double *kvecs, int64_t P,
               double weight, int full_sum, int with_short, int with_long,
               double *E_out, double *phi_out, double *dEdg, int nthreads)
{
    double *phi = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double *psi = (double *)calloc((size_t)(2 * (n > 0 ? n : 1)), sizeof(double));
    double dg_long[2] = {0.0, 0.0};
    (void)nthreads;
    if (with_short) {
        double *u = (double *)malloc(sizeof(double) * 3 * (size_t)(npairs > 0 ? npairs : 1));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 32) num_threads(nthreads > 0 ? nthreads : 1)
#endif
        for (int64_t t = 0; t < npairs; ++t) {
            int32_t i = pairs[2 * t], j = pairs[2 * t + 1];
            qo_pair_potential(p, xyz + 3 * i, xyz + 3 * j, u + 3 * t, u + 3 * t + 1);
        }
        for (int64_t t = 0; t < npairs; ++t) {
            int32_t i = pairs[2 * t], j = pairs[2 * t + 1];
            phi[i] += q[j] * u[3 * t];
            phi[j] += q[i] * u[3 * t];
            for (int w = 0; w < 2; ++w) {
                psi[2 * i + w] += q[j] * u[3 * t + 1 + w];
                psi[2 * j + w] += q[i] * u[3 * t + 1 + w];
            }
        }
        free(u);
        for (int64_t i = 0; i < n; ++i) {
            double s, ds[2];
            qo_self_potential(p, xyz + 3 * i, &s, ds);
            phi[i] += 2.0 * q[i] * s;
            psi[2 * i] += 2.0 * q[i] * ds[0];
            psi[2 * i + 1] += 2.0 * q[i] * ds[1];
        }
    }
    if (with_long && n > 0) {
        const double A = p->Lx * p->Ly * p->eps_0;
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = 0; j < n; ++j)
                if (j != i) phi[i] -= q[j] * fabs(xyz[3 * i + 2] - xyz[3 * j + 2]) / (2.0 * A);
        const double g1 = p->gamma_1, g2 = p->gamma_2, Lz = p->Lz;
        for (int64_t t = 0; t < P; ++t) {
            const double kx = kvecs[3 * t], ky = kvecs[3 * t + 1], k = kvecs[3 * t + 2];
            const double wk = full_sum ? exp(-(k * k) / (4.0 * p->alpha)) : weight;
            const double c = wk / (A * 2.0 * k);
            const double e2 = exp(-2.0 * k * Lz), x = g1 * g2 * e2;
            double dgk[2] = {0.0, 0.0};
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : dgk[:2])
#endif
            for (int64_t i = 0; i < n; ++i) {
                double acc = 0.0, a1 = 0.0, a2 = 0.0;
                const double thi = kx * xyz[3 * i] + ky * xyz[3 * i + 1], zi = xyz[3 * i + 2];
                for (int64_t j = 0; j < n; ++j) {
                    const double thj = kx * xyz[3 * j] + ky * xyz[3 * j + 1], zj = xyz[3 * j + 2];
                    const double cs = cos(thi - thj);
                    const double dz = zi - zj, zp = zi + zj;
                    const double R = g1 * exp(-k * zp) + g2 * exp(-k * (2.0 * Lz - zp)) +
                                     g1 * g2 * (exp(-k * (2.0 * Lz + dz)) + exp(-k * (2.0 * Lz - dz)));
                    const double dR1 = exp(-k * zp) + g2 * (exp(-k * (2.0 * Lz + dz)) + exp(-k * (2.0 * Lz - dz)));
                    const double dR2 = exp(-k * (2.0 * Lz - zp)) + g1 * (exp(-k * (2.0 * Lz + dz)) + exp(-k * (2.0 * Lz - dz)));
                    if (j != i) acc += q[j] * cs * exp(-k * fabs(dz));
                    acc += q[j] * cs * R / (1.0 - x);
                    a1 += q[j] * cs * (dR1 / (1.0 - x) + R * g2 * e2 / ((1.0 - x) * (1.0 - x)));
                    a2 += q[j] * cs * (dR2 / (1.0 - x) + R * g1 * e2 / ((1.0 - x) * (1.0 - x)));
                }
                phi[i] += c * acc;
                dgk[0] += 0.5 * c * q[i] * a1;
                dgk[1] += 0.5 * c * q[i] * a2;
            }
            dg_long[0] += dgk[0];
            dg_long[1] += dgk[1];
        }
    }
    double E = 0.0, d1 = 0.0, d2 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        E += 0.5 * q[i] * phi[i];
        d1 += 0.5 * q[i] * psi[2 * i];
        d2 += 0.5 * q[i] * psi[2 * i + 1];
    }
    if (E_out) *E_out = E;
    if (phi_out) memcpy(phi_out, phi, sizeof(double) * (size_t)n);
    if (dEdg) {
        dEdg[0] = d1 + dg_long[0];
        dEdg[1] = d2 + dg_long[1];
    }
    free(phi);
    free(psi);
}

/* ---- Substrate_LJ wall ---------------------------------------------------- */

/* substrate_LJ_force (src/Substrate_LJ.jl:19-22) and Molly.forces
 * (src/Substrate_LJ.jl:24-46): per atom dz = z - sub_pos; the @assert on
 * direction * dz > 0; F_z = direction * substrate_LJ_force(1/dz^2, sigma^2,
 * eps) when dz < cutoff (signed dz, as written). */
int qo_substrate_lj(double cutoff, double sigma, double eps, double sub_pos, int direction,
                    unsigned sem, int64_t n, const double *xyz, double *f, int accumulate)
{
    for (int64_t i = 0; i < n; ++i)
        if (!(direction * (xyz[3 * i + 2] - sub_pos) > 0.0)) return 1;
    const double sigma2 = sigma * sigma;
    for (int64_t i = 0; i < n; ++i) {
        const double dz = xyz[3 * i + 2] - sub_pos;
        const double dcut = (sem & 1u) ? fabs(dz) : dz;
        double fz = 0.0;
        if (dcut < cutoff) {
            const double invr2 = 1.0 / (dz * dz);
            const double six = pow(sigma2 * invr2, 3.0);
            fz = (24.0 * eps * invr2) * (2.0 * six * six - six);
            if (sem & 2u) fz *= fabs(dz);
            fz *= (double)direction;
        }
        if (accumulate) {
            f[3 * i + 2] += fz;
        } else {
            f[3 * i] = 0.0;
            f[3 * i + 1] = 0.0;
            f[3 * i + 2] = fz;
        }
    }
    return 0;
}
