// C ABI of libqem_hip.so: contexts, parameter checks, host-side tables
// (Gauss-Legendre rule, K set, alias sampler, Bessel table) and the entry
// points declared in include/qem.h.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <new>

#include "qem_device.h"

namespace qem {

hipError_t ensure(DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 1;
    if (b.cap >= bytes) return hipSuccess;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) return e;
    }
    size_t want = bytes + bytes / 4;  // headroom for slowly growing inputs
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return e;
    }
    b.cap = want;
    return hipSuccess;
}

void release(DevBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// Gauss-Legendre nodes and weights on [-1, 1], ascending (the rule that
// GaussQuadrature.legendre(n) returns, used at src/Gaussian_int.jl:11).
// Newton's method on P_n from Tricomi's initial guesses, long double.
GLRule make_gl_rule(int n)
{
    GLRule r;
    r.n = n;
    r.x.assign((size_t)n, 0.0);
    r.w.assign((size_t)n, 0.0);
    const long double PI = 3.141592653589793238462643383279502884L;
    const int half = (n + 1) / 2;
    for (int i = 1; i <= half; ++i) {
        // i-th largest root
        long double th = PI * (4.0L * i - 1.0L) / (4.0L * n + 2.0L);
        long double x = (1.0L - (n - 1.0L) / (8.0L * n * n * n)) * cosl(th);
        long double pn = 0.0L, dpn = 1.0L;
        for (int it = 0; it < 64; ++it) {
            long double pm1 = 1.0L, p = x;  // P_0, P_1
            for (int k = 2; k <= n; ++k) {
                long double pk = ((2 * k - 1) * x * p - (k - 1) * pm1) / k;
                pm1 = p;
                p = pk;
            }
            if (n == 1) pm1 = 1.0L;
            pn = p;
            dpn = n * (pm1 - x * p) / (1.0L - x * x);
            long double step = pn / dpn;
            x -= step;
            if (fabsl(step) <= 4e-20L * fabsl(x) + 1e-30L) break;
        }
        {
            long double pm1 = 1.0L, p = x;
            for (int k = 2; k <= n; ++k) {
                long double pk = ((2 * k - 1) * x * p - (k - 1) * pm1) / k;
                pm1 = p;
                p = pk;
            }
            if (n == 1) pm1 = 1.0L;
            dpn = n * (pm1 - x * p) / (1.0L - x * x);
        }
        const long double w = 2.0L / ((1.0L - x * x) * dpn * dpn);
        // ascending order: largest root last
        r.x[(size_t)(n - i)] = (double)x;
        r.w[(size_t)(n - i)] = (double)w;
        r.x[(size_t)(i - 1)] = (double)(-x);
        r.w[(size_t)(i - 1)] = (double)w;
    }
    if (n % 2 == 1) r.x[(size_t)(n / 2)] = 0.0;
    return r;
}

// J_order (order 0 or 1) on [m dx, (m+1) dx) as p(t) = c0 + c1 t + ... +
// cD t^D, t in [0, 1): least squares at 20 Chebyshev points of [0, 1]
// (normal equations in long double; J0/J1 from the C library's long double
// j0l/j1l).  Used at low degree (the mixed table), where the normal
// equations are well conditioned.  Two padding intervals past x_max.
std::vector<long double> fit_bessel_table(double x_max, double dx, int D, int order,
                                          int *n_intervals)
{
    constexpr int S = 20, DM = 8;
    if (D < 1 || D >= DM) return {};
    const int nm = (int)std::ceil(x_max / dx) + 2;
    std::vector<long double> tab((size_t)nm * (D + 1), 0.0L);
    long double ts[S];
    for (int i = 0; i < S; ++i)
        ts[i] = 0.5L - 0.5L * cosl(3.141592653589793238462643383279502884L * (i + 0.5L) / S);
    for (int m = 0; m < nm; ++m) {
        long double A[DM][DM + 1] = {};
        for (int i = 0; i < S; ++i) {
            const long double xi = ((long double)m + ts[i]) * (long double)dx;
            const long double f = order == 0 ? ::j0l(xi) : ::j1l(xi);
            long double pw[DM];
            pw[0] = 1.0L;
            for (int k = 1; k <= D; ++k) pw[k] = pw[k - 1] * ts[i];
            for (int r = 0; r <= D; ++r) {
                for (int k = 0; k <= D; ++k) A[r][k] += pw[r] * pw[k];
                A[r][D + 1] += pw[r] * f;
            }
        }
        // Gauss-Jordan with partial pivoting
        for (int col = 0; col <= D; ++col) {
            int piv = col;
            for (int r = col + 1; r <= D; ++r)
                if (fabsl(A[r][col]) > fabsl(A[piv][col])) piv = r;
            for (int k = 0; k <= D + 1; ++k) std::swap(A[col][k], A[piv][k]);
            for (int r = 0; r <= D; ++r) {
                if (r == col) continue;
                const long double f = A[r][col] / A[col][col];
                for (int k = col; k <= D + 1; ++k) A[r][k] -= f * A[col][k];
            }
        }
        for (int k = 0; k <= D; ++k) tab[(size_t)m * (D + 1) + k] = A[k][D + 1] / A[k][k];
    }
    *n_intervals = nm;
    return tab;
}

// J0 on [m dx, (m+1) dx) as p(t) = c0 + c1 t + ... + cD t^D, t in [0, 1),
// by interpolation at the D + 1 Chebyshev points of the interval (J0 from
// the C library's long double j0l): the Chebyshev coefficients by the
// discrete cosine sum, then re-expanded in powers of t, all in long double.
// Unlike the least-squares fit this stays well conditioned at degree 7, so
// the derivative (J1) keeps FP64 accuracy too.  Two padding intervals past
// x_max.
std::vector<long double> interp_j0_table(double x_max, double dx, int D, int *n_intervals)
{
    constexpr int DM = 12;
    if (D < 1 || D >= DM) return {};
    const long double pi = 3.141592653589793238462643383279502884L;
    const int N = D + 1;
    const int nm = (int)std::ceil(x_max / dx) + 2;
    // T_j(2t - 1) as coefficient rows in powers of t
    long double T[DM][DM] = {};
    T[0][0] = 1.0L;
    T[1][0] = -1.0L;
    T[1][1] = 2.0L;
    for (int j = 1; j + 1 < N; ++j)
        for (int k = 0; k < N; ++k) {
            long double v = -2.0L * T[j][k] - T[j - 1][k];
            if (k > 0) v += 4.0L * T[j][k - 1];
            T[j + 1][k] = v;
        }
    std::vector<long double> tab((size_t)nm * N, 0.0L);
    for (int m = 0; m < nm; ++m) {
        long double f[DM], a[DM];
        for (int i = 0; i < N; ++i) {
            const long double u = cosl(pi * (i + 0.5L) / N);
            f[i] = ::j0l(((long double)m + 0.5L * (u + 1.0L)) * (long double)dx);
        }
        for (int j = 0; j < N; ++j) {
            long double acc = 0.0L;
            for (int i = 0; i < N; ++i) acc += f[i] * cosl(pi * j * (i + 0.5L) / N);
            a[j] = (j == 0 ? 1.0L : 2.0L) * acc / N;
        }
        for (int k = 0; k < N; ++k) {
            long double c = 0.0L;
            for (int j = 0; j < N; ++j) c += a[j] * T[j][k];
            tab[(size_t)m * N + k] = c;
        }
    }
    *n_intervals = nm;
    return tab;
}

__global__ void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

__global__ void k_copy1(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    const size_t n16 = bytes / 16;
    if (n16 > 0) {
        size_t blocks = std::min<size_t>((n16 + 255) / 256, 256 * 32);
        k_copy16<<<(unsigned)blocks, 256, 0, s>>>(static_cast<const uint4 *>(src),
                                                  static_cast<uint4 *>(dst), n16);
    }
    const size_t tail = bytes - n16 * 16;
    if (tail > 0)
        k_copy1<<<1, 256, 0, s>>>(static_cast<const uint8_t *>(src) + n16 * 16,
                                  static_cast<uint8_t *>(dst) + n16 * 16, tail);
    return hipGetLastError();
}

// evaluates the FP64 table Bessel pair used by the FP64 kernels
__global__ void k_bessel_tab64(const double *__restrict__ x, int64_t n,
                               const double *__restrict__ tab, int ntab, double inv_dx,
                               double *__restrict__ j0o, double *__restrict__ j1o)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dev::bessel01_tab64(x[i], tab, ntab, inv_dx, j0o[i], j1o[i]);
}

// evaluates the FP32 table Bessel pair used by the mixed-precision kernels
__global__ void k_bessel_tab(const double *__restrict__ x, int64_t n,
                             const float4 *__restrict__ tab, int ntab, float inv_dx,
                             double *__restrict__ j0o, double *__restrict__ j1o)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float J0, J1;
    dev::bessel01_tab((float)x[i], tab, ntab, inv_dx, J0, J1);
    j0o[i] = J0;
    j1o[i] = J1;
}

}  // namespace qem

using namespace qem;

double qem::pole_eps_of(const qem_params &p)
{
    const double k0 = std::log(p.gamma_1 * p.gamma_2) / (2.0 * p.Lz);
    return p.pole_eps > 0.0 ? p.pole_eps : k0 / 20.0;
}

// The segment rule of src/python_lib/Force_short_norm_njit.py (node counts
// Force_short_range_njit.py:258-286): per node k_n the constant
// c_n = w_n k_n pf(k_n) R_n, R_n = 1/D(k_n) off the middle segment and the
// regular part 1/D + 1/x (green_norm_njit.py:13-32, the series of
// (e^{-x} - 1 + x)/x^2 to the 10th power over that of D/x) on it.
bool qem::make_pole_set(const qem_params &p, double eps, bool gauss, PoleSet &out)
{
    const double g12 = p.gamma_1 * p.gamma_2;
    if (!(g12 > 1.0) || p.N_t > 256) return false;
    const double two_lz = 2.0 * p.Lz, la = std::log(p.accuracy);
    const double k0 = std::log(g12) / two_lz;
    const double base = gauss ? std::sqrt(-4.0 * p.alpha * la) : -la / two_lz;
    const double kf = std::max(base, 2.1 * k0);
    const int n1 = (int)((double)p.N_t * k0 / kf), n3 = p.N_t - 2 * n1;
    if (n1 < 1 || n3 < 1 || !(eps > 0.0 && eps < k0)) return false;
    out = PoleSet();
    out.n1 = n1;
    out.n3 = n3;
    out.kf = kf;
    auto pf = [&](double k) {
        return gauss ? std::exp(-(k * k) / (4.0 * p.alpha)) : g12 * std::exp(-two_lz * k);
    };
    auto regular = [](double x) {
        double c[11], f = 2.0;
        for (int m = 0; m <= 10; ++m) {
            c[m] = 1.0 / f;
            f *= (double)(m + 3);
        }
        double u = 0.0;
        for (int m = 10; m >= 0; --m) u = u * (-x) + c[m];
        return u / (-1.0 + x * u);
    };
    const double seg[4][2] = {{0.0, k0 - eps}, {k0 - eps, k0 + eps}, {k0 + eps, 2.0 * k0},
                              {2.0 * k0, kf}};
    const int cnt[4] = {n1, 10, n1, n3};
    for (int sgi = 0; sgi < 4; ++sgi) {
        const GLRule r = make_gl_rule(cnt[sgi]);
        const double a = seg[sgi][0], b = seg[sgi][1];
        for (int i = 0; i < cnt[sgi]; ++i) {
            const double k = 0.5 * (b - a) * r.x[(size_t)i] + 0.5 * (b + a);
            const double w = 0.5 * (b - a) * r.w[(size_t)i];
            const double R = sgi == 1 ? regular(two_lz * (k - k0))
                                      : 1.0 / (g12 * std::exp(-two_lz * k) - 1.0);
            out.k.push_back(k);
            out.c.push_back(w * k * pf(k) * R);
            if (sgi == 1) out.Wm += w;
            if (sgi == 0 || sgi == 2) out.Wx += w / (two_lz * (k - k0));
        }
    }
    return true;
}

namespace {

constexpr double kTwoPi = 6.283185307179586;  // Float64(2pi)
// Bessel table spacing: 1/32 (cubic error ~1e-7 in J1), widened to 1/16
// only when the argument range needs more than kMaxTabIntervals
constexpr double kBesselDx = 1.0 / 32.0;
constexpr double kBesselDxMax = 1.0 / 16.0;
constexpr int kMaxTabIntervals = 4096;        // 64 KiB of LDS (16 B per interval)
constexpr double kBesselDx64 = 1.0 / 16.0;    // FP64 table spacing
constexpr int kMaxTab64Intervals = 910;       // 64 KiB of LDS (72 B per interval)

int set_err(qem_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

int hip_err(qem_ctx *ctx, hipError_t e, const char *what)
{
    return set_err(ctx, QEM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define API_HIP(ctx, expr)                                                    \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) return hip_err((ctx), e_, #expr);               \
    } while (0)

#define API_TRY(ctx, expr)                                                    \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            if ((ctx)->err.empty())                                           \
                (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_); \
            return QEM_ERR_HIP;                                               \
        }                                                                     \
    } while (0)

bool finite_all(const double *a, int64_t m)
{
    for (int64_t i = 0; i < m; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

int check_params(const qem_params *p, std::string &msg)
{
    auto bad = [&](const char *m) {
        msg = m;
        return QEM_ERR_INVALID_ARG;
    };
    const double vals[] = {p->Lx, p->Ly, p->Lz, p->r_cutoff, p->alpha,
                           p->gamma_1, p->gamma_2, p->eps_0, p->accuracy};
    for (double v : vals)
        if (!std::isfinite(v)) return bad("parameters must be finite");
    if (!(p->Lx > 0 && p->Ly > 0 && p->Lz > 0)) return bad("box lengths must be positive");
    if (!(p->r_cutoff > 0)) return bad("r_cutoff must be positive");
    if (!(2.0 * p->r_cutoff < std::min(p->Lx, p->Ly)))
        return bad("r_cutoff must be below min(Lx, Ly)/2 (one minimum image per pair)");
    if (!(p->alpha > 0)) return bad("alpha must be positive");
    if (!(p->accuracy > 0 && p->accuracy < 1)) return bad("accuracy must lie in (0, 1)");
    if (p->eps_0 == 0.0) return bad("eps_0 must be non-zero");
    if (p->N_t < 1 || p->N_t > 256) return bad("N_t must lie in 1..256");
    if (p->iter_period < 1) return bad("iter_period must be >= 1");
    if (p->rbe_p < 0) return bad("rbe_p must be >= 0");
    if (p->precision != QEM_PREC_FP64 && p->precision != QEM_PREC_MIXED)
        return bad("precision must be QEM_PREC_FP64 or QEM_PREC_MIXED");
    if (p->semantics & ~(QEM_SEM_CORRECTED | QEM_SEM_POLE_SPLIT))
        return bad("unknown semantics bits");
    if (!std::isfinite(p->pole_eps)) return bad("pole_eps must be finite");
    const double g12 = p->gamma_1 * p->gamma_2;
    if (g12 > 1.0) {
        // the denominator gamma_1 gamma_2 exp(-2 k Lz) - 1 vanishes at k0 > 0
        if (p->semantics & QEM_SEM_POLE_SPLIT) {
            PoleSet ps;
            if (!make_pole_set(*p, pole_eps_of(*p), false, ps) ||
                !make_pole_set(*p, pole_eps_of(*p), true, ps)) {
                msg = "the pole-splitting rule has no nodes for a segment: it needs "
                      "floor(N_t k0 / k_f) >= 1 and 0 < pole_eps < k0";
                return QEM_ERR_UNSUPPORTED;
            }
        } else if (p->precision == QEM_PREC_MIXED) {
            msg = "gamma_1 gamma_2 > 1 in mixed precision needs QEM_SEM_POLE_SPLIT: "
                  "the quadrature nodes run across the pole of 1/D";
            return QEM_ERR_UNSUPPORTED;
        }
    }
    return QEM_OK;
}


// positions finite, z strictly inside the slab, charges finite
int check_atoms(qem_ctx *ctx, int64_t n, const double *xyz, const double *q)
{
    if (n < 0) return set_err(ctx, QEM_ERR_INVALID_ARG, "n must be >= 0");
    if (n > 0 && (!xyz)) return set_err(ctx, QEM_ERR_INVALID_ARG, "null coordinate pointer");
    if (n > 0 && !finite_all(xyz, 3 * n))
        return set_err(ctx, QEM_ERR_DOMAIN, "coordinates must be finite");
    if (q && n > 0 && !finite_all(q, n))
        return set_err(ctx, QEM_ERR_DOMAIN, "charges must be finite");
    for (int64_t i = 0; i < n; ++i) {
        const double z = xyz[3 * i + 2];
        if (!(z > 0.0 && z < ctx->p.Lz)) {
            char buf[160];
            std::snprintf(buf, sizeof buf, "atom %lld: z = %.17g outside (0, Lz = %.17g)",
                          (long long)i, z, ctx->p.Lz);
            return set_err(ctx, QEM_ERR_DOMAIN, buf);
        }
    }
    return QEM_OK;
}

ShortConsts short_consts(const qem_ctx *ctx, int single)
{
    const qem_params &p = ctx->p;
    ShortConsts c;
    c.Lx = p.Lx;
    c.Ly = p.Ly;
    c.Lz = p.Lz;
    c.rc2 = p.r_cutoff * p.r_cutoff;
    c.alpha = p.alpha;
    c.g1 = p.gamma_1;
    c.g2 = p.gamma_2;
    c.g12 = p.gamma_1 * p.gamma_2;
    c.ln_acc = std::log(p.accuracy);
    c.inv2pieps = 1.0 / (kTwoPi * p.eps_0);
    c.sem = p.semantics;
    c.Nt = p.N_t;
    c.single = single;
    c.jt_n = ctx->sw.jt_n;
    c.jt_dx = ctx->sw.jt_dx;
    c.jt_inv_dx = ctx->sw.jt_dx > 0 ? 1.0f / ctx->sw.jt_dx : 0.0f;
    c.ncon = ctx->sw.ncon.as<float4>();
    c.nconi = ctx->sw.nconi.as<float>();
    c.ncon64 = ctx->sw.ncon64.as<double2>();
    {
        const double log2e = 1.4426950408889634;
        c.d_c2 = -2.0 * p.Lz * log2e;
        c.d_rG = (-log2e * (0.25 / p.alpha)) / (c.d_c2 * c.d_c2);
        c.d_hinvLz = 0.5 / p.Lz;
    }
    c.jd_n = ctx->sw.jd_n;
    c.jd_inv_dx = ctx->sw.jd_dx > 0 ? 1.0 / ctx->sw.jd_dx : 0.0;
    c.jdtab = ctx->sw.jdtab.as<double>();
    c.inv_Lx = 1.0 / p.Lx;
    c.inv_Ly = 1.0 / p.Ly;
    c.f_2Lz = (float)(2.0 * p.Lz);
    c.f_nlnacc = (float)(-c.ln_acc);
    c.f_2alpha = (float)(2.0 * p.alpha);
    c.f_m4alnacc = (float)(-4.0 * p.alpha * c.ln_acc);
    {
        // the same float operations the node loop used to do per pair
        const float log2e = 1.4426950408889634f;
        const float c2 = -2.0f * (float)p.Lz * log2e;
        const float cg = -log2e * (float)(0.25 / p.alpha);
        c.f_c2 = c2;
        c.f_rG = cg / (c2 * c2);
        c.f_hinvLz = (float)(0.5 / p.Lz);
    }
    c.pole = 0;
    c.pm[0] = c.pm[1] = 0;
    c.pnode = nullptr;
    c.pnodef = nullptr;
    c.pk0 = c.pdel = 0.0;
    for (int s = 0; s < 2; ++s) {
        c.pWx[s] = c.pWm[s] = c.pdpf[s] = 0.0;
        c.ppf[s][0] = c.ppf[s][1] = c.ppf[s][2] = 0.0;
    }
    if (pole_active(p)) {
        const double g12 = c.g12, two_lz = 2.0 * p.Lz;
        c.pole = 1;
        c.pm[0] = ctx->pole_m[0];
        c.pm[1] = ctx->pole_m[1];
        c.pnode = ctx->sw.pnode.as<double2>();
        c.pnodef = ctx->sw.pnodef.as<float2>();
        c.pk0 = std::log(g12) / two_lz;
        c.pdel = 0.01 * pole_eps_of(p);
        for (int s = 0; s < 2; ++s) {
            c.pWx[s] = ctx->pole_wx[s];
            c.pWm[s] = ctx->pole_wm[s];
            const double ks[3] = {c.pk0 - c.pdel, c.pk0, c.pk0 + c.pdel};
            for (int t = 0; t < 3; ++t)
                c.ppf[s][t] = s ? std::exp(-(ks[t] * ks[t]) / (4.0 * p.alpha))
                                : g12 * std::exp(-two_lz * ks[t]);
            c.pdpf[s] = s ? -(c.pk0 / (2.0 * p.alpha)) * c.ppf[s][1] : -two_lz * c.ppf[s][1];
        }
    }
    return c;
}

// K_set_generator (src/K_set_generator.jl:5-45): m_x outer, m_y inner,
// (0,0) skipped, k_x = m_x * 2pi / L_x, Float32 weights.
void build_kset(qem_ctx *ctx)
{
    const qem_params &p = ctx->p;
    const double k_c = 2.0 * std::sqrt(-4.0 * p.alpha * std::log(p.accuracy));
    const int64_t mxm = (int64_t)std::ceil(k_c * p.Lx / kTwoPi) + 1;
    const int64_t mym = (int64_t)std::ceil(k_c * p.Ly / kTwoPi) + 1;
    double Sx = 0.0, Sy = 0.0;
    for (int64_t m = -mxm; m <= mxm; ++m)
        if (m != 0) {
            const double kx = (double)m * kTwoPi / p.Lx;
            Sx += std::exp(-(kx * kx) / (4.0 * p.alpha));
        }
    for (int64_t m = -mym; m <= mym; ++m)
        if (m != 0) {
            const double ky = (double)m * kTwoPi / p.Ly;
            Sy += std::exp(-(ky * ky) / (4.0 * p.alpha));
        }
    const double sK = Sx + Sy + Sx * Sy;
    const size_t nk = (size_t)((2 * mxm + 1) * (2 * mym + 1) - 1);
    ctx->kset_mxm = mxm;
    ctx->kset_mym = mym;
    ctx->kset.resize(3 * nk);
    ctx->kprob.resize(nk);
    size_t t = 0;
    for (int64_t mx = -mxm; mx <= mxm; ++mx)
        for (int64_t my = -mym; my <= mym; ++my) {
            const double kx = (double)mx * kTwoPi / p.Lx;
            const double ky = (double)my * kTwoPi / p.Ly;
            const double k = std::sqrt(kx * kx + ky * ky);
            if (k != 0.0) {
                ctx->kset[3 * t] = kx;
                ctx->kset[3 * t + 1] = ky;
                ctx->kset[3 * t + 2] = k;
                ctx->kprob[t] = (float)(std::exp(-(k * k) / (4.0 * p.alpha)) / sK);
                ++t;
            }
        }
    ctx->sum_K = sK;
    // Vose alias table over prob[i] / sum(prob)
    double tot = 0.0;
    for (float v : ctx->kprob) tot += (double)v;
    const size_t n = nk;
    ctx->alias_prob.assign(n, 1.0);
    ctx->alias_idx.assign(n, 0);
    std::vector<double> sc(n);
    std::vector<int64_t> small, large;
    for (size_t i = 0; i < n; ++i) {
        sc[i] = (double)ctx->kprob[i] / tot * (double)n;
        ctx->alias_idx[i] = (int64_t)i;
        (sc[i] < 1.0 ? small : large).push_back((int64_t)i);
    }
    while (!small.empty() && !large.empty()) {
        const int64_t s = small.back(), l = large.back();
        small.pop_back();
        ctx->alias_prob[(size_t)s] = sc[(size_t)s];
        ctx->alias_idx[(size_t)s] = l;
        sc[(size_t)l] = (sc[(size_t)l] + sc[(size_t)s]) - 1.0;
        if (sc[(size_t)l] < 1.0) {
            large.pop_back();
            small.push_back(l);
        }
    }
    for (int64_t i : large) ctx->alias_prob[(size_t)i] = 1.0;
    for (int64_t i : small) ctx->alias_prob[(size_t)i] = 1.0;
}

int ensure_kset(qem_ctx *ctx)
{
    if (!ctx->kset.empty()) return QEM_OK;
    try {
        build_kset(ctx);
    } catch (const std::bad_alloc &) {
        return set_err(ctx, QEM_ERR_NOMEM, "K set: host allocation failed");
    }
    return QEM_OK;
}

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int upload(qem_ctx *ctx, DevBuf &b, const void *h, size_t bytes)
{
    API_HIP(ctx, ensure(b, bytes));
    if (bytes > 0) API_HIP(ctx, hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    return QEM_OK;
}

// Stages P k rows in the pinned buffer and copies them to lw.kv: (k_x, k_y,
// |k|, weight / (Lx Ly eps_0)); with_k0 marks the batch as carrying the
// k = 0 term (lw.k0_on), which has no row: launch_long_k0 computes it and
// the sums add it.
int stage_kvecs(qem_ctx *ctx, const double *kvecs, int64_t P, double weight, bool with_k0,
                hipStream_t s, int64_t *Ptot)
{
    const qem_params &p = ctx->p;
    const int64_t rows = P;
    ctx->lw.k0_on = with_k0;
    const size_t bytes = sizeof(double) * 4 * (size_t)(rows > 0 ? rows : 1);
    LongWork &lw = ctx->lw;
    // ring of pinned staging slots; a slot is rewritten only after the copy
    // that last read it has completed
    const int slot = lw.kv_slot;
    lw.kv_slot = (lw.kv_slot + 1) % LongWork::kSlots;
    if (lw.kv_ev[slot]) API_HIP(ctx, hipEventSynchronize(lw.kv_ev[slot]));
    else API_HIP(ctx, hipEventCreateWithFlags(&lw.kv_ev[slot], hipEventDisableTiming));
    if (lw.h_kv_cap[slot] < bytes) {
        if (lw.h_kv[slot]) (void)hipHostFree(lw.h_kv[slot]);
        lw.h_kv[slot] = nullptr;
        lw.h_kv_cap[slot] = 0;
        API_HIP(ctx, hipHostMalloc(&lw.h_kv[slot], bytes, hipHostMallocDefault));
        lw.h_kv_cap[slot] = bytes;
    }
    double *h = static_cast<double *>(lw.h_kv[slot]);
    const double inv_ae = 1.0 / (p.Lx * p.Ly * p.eps_0);
    for (int64_t t = 0; t < P; ++t) {
        const double kx = kvecs[3 * t], ky = kvecs[3 * t + 1], k = kvecs[3 * t + 2];
        if (!(std::isfinite(kx) && std::isfinite(ky) && std::isfinite(k) && k > 0.0))
            return set_err(ctx, QEM_ERR_DOMAIN, "k vectors must be finite with |k| > 0");
        const double wk = weight >= 0.0 ? weight : std::exp(-(k * k) / (4.0 * p.alpha));
        h[4 * t] = kx;
        h[4 * t + 1] = ky;
        h[4 * t + 2] = k;
        h[4 * t + 3] = wk * inv_ae;
    }
    API_HIP(ctx, ensure(lw.kv, bytes));
    if (rows > 0) API_HIP(ctx, hipMemcpyAsync(lw.kv.p, h, bytes, hipMemcpyHostToDevice, s));
    API_HIP(ctx, hipEventRecord(lw.kv_ev[slot], s));
    *Ptot = rows;
    return QEM_OK;
}

int kidx_to_vecs(qem_ctx *ctx, const int64_t *k_idx, int64_t P, std::vector<double> &out)
{
    int st = ensure_kset(ctx);
    if (st) return st;
    const int64_t nk = (int64_t)ctx->kprob.size();
    out.resize(3 * (size_t)P);
    for (int64_t t = 0; t < P; ++t) {
        const int64_t i = k_idx[t];
        if (i < 0 || i >= nk) {
            char buf[128];
            std::snprintf(buf, sizeof buf, "k index %lld out of range [0, %lld)", (long long)i,
                          (long long)nk);
            return set_err(ctx, QEM_ERR_INVALID_ARG, buf);
        }
        out[3 * t] = ctx->kset[3 * (size_t)i];
        out[3 * t + 1] = ctx->kset[3 * (size_t)i + 1];
        out[3 * t + 2] = ctx->kset[3 * (size_t)i + 2];
    }
    return QEM_OK;
}

hipStream_t pick_stream(qem_ctx *ctx, void *stream)
{
    if (stream == QEM_STREAM_NULL) return nullptr;   // HIP's null stream
    return stream ? static_cast<hipStream_t>(stream) : ctx->stream;
}

// phase timing: a start/stop event pair around each timed launch sequence
hipEvent_t take_event(qem_ctx *ctx)
{
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct PhaseScope {
    qem_ctx *ctx;
    int phase;
    hipStream_t s;
    hipEvent_t a = nullptr;
    PhaseScope(qem_ctx *c, int ph, hipStream_t st) : ctx(c), phase(ph), s(st)
    {
        if (ctx->timing && (a = take_event(ctx))) (void)hipEventRecord(a, s);
    }
    ~PhaseScope()
    {
        if (!a) return;
        hipEvent_t b = take_event(ctx);
        if (!b) {
            ctx->ev_pool.push_back(a);
            return;
        }
        (void)hipEventRecord(b, s);
        ctx->marks.push_back({phase, a, b});
    }
};

// One force call's step: the counter advances and the neighbour list is
// rebuilt when (n_steps - 1) % iter_period == 0 (src/QEM_short.jl:32-36),
// or when n changed or the list was installed by the caller.
int advance_step(qem_ctx *ctx, int64_t n, const double *d_xyz, hipStream_t s)
{
    ctx->n_steps += 1;
    const bool due = ((ctx->n_steps - 1) % ctx->p.iter_period) == 0;
    if (due || ctx->nl.n != n || ctx->nl.has_i_idx) {
        PhaseScope ps(ctx, QEM_PHASE_NEIGHBOR, s);
        hipError_t e = build_neighbor_list(ctx->nl, n, d_xyz, ctx->p.Lx, ctx->p.Ly,
                                           ctx->p.r_cutoff, s, ctx->err);
        if (e != hipSuccess) return hip_err(ctx, e, ctx->err.c_str());
    }
    return QEM_OK;
}

// device part of Molly.forces(::QEM_short) (src/QEM_short.jl:27-59)
// addz (optional, n doubles in atom order): added to the z forces in the
// final pass, after the short-range sum; after_pair (optional) is called
// once the pair kernel is queued, before the gather.  With deferred
// non-null and a slot-ordered list, the final pass is left to the caller:
// the forces stay in nl.sout (slot order, addz not added) and *deferred is
// set.
int short_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q, double *d_f,
              hipStream_t s, bool use_counter, const double *addz = nullptr,
              const std::function<int()> *after_pair = nullptr, bool *deferred = nullptr)
{
    if (n == 0) return QEM_OK;
    if (use_counter) {
        if (int st = advance_step(ctx, n, d_xyz, s)) return st;
    }
    const ShortConsts c = short_consts(ctx, 0);
    API_HIP(ctx, ensure(ctx->sw.fpair, pair_rec_bytes(ctx->nl.npairs, force_rec_bytes(c, ctx->p.precision))));
    NeighborList &nl = ctx->nl;
    const double *px = d_xyz, *pq = d_q;
    double *pf = d_f;
    {
        PhaseScope ps_pair(ctx, QEM_PHASE_PAIR, s);
        if (nl.slot_order) {
            API_HIP(ctx, ensure(nl.sxyz, sizeof(double) * 3 * (size_t)n));
            API_HIP(ctx, ensure(nl.sq, sizeof(double) * (size_t)n));
            API_HIP(ctx, ensure(nl.sout, sizeof(double) * 3 * (size_t)n));
            API_HIP(ctx, launch_atoms_to_slots(nl, n, d_xyz, d_q, nl.sxyz.as<double>(),
                                               nl.sq.as<double>(), s));
            px = nl.sxyz.as<double>();
            pq = nl.sq.as<double>();
            pf = nl.sout.as<double>();
        }
        API_HIP(ctx, launch_pair_forces(c, ctx->p.precision, ctx->sw, nl.pi.as<int32_t>(),
                                        nl.pj.as<int32_t>(), nl.jpos.as<int32_t>(), nl.npairs,
                                        px, pq,
                                        ctx->sw.fpair.as<double>(), s));
    }
    if (after_pair) {
        if (int st = (*after_pair)()) return st;
    }
    {
        PhaseScope ps_g(ctx, QEM_PHASE_GATHER, s);
        API_HIP(ctx, launch_gather_self(c, ctx->p.precision, ctx->sw, nl, n, px, pq,
                                        ctx->sw.fpair.as<double>(), pf, s));
        if (nl.slot_order && deferred) *deferred = true;
        else if (nl.slot_order) API_HIP(ctx, launch_rows_from_slots(nl, n, 3, pf, d_f, s, addz));
        else if (addz) API_HIP(ctx, launch_add_z(n, addz, d_f, s));
    }
    return QEM_OK;
}

// Slot-ordered copies of positions and charges for the current list (the
// inputs themselves for an installed list).
int slot_inputs(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                const double **px, const double **pq, hipStream_t s)
{
    NeighborList &nl = ctx->nl;
    *px = d_xyz;
    *pq = d_q;
    if (!nl.slot_order) return QEM_OK;
    API_HIP(ctx, ensure(nl.sxyz, sizeof(double) * 3 * (size_t)n));
    API_HIP(ctx, ensure(nl.sq, sizeof(double) * (size_t)n));
    API_HIP(ctx, ensure(nl.sout, sizeof(double) * 3 * (size_t)n));
    API_HIP(ctx, launch_atoms_to_slots(nl, n, d_xyz, d_q, nl.sxyz.as<double>(),
                                       nl.sq.as<double>(), s));
    *px = nl.sxyz.as<double>();
    *pq = nl.sq.as<double>();
    return QEM_OK;
}

// k rows per long-range pass: a longer k list (the full sum over a K set of
// ~10^6 vectors) runs in passes that add into the output, so the per-group
// buffers stay at (8192/64) x N x 3 doubles instead of growing with P
constexpr int64_t kLongBatch = 8192;

LongConsts long_consts(const qem_ctx *ctx, double weight, uint32_t parts);

// The long range for at most kLongBatch k vectors, first part: the k rows
// staged (*Ptot receives their number, P), the z sort, the k = 0 term when
// parts has it, and the per-k constants, on stream s.
int long_prep(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
              const double *kvecs, int64_t P, double weight, uint32_t parts, hipStream_t s,
              bool sorted, int64_t *Ptot)
{
    *Ptot = 0;
    if (n == 0) return QEM_OK;
    int st = stage_kvecs(ctx, kvecs, P, weight, (parts & QEM_LONG_K0) != 0, s, Ptot);
    if (st) return st;
    PhaseScope ps_long(ctx, QEM_PHASE_LONG, s);
    if (!sorted) {
        hipError_t e0 = launch_zsort(ctx->lw, n, d_xyz, d_q, s, ctx->err);
        if (e0 != hipSuccess) return hip_err(ctx, e0, "z sort");
    }
    const LongConsts c = long_consts(ctx, weight, parts);
    hipError_t e = ctx->lw.k0_on ? launch_long_k0(c, ctx->lw, n, s, ctx->err) : hipSuccess;
    if (e == hipSuccess) e = launch_long_krec(c, ctx->lw, *Ptot, s, ctx->err);
    if (e != hipSuccess) return hip_err(ctx, e, "long range");
    return QEM_OK;
}

// Second part: the passes up to the per-group partial forces (lw.partB), or
// (which) only their chunk sums and offsets, or only the forces from those.
int long_lanes(qem_ctx *ctx, int64_t n, int64_t Ptot, double weight, uint32_t parts,
               hipStream_t s, int which = kLongPassAll)
{
    if (n == 0) return QEM_OK;
    PhaseScope ps_long(ctx, QEM_PHASE_LONG, s);
    const LongConsts c = long_consts(ctx, weight, parts);
    hipError_t e = launch_long_passes(c, ctx->lw, n, Ptot, s, ctx->err, false, which);
    if (e != hipSuccess) return hip_err(ctx, e, "long range");
    return QEM_OK;
}

// The partial forces of long_passes (Ptot rows) summed into d_f (added with
// accumulate), on stream s.
int long_sum(qem_ctx *ctx, int64_t n, int64_t Ptot, double *d_f, int accumulate, hipStream_t s)
{
    if (n == 0) return QEM_OK;
    PhaseScope ps_long(ctx, QEM_PHASE_LONG, s);
    hipError_t e = launch_long_sum(ctx->lw, n, Ptot, d_f, accumulate, s, ctx->err);
    if (e != hipSuccess) return hip_err(ctx, e, "long range");
    return QEM_OK;
}

int long_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
             const double *kvecs, int64_t P, double weight, uint32_t parts, double *d_f,
             int accumulate, hipStream_t s, bool sorted = false)
{
    if (n == 0) return QEM_OK;
    if (P > kLongBatch) {
        // whole batches first, the k = 0 term with the last one; one z sort
        for (int64_t off = 0; off < P; off += kLongBatch) {
            const int64_t Pb = std::min(kLongBatch, P - off);
            const bool last = off + Pb >= P;
            const uint32_t pb = last ? parts : (parts & ~QEM_LONG_K0);
            int st = long_dev(ctx, n, d_xyz, d_q, kvecs + 3 * off, Pb, weight, pb, d_f,
                              accumulate || off > 0, s, sorted || off > 0);
            if (st) return st;
        }
        return QEM_OK;
    }
    int64_t Ptot = 0;
    int st = long_prep(ctx, n, d_xyz, d_q, kvecs, P, weight, parts, s, sorted, &Ptot);
    if (st == QEM_OK) st = long_lanes(ctx, n, Ptot, weight, parts, s);
    if (st) return st;
    return long_sum(ctx, n, Ptot, d_f, accumulate, s);
}

// QEM_short and QEM_long of one force call on stream s (addz as in
// short_dev).  In mixed precision the long range runs on the side stream
// alongside the short range: the k staging, the z sort (latency-bound, few
// blocks), the k = 0 term and the per-k constants beside the pair kernel,
// the k-lane passes (VALU) once the pair kernel is done, beside the gather
// (HBM); beside the pair kernel, which is VALU-bound, the passes slow it
// by more than their own time.  Their sum into d_f follows the
// short range on s, so d_f holds short + long in that order, as when both
// run on s.  The FP64 pair kernel keeps its 50 KB Bessel table in LDS,
// three blocks per CU; a long-range block beside them (25 KB of LDS for its
// lane sums) does not fit in the 160 KB and would take the place of one, so
// in FP64 the parts run one after the other.
int forces_core(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                const std::vector<double> &kv, int64_t P, double *d_f, hipStream_t s,
                const double *addz = nullptr)
{
    const double w = P > 0 ? ctx->sum_K / (double)P : 0.0;
    if (n == 0 || P > kLongBatch || ctx->p.precision != QEM_PREC_MIXED) {
        int st = short_dev(ctx, n, d_xyz, d_q, d_f, s, true, addz);
        if (st) return st;
        return long_dev(ctx, n, d_xyz, d_q, kv.data(), P, w, QEM_LONG_ALL, d_f, 1, s);
    }
    API_HIP(ctx, hipEventRecord(ctx->fork_ev, s));
    API_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
    int64_t Ptot = 0;
    // queued once the pair kernel is (so that a caller that waits for each
    // call does not see the GPU idle while the host queues this chain ahead
    // of it), and run beside it: k rows, z sort, k = 0 term, per-k
    // constants, then the k-lane passes.  The pair kernel leaves issue slots
    // that the passes' transcendental and FP64 work fills: measured at
    // N = 100k, the pair kernel +125 us and the tail after it (the gather,
    // the output pass) 267 -> 105 us against the passes after the pair
    // kernel beside the gather
    const std::function<int()> lanes = [&]() -> int {
        int st = long_prep(ctx, n, d_xyz, d_q, kv.data(), P, w, QEM_LONG_ALL, ctx->side, false,
                           &Ptot);
        if (st) return st;
        return long_lanes(ctx, n, Ptot, w, QEM_LONG_ALL, ctx->side, kLongPassAll);
    };
    bool deferred = false;
    int st = short_dev(ctx, n, d_xyz, d_q, d_f, s, true, addz, &lanes, &deferred);
    // joined whatever happened, so that s never runs ahead of the side work
    const hipError_t ej = hipEventRecord(ctx->join_ev, ctx->side);
    const hipError_t ew = hipStreamWaitEvent(s, ctx->join_ev, 0);
    if (st) return st;
    if (ej != hipSuccess) return hip_err(ctx, ej, "join event");
    if (ew != hipSuccess) return hip_err(ctx, ew, "join wait");
    if (!deferred) return long_sum(ctx, n, Ptot, d_f, 1, s);
    // the short range's output pass and the long range's sum in one pass
    PhaseScope ps_long(ctx, QEM_PHASE_LONG, s);
    NeighborList &nl = ctx->nl;
    hipError_t e = launch_long_sum_slots(ctx->lw, n, Ptot, nl.slot_of.as<int32_t>(),
                                         nl.sout.as<double>(), addz, d_f, s, ctx->err);
    if (e != hipSuccess) return hip_err(ctx, e, "long range");
    return QEM_OK;
}

// the current neighbour list, built if the context has none for n atoms
// (does not advance the step counter)
int ensure_list(qem_ctx *ctx, int64_t n, const double *d_xyz, hipStream_t s)
{
    if (ctx->nl.n == n && !ctx->nl.has_i_idx) return QEM_OK;
    PhaseScope ps(ctx, QEM_PHASE_NEIGHBOR, s);
    hipError_t e = build_neighbor_list(ctx->nl, n, d_xyz, ctx->p.Lx, ctx->p.Ly, ctx->p.r_cutoff,
                                       s, ctx->err);
    if (e != hipSuccess) return hip_err(ctx, e, ctx->err.c_str());
    return QEM_OK;
}

LongConsts long_consts(const qem_ctx *ctx, double weight, uint32_t parts)
{
    LongConsts c;
    c.Lz = ctx->p.Lz;
    c.g1 = ctx->p.gamma_1;
    c.g2 = ctx->p.gamma_2;
    c.g12 = c.g1 * c.g2;
    c.inv_area_eps = 1.0 / (ctx->p.Lx * ctx->p.Ly * ctx->p.eps_0);
    c.weight = weight;
    c.inv4alpha = 0.25 / ctx->p.alpha;
    c.parts = parts;
    c.precision = ctx->p.precision;
    return c;
}

// ctx->h_out with room for nd doubles (grown, never shrunk)
static int host_out(qem_ctx *ctx, size_t nd, double **out)
{
    if (ctx->h_out_cap < nd) {
        if (ctx->h_out) (void)hipHostFree(ctx->h_out);
        ctx->h_out = nullptr;
        ctx->h_out_cap = 0;
        API_HIP(ctx, hipHostMalloc(reinterpret_cast<void **>(&ctx->h_out), nd * sizeof(double),
                                   hipHostMallocDefault));
        ctx->h_out_cap = nd;
    }
    *out = ctx->h_out;
    return QEM_OK;
}

// E, phi = dE/dq (device, N) and dE/dgamma for the current list + k batch
// the energy and tangent kernels implement the image-series forms for
// |gamma_1 gamma_2| < 1 only
static int check_adjoint(qem_ctx *ctx)
{
    if (!(std::fabs(ctx->p.gamma_1 * ctx->p.gamma_2) < 1.0))
        return set_err(ctx, QEM_ERR_UNSUPPORTED,
                       "energy and Hessian-vector products need |gamma_1 gamma_2| < 1");
    return QEM_OK;
}

int energy_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
               const double *kvecs, int64_t P, double weight, double *d_phi, double *E,
               double *dEdg, hipStream_t s)
{
    if (int st0 = check_adjoint(ctx)) return st0;
    if (n == 0) {
        if (E) *E = 0.0;
        if (dEdg) dEdg[0] = dEdg[1] = 0.0;
        return QEM_OK;
    }
    int st = ensure_list(ctx, n, d_xyz, s);
    if (st) return st;
    const ShortConsts c = short_consts(ctx, 0);
    API_HIP(ctx, ensure(ctx->sw.fpair, pair_rec_bytes(ctx->nl.npairs, ctx->p.precision == QEM_PREC_MIXED ? 12 : 32)));
    API_HIP(ctx, ensure(ctx->d_adj, sizeof(double) * (2 * (size_t)n + 4 + half_dots_scratch())));
    double *psi = ctx->d_adj.as<double>();          // N x 2
    double *out3 = psi + 2 * n;                      // 3 (+1 pad)
    double *part = out3 + 4;                         // reduction scratch
    double *phi = d_phi;
    if (!phi) {
        API_HIP(ctx, ensure(ctx->d_f2, sizeof(double) * (size_t)n));
        phi = ctx->d_f2.as<double>();
    }
    // in mixed precision the long-range passes run on the side stream, as in
    // forces_core: k staging and z sort beside the pair kernel, the passes
    // beside the self term and the gather, the phi sum and the dE/dgamma
    // copy on s after the short range; a batch beyond kLongBatch k vectors
    // runs on s in passes of kLongBatch rows
    const bool overlap = P >= 0 && P <= kLongBatch && ctx->p.precision == QEM_PREC_MIXED;
    const LongConsts lc = long_consts(ctx, weight, QEM_LONG_ALL);
    int64_t Ptot = 0;
    int lst = QEM_OK;
    // queued once the pair kernel and the self term (on the side stream) are,
    // and run beside the pair kernel: k rows, z sort, the k = 0 term and the
    // potential passes (beside the pair kernel rather than after it: the
    // call 3.18 -> 3.08 ms at N = 100k)
    const std::function<void()> passes = [&]() {
        lst = stage_kvecs(ctx, kvecs, P, weight, true, ctx->side, &Ptot);
        if (lst) return;
        {
            PhaseScope ps(ctx, QEM_PHASE_LONG, ctx->side);
            hipError_t e = launch_zsort(ctx->lw, n, d_xyz, d_q, ctx->side, ctx->err);
            if (e == hipSuccess) e = launch_long_k0(lc, ctx->lw, n, ctx->side, ctx->err);
            if (e != hipSuccess) {
                lst = hip_err(ctx, e, "z sort");
                return;
            }
        }
        // the gather on s waits for the self term and passes A and S, queued
        // before it on the side stream
        const std::function<hipError_t()> sums = [&]() -> hipError_t {
            hipError_t r = hipEventRecord(ctx->sums_ev, ctx->side);
            return r == hipSuccess ? hipStreamWaitEvent(s, ctx->sums_ev, 0) : r;
        };
        PhaseScope ps(ctx, QEM_PHASE_LONG, ctx->side);
        const hipError_t e =
            launch_long_energy_passes(lc, ctx->lw, n, Ptot, ctx->side, ctx->err, &sums);
        if (e != hipSuccess) lst = hip_err(ctx, e, "long-range energy");
    };
    {
        PhaseScope ps(ctx, QEM_PHASE_PAIR, s);
        const double *px, *pq;
        if ((st = slot_inputs(ctx, n, d_xyz, d_q, &px, &pq, s))) return st;
        if (overlap) {
            // the side stream sees the slot-ordered inputs
            API_HIP(ctx, hipEventRecord(ctx->fork_ev, s));
            API_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
        }
        NeighborList &nl = ctx->nl;
        // slot order: phi and psi rows land in sout first
        double *phs = nl.slot_order ? nl.sout.as<double>() : phi;
        double *pss = nl.slot_order ? nl.sout.as<double>() + n : psi;
        // with the overlap the self term runs on the side stream beside the
        // pair kernel; the gather waits for the side stream's chunk sums,
        // queued after it
        API_HIP(ctx, launch_short_energy(c, ctx->p.precision, ctx->sw, nl, n, px, pq,
                                         ctx->sw.fpair.as<double>(), phs, pss, s,
                                         overlap ? &passes : nullptr,
                                         overlap ? ctx->side : nullptr));
        if (nl.slot_order) {
            API_HIP(ctx, launch_rows_from_slots(nl, n, 1, phs, phi, s));
            API_HIP(ctx, launch_rows_from_slots(nl, n, 2, pss, psi, s));
        }
    }
    double dg_long[2] = {0.0, 0.0};
    double *hout = nullptr;
    size_t nq0 = 0;   // per-k dE_k/dgamma values copied to hout + 4
    if (overlap) {
        API_HIP(ctx, hipEventRecord(ctx->join_ev, ctx->side));
        API_HIP(ctx, hipStreamWaitEvent(s, ctx->join_ev, 0));
        if (lst) return lst;
        nq0 = 2 * (size_t)Ptot;
        if ((st = host_out(ctx, 4 + nq0, &hout))) return st;
        PhaseScope ps(ctx, QEM_PHASE_LONG, s);
        hipError_t e = launch_long_energy_sum(ctx->lw, n, Ptot, phi, 1, hout + 4, s,
                                              ctx->err);
        if (e != hipSuccess) return hip_err(ctx, e, "long-range energy");
    } else if (P >= 0) {
        PhaseScope ps(ctx, QEM_PHASE_LONG, s);
        hipError_t e = launch_zsort(ctx->lw, n, d_xyz, d_q, s, ctx->err);
        if (e != hipSuccess) return hip_err(ctx, e, "z sort");
        // passes of at most kLongBatch k rows, so the per-row buffers stay
        // bounded; the k = 0 term comes with the last one
        for (int64_t off = 0;; off += kLongBatch) {
            const int64_t Pb = std::min(kLongBatch, P - off);
            const bool last = off + Pb >= P;
            if ((st = stage_kvecs(ctx, kvecs + 3 * off, Pb, weight, last, s, &Ptot))) return st;
            double dgb[2] = {0.0, 0.0};
            e = launch_long_energy(lc, ctx->lw, n, Ptot, phi, 1, dgb, s, ctx->err);
            if (e != hipSuccess) return hip_err(ctx, e, "long-range energy");
            dg_long[0] += dgb[0];
            dg_long[1] += dgb[1];
            if (last) break;
        }
    }
    if (!hout && (st = host_out(ctx, 4, &hout))) return st;
    API_HIP(ctx, launch_half_dots(n, d_q, phi, psi, out3, part, s));
    API_HIP(ctx, hipMemcpyAsync(hout, out3, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    API_HIP(ctx, hipStreamSynchronize(s));
    for (size_t p = 0; p < nq0; p += 2) {
        dg_long[0] += hout[4 + p];
        dg_long[1] += hout[4 + p + 1];
    }
    if (E) *E = hout[0];
    if (dEdg) {
        dEdg[0] = hout[1] + dg_long[0];
        dEdg[1] = hout[2] + dg_long[1];
    }
    return QEM_OK;
}

// hv = -(dF/dx) v for the current list + k batch
// saved: use the node sums of the last saving forces + energy call when they
// belong to this atom count and neighbour list (mixed precision); the caller
// passes the positions, charges and k rows of that call
int hvp_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q, const double *kvecs,
            int64_t P, double weight, const double *d_v, double *d_hv, hipStream_t s,
            bool saved = false)
{
    if (int st0 = check_adjoint(ctx)) return st0;
    if (n == 0) return QEM_OK;
    int st = ensure_list(ctx, n, d_xyz, s);
    if (st) return st;
    const float *sums =
        (saved && ctx->tsums_ok && ctx->tsums_n == n && ctx->tsums_np == ctx->nl.npairs &&
         ctx->tsums_builds == ctx->nl.builds && ctx->p.precision == QEM_PREC_MIXED)
            ? ctx->tsums.as<float>()
            : nullptr;
    const ShortConsts c = short_consts(ctx, 0);
    API_HIP(ctx, ensure(ctx->sw.fpair, pair_rec_bytes(ctx->nl.npairs, ctx->p.precision == QEM_PREC_MIXED ? 12 : 32)));
    const LongConsts lc = long_consts(ctx, weight, QEM_LONG_NONSORT | QEM_LONG_SORT | QEM_LONG_SELF);
    // in mixed precision the long-range tangent runs on the side stream, as
    // in forces_core: k staging and z sort beside the pair tangent kernel,
    // the tangent passes beside the self term and the gather, the sum into
    // d_hv on s after the short range; a batch beyond kLongBatch k vectors
    // runs on s in passes of kLongBatch rows
    const bool overlap = P > 0 && P <= kLongBatch && ctx->p.precision == QEM_PREC_MIXED;
    int64_t Ptot = 0;
    int lst = QEM_OK;
    // queued once the pair tangent kernel and the self term (on the side
    // stream) are, and run beside the pair kernel: k rows, z sort and the
    // tangent passes (beside the pair kernel rather than after it: the HVP
    // 4.35 -> 4.25 ms at N = 100k)
    const std::function<void()> lanes = [&]() {
        lst = stage_kvecs(ctx, kvecs, P, weight, false, ctx->side, &Ptot);
        if (lst) return;
        {
            PhaseScope ps(ctx, QEM_PHASE_LONG, ctx->side);
            hipError_t e = launch_zsort(ctx->lw, n, d_xyz, d_q, ctx->side, ctx->err);
            if (e != hipSuccess) {
                lst = hip_err(ctx, e, "z sort");
                return;
            }
        }
        // the gather on s waits for the self term and passes A and S, queued
        // before it on the side stream
        const std::function<hipError_t()> sums = [&]() -> hipError_t {
            hipError_t r = hipEventRecord(ctx->sums_ev, ctx->side);
            return r == hipSuccess ? hipStreamWaitEvent(s, ctx->sums_ev, 0) : r;
        };
        PhaseScope ps(ctx, QEM_PHASE_LONG, ctx->side);
        const hipError_t e = launch_long_tangent(lc, ctx->lw, n, Ptot, d_v, -1.0, d_hv, 1,
                                                 ctx->side, ctx->err, false, &sums);
        if (e != hipSuccess) lst = hip_err(ctx, e, "long-range tangent");
    };
    {
        PhaseScope ps(ctx, QEM_PHASE_PAIR, s);
        const double *px, *pq;
        if ((st = slot_inputs(ctx, n, d_xyz, d_q, &px, &pq, s))) return st;
        NeighborList &nl = ctx->nl;
        const double *pv = d_v;
        double *po = d_hv;
        if (nl.slot_order) {
            API_HIP(ctx, ensure(nl.sv, sizeof(double) * 3 * (size_t)n));
            API_HIP(ctx, launch_rows_to_slots(nl, n, 3, d_v, nl.sv.as<double>(), s));
            pv = nl.sv.as<double>();
            po = nl.sout.as<double>();
        }
        if (overlap) {
            // the side stream sees the slot-ordered inputs
            API_HIP(ctx, hipEventRecord(ctx->fork_ev, s));
            API_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
        }
        // with the overlap the self term runs on the side stream beside the
        // pair kernel; the gather waits for the side stream's chunk sums,
        // queued after it
        API_HIP(ctx, launch_short_tangent(c, ctx->p.precision, ctx->sw, nl, n, px, pq, pv,
                                          ctx->sw.fpair.as<double>(), -1.0, po, s,
                                          overlap ? &lanes : nullptr,
                                          overlap ? ctx->side : nullptr, sums));
        if (nl.slot_order) API_HIP(ctx, launch_rows_from_slots(nl, n, 3, po, d_hv, s));
    }
    if (overlap) {
        API_HIP(ctx, hipEventRecord(ctx->join_ev, ctx->side));
        API_HIP(ctx, hipStreamWaitEvent(s, ctx->join_ev, 0));
        if (lst) return lst;
        PhaseScope ps(ctx, QEM_PHASE_LONG, s);
        hipError_t e = launch_long_tangent_sum(ctx->lw, n, Ptot, -1.0, d_hv, 1, s, ctx->err);
        if (e != hipSuccess) return hip_err(ctx, e, "long-range tangent");
    } else if (P > 0) {
        PhaseScope ps(ctx, QEM_PHASE_LONG, s);
        hipError_t e = launch_zsort(ctx->lw, n, d_xyz, d_q, s, ctx->err);
        if (e != hipSuccess) return hip_err(ctx, e, "z sort");
        // passes of at most kLongBatch k rows, each added into d_hv
        for (int64_t off = 0; off < P; off += kLongBatch) {
            const int64_t Pb = std::min(kLongBatch, P - off);
            if ((st = stage_kvecs(ctx, kvecs + 3 * off, Pb, weight, false, s, &Ptot))) return st;
            e = launch_long_tangent(lc, ctx->lw, n, Ptot, d_v, -1.0, d_hv, 1, s, ctx->err);
            if (e != hipSuccess) return hip_err(ctx, e, "long-range tangent");
        }
    }
    return QEM_OK;
}

// Forces, energy, dE/dq and dE/dgamma of one step.  Mixed precision (batch
// of at most kLongBatch k vectors): one pair pass for both, the force's node
// loops also forming the energy sums; on the side stream the long range's
// passes A and S serve both the force pass B and the potential pass.  The
// forces are those of forces_core bit for bit.  Otherwise: forces_core,
// then energy_dev.  The step counter advances as for a force call.
int forces_energy_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                      const std::vector<double> &kv, int64_t P, double *d_f, double *d_phi,
                      double *E, double *dEdg, hipStream_t s, bool save = false)
{
    if (int st0 = check_adjoint(ctx)) return st0;
    ctx->tsums_ok = false;
    const double w = P > 0 ? ctx->sum_K / (double)P : 0.0;
    if (n == 0) {
        if (E) *E = 0.0;
        if (dEdg) dEdg[0] = dEdg[1] = 0.0;
        return QEM_OK;
    }
    if (ctx->p.precision != QEM_PREC_MIXED || P > kLongBatch || pole_active(ctx->p)) {
        int st = forces_core(ctx, n, d_xyz, d_q, kv, P, d_f, s);
        if (st) return st;
        return energy_dev(ctx, n, d_xyz, d_q, kv.data(), P, w, d_phi, E, dEdg, s);
    }
    if (int st = advance_step(ctx, n, d_xyz, s)) return st;
    NeighborList &nl = ctx->nl;
    const ShortConsts c = short_consts(ctx, 0);
    const LongConsts lc = long_consts(ctx, w, QEM_LONG_ALL);
    API_HIP(ctx, ensure(ctx->sw.fpair, pair_rec_bytes(nl.npairs, 16)));
    const int ngp = pair_fe_blocks(nl.npairs);
    API_HIP(ctx, ensure(ctx->d_adj, sizeof(double) * (2 * (size_t)n + 4 + half_dots_scratch() +
                                                      2 * (size_t)ngp)));
    double *psi = ctx->d_adj.as<double>();          // N x 2
    double *out3 = psi + 2 * n;                      // 3 (+1 pad)
    double *part = out3 + 4;                         // reduction scratch
    double *gpart = part + half_dots_scratch();      // the pair kernel's dE/dgamma per block
    double *phi = d_phi;
    if (!phi) {
        API_HIP(ctx, ensure(ctx->d_f2, sizeof(double) * (size_t)n));
        phi = ctx->d_f2.as<double>();
    }
    float *sums = nullptr;
    if (save && nl.npairs > 0) {
        API_HIP(ctx, ensure(ctx->tsums, tan_sums_bytes(nl.npairs)));
        sums = ctx->tsums.as<float>();
    }
    int64_t Ptot = 0;
    int lst = QEM_OK;
    int st = QEM_OK;
    const double *px = d_xyz, *pq = d_q;
    {
        PhaseScope ps(ctx, QEM_PHASE_PAIR, s);
        st = slot_inputs(ctx, n, d_xyz, d_q, &px, &pq, s);
        if (st == QEM_OK) {
            // the side stream sees the slot-ordered inputs
            hipError_t e = hipEventRecord(ctx->fork_ev, s);
            if (e == hipSuccess) e = hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0);
            if (e == hipSuccess) e = ensure(nl.sout2, sizeof(double) * 3 * (size_t)n);
            if (e == hipSuccess)
                e = launch_pair_forces_energy(c, ctx->sw, nl.pi.as<int32_t>(), nl.pj.as<int32_t>(),
                                              nl.jpos.as<int32_t>(), nl.npairs, px, pq,
                                              ctx->sw.fpair.as<double>(), gpart, s, sums);
            if (e != hipSuccess) st = hip_err(ctx, e, "fused pair kernel");
        }
    }
    // side stream, queued after the pair kernel and run beside it: the self
    // energies (phi and psi of every atom, in slot order), k rows, z sort,
    // the k = 0 term, per-k constants
    if (st == QEM_OK) {
        hipError_t e = launch_self_energy(c, QEM_PREC_MIXED, ctx->sw, n, px, pq,
                                          nl.sout2.as<double>(), nl.sout2.as<double>() + n,
                                          ctx->side);
        if (e != hipSuccess) lst = hip_err(ctx, e, "self energy");
    }
    if (st == QEM_OK && lst == QEM_OK)
        lst = long_prep(ctx, n, d_xyz, d_q, kv.data(), P, w, QEM_LONG_ALL, ctx->side, false,
                        &Ptot);
    // side stream, beside the pair kernel as well (measured at N = 100k:
    // the call 3.97 -> 3.94 ms against waiting for the pair kernel): passes
    // A and S, which the gather waits for, then pass B on their chunk sums,
    // forces and potential in one walk
    if (lst == QEM_OK && st == QEM_OK) {
        hipError_t e = hipSuccess;
        lst = long_lanes(ctx, n, Ptot, w, QEM_LONG_ALL, ctx->side, kLongPassAS);
        if (lst == QEM_OK) {
            e = hipEventRecord(ctx->sums_ev, ctx->side);
            if (e == hipSuccess) e = hipStreamWaitEvent(s, ctx->sums_ev, 0);
            if (e != hipSuccess) lst = hip_err(ctx, e, "sums event");
        }
        if (lst == QEM_OK) {
            PhaseScope ps(ctx, QEM_PHASE_LONG, ctx->side);
            e = launch_long_forces_energy_B(lc, ctx->lw, n, Ptot, ctx->side, ctx->err);
            if (e != hipSuccess) lst = hip_err(ctx, e, "long-range forces and energy");
        }
    }
    // call's stream, once passes A and S (queued after the self energies on
    // the side stream) are done: one gather for the forces (self force
    // included) and the pairs' share of dE/dq, in slot order
    if (st == QEM_OK) {
        PhaseScope ps(ctx, QEM_PHASE_GATHER, s);
        hipError_t e = launch_gather_forces_energy(c, ctx->sw, nl, n, px, pq,
                                                   ctx->sw.fpair.as<double>(),
                                                   nl.sout.as<double>(), nl.sout2.as<double>(),
                                                   s);
        if (e != hipSuccess) st = hip_err(ctx, e, "gather");
    }
    // joined whatever happened, so that s never runs ahead of the side work
    const hipError_t ej = hipEventRecord(ctx->join_ev, ctx->side);
    const hipError_t ew = hipStreamWaitEvent(s, ctx->join_ev, 0);
    if (st) return st;
    if (lst) return lst;
    if (ej != hipSuccess) return hip_err(ctx, ej, "join event");
    if (ew != hipSuccess) return hip_err(ctx, ew, "join wait");
    const size_t nq0 = 2 * (size_t)Ptot;   // per-k dE_k/dgamma values copied to hout + 4
    double *hout = nullptr;
    if (int st2 = host_out(ctx, 4 + nq0, &hout)) return st2;
    {
        PhaseScope ps(ctx, QEM_PHASE_LONG, s);
        hipError_t e = launch_long_sum_slots(ctx->lw, n, Ptot, nl.slot_of.as<int32_t>(),
                                             nl.sout.as<double>(), nullptr, d_f, s, ctx->err);
        if (e == hipSuccess) e = launch_rows_from_slots(nl, n, 1, nl.sout2.as<double>(), phi, s);
        if (e == hipSuccess)
            e = launch_rows_from_slots(nl, n, 2, nl.sout2.as<double>() + n, psi, s);
        if (e == hipSuccess)
            e = launch_long_energy_sum(ctx->lw, n, Ptot, phi, 1, hout + 4, s, ctx->err);
        if (e != hipSuccess) return hip_err(ctx, e, "output passes");
    }
    API_HIP(ctx, launch_half_dots(n, d_q, phi, psi, out3, part, s, gpart, ngp));
    API_HIP(ctx, hipMemcpyAsync(hout, out3, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    API_HIP(ctx, hipStreamSynchronize(s));
    double dg_long[2] = {0.0, 0.0};
    for (size_t q = 0; q < nq0; q += 2) {
        dg_long[0] += hout[4 + q];
        dg_long[1] += hout[4 + q + 1];
    }
    if (E) *E = hout[0];
    if (dEdg) {
        dEdg[0] = hout[1] + dg_long[0];
        dEdg[1] = hout[2] + dg_long[1];
    }
    if (save) {
        ctx->tsums_ok = true;
        ctx->tsums_n = n;
        ctx->tsums_np = nl.npairs;
        ctx->tsums_builds = nl.builds;
    }
    return QEM_OK;
}

#define CTX_GUARD(ctx)                                                        \
    do {                                                                      \
        if (!(ctx)) return QEM_ERR_INVALID_ARG;                               \
        (ctx)->err.clear();                                                   \
        hipError_t e0_ = hipSetDevice((ctx)->p.device);                       \
        if (e0_ != hipSuccess) return hip_err((ctx), e0_, "hipSetDevice");   \
    } while (0)

}  // namespace

extern "C" {

void qem_default_params(qem_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->Lx = p->Ly = p->Lz = 1.0;
    p->r_cutoff = 0.25;
    p->alpha = 1.0;
    p->gamma_1 = 0.0;
    p->gamma_2 = 0.0;
    p->eps_0 = 1.0;
    p->accuracy = 1e-6;
    p->N_t = 30;
    p->iter_period = 100;
    p->rbe_p = 30;
    p->precision = QEM_PREC_FP64;
    p->semantics = 0;
    p->device = 0;
}

const char *qem_status_string(int status)
{
    switch (status) {
    case QEM_OK: return "ok";
    case QEM_ERR_INVALID_ARG: return "invalid argument";
    case QEM_ERR_DOMAIN: return "input outside the domain";
    case QEM_ERR_HIP: return "HIP runtime error";
    case QEM_ERR_NOMEM: return "out of memory";
    case QEM_ERR_UNSUPPORTED: return "unsupported parameters";
    case QEM_ERR_CAPACITY: return "buffer too small";
    default: return "unknown status";
    }
}

int qem_inter_params(int64_t n_atoms, double Lx, double Ly, double s, int32_t rbe_mode,
                     double *alpha, double *r_cutoff, double *k_cutoff)
{
    if (!alpha || !r_cutoff || !k_cutoff || n_atoms <= 0 || !(Lx > 0) || !(Ly > 0) || !(s > 0))
        return QEM_ERR_INVALID_ARG;
    const double a = rbe_mode ? (double)n_atoms / (Lx * Ly) : std::sqrt((double)n_atoms) / (Lx * Ly);
    *alpha = a;
    *r_cutoff = s / std::sqrt(a);
    *k_cutoff = 2.0 * s * std::sqrt(a);
    return QEM_OK;
}

int qem_create(const qem_params *p, qem_ctx **out)
{
    if (!p || !out) return QEM_ERR_INVALID_ARG;
    *out = nullptr;
    std::string msg;
    int st = check_params(p, msg);
    if (st) return st;
    qem_ctx *ctx = new (std::nothrow) qem_ctx();
    if (!ctx) return QEM_ERR_NOMEM;
    ctx->p = *p;
    // room for the error messages, so that setting one does not allocate
    try {
        ctx->err.reserve(512);
    } catch (const std::bad_alloc &) {
        delete ctx;
        return QEM_ERR_NOMEM;
    }
    auto fail = [&](int code) {
        qem_destroy(ctx);
        return code;
    };
    if (hipSetDevice(p->device) != hipSuccess) return fail(QEM_ERR_HIP);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->sums_ev, hipEventDisableTiming) != hipSuccess)
        return fail(QEM_ERR_HIP);
    ctx->gl = make_gl_rule(p->N_t);
    std::vector<float> xf(ctx->gl.x.begin(), ctx->gl.x.end()), wf(ctx->gl.w.begin(), ctx->gl.w.end());
    std::vector<float> nc(4 * (size_t)p->N_t);
    for (int n = 0; n < p->N_t; ++n) {
        const double u = 1.0 + ctx->gl.x[(size_t)n];
        nc[4 * (size_t)n] = (float)u;
        nc[4 * (size_t)n + 1] = (float)(u * u);
        nc[4 * (size_t)n + 2] = (float)std::log2(ctx->gl.w[(size_t)n]);
        nc[4 * (size_t)n + 3] = (float)std::log2(ctx->gl.w[(size_t)n] * u);
    }
    if (upload(ctx, ctx->sw.ncon, nc.data(), sizeof(float) * nc.size())) return fail(QEM_ERR_HIP);
    {
        // FP64 node constants from long double: 1 + x_n and log2(w_n (1 + x_n))
        std::vector<double> n64(2 * (size_t)p->N_t);
        for (int n = 0; n < p->N_t; ++n) {
            const long double u = 1.0L + (long double)ctx->gl.x[(size_t)n];
            n64[2 * (size_t)n] = (double)u;
            n64[2 * (size_t)n + 1] = (double)log2l((long double)ctx->gl.w[(size_t)n] * u);
        }
        if (upload(ctx, ctx->sw.ncon64, n64.data(), sizeof(double) * n64.size()))
            return fail(QEM_ERR_HIP);
    }
    std::vector<float> nci((size_t)p->N_t);
    for (int n = 0; n < p->N_t; ++n) nci[(size_t)n] = (float)(1.0 / (1.0 + ctx->gl.x[(size_t)n]));
    if (upload(ctx, ctx->sw.nconi, nci.data(), sizeof(float) * nci.size())) return fail(QEM_ERR_HIP);
    if (upload(ctx, ctx->sw.glx, ctx->gl.x.data(), sizeof(double) * p->N_t) ||
        upload(ctx, ctx->sw.glw, ctx->gl.w.data(), sizeof(double) * p->N_t) ||
        upload(ctx, ctx->sw.glxf, xf.data(), sizeof(float) * p->N_t) ||
        upload(ctx, ctx->sw.glwf, wf.data(), sizeof(float) * p->N_t))
        return fail(QEM_ERR_HIP);
    // the pole rule's nodes (pair-independent), point set then gaussian set
    double kmax_pole = 0.0;
    if (pole_active(*p)) {
        PoleSet ps[2];
        const double eps = pole_eps_of(*p);
        if (!make_pole_set(*p, eps, false, ps[0]) || !make_pole_set(*p, eps, true, ps[1])) {
            qem_destroy(ctx);
            return QEM_ERR_UNSUPPORTED;
        }
        std::vector<double> nd;
        std::vector<float> ndf;
        for (int s = 0; s < 2; ++s) {
            ctx->pole_m[s] = (int)ps[s].k.size();
            ctx->pole_wx[s] = ps[s].Wx;
            ctx->pole_wm[s] = ps[s].Wm;
            kmax_pole = std::max(kmax_pole, ps[s].kf);
            for (size_t n = 0; n < ps[s].k.size(); ++n) {
                nd.push_back(ps[s].k[n]);
                nd.push_back(ps[s].c[n]);
                ndf.push_back((float)ps[s].k[n]);
                ndf.push_back((float)ps[s].c[n]);
            }
        }
        if (upload(ctx, ctx->sw.pnode, nd.data(), sizeof(double) * nd.size()) ||
            upload(ctx, ctx->sw.pnodef, ndf.data(), sizeof(float) * ndf.size()))
            return fail(QEM_ERR_HIP);
    }
    if (p->precision == QEM_PREC_MIXED) {
        // largest Bessel argument: r_cutoff times the largest limit
        const double la = -std::log(p->accuracy);
        const double kmax = std::max({std::sqrt(4.0 * p->alpha * la), la / (2.0 * p->Lz),
                                      kmax_pole});
        double dx = kBesselDx;
        while (p->r_cutoff * kmax * (1.0 + 1e-6) / dx + 4 > kMaxTabIntervals && dx < kBesselDxMax)
            dx *= 2.0;
        const double x_max = p->r_cutoff * kmax * (1.0 + 1e-6) + dx;
        int nm = 0;
        if (x_max / dx + 2 > kMaxTabIntervals) {
            qem_destroy(ctx);
            return QEM_ERR_UNSUPPORTED;
        }
        const std::vector<long double> fit = fit_bessel_table(x_max, dx, 3, 0, &nm);
        const std::vector<float> tab(fit.begin(), fit.end());
        ctx->sw.jt_n = nm;
        ctx->sw.jt_dx = (float)dx;
        if (upload(ctx, ctx->sw.jtab, tab.data(), sizeof(float) * tab.size()))
            return fail(QEM_ERR_HIP);
    } else {
        // FP64: J0 as a degree-7 interpolant per 1/16-wide interval, J1 from
        // its derivative (both ~5e-16 absolute); when the argument range needs
        // more than kMaxTab64Intervals the kernels call the device library's
        // j0/j1
        const double la = -std::log(p->accuracy);
        const double kmax = std::max({std::sqrt(4.0 * p->alpha * la), la / (2.0 * p->Lz),
                                      kmax_pole});
        const double x_max = p->r_cutoff * kmax * (1.0 + 1e-6) + kBesselDx64;
        if (x_max / kBesselDx64 + 2 <= kMaxTab64Intervals) {
            int nm = 0;
            const std::vector<long double> fit = interp_j0_table(x_max, kBesselDx64, 7, &nm);
            // rows of 9 doubles (8 coefficients + padding): consecutive rows
            // start 18 LDS banks apart, which spreads random lookups over the
            // banks (a 16-dword stride would put every row on 4 banks)
            std::vector<double> tab(9 * (size_t)nm, 0.0);
            for (size_t m = 0; m < (size_t)nm; ++m)
                for (int k = 0; k < 8; ++k) tab[9 * m + k] = (double)fit[8 * m + k];
            ctx->sw.jd_n = nm;
            ctx->sw.jd_dx = kBesselDx64;
            if (upload(ctx, ctx->sw.jdtab, tab.data(), sizeof(double) * tab.size()))
                return fail(QEM_ERR_HIP);
        }
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(QEM_ERR_HIP);
    if (p->rbe_p > 0 && ensure_kset(ctx)) return fail(QEM_ERR_NOMEM);
    *out = ctx;
    return QEM_OK;
}

static std::vector<DevBuf *> all_buffers(qem_ctx *ctx)
{
    return {&ctx->sw.ncon, &ctx->sw.nconi, &ctx->sw.ncon64, &ctx->nl.sout2, &ctx->lw.partE, &ctx->sw.glx, &ctx->sw.glw, &ctx->sw.glxf, &ctx->sw.glwf, &ctx->sw.jtab, &ctx->sw.jdtab,
                      &ctx->sw.pnode, &ctx->sw.pnodef,
                      &ctx->sw.fpair, &ctx->nl.pi, &ctx->nl.pj, &ctx->nl.i_start, &ctx->nl.i_end,
                      &ctx->nl.i_idx, &ctx->nl.j_start, &ctx->nl.j_end, &ctx->nl.j_idx, &ctx->nl.jpos,
                      &ctx->nl.slot_of,
                      &ctx->nl.sxyz, &ctx->nl.sq, &ctx->nl.sv, &ctx->nl.sout,
                      &ctx->nl.cell_key, &ctx->nl.cell_key_sorted,
                      &ctx->nl.atom_idx_sorted, &ctx->nl.cell_start, &ctx->nl.cell_end,
                      &ctx->nl.counts, &ctx->nl.row_ptr, &ctx->nl.tmp,
                      &ctx->lw.perm, &ctx->lw.tmp,
                      &ctx->lw.satom, &ctx->lw.k0tile, &ctx->lw.k0out,
                      &ctx->lw.kr, &ctx->lw.kv,
                      &ctx->lw.partA, &ctx->lw.offA, &ctx->lw.totA, &ctx->lw.q0part, &ctx->tsums,
                      &ctx->lw.partB, &ctx->lw.svx, &ctx->lw.svy,
                      &ctx->lw.svz, &ctx->lw.partT, &ctx->lw.offT, &ctx->lw.totT, &ctx->d_xyz,
                      &ctx->d_q, &ctx->d_f, &ctx->d_f2, &ctx->d_aux, &ctx->d_adj,
                      &ctx->d_flag, &ctx->d_wz};
}

int qem_device_bytes(qem_ctx *ctx, int64_t *bytes)
{
    if (!ctx || !bytes) return QEM_ERR_INVALID_ARG;
    size_t t = 0;
    for (DevBuf *b : all_buffers(ctx)) t += b->cap;
    *bytes = (int64_t)t;
    return QEM_OK;
}

int qem_destroy(qem_ctx *ctx)
{
    if (!ctx) return QEM_OK;
    (void)hipSetDevice(ctx->p.device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    for (DevBuf *b : all_buffers(ctx)) release(*b);
    for (int t = 0; t < LongWork::kSlots; ++t) {
        if (ctx->lw.h_kv[t]) (void)hipHostFree(ctx->lw.h_kv[t]);
        if (ctx->lw.kv_ev[t]) (void)hipEventDestroy(ctx->lw.kv_ev[t]);
    }
    if (ctx->h_flag) (void)hipHostFree(ctx->h_flag);
    if (ctx->h_out) (void)hipHostFree(ctx->h_out);
    if (ctx->wall_ev) (void)hipEventDestroy(ctx->wall_ev);
    for (auto &m : ctx->marks) {
        (void)hipEventDestroy(m.a);
        (void)hipEventDestroy(m.b);
    }
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
    if (ctx->sums_ev) (void)hipEventDestroy(ctx->sums_ev);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return QEM_OK;
}

const char *qem_last_error(const qem_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

int qem_get_params(const qem_ctx *ctx, qem_params *p)
{
    if (!ctx || !p) return QEM_ERR_INVALID_ARG;
    *p = ctx->p;
    return QEM_OK;
}

int qem_kset_size(qem_ctx *ctx, int64_t *n)
{
    if (!ctx || !n) return QEM_ERR_INVALID_ARG;
    int st = ensure_kset(ctx);
    if (st) return st;
    *n = (int64_t)ctx->kprob.size();
    return QEM_OK;
}

int qem_kset_copy(qem_ctx *ctx, double *kxkyk, float *prob, double *sum_K)
{
    if (!ctx) return QEM_ERR_INVALID_ARG;
    int st = ensure_kset(ctx);
    if (st) return st;
    if (kxkyk) std::memcpy(kxkyk, ctx->kset.data(), sizeof(double) * ctx->kset.size());
    if (prob) std::memcpy(prob, ctx->kprob.data(), sizeof(float) * ctx->kprob.size());
    if (sum_K) *sum_K = ctx->sum_K;
    return QEM_OK;
}

int qem_sample_k(qem_ctx *ctx, uint64_t seed, int64_t P, int64_t *k_idx)
{
    if (!ctx || P < 0 || (P > 0 && !k_idx)) return QEM_ERR_INVALID_ARG;
    int st = ensure_kset(ctx);
    if (st) return st;
    const uint64_t n = (uint64_t)ctx->kprob.size();
    uint64_t s = seed;
    for (int64_t t = 0; t < P; ++t) {
        const double u1 = (double)(splitmix64(s) >> 11) * 0x1.0p-53;
        const double u2 = (double)(splitmix64(s) >> 11) * 0x1.0p-53;
        uint64_t j = (uint64_t)(u1 * (double)n);
        if (j >= n) j = n - 1;
        k_idx[t] = u2 < ctx->alias_prob[j] ? (int64_t)j : ctx->alias_idx[j];
    }
    return QEM_OK;
}

int qem_neighbor_list(qem_ctx *ctx, int64_t n, const double *xyz, int64_t cap, int32_t *pairs,
                      int64_t *npairs)
{
    CTX_GUARD(ctx);
    if (!npairs || cap < 0 || (cap > 0 && !pairs))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad output arguments");
    int st = check_atoms(ctx, n, xyz, nullptr);
    if (st) return st;
    *npairs = 0;
    if (n == 0) return QEM_OK;
    st = upload(ctx, ctx->d_xyz, xyz, sizeof(double) * 3 * (size_t)n);
    if (st) return st;
    hipError_t e = build_neighbor_list(ctx->nl, n, ctx->d_xyz.as<double>(), ctx->p.Lx, ctx->p.Ly,
                                       ctx->p.r_cutoff, ctx->stream, ctx->err);
    if (e != hipSuccess) return hip_err(ctx, e, ctx->err.c_str());
    const int64_t np = ctx->nl.npairs;
    *npairs = np;
    const int64_t m = std::min(np, cap);
    if (m > 0) {
        // slot pairs to atom pairs (orientation i < j is kept by the builder)
        std::vector<int32_t> a((size_t)m), b((size_t)m), perm((size_t)n);
        API_HIP(ctx, hipMemcpyAsync(a.data(), ctx->nl.pi.p, sizeof(int32_t) * (size_t)m,
                                    hipMemcpyDeviceToHost, ctx->stream));
        API_HIP(ctx, hipMemcpyAsync(b.data(), ctx->nl.pj.p, sizeof(int32_t) * (size_t)m,
                                    hipMemcpyDeviceToHost, ctx->stream));
        API_HIP(ctx, hipMemcpyAsync(perm.data(), ctx->nl.atom_idx_sorted.p,
                                    sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost,
                                    ctx->stream));
        API_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int64_t t = 0; t < m; ++t) {
            pairs[2 * t] = perm[(size_t)a[(size_t)t]];
            pairs[2 * t + 1] = perm[(size_t)b[(size_t)t]];
        }
    }
    return np > cap ? set_err(ctx, QEM_ERR_CAPACITY, "pair buffer too small") : QEM_OK;
}

int qem_short_forces(qem_ctx *ctx, int64_t n, const double *xyz, const double *q, double *f)
{
    CTX_GUARD(ctx);
    if (n > 0 && (!q || !f)) return set_err(ctx, QEM_ERR_INVALID_ARG, "null pointer");
    int st = check_atoms(ctx, n, xyz, q);
    if (st) return st;
    if (n == 0) return QEM_OK;
    if ((st = upload(ctx, ctx->d_xyz, xyz, sizeof(double) * 3 * (size_t)n)) ||
        (st = upload(ctx, ctx->d_q, q, sizeof(double) * (size_t)n)))
        return st;
    API_HIP(ctx, ensure(ctx->d_f, sizeof(double) * 3 * (size_t)n));
    st = short_dev(ctx, n, ctx->d_xyz.as<double>(), ctx->d_q.as<double>(), ctx->d_f.as<double>(),
                   ctx->stream, true);
    if (st) return st;
    API_HIP(ctx, hipMemcpyAsync(f, ctx->d_f.p, sizeof(double) * 3 * (size_t)n,
                                hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QEM_OK;
}

int qem_short_forces_pairs(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                           int64_t npairs, const int32_t *pairs, double *f)
{
    CTX_GUARD(ctx);
    if ((n > 0 && (!q || !f)) || npairs < 0 || (npairs > 0 && !pairs))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "null pointer or negative size");
    int st = check_atoms(ctx, n, xyz, q);
    if (st) return st;
    if (npairs > INT32_MAX) return set_err(ctx, QEM_ERR_UNSUPPORTED, "more than 2^31-1 pairs");
    for (int64_t t = 0; t < npairs; ++t) {
        const int32_t i = pairs[2 * t], j = pairs[2 * t + 1];
        if (i < 0 || j < 0 || i >= n || j >= n || i == j)
            return set_err(ctx, QEM_ERR_INVALID_ARG, "pair endpoints must be distinct atoms in [0, n)");
    }
    if (n == 0) return QEM_OK;
    if ((st = upload(ctx, ctx->d_xyz, xyz, sizeof(double) * 3 * (size_t)n)) ||
        (st = upload(ctx, ctx->d_q, q, sizeof(double) * (size_t)n)))
        return st;
    hipError_t e = install_pairs(ctx->nl, n, pairs, npairs, ctx->stream, ctx->err);
    if (e != hipSuccess) return hip_err(ctx, e, ctx->err.c_str());
    API_HIP(ctx, ensure(ctx->d_f, sizeof(double) * 3 * (size_t)n));
    st = short_dev(ctx, n, ctx->d_xyz.as<double>(), ctx->d_q.as<double>(), ctx->d_f.as<double>(),
                   ctx->stream, false);
    if (st) return st;
    API_HIP(ctx, hipMemcpyAsync(f, ctx->d_f.p, sizeof(double) * 3 * (size_t)n,
                                hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QEM_OK;
}

int qem_short_pair(qem_ctx *ctx, int64_t m, const double *qi, const double *qj, const double *ri,
                   const double *rj, int32_t single, double *f_i, double *fz_j)
{
    CTX_GUARD(ctx);
    if (m < 0 || (m > 0 && (!qi || !qj || !ri || !rj || !f_i)))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "null pointer or negative size");
    int st;
    if ((st = check_atoms(ctx, m, ri, qi)) || (st = check_atoms(ctx, m, rj, qj))) return st;
    if (m == 0) return QEM_OK;
    const size_t b3 = sizeof(double) * 3 * (size_t)m, b1 = sizeof(double) * (size_t)m;
    API_HIP(ctx, ensure(ctx->d_aux, 2 * b3 + 2 * b1));
    char *base = ctx->d_aux.as<char>();
    API_HIP(ctx, hipMemcpyAsync(base, ri, b3, hipMemcpyHostToDevice, ctx->stream));
    API_HIP(ctx, hipMemcpyAsync(base + b3, rj, b3, hipMemcpyHostToDevice, ctx->stream));
    API_HIP(ctx, hipMemcpyAsync(base + 2 * b3, qi, b1, hipMemcpyHostToDevice, ctx->stream));
    API_HIP(ctx, hipMemcpyAsync(base + 2 * b3 + b1, qj, b1, hipMemcpyHostToDevice, ctx->stream));
    API_HIP(ctx, ensure(ctx->d_f2, sizeof(double) * 4 * (size_t)m));
    const ShortConsts c = short_consts(ctx, single ? 1 : 0);
    API_HIP(ctx, launch_pair_direct(c, ctx->p.precision, ctx->sw, m,
                                    reinterpret_cast<const double *>(base),
                                    reinterpret_cast<const double *>(base + b3),
                                    reinterpret_cast<const double *>(base + 2 * b3),
                                    reinterpret_cast<const double *>(base + 2 * b3 + b1),
                                    ctx->d_f2.as<double>(), ctx->stream));
    std::vector<double> h(4 * (size_t)m);
    API_HIP(ctx, hipMemcpyAsync(h.data(), ctx->d_f2.p, sizeof(double) * 4 * (size_t)m,
                                hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t t = 0; t < m; ++t) {
        f_i[3 * t] = h[4 * (size_t)t];
        f_i[3 * t + 1] = h[4 * (size_t)t + 1];
        f_i[3 * t + 2] = h[4 * (size_t)t + 2];
        if (fz_j) fz_j[t] = h[4 * (size_t)t + 3];
    }
    return QEM_OK;
}

int qem_short_self(qem_ctx *ctx, int64_t n, const double *q, const double *xyz, int32_t single,
                   double *f)
{
    CTX_GUARD(ctx);
    if (n > 0 && (!q || !f)) return set_err(ctx, QEM_ERR_INVALID_ARG, "null pointer");
    int st = check_atoms(ctx, n, xyz, q);
    if (st) return st;
    if (n == 0) return QEM_OK;
    if ((st = upload(ctx, ctx->d_xyz, xyz, sizeof(double) * 3 * (size_t)n)) ||
        (st = upload(ctx, ctx->d_q, q, sizeof(double) * (size_t)n)))
        return st;
    API_HIP(ctx, ensure(ctx->d_f, sizeof(double) * 3 * (size_t)n));
    const ShortConsts c = short_consts(ctx, single ? 1 : 0);
    API_HIP(ctx, launch_self_only(c, ctx->p.precision, ctx->sw, n, ctx->d_xyz.as<double>(),
                                  ctx->d_q.as<double>(), ctx->d_f.as<double>(), ctx->stream));
    API_HIP(ctx, hipMemcpyAsync(f, ctx->d_f.p, sizeof(double) * 3 * (size_t)n,
                                hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QEM_OK;
}

int qem_sortperm_z(qem_ctx *ctx, int64_t n, const double *xyz, int64_t *zlist)
{
    CTX_GUARD(ctx);
    if (n > 0 && !zlist) return set_err(ctx, QEM_ERR_INVALID_ARG, "null pointer");
    int st = check_atoms(ctx, n, xyz, nullptr);
    if (st) return st;
    if (n == 0) return QEM_OK;
    if ((st = upload(ctx, ctx->d_xyz, xyz, sizeof(double) * 3 * (size_t)n))) return st;
    API_HIP(ctx, ensure(ctx->d_q, sizeof(double) * (size_t)n));
    hipError_t e = launch_zsort(ctx->lw, n, ctx->d_xyz.as<double>(), ctx->d_q.as<double>(),
                                ctx->stream, ctx->err);
    if (e != hipSuccess) return hip_err(ctx, e, "z sort");
    std::vector<int32_t> h((size_t)n);
    API_HIP(ctx, hipMemcpyAsync(h.data(), ctx->lw.perm.p, sizeof(int32_t) * (size_t)n,
                                hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < n; ++i) zlist[i] = h[(size_t)i];
    return QEM_OK;
}

int qem_long_terms(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                   const double *kvecs, int64_t P, double weight, uint32_t parts, double *f)
{
    CTX_GUARD(ctx);
    if ((n > 0 && (!q || !f)) || P < 0 || (P > 0 && !kvecs) || (parts & ~QEM_LONG_ALL))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    int st = check_atoms(ctx, n, xyz, q);
    if (st) return st;
    if (n == 0) return QEM_OK;
    if ((st = upload(ctx, ctx->d_xyz, xyz, sizeof(double) * 3 * (size_t)n)) ||
        (st = upload(ctx, ctx->d_q, q, sizeof(double) * (size_t)n)))
        return st;
    API_HIP(ctx, ensure(ctx->d_f, sizeof(double) * 3 * (size_t)n));
    st = long_dev(ctx, n, ctx->d_xyz.as<double>(), ctx->d_q.as<double>(), kvecs, P, weight,
                  parts, ctx->d_f.as<double>(), 0, ctx->stream);
    if (st) return st;
    API_HIP(ctx, hipMemcpyAsync(f, ctx->d_f.p, sizeof(double) * 3 * (size_t)n,
                                hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QEM_OK;
}

int qem_long_forces(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                    const int64_t *k_idx, int64_t P, double *f)
{
    CTX_GUARD(ctx);
    if (P < 0 || (P > 0 && !k_idx)) return set_err(ctx, QEM_ERR_INVALID_ARG, "bad k indices");
    std::vector<double> kv;
    int st = kidx_to_vecs(ctx, k_idx, P, kv);
    if (st) return st;
    const double w = P > 0 ? ctx->sum_K / (double)P : 0.0;
    return qem_long_terms(ctx, n, xyz, q, kv.data(), P, w, QEM_LONG_ALL, f);
}

int qem_forces(qem_ctx *ctx, int64_t n, const double *xyz, const double *q, const int64_t *k_idx,
               int64_t P, double *f)
{
    CTX_GUARD(ctx);
    if ((n > 0 && (!q || !f)) || P < 0 || (P > 0 && !k_idx))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    int st = check_atoms(ctx, n, xyz, q);
    if (st) return st;
    std::vector<double> kv;
    if ((st = kidx_to_vecs(ctx, k_idx, P, kv))) return st;
    if (n == 0) return QEM_OK;
    if ((st = upload(ctx, ctx->d_xyz, xyz, sizeof(double) * 3 * (size_t)n)) ||
        (st = upload(ctx, ctx->d_q, q, sizeof(double) * (size_t)n)))
        return st;
    API_HIP(ctx, ensure(ctx->d_f, sizeof(double) * 3 * (size_t)n));
    st = forces_core(ctx, n, ctx->d_xyz.as<double>(), ctx->d_q.as<double>(), kv, P,
                     ctx->d_f.as<double>(), ctx->stream);
    if (st) return st;
    API_HIP(ctx, hipMemcpyAsync(f, ctx->d_f.p, sizeof(double) * 3 * (size_t)n,
                                hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QEM_OK;
}

int qem_short_forces_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                         double *d_f, void *stream)
{
    CTX_GUARD(ctx);
    if (n < 0 || (n > 0 && (!d_xyz || !d_q || !d_f)))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    return short_dev(ctx, n, d_xyz, d_q, d_f, pick_stream(ctx, stream), true);
}

int qem_long_forces_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                        const int64_t *k_idx, int64_t P, double *d_f, void *stream)
{
    CTX_GUARD(ctx);
    if (n < 0 || P < 0 || (P > 0 && !k_idx) || (n > 0 && (!d_xyz || !d_q || !d_f)))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    std::vector<double> kv;
    int st = kidx_to_vecs(ctx, k_idx, P, kv);
    if (st) return st;
    const double w = P > 0 ? ctx->sum_K / (double)P : 0.0;
    return long_dev(ctx, n, d_xyz, d_q, kv.data(), P, w, QEM_LONG_ALL, d_f, 0,
                    pick_stream(ctx, stream));
}

// the variance pass keeps one partial per 64 k rows for every atom
// (32 B each), so its batch is bounded: 65536 k vectors
constexpr int64_t kMaxVarianceP = 65536;

// batch variance of the long-range force (and optionally the force itself)
static int variance_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                        const std::vector<double> &kv, int64_t P, double *d_f, double *d_var,
                        hipStream_t s)
{
    if (n == 0) return QEM_OK;
    const double w = ctx->sum_K / (double)P;
    int64_t Ptot = 0;
    int st = stage_kvecs(ctx, kv.data(), P, w, false, s, &Ptot);
    if (st) return st;
    {
        PhaseScope ps(ctx, QEM_PHASE_LONG, s);
        hipError_t e = launch_zsort(ctx->lw, n, d_xyz, d_q, s, ctx->err);
        if (e != hipSuccess) return hip_err(ctx, e, "z sort");
        e = launch_long_variance(long_consts(ctx, w, QEM_LONG_NONSORT | QEM_LONG_SORT | QEM_LONG_SELF),
                                 ctx->lw, n, Ptot, P, d_var, s, ctx->err);
        if (e != hipSuccess) return hip_err(ctx, e, "long-range batch variance");
    }
    // the z order of these positions is already in place
    if (d_f) return long_dev(ctx, n, d_xyz, d_q, kv.data(), P, w, QEM_LONG_ALL, d_f, 0, s, true);
    return QEM_OK;
}

int qem_long_variance(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                      const int64_t *k_idx, int64_t P, double *f, double *var)
{
    CTX_GUARD(ctx);
    if (P < 2 || !k_idx) return set_err(ctx, QEM_ERR_INVALID_ARG, "the batch variance needs P >= 2");
    if (P > kMaxVarianceP)
        return set_err(ctx, QEM_ERR_UNSUPPORTED, "the batch variance takes at most 65536 k vectors");
    if (n > 0 && (!q || !var)) return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    int st = check_atoms(ctx, n, xyz, q);
    if (st) return st;
    std::vector<double> kv;
    if ((st = kidx_to_vecs(ctx, k_idx, P, kv))) return st;
    if (n == 0) return QEM_OK;
    if ((st = upload(ctx, ctx->d_xyz, xyz, sizeof(double) * 3 * (size_t)n)) ||
        (st = upload(ctx, ctx->d_q, q, sizeof(double) * (size_t)n)))
        return st;
    API_HIP(ctx, ensure(ctx->d_f, sizeof(double) * 3 * (size_t)n));
    API_HIP(ctx, ensure(ctx->d_aux, sizeof(double) * (size_t)n));
    st = variance_dev(ctx, n, ctx->d_xyz.as<double>(), ctx->d_q.as<double>(), kv, P,
                      f ? ctx->d_f.as<double>() : nullptr, ctx->d_aux.as<double>(), ctx->stream);
    if (st) return st;
    API_HIP(ctx, hipMemcpyAsync(var, ctx->d_aux.p, sizeof(double) * (size_t)n,
                                hipMemcpyDeviceToHost, ctx->stream));
    if (f)
        API_HIP(ctx, hipMemcpyAsync(f, ctx->d_f.p, sizeof(double) * 3 * (size_t)n,
                                    hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QEM_OK;
}

int qem_long_variance_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                          const int64_t *k_idx, int64_t P, double *d_f, double *d_var,
                          void *stream)
{
    CTX_GUARD(ctx);
    if (P < 2 || !k_idx) return set_err(ctx, QEM_ERR_INVALID_ARG, "the batch variance needs P >= 2");
    if (P > kMaxVarianceP)
        return set_err(ctx, QEM_ERR_UNSUPPORTED, "the batch variance takes at most 65536 k vectors");
    if (n < 0 || (n > 0 && (!d_xyz || !d_q || !d_var)))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    std::vector<double> kv;
    const int st = kidx_to_vecs(ctx, k_idx, P, kv);
    if (st) return st;
    return variance_dev(ctx, n, d_xyz, d_q, kv, P, d_f, d_var, pick_stream(ctx, stream));
}

int qem_long_terms_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                       const double *kvecs, int64_t P, double weight, uint32_t parts,
                       double *d_f, int32_t accumulate, void *stream)
{
    CTX_GUARD(ctx);
    if (n < 0 || P < 0 || (P > 0 && !kvecs) || (n > 0 && (!d_xyz || !d_q || !d_f)))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    if (parts & ~QEM_LONG_ALL) return set_err(ctx, QEM_ERR_INVALID_ARG, "unknown parts bits");
    if (!std::isfinite(weight)) return set_err(ctx, QEM_ERR_DOMAIN, "weight must be finite");
    return long_dev(ctx, n, d_xyz, d_q, kvecs, P, weight, parts, d_f, accumulate ? 1 : 0,
                    pick_stream(ctx, stream));
}

int qem_forces_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                   const int64_t *k_idx, int64_t P, double *d_f, void *stream)
{
    CTX_GUARD(ctx);
    if (n < 0 || P < 0 || (P > 0 && !k_idx) || (n > 0 && (!d_xyz || !d_q || !d_f)))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    std::vector<double> kv;
    int st = kidx_to_vecs(ctx, k_idx, P, kv);
    if (st) return st;
    return forces_core(ctx, n, d_xyz, d_q, kv, P, d_f, pick_stream(ctx, stream));
}

int qem_short_state(qem_ctx *ctx, int64_t *npairs, int64_t *n_builds, int64_t *n_steps)
{
    if (!ctx) return QEM_ERR_INVALID_ARG;
    if (npairs) *npairs = ctx->nl.npairs;
    if (n_builds) *n_builds = ctx->nl.builds;
    if (n_steps) *n_steps = ctx->n_steps;
    return QEM_OK;
}

int qem_energy(qem_ctx *ctx, int64_t n, const double *xyz, const double *q, const int64_t *k_idx,
               int64_t P, double *E, double *dE_dq, double *dE_dgamma)
{
    CTX_GUARD(ctx);
    if ((n > 0 && !q) || P < 0 || (P > 0 && !k_idx))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    int st = check_atoms(ctx, n, xyz, q);
    if (st) return st;
    std::vector<double> kv;
    if ((st = kidx_to_vecs(ctx, k_idx, P, kv))) return st;
    if (n == 0) return energy_dev(ctx, 0, nullptr, nullptr, nullptr, 0, 0.0, nullptr, E,
                                  dE_dgamma, ctx->stream);
    if ((st = upload(ctx, ctx->d_xyz, xyz, sizeof(double) * 3 * (size_t)n)) ||
        (st = upload(ctx, ctx->d_q, q, sizeof(double) * (size_t)n)))
        return st;
    API_HIP(ctx, ensure(ctx->d_f, sizeof(double) * 3 * (size_t)n));
    const double w = P > 0 ? ctx->sum_K / (double)P : 0.0;
    st = energy_dev(ctx, n, ctx->d_xyz.as<double>(), ctx->d_q.as<double>(), kv.data(), P, w,
                    ctx->d_f.as<double>(), E, dE_dgamma, ctx->stream);
    if (st) return st;
    if (dE_dq) {
        API_HIP(ctx, hipMemcpyAsync(dE_dq, ctx->d_f.p, sizeof(double) * (size_t)n,
                                    hipMemcpyDeviceToHost, ctx->stream));
        API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return QEM_OK;
}

int qem_energy_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                   const int64_t *k_idx, int64_t P, double *d_dE_dq, double *E,
                   double *dE_dgamma, void *stream)
{
    CTX_GUARD(ctx);
    if (n < 0 || P < 0 || (P > 0 && !k_idx) || (n > 0 && (!d_xyz || !d_q)))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    std::vector<double> kv;
    int st = kidx_to_vecs(ctx, k_idx, P, kv);
    if (st) return st;
    const double w = P > 0 ? ctx->sum_K / (double)P : 0.0;
    return energy_dev(ctx, n, d_xyz, d_q, kv.data(), P, w, d_dE_dq, E, dE_dgamma,
                      pick_stream(ctx, stream));
}

static int forces_energy_dev_entry(qem_ctx *ctx, int64_t n, const double *d_xyz,
                                   const double *d_q, const int64_t *k_idx, int64_t P,
                                   double *d_f, double *d_dE_dq, double *E, double *dE_dgamma,
                                   void *stream, bool save)
{
    CTX_GUARD(ctx);
    if (n < 0 || P < 0 || (P > 0 && !k_idx) || (n > 0 && (!d_xyz || !d_q || !d_f)))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    std::vector<double> kv;
    int st = kidx_to_vecs(ctx, k_idx, P, kv);
    if (st) return st;
    return forces_energy_dev(ctx, n, d_xyz, d_q, kv, P, d_f, d_dE_dq, E, dE_dgamma,
                             pick_stream(ctx, stream), save);
}

int qem_forces_energy_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                          const int64_t *k_idx, int64_t P, double *d_f, double *d_dE_dq,
                          double *E, double *dE_dgamma, void *stream)
{
    return forces_energy_dev_entry(ctx, n, d_xyz, d_q, k_idx, P, d_f, d_dE_dq, E, dE_dgamma,
                                   stream, false);
}

int qem_forces_energy_save_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                               const int64_t *k_idx, int64_t P, double *d_f, double *d_dE_dq,
                               double *E, double *dE_dgamma, void *stream)
{
    return forces_energy_dev_entry(ctx, n, d_xyz, d_q, k_idx, P, d_f, d_dE_dq, E, dE_dgamma,
                                   stream, true);
}

static int forces_energy_host(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                              const int64_t *k_idx, int64_t P, double *f, double *E,
                              double *dE_dq, double *dE_dgamma, bool save)
{
    CTX_GUARD(ctx);
    if ((n > 0 && (!q || !f)) || P < 0 || (P > 0 && !k_idx))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    int st = check_atoms(ctx, n, xyz, q);
    if (st) return st;
    std::vector<double> kv;
    if ((st = kidx_to_vecs(ctx, k_idx, P, kv))) return st;
    if (n == 0)
        return forces_energy_dev(ctx, 0, nullptr, nullptr, kv, P, nullptr, nullptr, E,
                                 dE_dgamma, ctx->stream);
    const size_t b3 = sizeof(double) * 3 * (size_t)n;
    if ((st = upload(ctx, ctx->d_xyz, xyz, b3)) ||
        (st = upload(ctx, ctx->d_q, q, sizeof(double) * (size_t)n)))
        return st;
    API_HIP(ctx, ensure(ctx->d_f, b3));
    API_HIP(ctx, ensure(ctx->d_aux, sizeof(double) * (size_t)n));
    st = forces_energy_dev(ctx, n, ctx->d_xyz.as<double>(), ctx->d_q.as<double>(), kv, P,
                           ctx->d_f.as<double>(), ctx->d_aux.as<double>(), E, dE_dgamma,
                           ctx->stream, save);
    if (st) return st;
    API_HIP(ctx, hipMemcpyAsync(f, ctx->d_f.p, b3, hipMemcpyDeviceToHost, ctx->stream));
    if (dE_dq)
        API_HIP(ctx, hipMemcpyAsync(dE_dq, ctx->d_aux.p, sizeof(double) * (size_t)n,
                                    hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QEM_OK;
}

int qem_forces_energy(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                      const int64_t *k_idx, int64_t P, double *f, double *E, double *dE_dq,
                      double *dE_dgamma)
{
    return forces_energy_host(ctx, n, xyz, q, k_idx, P, f, E, dE_dq, dE_dgamma, false);
}

int qem_forces_energy_save(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                           const int64_t *k_idx, int64_t P, double *f, double *E, double *dE_dq,
                           double *dE_dgamma)
{
    return forces_energy_host(ctx, n, xyz, q, k_idx, P, f, E, dE_dq, dE_dgamma, true);
}

static int hvp_host(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                    const int64_t *k_idx, int64_t P, const double *v, double *hv, bool saved)
{
    CTX_GUARD(ctx);
    if ((n > 0 && (!q || !v || !hv)) || P < 0 || (P > 0 && !k_idx))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    int st = check_atoms(ctx, n, xyz, q);
    if (st) return st;
    if (n > 0 && !finite_all(v, 3 * n)) return set_err(ctx, QEM_ERR_DOMAIN, "v must be finite");
    std::vector<double> kv;
    if ((st = kidx_to_vecs(ctx, k_idx, P, kv))) return st;
    if (n == 0) return QEM_OK;
    const size_t b3 = sizeof(double) * 3 * (size_t)n;
    if ((st = upload(ctx, ctx->d_xyz, xyz, b3)) ||
        (st = upload(ctx, ctx->d_q, q, sizeof(double) * (size_t)n)) ||
        (st = upload(ctx, ctx->d_aux, v, b3)))
        return st;
    API_HIP(ctx, ensure(ctx->d_f, b3));
    const double w = P > 0 ? ctx->sum_K / (double)P : 0.0;
    st = hvp_dev(ctx, n, ctx->d_xyz.as<double>(), ctx->d_q.as<double>(), kv.data(), P, w,
                 ctx->d_aux.as<double>(), ctx->d_f.as<double>(), ctx->stream, saved);
    if (st) return st;
    API_HIP(ctx, hipMemcpyAsync(hv, ctx->d_f.p, b3, hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QEM_OK;
}

int qem_hvp(qem_ctx *ctx, int64_t n, const double *xyz, const double *q, const int64_t *k_idx,
            int64_t P, const double *v, double *hv)
{
    return hvp_host(ctx, n, xyz, q, k_idx, P, v, hv, false);
}

int qem_hvp_saved(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                  const int64_t *k_idx, int64_t P, const double *v, double *hv)
{
    return hvp_host(ctx, n, xyz, q, k_idx, P, v, hv, true);
}

static int hvp_dev_entry(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                         const int64_t *k_idx, int64_t P, const double *d_v, double *d_hv,
                         void *stream, bool saved)
{
    CTX_GUARD(ctx);
    if (n < 0 || P < 0 || (P > 0 && !k_idx) || (n > 0 && (!d_xyz || !d_q || !d_v || !d_hv)))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    std::vector<double> kv;
    int st = kidx_to_vecs(ctx, k_idx, P, kv);
    if (st) return st;
    const double w = P > 0 ? ctx->sum_K / (double)P : 0.0;
    return hvp_dev(ctx, n, d_xyz, d_q, kv.data(), P, w, d_v, d_hv, pick_stream(ctx, stream),
                   saved);
}

int qem_hvp_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                const int64_t *k_idx, int64_t P, const double *d_v, double *d_hv, void *stream)
{
    return hvp_dev_entry(ctx, n, d_xyz, d_q, k_idx, P, d_v, d_hv, stream, false);
}

int qem_hvp_saved_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                      const int64_t *k_idx, int64_t P, const double *d_v, double *d_hv,
                      void *stream)
{
    return hvp_dev_entry(ctx, n, d_xyz, d_q, k_idx, P, d_v, d_hv, stream, true);
}

int qem_timing_enable(qem_ctx *ctx, int32_t on)
{
    if (!ctx) return QEM_ERR_INVALID_ARG;
    ctx->timing = on != 0;
    return QEM_OK;
}

int qem_timing_read(qem_ctx *ctx, double *ms, int64_t *calls, int32_t reset)
{
    CTX_GUARD(ctx);
    for (auto &m : ctx->marks) {
        API_HIP(ctx, hipEventSynchronize(m.b));
        float t = 0.f;
        API_HIP(ctx, hipEventElapsedTime(&t, m.a, m.b));
        ctx->phase_ms[m.phase] += (double)t;
        ctx->phase_calls[m.phase] += 1;
        ctx->ev_pool.push_back(m.a);
        ctx->ev_pool.push_back(m.b);
    }
    ctx->marks.clear();
    for (int i = 0; i < QEM_NPHASES; ++i) {
        if (ms) ms[i] = ctx->phase_ms[i];
        if (calls) calls[i] = ctx->phase_calls[i];
        if (reset) {
            ctx->phase_ms[i] = 0.0;
            ctx->phase_calls[i] = 0;
        }
    }
    return QEM_OK;
}

int qem_gauss_legendre(int32_t n, double *x, double *w)
{
    if (n < 1 || n > 256 || !x || !w) return QEM_ERR_INVALID_ARG;
    GLRule r = make_gl_rule(n);
    std::memcpy(x, r.x.data(), sizeof(double) * (size_t)n);
    std::memcpy(w, r.w.data(), sizeof(double) * (size_t)n);
    return QEM_OK;
}

int qem_bessel01_table(qem_ctx *ctx, int64_t n, const double *x, double *j0o, double *j1o)
{
    CTX_GUARD(ctx);
    if (n < 0 || (n > 0 && (!x || !j0o || !j1o)))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    const bool mixed = ctx->p.precision == QEM_PREC_MIXED;
    const int ntab = mixed ? ctx->sw.jt_n : ctx->sw.jd_n;
    if (ntab == 0)
        return set_err(ctx, QEM_ERR_UNSUPPORTED, "context has no Bessel table");
    const double xmax = ntab * (mixed ? (double)ctx->sw.jt_dx : ctx->sw.jd_dx);
    for (int64_t i = 0; i < n; ++i)
        if (!(x[i] >= 0.0 && x[i] <= xmax))
            return set_err(ctx, QEM_ERR_DOMAIN, "argument outside the table range");
    if (n == 0) return QEM_OK;
    const size_t b = sizeof(double) * (size_t)n;
    API_HIP(ctx, ensure(ctx->d_aux, 3 * b));
    char *base = ctx->d_aux.as<char>();
    API_HIP(ctx, hipMemcpyAsync(base, x, b, hipMemcpyHostToDevice, ctx->stream));
    const unsigned g = (unsigned)((n + 255) / 256);
    if (mixed)
        k_bessel_tab<<<g, 256, 0, ctx->stream>>>(
            reinterpret_cast<const double *>(base), n, ctx->sw.jtab.as<float4>(), ntab,
            1.0f / ctx->sw.jt_dx, reinterpret_cast<double *>(base + b),
            reinterpret_cast<double *>(base + 2 * b));
    else
        k_bessel_tab64<<<g, 256, 0, ctx->stream>>>(
            reinterpret_cast<const double *>(base), n, ctx->sw.jdtab.as<double>(), ntab,
            1.0 / ctx->sw.jd_dx, reinterpret_cast<double *>(base + b),
            reinterpret_cast<double *>(base + 2 * b));
    API_HIP(ctx, hipGetLastError());
    API_HIP(ctx, hipMemcpyAsync(j0o, base + b, b, hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipMemcpyAsync(j1o, base + 2 * b, b, hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QEM_OK;
}

int qem_dev_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return QEM_ERR_INVALID_ARG;
    return hipMalloc(ptr, bytes ? bytes : 1) == hipSuccess ? QEM_OK : QEM_ERR_NOMEM;
}

int qem_dev_free(void *ptr) { return hipFree(ptr) == hipSuccess ? QEM_OK : QEM_ERR_HIP; }

int qem_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? QEM_OK : QEM_ERR_HIP;
}

int qem_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? QEM_OK : QEM_ERR_HIP;
}

int qem_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? QEM_OK : QEM_ERR_HIP; }

int qem_dev_copy_kernel(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!dst || !src) return QEM_ERR_INVALID_ARG;
    hipStream_t hs = stream == QEM_STREAM_NULL ? nullptr : static_cast<hipStream_t>(stream);
    return launch_copy(dst, src, bytes, hs) == hipSuccess
               ? QEM_OK
               : QEM_ERR_HIP;
}

// ---- Substrate_LJ walls -------------------------------------------------------

static int check_wall(qem_ctx *ctx, const qem_wall *w)
{
    if (!w) return set_err(ctx, QEM_ERR_INVALID_ARG, "null wall");
    if (w->direction != 1 && w->direction != -1)
        return set_err(ctx, QEM_ERR_INVALID_ARG, "wall direction must be +1 or -1");
    if (w->semantics & ~(QEM_WALL_SEM_ABS_CUTOFF | QEM_WALL_SEM_TRUE_FORCE))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "unknown wall semantics bits");
    if (!(std::isfinite(w->cutoff) && std::isfinite(w->sigma) && std::isfinite(w->eps) &&
          std::isfinite(w->sub_pos)))
        return set_err(ctx, QEM_ERR_DOMAIN, "wall parameters must be finite");
    return QEM_OK;
}

// walls on device data, at most kMaxWallsPerLaunch per pass; reads the
// device domain flag once (synchronises s)
static int wall_dev(qem_ctx *ctx, const qem_wall *walls, int32_t nwalls, int64_t n,
                    const double *d_xyz, double *d_f, int accumulate, hipStream_t s)
{
    if (n == 0 || nwalls == 0) return QEM_OK;
    if (!ctx->d_flag.p) {
        API_HIP(ctx, ensure(ctx->d_flag, sizeof(uint32_t) * 2));
        API_HIP(ctx, launch_zero_flag(ctx->d_flag.as<uint32_t>(), s));
    }
    uint32_t *flag = ctx->d_flag.as<uint32_t>();
    if (++ctx->wall_epoch == 0) ++ctx->wall_epoch;   // 0 is the cleared value
    WallSet ws{};
    ws.epoch = ctx->wall_epoch;
    for (int32_t k0 = 0; k0 < nwalls; k0 += kMaxWallsPerLaunch) {
        ws.nw = std::min<int32_t>(kMaxWallsPerLaunch, nwalls - k0);
        for (int32_t k = 0; k < ws.nw; ++k) ws.w[k] = walls[k0 + k];
        API_HIP(ctx, launch_substrate_lj(ws, n, d_xyz, d_f, 3, accumulate || k0 > 0, flag, s));
    }
    uint32_t seen = 0;
    API_HIP(ctx, hipMemcpyAsync(&seen, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    API_HIP(ctx, hipStreamSynchronize(s));
    if (seen == ws.epoch)
        return set_err(ctx, QEM_ERR_DOMAIN,
                       "Substrate_LJ: an atom is on the wrong side of the wall "
                       "(direction * (z - sub_pos) <= 0)");
    return QEM_OK;
}

int qem_substrate_lj(qem_ctx *ctx, const qem_wall *w, int64_t n, const double *xyz, double *f,
                     int32_t accumulate)
{
    CTX_GUARD(ctx);
    int st = check_wall(ctx, w);
    if (st) return st;
    if (n < 0 || (n > 0 && (!xyz || !f))) return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    // the assertion is checked on the host first, so an error writes nothing
    for (int64_t i = 0; i < n; ++i) {
        if (!(std::isfinite(xyz[3 * i]) && std::isfinite(xyz[3 * i + 1]) &&
              std::isfinite(xyz[3 * i + 2])))
            return set_err(ctx, QEM_ERR_DOMAIN, "coordinates must be finite");
        if (!((double)w->direction * (xyz[3 * i + 2] - w->sub_pos) > 0.0))
            return set_err(ctx, QEM_ERR_DOMAIN,
                           "Substrate_LJ: an atom is on the wrong side of the wall "
                           "(direction * (z - sub_pos) <= 0)");
    }
    if (n == 0) return QEM_OK;
    const size_t b3 = sizeof(double) * 3 * (size_t)n;
    if ((st = upload(ctx, ctx->d_xyz, xyz, b3))) return st;
    API_HIP(ctx, ensure(ctx->d_f, b3));
    if (accumulate) API_HIP(ctx, hipMemcpyAsync(ctx->d_f.p, f, b3, hipMemcpyHostToDevice, ctx->stream));
    if ((st = wall_dev(ctx, w, 1, n, ctx->d_xyz.as<double>(), ctx->d_f.as<double>(), accumulate,
                       ctx->stream)))
        return st;
    API_HIP(ctx, hipMemcpyAsync(f, ctx->d_f.p, b3, hipMemcpyDeviceToHost, ctx->stream));
    API_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QEM_OK;
}

int qem_substrate_lj_dev(qem_ctx *ctx, const qem_wall *w, int64_t n, const double *d_xyz,
                         double *d_f, int32_t accumulate, void *stream)
{
    CTX_GUARD(ctx);
    int st = check_wall(ctx, w);
    if (st) return st;
    if (n < 0 || (n > 0 && (!d_xyz || !d_f)))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    return wall_dev(ctx, w, 1, n, d_xyz, d_f, accumulate, pick_stream(ctx, stream));
}

int qem_general_forces_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                           const int64_t *k_idx, int64_t P, const qem_wall *walls,
                           int32_t nwalls, double *d_f, void *stream)
{
    CTX_GUARD(ctx);
    if (nwalls < 0 || (nwalls > 0 && !walls))
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad wall list");
    for (int32_t i = 0; i < nwalls; ++i) {
        const int st = check_wall(ctx, walls + i);
        if (st) return st;
    }
    if (nwalls == 0 || n == 0) return qem_forces_dev(ctx, n, d_xyz, d_q, k_idx, P, d_f, stream);
    if (n < 0 || P < 0 || (P > 0 && !k_idx) || !d_xyz || !d_q || !d_f)
        return set_err(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    std::vector<double> kv;
    int st = kidx_to_vecs(ctx, k_idx, P, kv);
    if (st) return st;
    hipStream_t s = pick_stream(ctx, stream);
    // 1. the walls (the reference lists them first, test/QEM_test.jl:28) into
    //    d_wz, and the domain flag copied to pinned memory behind an event
    if (!ctx->d_flag.p) {
        API_HIP(ctx, ensure(ctx->d_flag, sizeof(uint32_t) * 2));
        API_HIP(ctx, launch_zero_flag(ctx->d_flag.as<uint32_t>(), s));
    }
    if (!ctx->h_flag) {
        API_HIP(ctx, hipHostMalloc(reinterpret_cast<void **>(&ctx->h_flag), sizeof(uint32_t),
                                   hipHostMallocDefault));
        API_HIP(ctx, hipEventCreateWithFlags(&ctx->wall_ev, hipEventDisableTiming));
    }
    API_HIP(ctx, ensure(ctx->d_wz, sizeof(double) * (size_t)n));
    uint32_t *flag = ctx->d_flag.as<uint32_t>();
    if (++ctx->wall_epoch == 0) ++ctx->wall_epoch;
    WallSet ws{};
    ws.epoch = ctx->wall_epoch;
    for (int32_t k0 = 0; k0 < nwalls; k0 += kMaxWallsPerLaunch) {
        ws.nw = std::min<int32_t>(kMaxWallsPerLaunch, nwalls - k0);
        for (int32_t k = 0; k < ws.nw; ++k) ws.w[k] = walls[k0 + k];
        API_HIP(ctx, launch_substrate_lj(ws, n, d_xyz, ctx->d_wz.as<double>(), 1, k0 > 0, flag, s));
    }
    API_HIP(ctx, hipMemcpyAsync(ctx->h_flag, flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    API_HIP(ctx, hipEventRecord(ctx->wall_ev, s));
    // 2. QEM_short (its final pass adds the wall z forces) and QEM_long
    if ((st = forces_core(ctx, n, d_xyz, d_q, kv, P, d_f, s, ctx->d_wz.as<double>()))) return st;
    // 3. the domain check waits for the wall pass only: the QEM kernels of
    //    this step are already queued behind it
    API_HIP(ctx, hipEventSynchronize(ctx->wall_ev));
    if (*ctx->h_flag == ws.epoch)
        return set_err(ctx, QEM_ERR_DOMAIN,
                       "Substrate_LJ: an atom is on the wrong side of the wall "
                       "(direction * (z - sub_pos) <= 0)");
    return QEM_OK;
}

}  // extern "C"
