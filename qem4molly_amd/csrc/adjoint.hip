// Short-range adjoint kernels on gfx950: pair and self energies with their
// gamma derivatives (for E, dE/dq, dE/dgamma), and forward-mode tangents of
// the pair and self forces (for Hessian-vector products).
//
// No reference function computes these: QEM4Molly has forces only.  The
// energy is the functional whose negative gradient at frozen quadrature
// limits is the force of src/Force_short.jl:25-107 (pair terms: the J0
// transforms of Gamma_2 and Gamma_1 e^{-k^2/4a} of src/greens_function.jl
// over the same nodes and limits; self terms: the images of the charge
// itself).  The tangents differentiate the force as computed, including the
// motion of the quadrature nodes through the z-dependent limits k_f1, k_f2
// (src/greens_function.jl:26-27).
//
// Same work mapping as the force kernels (one lane per pair, wave-uniform
// nodes, LDS Bessel table in mixed precision); nothing per pair is stored
// beyond one double4 of output, the quadrature is recomputed.
#include "qem_device.h"

namespace qem {
namespace {

using dev::bessel01_tab;
using dev::jl_sign;
using dev::kLog2e;

// ---- energy node sums -----------------------------------------------------
// kind 0 (point):   t = e^{-ka} E / D,              v += w t J0,  d += w t / D J0
// kind 1 (gauss):   t = e^{-ka - k^2/4a} / D,       v += w t J0,  d += w t E / D J0
// kind 2 (gauss, multiple reflections of the charge itself, a = 0):
//                   t = e^{-k^2/4a} E / D,          v += w t,     d += w t / D
// with E = e^{-2kLz}, D = g12 E - 1.  d is what d/dgamma of D brings down.
// Lane `lane` of `nl` lanes takes the nodes n = lane (mod nl); the nl
// shares add up to the sums (nl = 1 gives them whole).
template <int PREC, bool SELF>
__device__ __forceinline__ void nodes_energy(double K, double a, double rho, int kind,
                                             const ShortConsts &c,
                                             const double *__restrict__ glx,
                                             const double *__restrict__ glw,
                                             const float *__restrict__ glxf,
                                             const float *__restrict__ glwf,
                                             const float4 *__restrict__ tab, double &v, double &d,
                                             int lane = 0, int nl = 1)
{
    if (PREC != QEM_PREC_MIXED) {
        const double h = 0.5 * K, inv4a = 0.25 / c.alpha;
        double av = 0.0, ad = 0.0;
        for (int n = lane; n < c.Nt; n += nl) {
            const double k = h + h * glx[n];
            const double w = h * glw[n];
            const double E = exp(-2.0 * k * c.Lz);
            const double rD = 1.0 / (c.g12 * E - 1.0);
            double num, fd;
            if (kind == 0) { num = exp(-k * a) * E; fd = rD; }
            else if (kind == 1) { num = exp(-k * a - k * k * inv4a); fd = E * rD; }
            else { num = exp(-k * k * inv4a) * E; fd = rD; }
            double J = 1.0;
            if (!SELF) {
                if (PREC == QEM_PREC_FP64) {
                    double J1;
                    dev::bessel01_tab64(k * rho, reinterpret_cast<const double *>(tab), c.jd_n,
                                        c.jd_inv_dx, J, J1);
                } else {
                    J = j0(k * rho);
                }
            }
            const double wt = w * num * rD * J;
            av += wt;
            ad += wt * fd;
        }
        v = av;
        d = ad;
    } else {
        // node constants as in the force kernel, with u.z = log2 w_n (the
        // energy sums carry w, not w k): every exponent is one operation
        // with a wave-uniform constant, e = -2 Lz k log2(e) = (h c2) u.x
        const float4 *__restrict__ ncon = c.ncon;
        const float h = 0.5f * (float)K;
        const float c2 = -2.0f * (float)c.Lz * kLog2e;
        const float cg = -kLog2e * (float)(0.25 / c.alpha);
        const float g12 = (float)c.g12;
        const float hE = h * c2;
        const float rA = kind == 2 ? 1.0f : (float)a * (float)(0.5 / c.Lz);
        const float rP = 1.0f + (float)a * (float)(0.5 / c.Lz);
        const float rG = cg / (c2 * c2);
        const float hR = SELF ? 0.0f
                              : fminf(h * (float)rho * c.jt_inv_dx, 0.5f * (float)(c.jt_n - 1));
        float av = 0.f, ad = 0.f;
        for (int n = lane; n < c.Nt; n += nl) {
            const float4 u = ncon[n];
            const float e = hE * u.x;
            const float E = __builtin_amdgcn_exp2f(e);
            const float num = kind == 0 ? __builtin_amdgcn_exp2f(fmaf(e, rP, u.z))
                                        : __builtin_amdgcn_exp2f(fmaf(e, fmaf(e, rG, rA), u.z));
            const float rD = __builtin_amdgcn_rcpf(fmaf(g12, E, -1.0f));
            float wt = num * rD;
            if (!SELF) {
                const float xb = hR * u.x;
                const float tt = __builtin_amdgcn_fractf(xb);
                const float4 cc = tab[(int)xb];
                const float p = fmaf(fmaf(fmaf(cc.w, tt, cc.z), tt, cc.y), tt, cc.x);
                wt *= p;
            }
            av += wt;
            ad = fmaf(wt, kind == 1 ? E * rD : rD, ad);
        }
        v = h * av;
        d = h * ad;
    }
}

// d b_l / d gamma_w and d g12 / d gamma_w
__device__ __forceinline__ double db_dg(const ShortConsts &c, int l, int w)
{
    if (l == 1) return w == 0 ? 1.0 : 0.0;
    if (l == 2) return w == 1 ? 1.0 : 0.0;
    if (l == 3) return w == 0 ? c.g2 : c.g1;
    return 0.0;
}

// Pair potential u_ij (without q_i q_j) and du/dgamma_1, du/dgamma_2.
template <int PREC>
__device__ __forceinline__ double4 pair_energy(const ShortConsts &c, const double *__restrict__ glx,
                               const double *__restrict__ glw, const float *__restrict__ glxf,
                               const float *__restrict__ glwf, const float4 *__restrict__ tab,
                               double xi, double yi, double zi, double xj, double yj, double zj)
{
    double mx, my;
    const double d2 = dev::min_image_d2(c, xi - xj, yi - yj, mx, my);
    if (d2 > c.rc2) return make_double4(0.0, 0.0, 0.0, 0.0);
    const double rho = sqrt(d2);
    const double zn = fabs(zi - zj), zp = zi + zj, sg = jl_sign(zi - zj);
    const bool swap = (c.sem & QEM_SEM_SWAP_LIMITS) != 0;
    double u = 0.0, du0 = 0.0, du1 = 0.0;
    for (int l = 0; l < 4; ++l) {
        double a, b, s;
        dev::image(c, l, zn, zp, sg, a, b, s);
        const dev::Limits L = dev::limits(c, a);
        const double Kp = swap ? L.kf1 : L.kf2, Kg = swap ? L.kf2 : L.kf1;
        const double db0 = db_dg(c, l, 0), db1 = db_dg(c, l, 1);
        const double inv_r = 1.0 / sqrt(a * a + rho * rho);
        double Pv = 0.0, Pd = 0.0, Gv = 0.0, Gd = 0.0;
        if (b != 0.0 || db0 != 0.0 || db1 != 0.0) {
            nodes_energy<PREC, false>(Kp, a, rho, 0, c, glx, glw, glxf, glwf, tab, Pv, Pd);
            nodes_energy<PREC, false>(Kg, a, rho, 1, c, glx, glw, glxf, glwf, tab, Gv, Gd);
        }
        const double hb = 0.5 * b;
        u += hb * inv_r - hb * c.g12 * Pv + hb * Gv;
        du0 += 0.5 * db0 * (inv_r - c.g12 * Pv + Gv) + hb * c.g2 * (Pd - Gd);
        du1 += 0.5 * db1 * (inv_r - c.g12 * Pv + Gv) + hb * c.g1 * (Pd - Gd);
    }
    return make_double4(c.inv2pieps * u, c.inv2pieps * du0, c.inv2pieps * du1, 0.0);
}

// Self energy factor s_i (E_self = q_i^2 s_i) and its gamma derivatives:
// the images of the charge itself (image 1 keeps only its multiple
// reflections), split by exp(-k^2/4a) like the pair terms.  Lane `lane` of
// `nl` lanes takes the nodes n = lane (mod nl) of every node sum and lane 0
// the closed-form terms: the nl shares add up to the factor.
template <int PREC>
__device__ __forceinline__ double3 self_energy(const ShortConsts &c,
                                               const double *__restrict__ glx,
                                               const double *__restrict__ glw,
                                               const float *__restrict__ glxf,
                                               const float *__restrict__ glwf, double zi,
                                               int lane = 0, int nl = 1)
{
    const bool swap = (c.sem & QEM_SEM_SWAP_LIMITS) != 0;
    double s = 0.0, ds0 = 0.0, ds1 = 0.0;
    for (int l = 0; l < 4; ++l) {
        const double a = l == 0 ? 0.0 : (l == 1 ? 2.0 * zi : (l == 2 ? 2.0 * c.Lz - 2.0 * zi : 2.0 * c.Lz));
        const double b = l == 0 ? 1.0 : (l == 1 ? c.g1 : (l == 2 ? c.g2 : c.g12));
        const dev::Limits L = dev::limits(c, a);
        const double Kp = swap ? L.kf1 : L.kf2, Kg = swap ? L.kf2 : L.kf1;
        const double db0 = db_dg(c, l, 0), db1 = db_dg(c, l, 1);
        if (b == 0.0 && db0 == 0.0 && db1 == 0.0) continue;
        const double hb = 0.5 * b;
        double Pv, Pd, Gv, Gd;
        nodes_energy<PREC, true>(Kp, a, 0.0, 0, c, glx, glw, glxf, glwf, nullptr, Pv, Pd, lane,
                                 nl);
        s += -hb * c.g12 * Pv;
        ds0 += -0.5 * db0 * c.g12 * Pv + hb * c.g2 * Pd;
        ds1 += -0.5 * db1 * c.g12 * Pv + hb * c.g1 * Pd;
        if (l == 0) {
            nodes_energy<PREC, true>(Kg, 0.0, 0.0, 2, c, glx, glw, glxf, glwf, nullptr, Gv, Gd,
                                     lane, nl);
            s += hb * c.g12 * Gv;
            ds0 += -hb * c.g2 * Gd;
            ds1 += -hb * c.g1 * Gd;
        } else {
            nodes_energy<PREC, true>(Kg, a, 0.0, 1, c, glx, glw, glxf, glwf, nullptr, Gv, Gd,
                                     lane, nl);
            const double ia = lane == 0 ? 1.0 / a : 0.0;   // the closed form on lane 0
            s += hb * Gv + hb * ia;
            ds0 += 0.5 * db0 * (Gv + ia) - hb * c.g2 * Gd;
            ds1 += 0.5 * db1 * (Gv + ia) - hb * c.g1 * Gd;
        }
    }
    const double f = 0.5 * c.inv2pieps;
    return make_double3(f * s, f * ds0, f * ds1);
}

// ---- tangent node sums --------------------------------------------------------
// Value and first-order tangent of sum_n w k t(k) {J1, J0}(k rho), where the
// limit K (so every node k_n = K (1 + x_n)/2 and weight) moves with dK, the
// image distance a with da and rho with drho.
using dev::Dual4;
using dev::TanSums;

// Mixed-precision value and tangent node sums.  With the node constants
// u = (1 + x_n, (1 + x_n)^2, log2 w_n, log2(w_n (1 + x_n))) the node is
// k = h u.x (h = K/2) and the summand t J is taken as in the force kernel
// (t = w k e^{..}/D / h^2, J0 and J1 = -J0' from the cubic table).  Since k
// and the table argument are both multiples of u.x, the tangent is a fixed
// combination of six (eight for the gaussian set) node sums, TanSums, which
// dev::tangent_of_sums combines with the per-set factors (the derivation is
// there); the E/D-weighted sums are taken as 1/D-weighted ones: one product
// per node (u.x t times 1/D) instead of two.
__device__ __forceinline__ Dual4 nodes_tangent_mixed(float K, float dK, float a, float da,
                                                     float rho, float drho, bool gauss,
                                                     bool self, const ShortConsts &c,
                                                     const float4 *__restrict__ tab)
{
    const float4 *__restrict__ ncon = c.ncon;
    const float h = 0.5f * K;
    const float Lz2 = 2.0f * (float)c.Lz;
    const float c2 = -Lz2 * kLog2e;
    const float cg = -kLog2e * (float)(0.25 / c.alpha);
    const float g12 = (float)c.g12;
    const float hE = h * c2;
    const float rA = a * (float)(0.5 / c.Lz);
    const float rP = 1.0f + rA;
    const float rG = cg / (c2 * c2);
    const float hR = self ? 0.0f : fminf(h * rho * c.jt_inv_dx, 0.5f * (float)(c.jt_n - 1));
    // S0, S1 = sum t p, t dp; U = u.x-weighted; Q = u.x E/D-weighted; R = u.x^2-weighted
    float S0 = 0.f, S1 = 0.f, U0 = 0.f, U1 = 0.f, Q0 = 0.f, Q1 = 0.f, R0 = 0.f, R1 = 0.f;
    // two nodes per iteration: their scalar and LDS loads overlap (measured
    // at N = 100k: HVP 4.57 -> 4.35 ms; four per iteration 4.42)
#pragma unroll 2
    for (int n = 0; n < c.Nt; ++n) {
        const float4 u = ncon[n];
        const float e = hE * u.x;
        const float E = __builtin_amdgcn_exp2f(e);
        const float num = gauss ? __builtin_amdgcn_exp2f(fmaf(e, fmaf(e, rG, rA), u.w))
                                : __builtin_amdgcn_exp2f(fmaf(e, rP, u.w));
        const float rD = __builtin_amdgcn_rcpf(fmaf(g12, E, -1.0f));
        const float t = num * rD;
        const float tu = t * u.x, tq = tu * rD;   // u.x t, u.x t / D
        if (self) {  // J0 = 1, J1 = 0
            S0 += t;
            U0 += tu;
            Q0 += tq;
            if (gauss) R0 = fmaf(u.x, tu, R0);
            continue;
        }
        const float xb = hR * u.x;
        const float tt = __builtin_amdgcn_fractf(xb);
        const float4 cc = tab[(int)xb];
        float p = fmaf(cc.w, tt, cc.z);
        float dp = cc.w;
        dp = fmaf(dp, tt, p); p = fmaf(p, tt, cc.y);
        dp = fmaf(dp, tt, p); p = fmaf(p, tt, cc.x);
        // p = J0, dp = -J1 dx
        S0 = fmaf(t, p, S0);
        S1 = fmaf(t, dp, S1);
        U0 = fmaf(tu, p, U0);
        U1 = fmaf(tu, dp, U1);
        Q0 = fmaf(tq, p, Q0);
        Q1 = fmaf(tq, dp, Q1);
        if (gauss) {
            const float tr = tu * u.x;   // u.x^2 t
            R0 = fmaf(tr, p, R0);
            R1 = fmaf(tr, dp, R1);
        }
    }
    return dev::tangent_of_sums(TanSums{S0, S1, U0, U1, Q0, Q1, R0, R1}, K, dK, a, da, rho,
                                drho, gauss, self, c);
}

template <int PREC>
__device__ __forceinline__ Dual4 nodes_tangent(double K, double dK, double a, double da,
                                               double rho, double drho, bool gauss,
                                               const ShortConsts &c,
                                               const double *__restrict__ glx,
                                               const double *__restrict__ glw,
                                               const float *__restrict__ glxf,
                                               const float *__restrict__ glwf,
                                               const float4 *__restrict__ tab, bool self)
{
    Dual4 o{0.0, 0.0, 0.0, 0.0};
    if (PREC != QEM_PREC_MIXED) {
        const double h = 0.5 * K, dh = 0.5 * dK, inv4a = 0.25 / c.alpha;
        for (int n = 0; n < c.Nt; ++n) {
            const double k = h + h * glx[n], dk = dh + dh * glx[n];
            const double w = h * glw[n], dw = dh * glw[n];
            const double E = exp(-2.0 * k * c.Lz);
            const double rD = 1.0 / (c.g12 * E - 1.0);
            const double dE = -2.0 * c.Lz * E * dk;
            const double drD = -c.g12 * dE * rD * rD;
            double num, dnum;
            if (gauss) {
                num = exp(-k * a - k * k * inv4a);
                dnum = num * (-a * dk - k * da - 2.0 * k * inv4a * dk);
            } else {
                num = exp(-k * a) * E;
                dnum = num * (-a * dk - k * da - 2.0 * c.Lz * dk);
            }
            const double t = num * rD, dt = dnum * rD + num * drD;
            const double wk = w * k, dwk = dw * k + w * dk;
            const double wkt = wk * t, dwkt = dwk * t + wk * dt;
            double J0 = 1.0, J1 = 0.0, dJ0 = 0.0, dJ1 = 0.0;
            if (!self) {
                const double x = k * rho, dx = dk * rho + k * drho;
                if (PREC == QEM_PREC_FP64) {
                    dev::bessel01_tab64(x, reinterpret_cast<const double *>(tab), c.jd_n,
                                        c.jd_inv_dx, J0, J1);
                } else {
                    J0 = j0(x);
                    J1 = j1(x);
                }
                const double j1x = x > 1e-8 ? J1 / x : 0.5;
                dJ0 = -J1 * dx;
                dJ1 = (J0 - j1x) * dx;
            }
            o.sr += wkt * J1;
            o.sz += wkt * J0;
            o.dsr += dwkt * J1 + wkt * dJ1;
            o.dsz += dwkt * J0 + wkt * dJ0;
        }
    } else {
        return nodes_tangent_mixed((float)K, (float)dK, (float)a, (float)da, (float)rho,
                                   (float)drho, gauss, self, c, tab);
    }
    return o;
}

// Tangent of F_short_ij (force on i; z force on j) along (vi, vj).
// SAVED (mixed precision): the node sums of every set come from the saved
// arrays sv (np pairs, this pair at sp) that the fused forces + energy
// kernel wrote at the same positions; rho and the limits are then formed
// as that kernel formed them, so the sums belong to these nodes.
template <int PREC, bool SAVED = false>
__device__ __forceinline__ double4 pair_force_tangent(const ShortConsts &c, const double *__restrict__ glx,
                                      const double *__restrict__ glw,
                                      const float *__restrict__ glxf,
                                      const float *__restrict__ glwf,
                                      const float4 *__restrict__ tab, double qi, double qj,
                                      const double *ri, const double *rj, const double *vi,
                                      const double *vj, const float4 *__restrict__ sv = nullptr,
                                      int64_t np = 0, int64_t sp = 0)
{
    static_assert(!SAVED || PREC == QEM_PREC_MIXED, "the saved sums are mixed precision");
    const double dx = ri[0] - rj[0], dy = ri[1] - rj[1];
    double mx, my;
    const double d2 = dev::min_image_d2(c, dx, dy, mx, my);
    if (d2 > c.rc2) return make_double4(0.0, 0.0, 0.0, 0.0);
    const double ddx = vi[0] - vj[0], ddy = vi[1] - vj[1];
    double rho, irho = 0.0;
    if constexpr (SAVED) {
        dev::mixed_rho(d2, rho, irho);
    } else {
        rho = sqrt(d2);
    }
    double drho;
    if constexpr (SAVED)
        drho = (mx * ddx + my * ddy) * irho;
    else
        drho = rho > 0.0 ? (mx * ddx + my * ddy) / rho : 0.0;
    const double zn = fabs(ri[2] - rj[2]), zp = ri[2] + rj[2], sg = jl_sign(ri[2] - rj[2]);
    const double dzn = sg * (vi[2] - vj[2]), dzp = vi[2] + vj[2];
    const bool swap = (c.sem & QEM_SEM_SWAP_LIMITS) != 0;
    double Fsr = 0.0, dFsr = 0.0, dFsz = 0.0;
    double dFszj = 0.0;   // the z tangent on j of the as-written Newton semantics
    for (int l = 0; l < 4; ++l) {
        double a, b, s;
        dev::image(c, l, zn, zp, sg, a, b, s);
        const double da = l == 0 ? dzn : (l == 1 ? dzp : (l == 2 ? -dzp : -dzn));
        double Kp, Kg, dKp, dKg, inv32;
        const double r2 = a * a + rho * rho;
        if constexpr (SAVED) {
            // FP32 limits and limit derivatives, 1/r^3 from the reciprocal
            // square root and two Newton steps
            float kp, kg, dkp, dkg;
            dev::mixed_limits_d(c, a, swap, kp, kg, dkp, dkg);
            Kp = kp;
            Kg = kg;
            dKp = (double)dkp * da;
            dKg = (double)dkg * da;
            double y = __builtin_amdgcn_rsq(r2);
            y = y * fma(-0.5 * r2 * y, y, 1.5);
            y = y * fma(-0.5 * r2 * y, y, 1.5);
            inv32 = y * y * y;
        } else {
            const dev::Limits L = dev::limits(c, a);
            Kp = swap ? L.kf1 : L.kf2;
            Kg = swap ? L.kf2 : L.kf1;
            dKp = (swap ? L.dkf1 : L.dkf2) * da;
            dKg = (swap ? L.dkf2 : L.dkf1) * da;
            inv32 = 1.0 / (r2 * sqrt(r2));
        }
        const double dinv32 = -3.0 * inv32 * (a * da + rho * drho) / r2;
        const double hb = 0.5 * b;
        const double cr = hb * rho * inv32, dcr = hb * (drho * inv32 + rho * dinv32);
        const double dcz = hb * s * (da * inv32 + a * dinv32);
        Dual4 P{0.0, 0.0, 0.0, 0.0}, G{0.0, 0.0, 0.0, 0.0};
        if (b != 0.0) {
            if constexpr (SAVED) {
                P = dev::tangent_of_sums(dev::load_tan_sums(sv, np, sp, l, 0), (float)Kp,
                                         (float)dKp, (float)a, (float)da, (float)rho,
                                         (float)drho, false, false, c);
                if (!c.single)
                    G = dev::tangent_of_sums(dev::load_tan_sums(sv, np, sp, l, 1), (float)Kg,
                                             (float)dKg, (float)a, (float)da, (float)rho,
                                             (float)drho, true, false, c);
            } else {
                P = nodes_tangent<PREC>(Kp, dKp, a, da, rho, drho, false, c, glx, glw, glxf,
                                        glwf, tab, false);
                if (!c.single)
                    G = nodes_tangent<PREC>(Kg, dKg, a, da, rho, drho, true, c, glx, glw, glxf,
                                            glwf, tab, false);
            }
        }
        Fsr += -(hb * c.g12) * P.sr + cr + hb * G.sr;
        dFsr += -(hb * c.g12) * P.dsr + dcr + hb * G.dsr;
        const double dpart = (hb * c.g12 * s) * P.dsz - dcz - (hb * s) * G.dsz;
        dFsz += dpart;
        // images 0 and 3 flip sign on the second atom, 1 and 2 do not
        dFszj += (l == 0 || l == 3) ? -dpart : dpart;
    }
    const double pref = qi * qj * c.inv2pieps;
    double4 o = make_double4(0.0, 0.0, pref * dFsz, 0.0);
    if (rho != 0.0) {
        const bool mi = (c.sem & QEM_SEM_MINIMAGE_DIR) != 0;
        const double ux = mi ? mx : dx, uy = mi ? my : dy;
        double ir;
        if constexpr (SAVED)
            ir = irho;
        else
            ir = 1.0 / rho;
        o.x = pref * (dFsr * ux * ir + Fsr * (ddx * ir - ux * drho * ir * ir));
        o.y = pref * (dFsr * uy * ir + Fsr * (ddy * ir - uy * drho * ir * ir));
    }
    if (c.sem & QEM_SEM_NEWTON_IMAGES)
        o.w = pref * dFszj;
    else
        o.w = -o.z;
    return o;
}

// Tangent of the self-image z force F_short_i along v_z.
template <int PREC>
__device__ double self_force_tangent(const ShortConsts &c, const double *__restrict__ glx,
                                     const double *__restrict__ glw,
                                     const float *__restrict__ glxf,
                                     const float *__restrict__ glwf, double qi, double zi,
                                     double vz)
{
    const bool swap = (c.sem & QEM_SEM_SWAP_LIMITS) != 0;
    double dP1 = 0.0, dP2 = 0.0, dG = 0.0;
    for (int l = 1; l <= 2; ++l) {
        const double a = l == 1 ? 2.0 * zi : 2.0 * c.Lz - 2.0 * zi;
        const double da = l == 1 ? 2.0 * vz : -2.0 * vz;
        const double b = l == 1 ? c.g1 : c.g2;
        const double s = l == 1 ? -1.0 : 1.0;
        if (b == 0.0) continue;
        const dev::Limits L = dev::limits(c, a);
        const double Kp = swap ? L.kf1 : L.kf2, Kg = swap ? L.kf2 : L.kf1;
        const double dKp = (swap ? L.dkf1 : L.dkf2) * da, dKg = (swap ? L.dkf2 : L.dkf1) * da;
        const double hb = 0.5 * b;
        const Dual4 P = nodes_tangent<PREC>(Kp, dKp, a, da, 0.0, 0.0, false, c, glx, glw, glxf,
                                            glwf, nullptr, true);
        dP1 += (hb * c.g12 * s) * P.dsz;
        dP2 += b * s * (-2.0 * da / (a * a * a));
        if (!c.single) {
            const Dual4 G = nodes_tangent<PREC>(Kg, dKg, a, da, 0.0, 0.0, true, c, glx, glw,
                                                glxf, glwf, nullptr, true);
            dG += (hb * s) * G.dsz;
        }
    }
    dP2 *= 0.5;
    return -(qi * qi) * c.inv2pieps * (-dP1 + dP2 + dG);
}

// ---- kernels ------------------------------------------------------------------------

// Per-pair records of the energy and tangent passes: float3 in mixed
// precision (the values come from FP32 node sums; the per-atom sums of the
// gathers stay FP64), double4 in FP64 (w unused).
template <int PREC>
using TanRec = typename std::conditional<PREC == QEM_PREC_MIXED, float3, double4>::type;


// Potential records (endpoint records, placed as the force records): the
// first atom's is q_j (G, dG/dgamma_1, dG/dgamma_2), the second's
// q_i (...), so the gather reads one record per pair and endpoint and no
// charges.
template <int PREC>
__global__ __launch_bounds__(256) void k_pair_energy(
    ShortConsts c, const double *__restrict__ glx, const double *__restrict__ glw,
    const float *__restrict__ glxf, const float *__restrict__ glwf,
    const float4 *__restrict__ gtab, const int32_t *__restrict__ pi,
    const int32_t *__restrict__ pj, const int32_t *__restrict__ jpos, int64_t npairs,
    const double *__restrict__ xyz, const double *__restrict__ q, TanRec<PREC> *__restrict__ out)
{
    extern __shared__ float4 lds_tab[];
    const float4 *tab = dev::stage_tables<PREC>(c, gtab, lds_tab);
    dev::walk_pairs(pi, pj, jpos, npairs, xyz, q, [&](int64_t p, const dev::PairOps &o) {
        const double4 u =
            pair_energy<PREC>(c, glx, glw, glxf, glwf, tab, o.xi, o.yi, o.zi, o.xj, o.yj, o.zj);
        dev::store4(out + p, o.qj * u.x, o.qj * u.y, o.qj * u.z, 0.0);
        dev::store4(out + npairs + o.jp, o.qi * u.x, o.qi * u.y, o.qi * u.z, 0.0);
    });
}

// Tangent records: the first atom's (dfx, dfy, dfz_i), the second's
// (-dfx, -dfy, dfz_j), placed as the force records.
template <int PREC>
__global__ __launch_bounds__(256) void k_pair_tangent(
    ShortConsts c, const double *__restrict__ glx, const double *__restrict__ glw,
    const float *__restrict__ glxf, const float *__restrict__ glwf,
    const float4 *__restrict__ gtab, const int32_t *__restrict__ pi,
    const int32_t *__restrict__ pj, const int32_t *__restrict__ jpos, int64_t npairs,
    const double *__restrict__ xyz, const double *__restrict__ q, const double *__restrict__ v,
    TanRec<PREC> *__restrict__ out)
{
    extern __shared__ float4 lds_tab[];
    const float4 *tab = dev::stage_tables<PREC>(c, gtab, lds_tab);
    dev::walk_pairs(pi, pj, jpos, npairs, xyz, q, [&](int64_t p, const dev::PairOps &o) {
        const int64_t i = o.i, j = o.j;
        const double4 d = pair_force_tangent<PREC>(c, glx, glw, glxf, glwf, tab, o.qi, o.qj,
                                                   xyz + 3 * i, xyz + 3 * j, v + 3 * i, v + 3 * j);
        dev::store4(out + p, d.x, d.y, d.z, 0.0);
        dev::store4(out + npairs + o.jp, -d.x, -d.y, d.w, 0.0);
    });
}

// k_pair_tangent (mixed precision) from the node sums that the fused forces
// + energy kernel kept for these pairs (no node loop, no Bessel table)
__global__ __launch_bounds__(256) void k_pair_tangent_saved(
    ShortConsts c, const float4 *__restrict__ sv, const int32_t *__restrict__ pi,
    const int32_t *__restrict__ pj, const int32_t *__restrict__ jpos, int64_t npairs,
    const double *__restrict__ xyz, const double *__restrict__ q, const double *__restrict__ v,
    float3 *__restrict__ out)
{
    dev::walk_pairs(pi, pj, jpos, npairs, xyz, q, [&](int64_t p, const dev::PairOps &o) {
        const int64_t i = o.i, j = o.j;
        const double4 d = pair_force_tangent<QEM_PREC_MIXED, true>(
            c, nullptr, nullptr, nullptr, nullptr, nullptr, o.qi, o.qj, xyz + 3 * i, xyz + 3 * j,
            v + 3 * i, v + 3 * j, sv, npairs, p);
        dev::store4(out + p, d.x, d.y, d.z, 0.0);
        dev::store4(out + npairs + o.jp, -d.x, -d.y, d.w, 0.0);
    });
}

// self terms per atom: phi = 2 q s_i, psi_w = 2 q ds_i/dgamma_w (overwrite)
template <int PREC>
__global__ __launch_bounds__(256) void k_self_energy(
    ShortConsts c, const double *__restrict__ glx, const double *__restrict__ glw,
    const float *__restrict__ glxf, const float *__restrict__ glwf, int64_t n,
    const double *__restrict__ xyz, const double *__restrict__ q, double *__restrict__ phi,
    double *__restrict__ psi)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const double3 se = self_energy<PREC>(c, glx, glw, glxf, glwf, xyz[3 * a + 2]);
    const double qa = q[a];
    phi[a] = 2.0 * qa * se.x;
    psi[2 * a] = 2.0 * qa * se.y;
    psi[2 * a + 1] = 2.0 * qa * se.z;
}

// per atom: phi (dE/dq) and psi_1,2 (d phi / d gamma) from the potential
// records of its pairs (both endpoints), added to the self terms already
// there.  Sixteen lanes per atom, contiguous strided reads of the atom's two
// record runs, fixed-order row sums, and the XCD-aware atom order of the
// force gather.
template <typename Rec>
__global__ __launch_bounds__(256) void k_gather_energy(
    int64_t n, int64_t npairs, const Rec *__restrict__ pot, const int32_t *__restrict__ i_start,
    const int32_t *__restrict__ i_end, const int32_t *__restrict__ i_idx,
    const int32_t *__restrict__ j_start, const int32_t *__restrict__ j_end,
    double *__restrict__ phi, double *__restrict__ psi)
{
    const int64_t a = dev::gather_atom();
    const int l = threadIdx.x & 15;
    const bool live = a < n;
    double acc[3] = {0.0, 0.0, 0.0};
    if (live)
        dev::add_endpoint_runs<Rec, 3>(acc, pot, npairs, a, i_start, i_end, i_idx, j_start, j_end,
                                       l);
    const double f = dev::row_sum16(acc[0]);
    const double s0 = dev::row_sum16(acc[1]);
    const double s1 = dev::row_sum16(acc[2]);
    if (live && l == 0) {
        phi[a] += f;
        psi[2 * a] += s0;
        psi[2 * a + 1] += s1;
    }
}

// The fused forces + energy call's gather (mixed precision): per atom the
// force records of its pairs (x, y, z: the force; w: the pair potentials
// times the other charge) and the self force, split over the atom's 16
// lanes by nodes.  Writes the force and adds the pairs' share of phi =
// dE/dq to the self energy's, which k_self_energy has written (the pairs'
// share of dE/dgamma is summed in the pair kernel).
__global__ __launch_bounds__(256) void k_gather_fe(
    int64_t n, int64_t npairs, const float4 *__restrict__ rec,
    const int32_t *__restrict__ i_start, const int32_t *__restrict__ i_end,
    const int32_t *__restrict__ i_idx, const int32_t *__restrict__ j_start,
    const int32_t *__restrict__ j_end, ShortConsts c, const double *__restrict__ glx,
    const double *__restrict__ glw, const float *__restrict__ glxf,
    const float *__restrict__ glwf, const double *__restrict__ xyz,
    const double *__restrict__ q, double *__restrict__ f, double *__restrict__ phi)
{
    const int64_t a = dev::gather_atom();
    const int l = threadIdx.x & 15;
    const bool live = a < n;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (live) {
        dev::add_endpoint_runs<float4, 4>(acc, rec, npairs, a, i_start, i_end, i_idx, j_start,
                                          j_end, l);
        acc[2] += dev::self_force_z<QEM_PREC_MIXED>(c, glx, glw, glxf, glwf, q[a],
                                                    xyz[3 * a + 2], l, 16);
    }
    const double fx = dev::row_sum16(acc[0]);
    const double fy = dev::row_sum16(acc[1]);
    const double fz = dev::row_sum16(acc[2]);
    const double ph = dev::row_sum16(acc[3]);
    if (live && l == 0) {
        f[3 * a] = fx;
        f[3 * a + 1] = fy;
        f[3 * a + 2] = fz;
        phi[a] += ph;
    }
}

// self term of the force tangent per atom: out = scale (0, 0, dF_self,z)
template <int PREC>
__global__ __launch_bounds__(256) void k_self_tangent(
    ShortConsts c, const double *__restrict__ glx, const double *__restrict__ glw,
    const float *__restrict__ glxf, const float *__restrict__ glwf, int64_t n,
    const double *__restrict__ xyz, const double *__restrict__ q, const double *__restrict__ v,
    double scale, double *__restrict__ out)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    out[3 * a] = 0.0;
    out[3 * a + 1] = 0.0;
    out[3 * a + 2] =
        scale * self_force_tangent<PREC>(c, glx, glw, glxf, glwf, q[a], xyz[3 * a + 2], v[3 * a + 2]);
}

// per atom: scale * (the tangent records of its pairs, both endpoints),
// added to the self term already there; sixteen lanes per atom and the
// XCD-aware atom order of the force gather
template <typename Rec>
__global__ __launch_bounds__(256) void k_gather_tangent(
    int64_t n, int64_t npairs, const Rec *__restrict__ tp, const int32_t *__restrict__ i_start,
    const int32_t *__restrict__ i_end, const int32_t *__restrict__ i_idx,
    const int32_t *__restrict__ j_start, const int32_t *__restrict__ j_end, double scale,
    double *__restrict__ out)
{
    const int64_t a = dev::gather_atom();
    const int l = threadIdx.x & 15;
    const bool live = a < n;
    double acc[3] = {0.0, 0.0, 0.0};
    if (live)
        dev::add_endpoint_runs<Rec, 3>(acc, tp, npairs, a, i_start, i_end, i_idx, j_start, j_end,
                                       l);
    const double fx = dev::row_sum16(acc[0]);
    const double fy = dev::row_sum16(acc[1]);
    const double fz = dev::row_sum16(acc[2]);
    if (live && l == 0) {
        out[3 * a] = scale * fx;
        out[3 * a + 1] = scale * fy;
        out[3 * a + 2] += scale * fz;
    }
}

// 1/2 sum_i q_i phi_i and 1/2 sum_i q_i psi_i,w in two fixed-order passes:
// block b sums atoms [b*chunk, (b+1)*chunk) into part[b], one block sums the
// parts in block order
constexpr int kDotBlocks = 128;
// Block b: the sums over atoms [b*chunk, (b+1)*chunk) of q phi and q psi_w
// and, with gpart, the sums of its entries [b*gchunk, (b+1)*gchunk), each
// thread a strided subset in order, the block in a fixed order.
__global__ __launch_bounds__(256) void k_half_dots_part(int64_t n, const double *__restrict__ q,
                                                        const double *__restrict__ phi,
                                                        const double *__restrict__ psi,
                                                        const double *__restrict__ gpart, int ng,
                                                        double *__restrict__ part)
{
    __shared__ double s[5][256];
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    double a = 0.0, b = 0.0, d = 0.0, g1 = 0.0, g2 = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        a += q[i] * phi[i];
        if (psi) {
            b += q[i] * psi[2 * i];
            d += q[i] * psi[2 * i + 1];
        }
    }
    if (gpart) {
        const int gchunk = (ng + gridDim.x - 1) / gridDim.x;
        const int k0 = blockIdx.x * gchunk, k1 = min(ng, k0 + gchunk);
        for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
            g1 += gpart[2 * k];
            g2 += gpart[2 * k + 1];
        }
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    s[2][threadIdx.x] = d;
    s[3][threadIdx.x] = g1;
    s[4][threadIdx.x] = g2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int r = 0; r < 5; ++r) s[r][threadIdx.x] += s[r][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int r = 0; r < 5; ++r) part[5 * blockIdx.x + r] = s[r][0];
}

// out3 = (1/2 sum q phi, 1/2 sum q psi_1 + sum gpart_1, 1/2 sum q psi_2 +
// sum gpart_2) from the block parts: one wave, lane l adding parts l,
// l + 64, ... in order, then a butterfly over the 64 lanes (a fixed order;
// every lane forms the same sums).  One thread summing all 5 nb parts took
// 25 us of dependent loads.
__global__ __launch_bounds__(64) void k_half_dots_final(int nb, const double *__restrict__ part,
                                                        double *__restrict__ out3)
{
    const int l = threadIdx.x;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = l; i < nb; i += 64)
#pragma unroll
        for (int r = 0; r < 5; ++r) v[r] += part[5 * i + r];
#pragma unroll
    for (int r = 0; r < 5; ++r)
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v[r] += __shfl_xor(v[r], d);
    if (l == 0) {
        out3[0] = 0.5 * v[0];
        out3[1] = 0.5 * v[1] + v[3];
        out3[2] = 0.5 * v[2] + v[4];
    }
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

using dev::pair_grid;

}  // namespace

hipError_t launch_short_energy(const ShortConsts &c, int precision, const ShortWork &ws,
                               const NeighborList &nl, int64_t n, const double *d_xyz,
                               const double *d_q, double *d_pot, double *d_phi,
                               double *d_psi, hipStream_t s,
                               const std::function<void()> *after_pair, hipStream_t self_s)
{
    if (n <= 0) return hipSuccess;
    if (!self_s) self_s = s;
    const int64_t np = nl.npairs;
    const int32_t *iidx = nl.has_i_idx ? nl.i_idx.as<int32_t>() : nullptr;
    const int32_t *jpos = nl.jpos.as<int32_t>();
    double4 *pot = reinterpret_cast<double4 *>(d_pot);
    float3 *potf = reinterpret_cast<float3 *>(d_pot);
    const size_t lds = sizeof(float4) * (size_t)ws.jt_n;
    if (precision == QEM_PREC_MIXED) {
        if (np > 0)
            k_pair_energy<QEM_PREC_MIXED><<<pair_grid(np), 256, lds, s>>>(
                c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
                ws.glwf.as<float>(), ws.jtab.as<float4>(), nl.pi.as<int32_t>(),
                nl.pj.as<int32_t>(), jpos, np, d_xyz, d_q, potf);
        k_self_energy<QEM_PREC_MIXED><<<nblk(n, 256), 256, 0, self_s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
            ws.glwf.as<float>(), n, d_xyz, d_q, d_phi, d_psi);
        if (after_pair) (*after_pair)();
    } else {
        if (np > 0) {
            if (c.jd_n > 0)
                k_pair_energy<QEM_PREC_FP64><<<pair_grid(np), 256, dev::lds_bytes64(c), s>>>(
                    c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
                    ws.glwf.as<float>(), nullptr, nl.pi.as<int32_t>(), nl.pj.as<int32_t>(), jpos,
                    np, d_xyz, d_q, pot);
            else
                k_pair_energy<kPrecFP64Lib><<<pair_grid(np), 256, 0, s>>>(
                    c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
                    ws.glwf.as<float>(), nullptr, nl.pi.as<int32_t>(), nl.pj.as<int32_t>(), jpos,
                    np, d_xyz, d_q, pot);
        }
        k_self_energy<QEM_PREC_FP64><<<nblk(n, 256), 256, 0, self_s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
            ws.glwf.as<float>(), n, d_xyz, d_q, d_phi, d_psi);
        if (after_pair) (*after_pair)();
    }
    const unsigned nb = (nblk(16 * n, 256) + 7) / 8 * 8;
    if (precision == QEM_PREC_MIXED)
        k_gather_energy<float3><<<nb, 256, 0, s>>>(
            n, np, potf, nl.i_start.as<int32_t>(), nl.i_end.as<int32_t>(), iidx,
            nl.j_start.as<int32_t>(), nl.j_end.as<int32_t>(), d_phi, d_psi);
    else
        k_gather_energy<double4><<<nb, 256, 0, s>>>(
            n, np, pot, nl.i_start.as<int32_t>(), nl.i_end.as<int32_t>(), iidx,
            nl.j_start.as<int32_t>(), nl.j_end.as<int32_t>(), d_phi, d_psi);
    return hipGetLastError();
}

hipError_t launch_self_energy(const ShortConsts &c, int precision, const ShortWork &ws,
                              int64_t n, const double *d_xyz, const double *d_q, double *d_phi,
                              double *d_psi, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (precision == QEM_PREC_MIXED)
        k_self_energy<QEM_PREC_MIXED><<<nblk(n, 256), 256, 0, s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
            ws.glwf.as<float>(), n, d_xyz, d_q, d_phi, d_psi);
    else
        k_self_energy<QEM_PREC_FP64><<<nblk(n, 256), 256, 0, s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
            ws.glwf.as<float>(), n, d_xyz, d_q, d_phi, d_psi);
    return hipGetLastError();
}

hipError_t launch_gather_forces_energy(const ShortConsts &c, const ShortWork &ws,
                                       const NeighborList &nl, int64_t n, const double *d_xyz,
                                       const double *d_q, const double *d_rec, double *d_f,
                                       double *d_phi, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int32_t *iidx = nl.has_i_idx ? nl.i_idx.as<int32_t>() : nullptr;
    const unsigned nb = (nblk(16 * n, 256) + 7) / 8 * 8;
    k_gather_fe<<<nb, 256, 0, s>>>(
        n, nl.npairs, reinterpret_cast<const float4 *>(d_rec), nl.i_start.as<int32_t>(),
        nl.i_end.as<int32_t>(), iidx, nl.j_start.as<int32_t>(), nl.j_end.as<int32_t>(), c,
        ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(), ws.glwf.as<float>(),
        d_xyz, d_q, d_f, d_phi);
    return hipGetLastError();
}

hipError_t launch_short_tangent(const ShortConsts &c, int precision, const ShortWork &ws,
                                const NeighborList &nl, int64_t n, const double *d_xyz,
                                const double *d_q, const double *d_v, double *d_tp,
                                double scale, double *d_out, hipStream_t s,
                                const std::function<void()> *after_pair, hipStream_t self_s,
                                const float *d_sums)
{
    if (n <= 0) return hipSuccess;
    if (!self_s) self_s = s;
    const int64_t np = nl.npairs;
    const int32_t *iidx = nl.has_i_idx ? nl.i_idx.as<int32_t>() : nullptr;
    const int32_t *jpos = nl.jpos.as<int32_t>();
    double4 *tp = reinterpret_cast<double4 *>(d_tp);
    float3 *tpf = reinterpret_cast<float3 *>(d_tp);
    const size_t lds = sizeof(float4) * (size_t)ws.jt_n;
    if (precision == QEM_PREC_MIXED) {
        if (np > 0 && d_sums)
            k_pair_tangent_saved<<<pair_grid(np), 256, 0, s>>>(
                c, reinterpret_cast<const float4 *>(d_sums), nl.pi.as<int32_t>(),
                nl.pj.as<int32_t>(), jpos, np, d_xyz, d_q, d_v, tpf);
        else if (np > 0)
            k_pair_tangent<QEM_PREC_MIXED><<<pair_grid(np), 256, lds, s>>>(
                c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
                ws.glwf.as<float>(), ws.jtab.as<float4>(), nl.pi.as<int32_t>(),
                nl.pj.as<int32_t>(), jpos, np, d_xyz, d_q, d_v, tpf);
        k_self_tangent<QEM_PREC_MIXED><<<nblk(n, 256), 256, 0, self_s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
            ws.glwf.as<float>(), n, d_xyz, d_q, d_v, scale, d_out);
        if (after_pair) (*after_pair)();
    } else {
        if (np > 0) {
            if (c.jd_n > 0)
                k_pair_tangent<QEM_PREC_FP64><<<pair_grid(np), 256, dev::lds_bytes64(c), s>>>(
                    c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
                    ws.glwf.as<float>(), nullptr, nl.pi.as<int32_t>(), nl.pj.as<int32_t>(), jpos,
                    np, d_xyz, d_q, d_v, tp);
            else
                k_pair_tangent<kPrecFP64Lib><<<pair_grid(np), 256, 0, s>>>(
                    c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
                    ws.glwf.as<float>(), nullptr, nl.pi.as<int32_t>(), nl.pj.as<int32_t>(), jpos,
                    np, d_xyz, d_q, d_v, tp);
        }
        k_self_tangent<QEM_PREC_FP64><<<nblk(n, 256), 256, 0, self_s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
            ws.glwf.as<float>(), n, d_xyz, d_q, d_v, scale, d_out);
        if (after_pair) (*after_pair)();
    }
    const unsigned nb = (nblk(16 * n, 256) + 7) / 8 * 8;
    if (precision == QEM_PREC_MIXED)
        k_gather_tangent<float3><<<nb, 256, 0, s>>>(
            n, np, tpf, nl.i_start.as<int32_t>(), nl.i_end.as<int32_t>(), iidx,
            nl.j_start.as<int32_t>(), nl.j_end.as<int32_t>(), scale, d_out);
    else
        k_gather_tangent<double4><<<nb, 256, 0, s>>>(
            n, np, tp, nl.i_start.as<int32_t>(), nl.i_end.as<int32_t>(), iidx,
            nl.j_start.as<int32_t>(), nl.j_end.as<int32_t>(), scale, d_out);
    return hipGetLastError();
}

hipError_t launch_half_dots(int64_t n, const double *d_q, const double *d_phi,
                            const double *d_psi, double *d_out3, double *d_part, hipStream_t s,
                            const double *d_gpart, int ngpart)
{
    k_half_dots_part<<<kDotBlocks, 256, 0, s>>>(n, d_q, d_phi, d_psi, d_gpart, ngpart, d_part);
    k_half_dots_final<<<1, 64, 0, s>>>(kDotBlocks, d_part, d_out3);
    return hipGetLastError();
}

int half_dots_scratch() { return 5 * kDotBlocks; }

}  // namespace qem
