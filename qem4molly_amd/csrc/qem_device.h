// Device helpers shared by the short-range force, energy and tangent kernels.
#pragma once

#include "qem_internal.h"

namespace qem {
namespace dev {

constexpr float kLog2e = 1.4426950408889634f;

__device__ __forceinline__ double jl_sign(double v) { return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0); }

// FP32 J0/J1 from the LDS table: per interval of width dx (= 1/inv_dx) a
// cubic p(t) ~ J0(x0 + t dx), t in [0, 1), fitted at Chebyshev points and
// stored as one float4 (c0..c3).  J1 = -J0' = -p'(t) / dx from the same
// polynomial.  The index is clamped to the last interval.
__device__ __forceinline__ void bessel01_tab(float x, const float4 *__restrict__ tab, int ntab,
                                             float inv_dx, float &J0, float &J1)
{
    const float xb = x * inv_dx;
    int m = (int)xb;
    m = m < ntab - 1 ? m : ntab - 1;
    const float t = __builtin_amdgcn_fractf(xb);
    const float4 c = tab[m];
    float p = fmaf(c.w, t, c.z);
    float dp = c.w;
    dp = fmaf(dp, t, p); p = fmaf(p, t, c.y);
    dp = fmaf(dp, t, p); p = fmaf(p, t, c.x);
    J0 = p;
    J1 = -dp * inv_dx;
}

// 2^y in FP64: y = m + f with m = rint(y) and |f| <= 1/2 (exact), 2^f from
// the degree-10 polynomial interpolating 2^f at the 11 Chebyshev points of
// [-1/2, 1/2] (near-minimax; coefficients from csrc/gen_exp2_coeffs.py,
// relative error 4.7e-16 in double Horner), scaled by 2^m through the
// exponent (v_ldexp_f64; results below the FP64 range go to zero).
constexpr double kExp2C[11] = {
    1.0,
    0.69314718055995,
    0.24022650695910097,
    0.05550410866444772,
    0.009618129107606888,
    0.0013333558230164974,
    0.0001540353044173605,
    1.5252657260200837e-05,
    1.321544258792169e-06,
    1.0208690299958306e-07,
    7.072585949269223e-09};

__device__ __forceinline__ double exp2_fp64(double y)
{
    const double m = rint(y);
    const double f = y - m;
    double p = kExp2C[10];
#pragma unroll
    for (int i = 9; i >= 0; --i) p = fma(p, f, kExp2C[i]);
    return ldexp(p, (int)m);
}

// 1/d in FP64: the hardware reciprocal refined by two Newton steps
// r <- r + r (1 - d r) (for |d| >= 1 - |gamma_1 gamma_2| > 0, the kernels'
// denominators, no special cases arise).
__device__ __forceinline__ double rcp_fp64(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// FP64 J0/J1 from the FP64 table: per interval of width dx a degree-7
// polynomial p(t) ~ J0(x0 + t dx) (8 doubles in a row of 9), J1 = -p'(t) / dx
// from the same polynomial.  The index is clamped to the last interval.
__device__ __forceinline__ void bessel01_tab64(double x, const double *__restrict__ tab, int ntab,
                                               double inv_dx, double &J0, double &J1)
{
    const double xb = x * inv_dx;
    int m = (int)xb;
    m = m < ntab - 1 ? m : ntab - 1;
    const double t = xb - (double)m;
    const double *__restrict__ c = tab + 9 * m;
    double p = c[7], dp = 0.0;
#pragma unroll
    for (int k = 6; k >= 0; --k) {
        dp = fma(dp, t, p);
        p = fma(p, t, c[k]);
    }
    J0 = p;
    J1 = -dp * inv_dx;
}

// Copies the FP64 Bessel table into this block's LDS (9 doubles per interval).
__device__ __forceinline__ const double *stage_table64(int jd_n, const double *__restrict__ gtab,
                                                       void *lds)
{
    double *d = static_cast<double *>(lds);
    for (int t = threadIdx.x; t < 9 * jd_n; t += blockDim.x) d[t] = gtab[t];
    __syncthreads();
    return d;
}

// Stages the kernel's Bessel table in LDS: the FP32 cubic table in mixed
// precision, the FP64 table in FP64 (returned through the same pointer,
// to be read as doubles), or null when an FP64 context has no table.
template <int PREC>
__device__ __forceinline__ const float4 *stage_tables(const ShortConsts &c,
                                                      const float4 *__restrict__ gtab,
                                                      float4 *lds);

// LDS bytes of the FP64 table (0 without one)
inline size_t lds_bytes64(const ShortConsts &c) { return (size_t)c.jd_n * 9 * sizeof(double); }

// Copies the Bessel table into this block's LDS.
__device__ __forceinline__ const float4 *stage_table(int jt_n, const float4 *__restrict__ gtab,
                                                     float4 *lds)
{
    for (int t = threadIdx.x; t < jt_n; t += blockDim.x) lds[t] = gtab[t];
    __syncthreads();
    return lds;
}

template <int PREC>
__device__ __forceinline__ const float4 *stage_tables(const ShortConsts &c,
                                                      const float4 *__restrict__ gtab,
                                                      float4 *lds)
{
    if (PREC == QEM_PREC_MIXED) return stage_table(c.jt_n, gtab, lds);
    if (PREC == kPrecFP64Lib || c.jd_n == 0) return nullptr;
    return reinterpret_cast<const float4 *>(stage_table64(c.jd_n, c.jdtab, lds));
}

// Quadrature limits of one image (src/greens_function.jl:26-27) and their
// derivatives with respect to the image distance a.
struct Limits {
    double kf1, kf2, dkf1, dkf2;
};

__device__ __forceinline__ Limits limits(const ShortConsts &c, double a)
{
    Limits L;
    L.kf1 = -c.ln_acc / (2.0 * c.Lz + a);
    const double root = sqrt(4.0 * c.alpha * c.alpha * a * a - 4.0 * c.alpha * c.ln_acc);
    L.kf2 = root - 2.0 * c.alpha * a;
    L.dkf1 = -L.kf1 / (2.0 * c.Lz + a);
    L.dkf2 = 4.0 * c.alpha * c.alpha * a / root - 2.0 * c.alpha;
    return L;
}

// image l of a pair: distance a, factor b, sign s (src/greens_function.jl:17-24)
__device__ __forceinline__ void image(const ShortConsts &c, int l, double zn, double zp, double sg,
                                      double &a, double &b, double &s)
{
    a = l == 0 ? zn : (l == 1 ? zp : (l == 2 ? 2.0 * c.Lz - zp : 2.0 * c.Lz - zn));
    b = l == 0 ? 1.0 : (l == 1 ? c.g1 : (l == 2 ? c.g2 : c.g12));
    s = l == 0 ? -sg : (l == 1 ? -1.0 : (l == 2 ? 1.0 : sg));
}

// minimum-image xy separation (no contraction: the cutoff test must match
// the neighbour-list builder bit for bit)
__device__ __forceinline__ double min_image_d2(const ShortConsts &c, double dx, double dy,
                                               double &mx, double &my)
{
#pragma clang fp contract(off)
    mx = dx - c.Lx * rint(dx / c.Lx);
    my = dy - c.Ly * rint(dy / c.Ly);
    return mx * mx + my * my;
}

// lane permutation of a double (every lane of every row written)
template <int CTRL>
__device__ __forceinline__ double dpp_perm_f64(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// row broadcast of a double; rows outside ROW_MASK read 0
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_bcast_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over each row of 16 lanes in a fixed order (pairs, quads, half rows,
// rows); every lane of the row receives its row's sum.
__device__ __forceinline__ double row_sum16(double v)
{
    v += dpp_perm_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_perm_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_perm_f64<0x141>(v);  // row_half_mirror
    v += dpp_perm_f64<0x140>(v);  // row_mirror
    return v;
}

}  // namespace dev
}  // namespace qem
