// Pole splitting on gfx950: the short-range forces for gamma_1 gamma_2 > 1,
// where D(k) = g12 e^{-2 k Lz} - 1 vanishes at k0 = ln(g12) / (2 Lz) inside
// every integration range.
//
// Reference: the principal-value rule of src/python_lib/Force_short_norm_njit.py
// (segment sums 75-93, 186-229; the intcore functions 23-73, 138-184; node
// counts Force_short_range_njit.py:258-286) and the regular part of 1/D,
// green_norm_njit.py:13-32, applied to the Julia integrands
// (src/Force_short.jl:4-22) with the four images summed, and assembled as
// F_short_ij / F_short_i (src/Force_short.jl:25-107).
//
// The rule's nodes do not depend on the pair: per set (point, gaussian) the
// host gives each node k_n one constant c_n = w_n k_n pf(k_n) R_n (pf the
// set's factor g12 e^{-2kLz} or e^{-k^2/4alpha}, R_n = 1/D(k_n), or
// 1/D + 1/x on the middle segment), which weights all three integrands:
//   rho:        c_n gu(k) J1(k rho),         gu = 1/2 sum_l b_l e^{-k a_l}
//   z, odd:     c_n go(k) J0(k rho),         go = 1/2 sum_{l=1,4} s_l b_l e^{-k a_l}
//   z, even:    c_n ge(k) J0(k rho),         ge = 1/2 sum_{l=2,3} s_l b_l e^{-k a_l}
// (the odd images change sign under z_i <-> z_j, which gives the j-side z
// force).  Per pair the pole terms N(k0) W_x - W_m N'(k0) / (2 Lz) are
// added, N the integrand times D: in FP64 N' is the reference's central
// difference with step 0.01 eps, in mixed precision its closed form (the
// difference of two FP32 values would lose every digit).
//
// One lane per pair, as in the pair kernel; 2 (N_t + 10) nodes per pair.
#include "qem_device.h"

namespace qem {
namespace {

using dev::bessel01_tab;
using dev::bessel01_tab64;
using dev::jl_sign;
using dev::kLog2e;

// per-pair image data (greens_element_ij_init, src/greens_function.jl:17-30)
struct PoleImg {
    double a[4], hb[4], hs[4];  // a_l, b_l / 2, s_l b_l / 2
    double rho;
};

// J0, J1 at x (x = 0 exactly gives 1, 0)
template <int PREC>
__device__ __forceinline__ void bessel_pair(double x, const float4 *__restrict__ tab,
                                            const ShortConsts &c, double &J0, double &J1)
{
    if (x == 0.0) {
        J0 = 1.0;
        J1 = 0.0;
    } else if (PREC == QEM_PREC_MIXED) {
        float j0, j1;
        bessel01_tab((float)x, tab, c.jt_n, c.jt_inv_dx, j0, j1);
        J0 = j0;
        J1 = j1;
    } else if (PREC == QEM_PREC_FP64) {
        bessel01_tab64(x, reinterpret_cast<const double *>(tab), c.jd_n, c.jd_inv_dx, J0, J1);
    } else {
        J0 = j0(x);
        J1 = j1(x);
    }
}

// the three integrand factors at k, without the set's factor and 1/D:
// (k gu J1, k go J0, k ge J0)
template <int PREC>
__device__ __forceinline__ double3 pole_core(const PoleImg &m, double k,
                                             const float4 *__restrict__ tab,
                                             const ShortConsts &c)
{
    double E[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) E[l] = exp(-k * m.a[l]);
    const double gu = m.hb[0] * E[0] + m.hb[1] * E[1] + m.hb[2] * E[2] + m.hb[3] * E[3];
    const double go = m.hs[0] * E[0] + m.hs[3] * E[3];
    const double ge = m.hs[1] * E[1] + m.hs[2] * E[2];
    double J0, J1;
    bessel_pair<PREC>(k * m.rho, tab, c, J0, J1);
    return make_double3(k * gu * J1, k * go * J0, k * ge * J0);
}

// node sums of one set: (sum c gu J1, sum c go J0, sum c ge J0)
template <int PREC>
__device__ __forceinline__ double3 pole_nodes(const PoleImg &m, int set,
                                              const float4 *__restrict__ tab,
                                              const ShortConsts &c)
{
    const int off = set ? c.pm[0] : 0, M = c.pm[set];
    if (PREC != QEM_PREC_MIXED) {
        double r = 0.0, o = 0.0, e = 0.0;
        for (int n = 0; n < M; ++n) {
            const double2 nd = c.pnode[off + n];
            const double k = nd.x;
            double E[4];
#pragma unroll
            for (int l = 0; l < 4; ++l) E[l] = exp(-k * m.a[l]);
            const double gu = m.hb[0] * E[0] + m.hb[1] * E[1] + m.hb[2] * E[2] + m.hb[3] * E[3];
            const double go = m.hs[0] * E[0] + m.hs[3] * E[3];
            const double ge = m.hs[1] * E[1] + m.hs[2] * E[2];
            double J0, J1;
            bessel_pair<PREC>(k * m.rho, tab, c, J0, J1);
            r += nd.y * (gu * J1);
            o += nd.y * (go * J0);
            e += nd.y * (ge * J0);
        }
        return make_double3(r, o, e);
    } else {
        float ca[4], hb[4], hs[4];
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            ca[l] = (float)(-m.a[l]) * kLog2e;
            hb[l] = (float)m.hb[l];
            hs[l] = (float)m.hs[l];
        }
        const float rho = (float)m.rho;
        float r = 0.f, o = 0.f, e = 0.f;
        for (int n = 0; n < M; ++n) {
            const float2 nd = c.pnodef[off + n];
            const float k = nd.x;
            float E[4];
#pragma unroll
            for (int l = 0; l < 4; ++l) E[l] = __builtin_amdgcn_exp2f(k * ca[l]);
            const float gu = fmaf(hb[3], E[3], fmaf(hb[2], E[2], fmaf(hb[1], E[1], hb[0] * E[0])));
            const float go = fmaf(hs[3], E[3], hs[0] * E[0]);
            const float ge = fmaf(hs[2], E[2], hs[1] * E[1]);
            float J0 = 1.f, J1 = 0.f;
            if (rho != 0.f) bessel01_tab(k * rho, tab, c.jt_n, c.jt_inv_dx, J0, J1);
            const float cj0 = nd.y * J0;
            r = fmaf(nd.y * gu, J1, r);
            o = fmaf(cj0, go, o);
            e = fmaf(cj0, ge, e);
        }
        return make_double3(r, o, e);
    }
}

// both sets with their pole terms: out[set] = (rho, z odd, z even) integrals
template <int PREC>
__device__ void pole_integrals(const PoleImg &m, const float4 *__restrict__ tab,
                               const ShortConsts &c, bool with_gauss, double3 out[2])
{
    const double k0 = c.pk0, two_lz = 2.0 * c.Lz;
    if (PREC != QEM_PREC_MIXED) {
        // the reference's central difference (Force_short_norm_njit.py:29-32)
        const double k1 = k0 - c.pdel, k2 = k0 + c.pdel;
        const double3 cm = pole_core<PREC>(m, k1, tab, c);
        const double3 c0 = pole_core<PREC>(m, k0, tab, c);
        const double3 cp = pole_core<PREC>(m, k2, tab, c);
        for (int s = 0; s < (with_gauss ? 2 : 1); ++s) {
            const double3 A = pole_nodes<PREC>(m, s, tab, c);
            const double fm = c.ppf[s][0], f0 = c.ppf[s][1], fp = c.ppf[s][2];
            const double dk = k2 - k1;
            const double dr = (fp * cp.x - fm * cm.x) / dk;
            const double dO = (fp * cp.y - fm * cm.y) / dk;
            const double de = (fp * cp.z - fm * cm.z) / dk;
            out[s].x = A.x + (f0 * c0.x) * c.pWx[s] - c.pWm[s] * (dr / two_lz);
            out[s].y = A.y + (f0 * c0.y) * c.pWx[s] - c.pWm[s] * (dO / two_lz);
            out[s].z = A.z + (f0 * c0.z) * c.pWx[s] - c.pWm[s] * (de / two_lz);
        }
    } else {
        // N and N' at k0 in closed form (FP64 arithmetic, table Bessel values)
        double E[4], dE[4];
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            E[l] = exp(-k0 * m.a[l]);
            dE[l] = -m.a[l] * E[l];
        }
        const double gu = m.hb[0] * E[0] + m.hb[1] * E[1] + m.hb[2] * E[2] + m.hb[3] * E[3];
        const double du = m.hb[0] * dE[0] + m.hb[1] * dE[1] + m.hb[2] * dE[2] + m.hb[3] * dE[3];
        const double go = m.hs[0] * E[0] + m.hs[3] * E[3];
        const double dgo = m.hs[0] * dE[0] + m.hs[3] * dE[3];
        const double ge = m.hs[1] * E[1] + m.hs[2] * E[2];
        const double dge = m.hs[1] * dE[1] + m.hs[2] * dE[2];
        double J0, J1;
        bessel_pair<PREC>(k0 * m.rho, tab, c, J0, J1);
        // k-derivatives of k gu J1 and k g J0 (J1' = J0 - J1/x, J0' = -J1)
        const double nr = k0 * gu * J1;
        // d(k gu J1)/dk = gu J1 + k gu' J1 + k gu rho (J0 - J1/(k rho)) = k (gu' J1 + gu rho J0)
        const double dnr = k0 * (du * J1 + gu * m.rho * J0);
        const double no = k0 * go * J0, ne = k0 * ge * J0;
        const double dno = (go + k0 * dgo) * J0 - k0 * go * m.rho * J1;
        const double dne = (ge + k0 * dge) * J0 - k0 * ge * m.rho * J1;
        for (int s = 0; s < (with_gauss ? 2 : 1); ++s) {
            const double3 A = pole_nodes<PREC>(m, s, tab, c);
            const double f0 = c.ppf[s][1], df = c.pdpf[s];
            // d(pf N)/dk = pf' N + pf N'
            const double dr = df * nr + f0 * dnr;
            const double dO = df * no + f0 * dno;
            const double de = df * ne + f0 * dne;
            out[s].x = A.x + (f0 * nr) * c.pWx[s] - c.pWm[s] * (dr / two_lz);
            out[s].y = A.y + (f0 * no) * c.pWx[s] - c.pWm[s] * (dO / two_lz);
            out[s].z = A.z + (f0 * ne) * c.pWx[s] - c.pWm[s] * (de / two_lz);
        }
    }
}

struct PolePairForce {
    double fx, fy, fz, fzj;
};

// F_short_ij with the pole rule (src/Force_short.jl:25-78 assembly)
template <int PREC>
__device__ __forceinline__ PolePairForce pole_pair_force(const ShortConsts &c, const float4 *__restrict__ tab,
                                         double qi, double qj, double xi, double yi, double zi,
                                         double xj, double yj, double zj)
{
    PolePairForce out{0.0, 0.0, 0.0, 0.0};
    const double dx = xi - xj, dy = yi - yj;
    double mx, my, d2;
    {
#pragma clang fp contract(off)
        mx = dx - c.Lx * rint(dx / c.Lx);
        my = dy - c.Ly * rint(dy / c.Ly);
        d2 = mx * mx + my * my;
    }
    if (d2 > c.rc2) return out;
    PoleImg m;
    m.rho = sqrt(d2);
    const double zn = fabs(zi - zj), zp = zi + zj, sg = jl_sign(zi - zj);
    const double b[4] = {1.0, c.g1, c.g2, c.g12};
    const double s[4] = {-sg, -1.0, 1.0, sg};
    m.a[0] = zn;
    m.a[1] = zp;
    m.a[2] = 2.0 * c.Lz - zp;
    m.a[3] = 2.0 * c.Lz - zn;
    double cr = 0.0, cz[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        m.hb[l] = 0.5 * b[l];
        m.hs[l] = 0.5 * s[l] * b[l];
        const double r2 = m.a[l] * m.a[l] + d2;
        const double inv32 = 1.0 / (r2 * sqrt(r2));
        cr += 0.5 * b[l] * m.rho * inv32;
        cz[l] = 0.5 * b[l] * m.a[l] * s[l] * inv32;
    }
    double3 I[2];
    I[1] = make_double3(0.0, 0.0, 0.0);
    pole_integrals<PREC>(m, tab, c, !c.single, I);
    const double Fsr = -I[0].x + cr + I[1].x;
    const double odd = I[0].y - (cz[0] + cz[3]) - I[1].y;
    const double even = I[0].z - (cz[1] + cz[2]) - I[1].z;
    const double pref = qi * qj * c.inv2pieps;
    if (m.rho != 0.0) {
        const bool mi = (c.sem & QEM_SEM_MINIMAGE_DIR) != 0;
        const double ux = mi ? mx : dx, uy = mi ? my : dy;
        out.fx = pref * (Fsr * ux / m.rho);
        out.fy = pref * (Fsr * uy / m.rho);
    }
    out.fz = pref * (odd + even);
    out.fzj = (c.sem & QEM_SEM_NEWTON_IMAGES) ? pref * (even - odd) : -out.fz;
    return out;
}

// F_short_i with the pole rule (src/Force_short.jl:82-107 assembly)
template <int PREC>
__device__ double pole_self_z(const ShortConsts &c, double qi, double zi)
{
    PoleImg m;
    m.rho = 0.0;
    const double b[4] = {1.0, c.g1, c.g2, c.g12};
    const double s[4] = {0.0, -1.0, 1.0, 0.0};
    m.a[0] = 0.0;
    m.a[1] = 2.0 * zi;
    m.a[2] = 2.0 * c.Lz - 2.0 * zi;
    m.a[3] = 2.0 * c.Lz;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        m.hb[l] = 0.5 * b[l];
        m.hs[l] = 0.5 * s[l] * b[l];
    }
    double3 I[2];
    I[1] = make_double3(0.0, 0.0, 0.0);
    pole_integrals<PREC>(m, nullptr, c, !c.single, I);
    double P2 = 0.0;
#pragma unroll
    for (int l = 1; l <= 2; ++l) P2 += b[l] * m.a[l] * s[l] / (m.a[l] * m.a[l] * m.a[l]);
    P2 *= 0.5;
    const double Fsz = -(I[0].y + I[0].z) + P2 + (I[1].y + I[1].z);
    return -(qi * qi) * c.inv2pieps * Fsz;
}

template <int PREC>
__global__ __launch_bounds__(256) void k_pair_list_pole(
    ShortConsts c, const float4 *__restrict__ gtab, const int32_t *__restrict__ pi,
    const int32_t *__restrict__ pj, const int32_t *__restrict__ jpos, int64_t npairs,
    const double *__restrict__ xyz, const double *__restrict__ q, double4 *__restrict__ fpair)
{
    extern __shared__ float4 lds_tab[];
    const float4 *tab = dev::stage_tables<PREC>(c, gtab, lds_tab);
    dev::walk_pairs(pi, pj, jpos, npairs, xyz, q, [&](int64_t p, const dev::PairOps &o) {
        const PolePairForce f = pole_pair_force<PREC>(c, tab, o.qi, o.qj, o.xi, o.yi, o.zi, o.xj,
                                                      o.yj, o.zj);
        // endpoint records: (fx, fy, fz) for the first atom, (-fx, -fy, fz_j)
        // for the second, at its place in the second-endpoint grouping
        fpair[p] = make_double4(f.fx, f.fy, f.fz, 0.0);
        fpair[npairs + o.jp] = make_double4(-f.fx, -f.fy, f.fzj, 0.0);
    });
}

template <int PREC>
__global__ __launch_bounds__(256) void k_pair_direct_pole(
    ShortConsts c, const float4 *__restrict__ gtab, int64_t m, const double *__restrict__ ri,
    const double *__restrict__ rj, const double *__restrict__ qi, const double *__restrict__ qj,
    double4 *__restrict__ fpair)
{
    extern __shared__ float4 lds_tab[];
    const float4 *tab = dev::stage_tables<PREC>(c, gtab, lds_tab);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m; p += stride) {
        const PolePairForce f = pole_pair_force<PREC>(c, tab, qi[p], qj[p], ri[3 * p],
                                                      ri[3 * p + 1], ri[3 * p + 2], rj[3 * p],
                                                      rj[3 * p + 1], rj[3 * p + 2]);
        fpair[p] = make_double4(f.fx, f.fy, f.fz, f.fzj);
    }
}

template <int PREC>
__global__ __launch_bounds__(256) void k_self_pole(ShortConsts c, int64_t n,
                                                   const double *__restrict__ xyz,
                                                   const double *__restrict__ q, int zero_xy,
                                                   double *__restrict__ f)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    f[3 * a + 2] = pole_self_z<PREC>(c, q[a], xyz[3 * a + 2]);
    if (zero_xy) {
        f[3 * a] = 0.0;
        f[3 * a + 1] = 0.0;
    }
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

inline unsigned grid_for(int64_t m)
{
    const int64_t want = (m + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, 256 * 16));
}

}  // namespace

hipError_t launch_pair_pole(const ShortConsts &c, int precision, const ShortWork &ws,
                            const int32_t *d_pi, const int32_t *d_pj, const int32_t *d_jpos,
                            int64_t npairs, const double *d_xyz, const double *d_q,
                            double *d_fpair, hipStream_t s)
{
    if (npairs <= 0) return hipSuccess;
    double4 *out = reinterpret_cast<double4 *>(d_fpair);
    if (precision == QEM_PREC_MIXED)
        k_pair_list_pole<QEM_PREC_MIXED><<<dev::pair_grid(npairs), 256, sizeof(float4) * ws.jt_n, s>>>(
            c, ws.jtab.as<float4>(), d_pi, d_pj, d_jpos, npairs, d_xyz, d_q, out);
    else if (c.jd_n > 0)
        k_pair_list_pole<QEM_PREC_FP64><<<dev::pair_grid(npairs), 256, dev::lds_bytes64(c), s>>>(
            c, nullptr, d_pi, d_pj, d_jpos, npairs, d_xyz, d_q, out);
    else
        k_pair_list_pole<kPrecFP64Lib><<<dev::pair_grid(npairs), 256, 0, s>>>(
            c, nullptr, d_pi, d_pj, d_jpos, npairs, d_xyz, d_q, out);
    return hipGetLastError();
}

hipError_t launch_pair_direct_pole(const ShortConsts &c, int precision, const ShortWork &ws,
                                   int64_t m, const double *d_ri, const double *d_rj,
                                   const double *d_qi, const double *d_qj, double *d_fpair,
                                   hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    double4 *out = reinterpret_cast<double4 *>(d_fpair);
    if (precision == QEM_PREC_MIXED)
        k_pair_direct_pole<QEM_PREC_MIXED><<<grid_for(m), 256, sizeof(float4) * ws.jt_n, s>>>(
            c, ws.jtab.as<float4>(), m, d_ri, d_rj, d_qi, d_qj, out);
    else if (c.jd_n > 0)
        k_pair_direct_pole<QEM_PREC_FP64><<<grid_for(m), 256, dev::lds_bytes64(c), s>>>(
            c, nullptr, m, d_ri, d_rj, d_qi, d_qj, out);
    else
        k_pair_direct_pole<kPrecFP64Lib><<<grid_for(m), 256, 0, s>>>(
            c, nullptr, m, d_ri, d_rj, d_qi, d_qj, out);
    return hipGetLastError();
}

hipError_t launch_self_pole(const ShortConsts &c, int precision, int64_t n, const double *d_xyz,
                            const double *d_q, int zero_xy, double *d_f, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (precision == QEM_PREC_MIXED)
        k_self_pole<QEM_PREC_MIXED><<<nblk(n, 256), 256, 0, s>>>(c, n, d_xyz, d_q, zero_xy, d_f);
    else
        k_self_pole<kPrecFP64Lib><<<nblk(n, 256), 256, 0, s>>>(c, n, d_xyz, d_q, zero_xy, d_f);
    return hipGetLastError();
}

}  // namespace qem
