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

// Mixed precision: rho and 1/rho from the FP64 reciprocal square root and
// two Newton steps (no FP64 square root or division)
__device__ __forceinline__ void mixed_rho(double d2, double &rho, double &irho)
{
    rho = 0.0;
    irho = 0.0;
    if (d2 > 0.0) {
        double y = __builtin_amdgcn_rsq(d2);
        y = y * fma(-0.5 * d2 * y, y, 1.5);
        y = y * fma(-0.5 * d2 * y, y, 1.5);
        irho = y;
        rho = d2 * y;
    }
}

// Mixed precision: the point-charge and gaussian limits Kp, Kg of image
// distance a in FP32 (they only place FP32 nodes), with k_f2 in the form
// Y / (sqrt(X^2 + Y) + X) = sqrt(X^2 + Y) - X, X = 2 alpha a,
// Y = -4 alpha ln(acc), which does not cancel; swap: the limits exchanged
// (greens_element_ij_init, src/greens_function.jl:26-27)
__device__ __forceinline__ void mixed_limits(const ShortConsts &c, double a, bool swap, float &Kp,
                                             float &Kg)
{
    const float af = (float)a, X = c.f_2alpha * af;
    const float kf1 = c.f_nlnacc * __builtin_amdgcn_rcpf(c.f_2Lz + af);
    const float kf2 =
        c.f_m4alnacc * __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(fmaf(X, X, c.f_m4alnacc)) + X);
    Kp = swap ? kf1 : kf2;
    Kg = swap ? kf2 : kf1;
}

// mixed_limits with the limits' derivatives in a (FP32, the tangent of the
// saved node sums only uses them in FP32): d k_f1/da = -k_f1/(2Lz + a) and
// d k_f2/da = -2 alpha k_f2 / sqrt(X^2 + Y), the non-cancelling form of
// 2 alpha (X / sqrt(X^2 + Y) - 1).  Kp, Kg are those of mixed_limits.
__device__ __forceinline__ void mixed_limits_d(const ShortConsts &c, double a, bool swap,
                                               float &Kp, float &Kg, float &dKp, float &dKg)
{
    mixed_limits(c, a, swap, Kp, Kg);
    const float af = (float)a, X = c.f_2alpha * af;
    const float kf1 = swap ? Kp : Kg, kf2 = swap ? Kg : Kp;
    const float dkf1 = -kf1 * __builtin_amdgcn_rcpf(c.f_2Lz + af);
    const float dkf2 = -c.f_2alpha * kf2 *
                       __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(fmaf(X, X, c.f_m4alnacc)));
    dKp = swap ? dkf1 : dkf2;
    dKg = swap ? dkf2 : dkf1;
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

// ---- endpoint records --------------------------------------------------------
// A pair kernel over a list writes, for pair p, the first atom's record at
// rec[p] and the second atom's at rec[npairs + jpos[p]] (jpos: the pair's
// place in the second-endpoint grouping).  Each record holds what its atom
// adds up, so a gather reads every atom's records as two contiguous runs
// and never follows a pair index.
__device__ __forceinline__ void store4(float4 *r, double x, double y, double z, double w)
{
    *r = make_float4((float)x, (float)y, (float)z, (float)w);
}
__device__ __forceinline__ void store4(double4 *r, double x, double y, double z, double w)
{
    *r = make_double4(x, y, z, w);
}
// three-component records (w is not kept)
__device__ __forceinline__ void store4(float3 *r, double x, double y, double z, double)
{
    *r = make_float3((float)x, (float)y, (float)z);
}

template <typename Rec, int NC>
__device__ __forceinline__ void add_rec(double *acc, const Rec &r)
{
    acc[0] += (double)r.x;
    acc[1] += (double)r.y;
    acc[2] += (double)r.z;
    if constexpr (NC > 3) acc[3] += (double)r.w;
}

// Lane l of an atom's L lanes adds records t = s + l, s + l + L, ... of
// run [s, e) in that order, four loads in flight.
template <typename Rec, int NC, int L = 16>
__device__ __forceinline__ void add_run(double *acc, const Rec *__restrict__ rec, int32_t s,
                                        int32_t e, int l)
{
    int32_t t = s + l;
    for (; t + 3 * L < e; t += 4 * L) {
        const Rec r0 = rec[t], r1 = rec[t + L], r2 = rec[t + 2 * L], r3 = rec[t + 3 * L];
        add_rec<Rec, NC>(acc, r0);
        add_rec<Rec, NC>(acc, r1);
        add_rec<Rec, NC>(acc, r2);
        add_rec<Rec, NC>(acc, r3);
    }
    for (; t < e; t += L) add_rec<Rec, NC>(acc, rec[t]);
}

// This lane's share of atom a's records: its first-endpoint run (through
// i_idx for a list that is not grouped by first atom), then its
// second-endpoint run in the second half.
template <typename Rec, int NC, int L = 16>
__device__ __forceinline__ void add_endpoint_runs(double *acc, const Rec *__restrict__ rec,
                                                  int64_t npairs, int64_t a,
                                                  const int32_t *__restrict__ i_start,
                                                  const int32_t *__restrict__ i_end,
                                                  const int32_t *__restrict__ i_idx,
                                                  const int32_t *__restrict__ j_start,
                                                  const int32_t *__restrict__ j_end, int l)
{
    if (i_idx) {
        const int32_t e = i_end[a];
        for (int32_t t = i_start[a] + l; t < e; t += L) add_rec<Rec, NC>(acc, rec[i_idx[t]]);
    } else {
        add_run<Rec, NC, L>(acc, rec, i_start[a], i_end[a], l);
    }
    add_run<Rec, NC, L>(acc, rec + npairs, j_start[a], j_end[a], l);
}

// Sum over each group of L (8 or 16) lanes in a fixed order (pairs, quads,
// half rows[, rows]); every lane of the group receives its group's sum.
template <int L>
__device__ __forceinline__ double lane_group_sum(double v)
{
    static_assert(L == 8 || L == 16, "lanes per atom");
    v += dpp_perm_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_perm_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_perm_f64<0x141>(v);  // row_half_mirror
    if (L == 16) v += dpp_perm_f64<0x140>(v);  // row_mirror
    return v;
}

// XCD-aware atom order of the gathers: the hardware deals blocks to the 8
// XCDs round robin, so block b runs on XCD b % 8; each XCD takes a
// contiguous range of atoms, whose records then share that XCD's L2
// (gridDim.x a multiple of 8).  L lanes per atom.
template <int L = 16>
__device__ __forceinline__ int64_t gather_atom()
{
    const int64_t per_xcd = gridDim.x / 8;
    const int64_t blk = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    return (blk * blockDim.x + threadIdx.x) / L;
}

// One pair's endpoints, positions and charges.
struct PairOps {
    int32_t i, j;
    int32_t jp;   // the pair's place in the second-endpoint grouping
    double qi, qj, xi, yi, zi, xj, yj, zj;
};

__device__ __forceinline__ void load_ends(PairOps &o, const int32_t *__restrict__ pi,
                                          const int32_t *__restrict__ pj,
                                          const int32_t *__restrict__ jpos, int64_t p)
{
    o.i = pi[p];
    o.j = pj[p];
    o.jp = jpos[p];
}

__device__ __forceinline__ void load_atoms(PairOps &o, const double *__restrict__ xyz,
                                           const double *__restrict__ q)
{
    const int64_t i = o.i, j = o.j;
    o.qi = q[i];
    o.qj = q[j];
    o.xi = xyz[3 * i];
    o.yi = xyz[3 * i + 1];
    o.zi = xyz[3 * i + 2];
    o.xj = xyz[3 * j];
    o.yj = xyz[3 * j + 1];
    o.zj = xyz[3 * j + 2];
}

// Walk over the pair list, one pair per lane at a time, in an XCD-aware
// order: the hardware deals blocks to the 8 XCDs round robin (block b runs
// on XCD b % 8), and XCD x takes the x-th eighth of the list, which its
// blocks stride through in order.  The pairs of nearby first atoms then run
// on one XCD, and so do most writes into a second atom's record run: each
// L2 gathers whole lines of second-endpoint records instead of eight L2s
// holding pieces of the same lines.
// body(p, ops) evaluates pair p.
template <class Body>
__device__ __forceinline__ void walk_pairs(const int32_t *__restrict__ pi,
                                           const int32_t *__restrict__ pj,
                                           const int32_t *__restrict__ jpos, int64_t npairs,
                                           const double *__restrict__ xyz,
                                           const double *__restrict__ q, Body body)
{
    const int64_t per_xcd = (int64_t)(gridDim.x / 8) * blockDim.x;
    int64_t p, e, stride;
    if (per_xcd == 0) {
        // fewer than 8 blocks: a plain grid stride
        p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        e = npairs;
        stride = (int64_t)gridDim.x * blockDim.x;
    } else {
        const int x = blockIdx.x % 8;
        const int64_t share = (npairs + 7) / 8;
        const int64_t b = x * share;
        e = b + share < npairs ? b + share : npairs;
        // blocks past the last multiple of 8 join the last XCD's range
        const int64_t local = blockIdx.x < (gridDim.x / 8) * 8 ? blockIdx.x / 8 : -1;
        if (local < 0) return;
        p = b + local * blockDim.x + threadIdx.x;
        stride = per_xcd;
    }
    for (; p < e; p += stride) {
        PairOps o;
        load_ends(o, pi, pj, jpos, p);
        load_atoms(o, xyz, q);
        body(p, o);
    }
}


// Grid of a pair-list kernel (blocks of `block` lanes, about two pairs per
// lane at the largest lists): small blocks let the last ones end together
// (a block of ~15 pairs per lane ran ~0.75 ms, and the last round of them
// left CUs idle); a multiple of 8 from 8 blocks on, since walk_pairs gives
// each XCD an eighth of them.
inline unsigned pair_grid(int64_t m, int block = 256)
{
    const int64_t want = std::min<int64_t>((m + block - 1) / block, (int64_t)256 * 128 * 256 / block);
    return (unsigned)(want >= 8 ? want / 8 * 8 : std::max<int64_t>(1, want));
}

// Self-image z force F_short_i (src/Force_short.jl:82-107).  Images 1 and 4
// have sign_a = 0 and contribute exactly zero; they are skipped.  Lane
// `lane` of `nl` lanes takes the nodes n = lane (mod nl) of every node sum,
// and lane 0 the closed-form image sum: the nl shares add up to the force
// (nl = 1 gives it whole).
template <int PREC>
__device__ __forceinline__ double self_force_z(const ShortConsts &c, const double *__restrict__ glx,
                               const double *__restrict__ glw, const float *__restrict__ glxf,
                               const float *__restrict__ glwf, double qi, double zi,
                               int lane = 0, int nl = 1)
{
    const bool swap = (c.sem & QEM_SEM_SWAP_LIMITS) != 0;
    double P1 = 0.0, P2 = 0.0, G = 0.0;
#pragma unroll
    for (int l = 1; l <= 2; ++l) {
        const double a = l == 1 ? 2.0 * zi : 2.0 * c.Lz - 2.0 * zi;
        const double b = l == 1 ? c.g1 : c.g2;
        const double s = l == 1 ? -1.0 : 1.0;
        const double kf1 = -c.ln_acc / (2.0 * c.Lz + a);
        const double kf2 = sqrt(4.0 * c.alpha * c.alpha * a * a - 4.0 * c.alpha * c.ln_acc) -
                           2.0 * c.alpha * a;
        const double Kp = swap ? kf1 : kf2;
        const double Kg = swap ? kf2 : kf1;
        double pz = 0.0, gz = 0.0;
        if (b != 0.0) {
            // J0(0) = 1: the node sums reduce to sum w k num / D
            for (int pass = 0; pass < (c.single ? 1 : 2); ++pass) {
                const bool gauss = pass == 1;
                const double K = gauss ? Kg : Kp;
                double acc = 0.0;
                if (PREC != QEM_PREC_MIXED) {
                    // the folded FP64 node terms of nodes_fp64 with J0(0) = 1
                    const double h = 0.5 * K;
                    const double hE = h * c.d_c2, rA = a * c.d_hinvLz, rP = 1.0 + rA;
                    double ad = 0.0;
                    for (int n = lane; n < c.Nt; n += nl) {
                        const double2 u = c.ncon64[n];
                        const double e = hE * u.x;
                        const double E = exp2_fp64(e);
                        const double y = gauss ? fma(e, fma(e, c.d_rG, rA), u.y) : fma(e, rP, u.y);
                        ad = fma(exp2_fp64(y), rcp_fp64(fma(c.g12, E, -1.0)), ad);
                    }
                    acc = h * h * ad;
                } else {
                    const float h = 0.5f * (float)K;
                    const float c2 = -2.0f * (float)c.Lz * kLog2e;
                    const float ca = -(float)a * kLog2e;
                    const float cg = -kLog2e * (float)(0.25 / c.alpha);
                    const float g12 = (float)c.g12;
                    float af = 0.f;
                    for (int n = lane; n < c.Nt; n += nl) {
                        const float k = fmaf(h, glxf[n], h);
                        const float w = h * glwf[n];
                        const float E = __builtin_amdgcn_exp2f(k * c2);
                        const float rD = __builtin_amdgcn_rcpf(fmaf(g12, E, -1.0f));
                        const float num = gauss ? __builtin_amdgcn_exp2f(k * fmaf(k, cg, ca))
                                                : __builtin_amdgcn_exp2f(k * ca) * E;
                        af = fmaf(w * k, num * rD, af);
                    }
                    acc = af;
                }
                if (gauss) gz = acc; else pz = acc;
            }
        }
        const double hb = 0.5 * b;
        P1 += (hb * c.g12 * s) * pz;
        if (lane == 0) P2 += b * a * s / (a * a * a);
        G += (hb * s) * gz;
    }
    P2 *= 0.5;
    const double Fsz = -P1 + P2 + G;
    return -(qi * qi) * c.inv2pieps * Fsz;
}


// Sums of a and b over the block in a fixed order (rows of 16 lanes, the
// four rows of a wave, the waves in order); the results land in thread 0.
// red holds two doubles per wave; every thread of the block must call.
__device__ __forceinline__ void block_sum2(double &a, double &b, double *red)
{
    a = row_sum16(a);
    b = row_sum16(b);
    auto lane_d = [](double v, int lane) {
        const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
        return __hiloint2double(hi, lo);
    };
    const double wa = ((lane_d(a, 0) + lane_d(a, 16)) + lane_d(a, 32)) + lane_d(a, 48);
    const double wb = ((lane_d(b, 0) + lane_d(b, 16)) + lane_d(b, 32)) + lane_d(b, 48);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[2 * w] = wa;
        red[2 * w + 1] = wb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = red[0];
        b = red[1];
        for (int k = 1; k < nw; ++k) {
            a += red[2 * k];
            b += red[2 * k + 1];
        }
    }
}

// ---- force tangent of one node set from its node sums ------------------------
// Value and first-order tangent of sum_n w k t(k) {J1, J0}(k rho) (mixed
// precision).
struct Dual4 {
    double sr, sz, dsr, dsz;
};

// The direction-independent node sums of one node set (mixed precision)
// that its force tangent combines: with u = (1 + x_n, ...) the node
// constants, t the summand of the force loop and p = J0, dp = -J1 dx from
// the table, S = sum t (p, dp), U = sum u.x t (p, dp), Q = sum u.x t / D
// (p, dp), R = sum u.x^2 t (p, dp) (R only for the gaussian set).
struct TanSums {
    float S0, S1, U0, U1, Q0, Q1, R0, R1;
};

// The node set's value and tangent from its sums.  Along the tangent the
// nodes move with K (dk = r k, r = dK/K), the image distance with da and
// rho with drho:
//   dt = t k (A + B E/D [+ C k]),   A = -r (a [+ 2 Lz]) - da,
//   B = 2 Lz g12 r,  C = -r / (2 alpha) (gaussian set),
//   dJ0 = -J1 x s,  dJ1 = (x J0 - J1) s,  s = r + drho/rho,  x = k rho,
// and d(h^2 S) = h^2 (2 r S + dS), h = K/2.  Since g12 E = D + 1,
// B E/D = 2 Lz r (1 + 1/D) (g12 = 0 included): the E/D-weighted sums are
// the 1/D-weighted Q and the coefficient of U becomes A + 2 Lz r.
__device__ __forceinline__ Dual4 tangent_of_sums(const TanSums &t, float K, float dK, float a,
                                                 float da, float rho, float drho, bool gauss,
                                                 bool self, const ShortConsts &c)
{
    const float h = 0.5f * K;
    const float r = K != 0.0f ? dK / K : 0.0f;
    const float Lz2 = 2.0f * (float)c.Lz;
    const float hR = self ? 0.0f : fminf(h * rho * c.jt_inv_dx, 0.5f * (float)(c.jt_n - 1));
    const float A = gauss ? fmaf(-r, a - Lz2, -da) : -r * a - da;
    const float B = Lz2 * r;
    const float C = gauss ? -r * (float)(0.5 / c.alpha) : 0.0f;
    const float sg = (!self && rho > 0.0f) ? r + drho / rho : 0.0f;
    const float sg2 = sg * c.jt_dx * c.jt_dx;
    const float hA = h * A, hB = h * B, h2C = h * h * C;
    // tangents of the sums over nodes, before the h^2 factor
    const float dS0 = hA * t.U0 + hB * t.Q0 + h2C * t.R0 + sg * hR * t.U1;
    const float dS1 = hA * t.U1 + hB * t.Q1 + h2C * t.R1 - sg * t.S1 - sg2 * hR * t.U0;
    const float h2 = h * h;
    Dual4 o;
    o.sr = -h2 * c.jt_inv_dx * t.S1;
    o.sz = h2 * t.S0;
    o.dsr = -h2 * c.jt_inv_dx * fmaf(2.0f * r, t.S1, dS1);
    o.dsz = h2 * fmaf(2.0f * r, t.S0, dS0);
    return o;
}

// Saved node sums, 224 B per pair: per image l three float4 arrays,
// 3 l + 0 = (S0, S1, U0, U1) of the point set, 3 l + 1 = the same of the
// gaussian set, 3 l + 2 = (Q0, Q1, R0, R1) of the gaussian set, then after
// the twelve float4 arrays four float2 arrays, l = (Q0, Q1) of the point set
// (whose R0, R1 are zero: its tangent has no u.x^2 term); element p of each
// array at [array npairs + p]; tan_sums_bytes(npairs) in all.
__device__ __forceinline__ void store_tan_sums(float4 *__restrict__ sv, int64_t npairs,
                                               int64_t p, int l, int set, const TanSums &t)
{
    sv[(3 * l + set) * npairs + p] = make_float4(t.S0, t.S1, t.U0, t.U1);
    if (set)
        sv[(3 * l + 2) * npairs + p] = make_float4(t.Q0, t.Q1, t.R0, t.R1);
    else
        reinterpret_cast<float2 *>(sv + 12 * npairs)[l * npairs + p] = make_float2(t.Q0, t.Q1);
}
__device__ __forceinline__ TanSums load_tan_sums(const float4 *__restrict__ sv, int64_t npairs,
                                                 int64_t p, int l, int set)
{
    const float4 x = sv[(3 * l + set) * npairs + p];
    if (set) {
        const float4 y = sv[(3 * l + 2) * npairs + p];
        return TanSums{x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w};
    }
    const float2 y = reinterpret_cast<const float2 *>(sv + 12 * npairs)[l * npairs + p];
    return TanSums{x.x, x.y, x.z, x.w, y.x, y.y, 0.0f, 0.0f};
}

}  // namespace dev
}  // namespace qem
