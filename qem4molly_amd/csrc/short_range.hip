// QEM short range on gfx950: per-pair Gauss-Legendre quadrature of the
// image-charge Sommerfeld integrals, the self-image term, and a
// deterministic per-atom gather.
//
// Reference: F_short_ij (src/Force_short.jl:25-78), integrands
// (src/Force_short.jl:4-22), Gamma_1/2 and dz_Gamma_1/2
// (src/greens_function.jl:68-116), element set-up (17-48), Gauss_Legendre
// (src/Gaussian_int.jl:9-28), F_short_i (src/Force_short.jl:82-107) and the
// accumulation of Molly.forces(::QEM_short) (src/QEM_short.jl:41-56).
//
// Work mapping: one lane per pair.  All pairs cost the same (8 node sets of
// N_t nodes), so a wave runs in lockstep with no divergence: quadrature
// nodes and weights are wave-uniform (scalar loads), the per-pair limits
// and image factors live in registers.  Per node the integrands of the
// rho-force (k Gamma J1) and z-force (dz Gamma J0) share k, the exponentials,
// the denominator and the Bessel pair, so each node set is one loop.
//
// Precision modes
//   FP64:  everything double; J0/J1 from the device math library.
//   MIXED: node arithmetic, exponentials (v_exp_f32), reciprocal
//          (v_rcp_f32) and J0/J1 in FP32; J0/J1 come from a table of cubic
//          polynomials of J0 fitted on 1/32-wide intervals (1/16 when the
//          argument range needs more than 4096 intervals; LDS resident;
//          J1 = -J0' from the same polynomial, p and p' by one Horner
//          pass).  Per-pair set-up, closed-form image terms and all sums
//          over images, pairs and atoms stay FP64.
#include "qem_device.h"

namespace qem {
namespace {

using dev::bessel01_tab;
using dev::jl_sign;
using dev::kLog2e;

// ---------------------------------------------------------------------------
// Node sums for one image l and one integral type.
//   point: sum_n w k [e^{-k a} e^{-2kLz} / D(k)] {J1, J0}(k rho)
//   gauss: sum_n w k [e^{-k a - k^2/(4 alpha)} / D(k)] {J1, J0}(k rho)
// with D(k) = g12 e^{-2kLz} - 1 and the nodes mapped to [0, K].
// The constant factors 0.5 b_l (g12 for point) and s_l are applied by the
// caller.
// FP64 with the node constants folded in, as in the mixed loop below: per
// node u = (1 + x_n, log2(w_n (1 + x_n))) (from long double), k = h u.x,
// w k = h^2 w_n u.x, and every exponent goes through e = -2 Lz k log2(e):
//   point: e^{-k (a + 2 Lz)} w k / h^2 = 2^{e (1 + a/2Lz) + u.y}
//   gauss: e^{-k a - k^2/4 alpha} w k / h^2 = 2^{e (a/2Lz + e cg/c2^2) + u.y}
// with the FP64 2^y and 1/D of qem_device.h.
// TAB: J0/J1 from the FP64 Bessel table staged in LDS (tab64), else from
// the device library (argument ranges beyond the table's capacity).
template <bool TAB>
__device__ __forceinline__ void nodes_fp64(double K, double a, double rho, bool gauss,
                                           const ShortConsts &c, const double *__restrict__ glx,
                                           const double *__restrict__ glw,
                                           const double *__restrict__ tab64, double &sr,
                                           double &sz)
{
    (void)glx;
    (void)glw;
    const double2 *__restrict__ nc = c.ncon64;
    const double h = 0.5 * K;
    const double hE = h * c.d_c2;                  // e = hE u.x
    const double rA = a * c.d_hinvLz;              // a / 2Lz
    const double rP = 1.0 + rA;
    const double rG = c.d_rG;
    // Bessel argument in table units (TAB; clamped once per set, u.x < 2
    // keeps the interval below jd_n - 1, as in the mixed loop) or the
    // argument itself
    const double hR = TAB ? fmin(h * rho * c.jd_inv_dx, 0.5 * (double)(c.jd_n - 1)) : h * rho;
    double ar = 0.0, az = 0.0;
    for (int n = 0; n < c.Nt; ++n) {
        const double2 u = nc[n];
        const double e = hE * u.x;
        const double E = dev::exp2_fp64(e);
        const double y = gauss ? fma(e, fma(e, rG, rA), u.y) : fma(e, rP, u.y);
        const double t = dev::exp2_fp64(y) * dev::rcp_fp64(fma(c.g12, E, -1.0));
        const double xb = hR * u.x;
        double J0, J1;
        if (TAB) {
            const double tt = __builtin_amdgcn_fract(xb);
            const double *__restrict__ cc = tab64 + 9 * (int)xb;
            double p = cc[7], dp = 0.0;
#pragma unroll
            for (int k = 6; k >= 0; --k) {
                dp = fma(dp, tt, p);
                p = fma(p, tt, cc[k]);
            }
            J0 = p;
            J1 = dp;   // -J1 dx, scaled below
        } else {
            J0 = j0(xb);
            J1 = j1(xb);
        }
        ar = fma(t, J1, ar);
        az = fma(t, J0, az);
    }
    const double h2 = h * h;
    sr = TAB ? -h2 * c.jd_inv_dx * ar : h2 * ar;
    sz = h2 * az;
}

// Mixed-precision node sums with the node constants folded in: for the
// rule's node n, u = (1 + x_n, (1 + x_n)^2, log2 w_n, log2(w_n (1 + x_n)));
// with h = K/2 the node is k = h u.x and w k = h^2 w_n u.x.  Every exponent is
// written through e = -2 Lz k log2(e) (E = 2^e):
//   point:  e^{-k (a + 2 Lz)} w k / h^2 = 2^{e (1 + a/2Lz) + u.w}
//   gauss:  e^{-k a - k^2/4 alpha} w k / h^2 = 2^{e (a/2Lz + e cg/c2^2) + u.w}
// so each VALU op reads at most one wave-uniform (scalar) operand.
// Returns sum w k t J1 and sum w k t J0 as above.
//
// Table bound: k < K and rho <= r_cutoff, and the table spans
// r_cutoff * max(K) plus two intervals, so x / dx stays below the last
// interval; hR is clamped once per set so that holds for any input.
//
// With ENERGY the same loop also forms the energy node sums of the set
// (the J0 transforms of the image factors, weight w instead of w k): the
// summand of sz divided by u.x (a wave-uniform node constant, 1/(1 + x_n))
// is w e^{..} / D J0 / h, so
//   ev = h sum (t / u.x) J0,   ed = h sum (t / u.x) J0 {1/D point, E/D gauss},
// ed being what d/dgamma of 1/D brings down; the force operations are the
// same, so the forces are the same bits with or without it.
// With SAVE the loop also forms the node sums of the set's force tangent
// (dev::TanSums: the u.x-, u.x/D- and u.x^2-weighted sums beside the
// force's own), into *ts, so that a later Hessian-vector product at the same
// positions combines them instead of running the quadrature again.
template <bool ENERGY, bool SAVE = false>
__device__ __forceinline__ void nodes_mixed(float K, float a, float rho, bool gauss,
                                            const ShortConsts &c, const float4 *__restrict__ ncon,
                                            const float4 *__restrict__ tab, float &sr, float &sz,
                                            float &ev, float &ed, dev::TanSums *ts = nullptr)
{
    const float h = 0.5f * K;
    const float c2 = c.f_c2;                       // -2 Lz log2(e)
    const float g12 = (float)c.g12;
    const float hE = h * c2;                       // e = hE u.x
    const float rA = a * c.f_hinvLz;               // a / 2Lz
    const float rP = 1.0f + rA;
    const float rG = c.f_rG;                       // cg / c2^2, cg = -log2(e) / (4 alpha)
    // Bessel argument in table units; u.x < 2 keeps m <= jt_n - 1
    const float hR = fminf(h * rho * c.jt_inv_dx, 0.5f * (float)(c.jt_n - 1));
    float ar = 0.f, az = 0.f, av = 0.f, ad = 0.f;
    float U0 = 0.f, U1 = 0.f, Q0 = 0.f, Q1 = 0.f, R0 = 0.f, R1 = 0.f;
    for (int n = 0; n < c.Nt; ++n) {
        const float4 u = ncon[n];
        const float e = hE * u.x;
        const float E = __builtin_amdgcn_exp2f(e);
        const float num = gauss ? __builtin_amdgcn_exp2f(fmaf(e, fmaf(e, rG, rA), u.w))
                                : __builtin_amdgcn_exp2f(fmaf(e, rP, u.w));
        const float rD = __builtin_amdgcn_rcpf(fmaf(g12, E, -1.0f));
        const float t = num * rD;
        const float xb = hR * u.x;
        const float tt = __builtin_amdgcn_fractf(xb);
        const float4 cc = tab[(int)xb];
        float p = fmaf(cc.w, tt, cc.z);
        float dp = cc.w;
        dp = fmaf(dp, tt, p); p = fmaf(p, tt, cc.y);
        dp = fmaf(dp, tt, p); p = fmaf(p, tt, cc.x);
        ar = fmaf(t, dp, ar);   // J1 = -dp / dx
        az = fmaf(t, p, az);
        if (ENERGY) {
            const float tp = (t * c.nconi[n]) * p;
            av += tp;
            ad = fmaf(tp, gauss ? E * rD : rD, ad);
        }
        if constexpr (SAVE) {
            const float tu = t * u.x, tq = tu * rD;   // u.x t, u.x t / D
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
    }
    if constexpr (SAVE) *ts = dev::TanSums{az, ar, U0, U1, Q0, Q1, R0, R1};
    const float h2 = h * h;
    sr = -h2 * c.jt_inv_dx * ar;
    sz = h2 * az;
    if (ENERGY) {
        ev = h * av;
        ed = h * ad;
    }
}

struct PairForce {
    double fx, fy, fz, fzj;
};

// The per-pair record of the force path: float3 in mixed precision (the
// pair values come from FP32 node sums; the per-atom sums of the gather
// stay FP64), double4 in FP64.  12 bytes per endpoint for the gather to read.
template <int PREC>
using PairRec = typename std::conditional<PREC == QEM_PREC_MIXED, float3, double4>::type;

// The pair's two endpoint records: the first atom adds (fx, fy, fz), the
// second (-fx, -fy, fz_j).
template <typename Rec>
__device__ __forceinline__ void store_recs(Rec *rec, int64_t npairs, int64_t p, int32_t jp,
                                           const PairForce &f)
{
    dev::store4(rec + p, f.fx, f.fy, f.fz, 0.0);
    dev::store4(rec + npairs + jp, -f.fx, -f.fy, f.fzj, 0.0);
}

// d b_l / d gamma_w (w = 0: gamma_1, 1: gamma_2) of image l
__device__ __forceinline__ double db_dgamma(const ShortConsts &c, int l, int w)
{
    if (l == 1) return w == 0 ? 1.0 : 0.0;
    if (l == 2) return w == 1 ? 1.0 : 0.0;
    if (l == 3) return w == 0 ? c.g2 : c.g1;
    return 0.0;
}

// F_short_ij for one pair (src/Force_short.jl:25-78).  Returns zeros when
// the minimum-image distance exceeds r_cutoff (the reference's per-pair
// 2-point neighbour list is empty then, line 36-41).
// With ENERGY (mixed precision) *pot also receives the pair potential
// without q_i q_j and its gamma derivatives, (G, dG/dgamma_1, dG/dgamma_2),
// from the same node loops: per image the closed form b/(2 sqrt(a^2 +
// rho^2)) minus the point-charge J0 transform of Gamma_2 plus the gaussian
// one of Gamma_1 e^{-k^2/4 alpha}; the derivatives take d b / d gamma and
// d(1/D)/d gamma = -g_other E / D^2 (src/greens_function.jl:68-88).
// With SAVE (mixed precision) the node sums of the pair's force tangent go
// to the saved arrays sav (dev::store_tan_sums, snp pairs, this pair at sp;
// zeros for a set that is not evaluated).
template <int PREC, bool ENERGY = false, bool SAVE = false>
__device__ __forceinline__ PairForce pair_force(const ShortConsts &c, const double *__restrict__ glx,
                                const double *__restrict__ glw, const float4 *__restrict__ ncon,
                                const float4 *__restrict__ tab,
                                double qi, double qj, double xi, double yi, double zi, double xj,
                                double yj, double zj, double4 *pot = nullptr,
                                float4 *__restrict__ sav = nullptr, int64_t snp = 0,
                                int64_t sp = 0)
{
    static_assert(!ENERGY || PREC == QEM_PREC_MIXED, "the fused energy is mixed precision");
    static_assert(!SAVE || PREC == QEM_PREC_MIXED, "the saved tangent sums are mixed precision");
    PairForce out{0.0, 0.0, 0.0, 0.0};
    if (ENERGY) *pot = make_double4(0.0, 0.0, 0.0, 0.0);
    const double dx = xi - xj, dy = yi - yj;
    double mx, my, d2;
    {
        // minimum image; dx * (1/Lx) can round differently from dx / Lx only
        // next to a half box, where both images lie beyond r_cutoff < L/2
#pragma clang fp contract(off)
        mx = dx - c.Lx * rint(dx * c.inv_Lx);
        my = dy - c.Ly * rint(dy * c.inv_Ly);
        d2 = mx * mx + my * my;
    }
    if (d2 > c.rc2) return out;
    // rho and 1/rho: in mixed precision from the FP64 reciprocal square root
    // and two Newton steps (no FP64 square root or division in the set-up)
    double rho, irho = 0.0;
    if (PREC == QEM_PREC_MIXED)
        dev::mixed_rho(d2, rho, irho);
    else
        rho = sqrt(d2);
    const double zn = fabs(zi - zj), zp = zi + zj;
    const double sg = jl_sign(zi - zj);
    const bool swap = (c.sem & QEM_SEM_SWAP_LIMITS) != 0;

    double Fsr = 0.0, Fsz = 0.0;
    double Fszj = 0.0;   // the z force on j of the as-written Newton semantics
    double Eu = 0.0, Ed0 = 0.0, Ed1 = 0.0;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const double a = l == 0 ? zn : (l == 1 ? zp : (l == 2 ? 2.0 * c.Lz - zp : 2.0 * c.Lz - zn));
        const double b = l == 0 ? 1.0 : (l == 1 ? c.g1 : (l == 2 ? c.g2 : c.g12));
        const double s = l == 0 ? -sg : (l == 1 ? -1.0 : (l == 2 ? 1.0 : sg));
        // greens_element_ij_init limits (src/greens_function.jl:26-27)
        double Kp, Kg, inv32, inv_r = 0.0;
        const double r2 = a * a + (PREC == QEM_PREC_MIXED ? d2 : rho * rho);
        if (PREC == QEM_PREC_MIXED) {
            // the limits only place FP32 nodes: FP32
            float kp, kg;
            dev::mixed_limits(c, a, swap, kp, kg);
            Kp = kp;
            Kg = kg;
            // r2^{-3/2} from the FP64 reciprocal square root and two Newton steps
            double y = __builtin_amdgcn_rsq(r2);
            y = y * fma(-0.5 * r2 * y, y, 1.5);
            y = y * fma(-0.5 * r2 * y, y, 1.5);
            inv32 = y * y * y;
            inv_r = y;
        } else {
            const double kf1 = -c.ln_acc / (2.0 * c.Lz + a);
            const double kf2 = sqrt(4.0 * c.alpha * c.alpha * a * a - 4.0 * c.alpha * c.ln_acc) -
                               2.0 * c.alpha * a;
            Kp = swap ? kf1 : kf2;  // point-charge integral limit
            Kg = swap ? kf2 : kf1;  // gaussian integral limit
            inv32 = 1.0 / (r2 * sqrt(r2));
        }
        // closed-form image sums (src/Force_short.jl:49,65)
        const double cr = 0.5 * b * rho * inv32;
        const double cz = 0.5 * b * a * s * inv32;
        double pr = 0.0, pz = 0.0, gr = 0.0, gz = 0.0;
        double Pv = 0.0, Pd = 0.0, Gv = 0.0, Gd = 0.0;
        const double db0 = ENERGY ? db_dgamma(c, l, 0) : 0.0;
        const double db1 = ENERGY ? db_dgamma(c, l, 1) : 0.0;
        if (b != 0.0 || db0 != 0.0 || db1 != 0.0) {
            if (PREC != QEM_PREC_MIXED) {
                constexpr bool TAB = PREC == QEM_PREC_FP64;
                const double *tab64 = reinterpret_cast<const double *>(tab);
                nodes_fp64<TAB>(Kp, a, rho, false, c, glx, glw, tab64, pr, pz);
                if (!c.single) nodes_fp64<TAB>(Kg, a, rho, true, c, glx, glw, tab64, gr, gz);
            } else {
                float fr, fz, fv = 0.f, fd = 0.f;
                dev::TanSums ts{};
                nodes_mixed<ENERGY, SAVE>((float)Kp, (float)a, (float)rho, false, c, ncon, tab,
                                          fr, fz, fv, fd, &ts);
                if constexpr (SAVE) dev::store_tan_sums(sav, snp, sp, l, 0, ts);
                pr = fr;
                pz = fz;
                Pv = fv;
                Pd = fd;
                if (!c.single) {
                    nodes_mixed<ENERGY, SAVE>((float)Kg, (float)a, (float)rho, true, c, ncon, tab,
                                              fr, fz, fv, fd, &ts);
                    gr = fr;
                    gz = fz;
                    Gv = fv;
                    Gd = fd;
                } else {
                    ts = dev::TanSums{};
                }
                if constexpr (SAVE) dev::store_tan_sums(sav, snp, sp, l, 1, ts);
            }
        } else if constexpr (SAVE) {
            dev::store_tan_sums(sav, snp, sp, l, 0, dev::TanSums{});
            dev::store_tan_sums(sav, snp, sp, l, 1, dev::TanSums{});
        }
        const double hb = 0.5 * b;
        // F_sr: -sum GL[k Gamma_2 J1] + closed + sum GL[k Gamma_1 e^{-k^2/4a} J1]
        Fsr += -(hb * c.g12) * pr + cr + hb * gr;
        // F_sz: +sum GL[dz Gamma_2 J0] - closed - sum GL[dz Gamma_1 e^{..} J0]
        const double part = (hb * c.g12 * s) * pz - cz - (hb * s) * gz;
        Fsz += part;
        // images 0 and 3 flip sign on the second atom, 1 and 2 do not
        Fszj += (l == 0 || l == 3) ? -part : part;
        if (ENERGY) {
            const double img = inv_r - c.g12 * Pv + Gv;
            Eu += hb * img;
            Ed0 += 0.5 * db0 * img + hb * c.g2 * (Pd - Gd);
            Ed1 += 0.5 * db1 * img + hb * c.g1 * (Pd - Gd);
        }
    }
    if (ENERGY) *pot = make_double4(c.inv2pieps * Eu, c.inv2pieps * Ed0, c.inv2pieps * Ed1, 0.0);
    const double pref = qi * qj * c.inv2pieps;
    if (rho != 0.0) {
        const bool mi = (c.sem & QEM_SEM_MINIMAGE_DIR) != 0;
        const double ux = mi ? mx : dx, uy = mi ? my : dy;
        if (PREC == QEM_PREC_MIXED) {
            out.fx = pref * (Fsr * ux * irho);
            out.fy = pref * (Fsr * uy * irho);
        } else {
            out.fx = pref * (Fsr * ux / rho);
            out.fy = pref * (Fsr * uy / rho);
        }
    }
    out.fz = pref * Fsz;
    if (c.sem & QEM_SEM_NEWTON_IMAGES)
        out.fzj = pref * Fszj;
    else
        out.fzj = -out.fz;
    return out;
}

using dev::self_force_z;

// FP64 blocks are 512 threads: the FP64 Bessel table (~47 KB at the 100k
// slab) is staged once per block, so two blocks of eight waves fill a CU to
// the VGPR limit (4 waves per SIMD) where three blocks of four waves would
// leave it at 3.
template <int PREC>
constexpr int pair_block() { return PREC == QEM_PREC_FP64 ? 512 : 256; }

// Waves per SIMD: the mixed kernel fits five in 96 VGPRs without spilling
// (left to itself the compiler takes 115, four waves), 1.3% faster; the
// FP64 kernels keep what the compiler gives them.
template <int PREC>
__global__ __launch_bounds__(pair_block<PREC>())
__attribute__((amdgpu_waves_per_eu(PREC == QEM_PREC_MIXED ? 5 : 1))) void k_pair_list(
    ShortConsts c, const double *__restrict__ glx, const double *__restrict__ glw,
    const float4 *__restrict__ ncon, const float4 *__restrict__ gtab, const int32_t *__restrict__ pi,
    const int32_t *__restrict__ pj, const int32_t *__restrict__ jpos, int64_t npairs,
    const double *__restrict__ xyz, const double *__restrict__ q,
    PairRec<PREC> *__restrict__ fpair)
{
    extern __shared__ float4 lds_tab[];
    const float4 *tab = dev::stage_tables<PREC>(c, gtab, lds_tab);
    dev::walk_pairs(pi, pj, jpos, npairs, xyz, q, [&](int64_t p, const dev::PairOps &o) {
        const PairForce f = pair_force<PREC>(c, glx, glw, ncon, tab, o.qi, o.qj, o.xi, o.yi,
                                             o.zi, o.xj, o.yj, o.zj);
        store_recs(fpair, npairs, p, o.jp, f);
    });
}

// The force pair kernel with the pair potentials (mixed precision).  The
// force records of k_pair_list carry in w what the pair adds to the
// endpoint's dE/dq: q_j G for the first atom, q_i G for the second (G the
// pair potential without the charges).  The pair's share of dE/dgamma_1,2,
// q_i q_j dG/dgamma, is summed per lane, then per block in a fixed order,
// into gpart[blockIdx.x].
// Four waves per SIMD: at 128 VGPRs the compiler spills 28 registers to
// scratch, and the kernel still runs 4% faster than at three waves (156).
// SAVE: the pairs' tangent node sums into sums as well (dev::store_tan_sums);
// 44 registers spill at four waves, and three waves (156 VGPRs, no spill)
// run the saving call 4% slower.
template <bool SAVE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_pair_list_fe(
    ShortConsts c, const double *__restrict__ glx, const double *__restrict__ glw,
    const float4 *__restrict__ ncon, const float4 *__restrict__ gtab, const int32_t *__restrict__ pi,
    const int32_t *__restrict__ pj, const int32_t *__restrict__ jpos, int64_t npairs,
    const double *__restrict__ xyz, const double *__restrict__ q, float4 *__restrict__ fpair,
    double2 *__restrict__ gpart, float4 *__restrict__ sums)
{
    extern __shared__ float4 lds_tab[];
    __shared__ double red[8];
    const float4 *tab = dev::stage_tables<QEM_PREC_MIXED>(c, gtab, lds_tab);
    double g1 = 0.0, g2 = 0.0;
    dev::walk_pairs(pi, pj, jpos, npairs, xyz, q, [&](int64_t p, const dev::PairOps &o) {
        double4 u;
        const PairForce f = pair_force<QEM_PREC_MIXED, true, SAVE>(
            c, glx, glw, ncon, tab, o.qi, o.qj, o.xi, o.yi, o.zi, o.xj, o.yj, o.zj, &u, sums,
            npairs, p);
        dev::store4(fpair + p, f.fx, f.fy, f.fz, o.qj * u.x);
        dev::store4(fpair + npairs + o.jp, -f.fx, -f.fy, f.fzj, o.qi * u.x);
        const double qq = o.qi * o.qj;
        g1 = fma(qq, u.y, g1);
        g2 = fma(qq, u.z, g2);
    });
    dev::block_sum2(g1, g2, red);
    if (threadIdx.x == 0) gpart[blockIdx.x] = make_double2(g1, g2);
}

template <int PREC>
__global__ __launch_bounds__(256) void k_pair_direct(
    ShortConsts c, const double *__restrict__ glx, const double *__restrict__ glw,
    const float4 *__restrict__ ncon, const float4 *__restrict__ gtab, int64_t m, const double *__restrict__ ri,
    const double *__restrict__ rj, const double *__restrict__ qi,
    const double *__restrict__ qj, double4 *__restrict__ fpair)
{
    extern __shared__ float4 lds_tab[];
    const float4 *tab = dev::stage_tables<PREC>(c, gtab, lds_tab);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m; p += stride) {
        PairForce f = pair_force<PREC>(c, glx, glw, ncon, tab, qi[p], qj[p], ri[3 * p],
                                       ri[3 * p + 1], ri[3 * p + 2], rj[3 * p], rj[3 * p + 1],
                                       rj[3 * p + 2]);
        fpair[p] = make_double4(f.fx, f.fy, f.fz, f.fzj);
    }
}

// Self-image z force per atom (src/QEM_short.jl:41-56, F_short_i) into
// f[3a + 2]; with zero_xy the x, y entries are cleared as well.
template <int PREC>
__global__ __launch_bounds__(256) void k_self_z(
    ShortConsts c, const double *__restrict__ glx, const double *__restrict__ glw,
    const float *__restrict__ glxf, const float *__restrict__ glwf, int64_t n,
    const double *__restrict__ xyz, const double *__restrict__ q, int zero_xy,
    double *__restrict__ f)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    f[3 * a + 2] = self_force_z<PREC>(c, glx, glw, glxf, glwf, q[a], xyz[3 * a + 2]);
    if (zero_xy) {
        f[3 * a] = 0.0;
        f[3 * a + 1] = 0.0;
    }
}

// Per atom: the force records of its pairs, first endpoint (+F_ij) and
// second (-F_ij in xy; in z what the semantics give the second atom)
// (src/QEM_short.jl:41-56), plus the self term.  Sixteen lanes per atom
// read the atom's two record runs in contiguous strides; each lane keeps a
// strided partial sum and the row of 16 lanes is summed in a fixed order.
// No atomics.  With SELF the kernel also evaluates the self term
// (precision SPREC), each of the 16 lanes a share of its nodes, which the
// same row sum adds up: the node loops run while the records stream in.
// Without it the self term is already in f[3a + 2].
constexpr int kGatherLanes = 16;   // lanes per atom in the force gather

template <typename Rec, int SPREC, bool SELF>
__global__ __launch_bounds__(256) void k_gather_pairs(
    int64_t n, int64_t npairs, const Rec *__restrict__ fpair,
    const int32_t *__restrict__ i_start, const int32_t *__restrict__ i_end,
    const int32_t *__restrict__ i_idx, const int32_t *__restrict__ j_start,
    const int32_t *__restrict__ j_end, ShortConsts c, const double *__restrict__ glx,
    const double *__restrict__ glw, const float *__restrict__ glxf,
    const float *__restrict__ glwf, const double *__restrict__ xyz,
    const double *__restrict__ q, double *__restrict__ f)
{
    const int64_t a = dev::gather_atom<kGatherLanes>();
    const int l = threadIdx.x & (kGatherLanes - 1);
    const bool live = a < n;
    double acc[3] = {0.0, 0.0, 0.0};
    if (live) {
        dev::add_endpoint_runs<Rec, 3, kGatherLanes>(acc, fpair, npairs, a, i_start, i_end, i_idx,
                                                     j_start, j_end, l);
        if (SELF)
            acc[2] += self_force_z<SPREC>(c, glx, glw, glxf, glwf, q[a], xyz[3 * a + 2], l,
                                          kGatherLanes);
    }
    // every lane takes part in the group sums
    const double fx = dev::lane_group_sum<kGatherLanes>(acc[0]);
    const double fy = dev::lane_group_sum<kGatherLanes>(acc[1]);
    const double fz = dev::lane_group_sum<kGatherLanes>(acc[2]);
    if (live && l == 0) {
        f[3 * a] = fx;
        f[3 * a + 1] = fy;
        f[3 * a + 2] = SELF ? fz : fz + f[3 * a + 2];
    }
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// grid for grid-stride pair kernels: enough blocks to fill 256 CUs several
// times over, without relaunching the LDS table stage per 256 pairs
using dev::pair_grid;

}  // namespace

hipError_t launch_pair_forces(const ShortConsts &c, int precision, const ShortWork &ws,
                              const int32_t *d_pi, const int32_t *d_pj, const int32_t *d_jpos,
                              int64_t npairs, const double *d_xyz, const double *d_q,
                              double *d_fpair, hipStream_t s)
{
    if (npairs <= 0) return hipSuccess;
    if (c.pole)
        return launch_pair_pole(c, precision, ws, d_pi, d_pj, d_jpos, npairs, d_xyz, d_q,
                                d_fpair, s);
    if (precision == QEM_PREC_MIXED) {
        const size_t lds = sizeof(float4) * (size_t)ws.jt_n;
        k_pair_list<QEM_PREC_MIXED><<<pair_grid(npairs), 256, lds, s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.ncon.as<float4>(),
            ws.jtab.as<float4>(), d_pi, d_pj, d_jpos, npairs, d_xyz, d_q,
            reinterpret_cast<float3 *>(d_fpair));
    } else {
        if (c.jd_n > 0)
            k_pair_list<QEM_PREC_FP64><<<pair_grid(npairs, 512), 512, dev::lds_bytes64(c), s>>>(
                c, ws.glx.as<double>(), ws.glw.as<double>(), ws.ncon.as<float4>(), nullptr, d_pi,
                d_pj, d_jpos, npairs, d_xyz, d_q,
                reinterpret_cast<double4 *>(d_fpair));
        else
            k_pair_list<kPrecFP64Lib><<<pair_grid(npairs), 256, 0, s>>>(
                c, ws.glx.as<double>(), ws.glw.as<double>(), ws.ncon.as<float4>(), nullptr, d_pi,
                d_pj, d_jpos, npairs, d_xyz, d_q,
                reinterpret_cast<double4 *>(d_fpair));
    }
    return hipGetLastError();
}

int pair_fe_blocks(int64_t npairs) { return npairs > 0 ? (int)pair_grid(npairs) : 0; }

hipError_t launch_pair_forces_energy(const ShortConsts &c, const ShortWork &ws,
                                     const int32_t *d_pi, const int32_t *d_pj,
                                     const int32_t *d_jpos, int64_t npairs,
                                     const double *d_xyz, const double *d_q, double *d_fpair,
                                     double *d_gpart, hipStream_t s, float *d_sums)
{
    if (npairs <= 0) return hipSuccess;
    if (c.pole || !c.nconi) return hipErrorInvalidValue;
    const size_t lds = sizeof(float4) * (size_t)ws.jt_n;
    float4 *sums = reinterpret_cast<float4 *>(d_sums);
    if (sums)
        k_pair_list_fe<true><<<pair_grid(npairs), 256, lds, s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.ncon.as<float4>(),
            ws.jtab.as<float4>(), d_pi, d_pj, d_jpos, npairs, d_xyz, d_q,
            reinterpret_cast<float4 *>(d_fpair), reinterpret_cast<double2 *>(d_gpart), sums);
    else
        k_pair_list_fe<false><<<pair_grid(npairs), 256, lds, s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.ncon.as<float4>(),
            ws.jtab.as<float4>(), d_pi, d_pj, d_jpos, npairs, d_xyz, d_q,
            reinterpret_cast<float4 *>(d_fpair), reinterpret_cast<double2 *>(d_gpart), nullptr);
    return hipGetLastError();
}

hipError_t launch_pair_direct(const ShortConsts &c, int precision, const ShortWork &ws,
                              int64_t m, const double *d_ri, const double *d_rj,
                              const double *d_qi, const double *d_qj, double *d_fpair,
                              hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    if (c.pole)
        return launch_pair_direct_pole(c, precision, ws, m, d_ri, d_rj, d_qi, d_qj, d_fpair, s);
    if (precision == QEM_PREC_MIXED) {
        const size_t lds = sizeof(float4) * (size_t)ws.jt_n;
        k_pair_direct<QEM_PREC_MIXED><<<pair_grid(m), 256, lds, s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.ncon.as<float4>(),
            ws.jtab.as<float4>(), m, d_ri, d_rj, d_qi, d_qj,
            reinterpret_cast<double4 *>(d_fpair));
    } else {
        if (c.jd_n > 0)
            k_pair_direct<QEM_PREC_FP64><<<pair_grid(m), 256, dev::lds_bytes64(c), s>>>(
                c, ws.glx.as<double>(), ws.glw.as<double>(), ws.ncon.as<float4>(), nullptr, m, d_ri,
                d_rj, d_qi, d_qj,
                reinterpret_cast<double4 *>(d_fpair));
        else
            k_pair_direct<kPrecFP64Lib><<<pair_grid(m), 256, 0, s>>>(
                c, ws.glx.as<double>(), ws.glw.as<double>(), ws.ncon.as<float4>(), nullptr, m, d_ri,
                d_rj, d_qi, d_qj,
                reinterpret_cast<double4 *>(d_fpair));
    }
    return hipGetLastError();
}

static hipError_t launch_self_z(const ShortConsts &c, int precision, const ShortWork &ws,
                                int64_t n, const double *d_xyz, const double *d_q, int zero_xy,
                                double *d_f, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (c.pole) return launch_self_pole(c, precision, n, d_xyz, d_q, zero_xy, d_f, s);
    if (precision == QEM_PREC_MIXED)
        k_self_z<QEM_PREC_MIXED><<<nblk(n, 256), 256, 0, s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
            ws.glwf.as<float>(), n, d_xyz, d_q, zero_xy, d_f);
    else
        k_self_z<QEM_PREC_FP64><<<nblk(n, 256), 256, 0, s>>>(
            c, ws.glx.as<double>(), ws.glw.as<double>(), ws.glxf.as<float>(),
            ws.glwf.as<float>(), n, d_xyz, d_q, zero_xy, d_f);
    return hipGetLastError();
}

hipError_t launch_gather_self(const ShortConsts &c, int precision, const ShortWork &ws,
                              const NeighborList &nl, int64_t n, const double *d_xyz,
                              const double *d_q, const double *d_fpair, double *d_f,
                              hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int32_t *iidx = nl.has_i_idx ? nl.i_idx.as<int32_t>() : nullptr;
    const unsigned nb = (nblk(kGatherLanes * n, 256) + 7) / 8 * 8;
    const int32_t *is = nl.i_start.as<int32_t>(), *ie = nl.i_end.as<int32_t>();
    const int32_t *js = nl.j_start.as<int32_t>(), *je = nl.j_end.as<int32_t>();
    const int64_t np = nl.npairs;
    const double *glx = ws.glx.as<double>(), *glw = ws.glw.as<double>();
    const float *glxf = ws.glxf.as<float>(), *glwf = ws.glwf.as<float>();
    if (c.pole) {
        // the pole-splitting self term has its own kernel
        hipError_t e = launch_self_z(c, precision, ws, n, d_xyz, d_q, 0, d_f, s);
        if (e != hipSuccess) return e;
        k_gather_pairs<double4, QEM_PREC_FP64, false><<<nb, 256, 0, s>>>(
            n, np, reinterpret_cast<const double4 *>(d_fpair), is, ie, iidx, js, je, c, glx,
            glw, glxf, glwf, d_xyz, d_q, d_f);
    } else if (pair_records_f32(c, precision)) {
        k_gather_pairs<float3, QEM_PREC_MIXED, true><<<nb, 256, 0, s>>>(
            n, np, reinterpret_cast<const float3 *>(d_fpair), is, ie, iidx, js, je, c, glx,
            glw, glxf, glwf, d_xyz, d_q, d_f);
    } else {
        k_gather_pairs<double4, QEM_PREC_FP64, true><<<nb, 256, 0, s>>>(
            n, np, reinterpret_cast<const double4 *>(d_fpair), is, ie, iidx, js, je, c, glx,
            glw, glxf, glwf, d_xyz, d_q, d_f);
    }
    return hipGetLastError();
}

hipError_t launch_self_only(const ShortConsts &c, int precision, const ShortWork &ws,
                            int64_t n, const double *d_xyz, const double *d_q, double *d_f,
                            hipStream_t s)
{
    return launch_self_z(c, precision, ws, n, d_xyz, d_q, 1, d_f, s);
}

}  // namespace qem
