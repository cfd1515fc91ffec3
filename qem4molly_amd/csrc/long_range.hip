// QEM long range on gfx950: random-batch k-space sum with z-sorted prefix
// sums and the k = 0 term; energy and tangent passes for the adjoint.
//
// Reference: Molly.forces(::QEM_long) (src/QEM_long.jl:32-44: sortperm of z,
// then F_long), F_long (src/Force_long.jl:201-242, random-batch branch),
// F_l_total (186-199), Fl_non_sort (5-70), Fl_sort (72-127), Flz_sort_k0
// (130-161), Flz_self (164-184).
//
// Passes (atoms in ascending-z order):
//   sort   stable radix sort of z with the atom index as payload
//          (radix_sort.hip), whose last pass writes (x, y, z, q) in z order.
//   krec   per k row: derived constants (scaled k, weights, the non-sort and
//          self factors with their exponentials and divisions).
//   A      per (chunk, group of 64 k rows; one row per lane): the four chunk
//          sums q cos e^{+kz}, q sin e^{+kz}, q cos e^{-kz}, q sin e^{-kz},
//          accumulated along the chunk in ascending z.
//   S      per k: exclusive prefix over chunks of the e^{+kz} pair, exclusive
//          suffix of the e^{-kz} pair (accumulated from the top, so no
//          subtraction from a total), and the four structure factors.
//   B      per (chunk, group): walk the chunk again; running sums plus chunk
//          offsets give A, B, C, D of Fl_sort; the structure factors minus
//          the atom's own term give Fl_non_sort; Flz_self; each atom's force
//          is summed over the group's rows (fixed DPP order).
//   R      per atom: sum over groups, scatter back to input order.
// The k = 0 term (Flz_sort_k0) is not a k lane: it is a prefix sum of the
// charges in z order (tile sums, a scan, a per-atom pass; k_k0_*), run
// right after the sort and added in R.  Nothing of size N x P is stored.
#include "qem_device.h"

namespace qem {
namespace {

constexpr int CH = 256;  // threads of the chunk-scan block (k_long_S)
constexpr double kInv2Pi = 0.15915494309189535;
constexpr double kLog2eD = 1.4426950408889634;

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__global__ void k_fill_f64(double *__restrict__ p, int64_t n, double v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// exclusive scan of (x, y) over the CH threads of the block in thread
// order; *tot receives the block total (the same value in every thread)
__device__ __forceinline__ double2 block_excl_scan(double2 v, double2 *tot,
                                                   double2 *__restrict__ sh)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double2 inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double x = __shfl_up(inc.x, d), y = __shfl_up(inc.y, d);
        if (lane >= d) {
            inc.x += x;
            inc.y += y;
        }
    }
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    double2 base = make_double2(0.0, 0.0), all = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < CH / 64; ++i) {
        if (i < w) {
            base.x += sh[i].x;
            base.y += sh[i].y;
        }
        all.x += sh[i].x;
        all.y += sh[i].y;
    }
    __syncthreads();
    double2 ex;
    ex.x = __shfl_up(inc.x, 1);
    ex.y = __shfl_up(inc.y, 1);
    if (lane == 0) ex = make_double2(0.0, 0.0);
    *tot = all;
    return make_double2(base.x + ex.x, base.y + ex.y);
}

// Pass S: per k row, chunk offsets and totals.  One block per row; the
// chunks go through block scans in tiles of CH (a fixed summation order).
__global__ __launch_bounds__(CH) void k_long_S(const double4 *__restrict__ partA, int64_t Ptot,
                                               int nchunk, double4 *__restrict__ offA,
                                               double4 *__restrict__ totA)
{
    __shared__ double2 sh[CH / 64];
    const int64_t p = blockIdx.x;
    const double4 *pa = partA + p * nchunk;
    double4 *oa = offA + p * nchunk;
    // exclusive prefix of the e^{+kz} pair in ascending chunk order, tiles
    // of CH chunks
    double2 run = make_double2(0.0, 0.0);
    for (int b = 0; b < nchunk; b += CH) {
        const int m = b + (int)threadIdx.x;
        double2 v = make_double2(0.0, 0.0);
        if (m < nchunk) v = make_double2(pa[m].x, pa[m].y);
        double2 tot;
        const double2 ex = block_excl_scan(v, &tot, sh);
        if (m < nchunk) {
            oa[m].x = run.x + ex.x;
            oa[m].y = run.y + ex.y;
        }
        run.x += tot.x;
        run.y += tot.y;
    }
    // exclusive suffix of the e^{-kz} pair: the same from the top
    double2 runs = make_double2(0.0, 0.0);
    for (int b = 0; b < nchunk; b += CH) {
        const int m = nchunk - 1 - (b + (int)threadIdx.x);
        double2 v = make_double2(0.0, 0.0);
        if (m >= 0) v = make_double2(pa[m].z, pa[m].w);
        double2 tot;
        const double2 ex = block_excl_scan(v, &tot, sh);
        if (m >= 0) {
            oa[m].z = runs.x + ex.x;
            oa[m].w = runs.y + ex.y;
        }
        runs.x += tot.x;
        runs.y += tot.y;
    }
    if (threadIdx.x == 0) totA[p] = make_double4(run.x, run.y, runs.x, runs.y);
}

// ---------------------------------------------------------------------------
// Force passes, k-lane layout.  A wave owns one chunk of chk consecutive
// z-ordered atoms and 64 k rows, one row per lane.  It walks the chunk's
// atoms in ascending z with the atom read once per wave (uniform loads);
// every lane keeps its row's constants, its running prefix sums and its
// share of the force in registers.  The per-atom force is summed over the
// wave's 64 rows with a fixed DPP pattern; the row groups are summed in
// pass R.  Sums over atoms run in the reference's sequential order.

// derived constants of one k row (k_make_krec)
struct KRec {
    double kx, ky, k, pad;    // k vector (|k| > 0); pad keeps 16 doubles
    double kxt, kyt, kl, w;   // k_x / 2pi, k_y / 2pi, |k| log2 e; z weight
    double wx, wy;            // w k_x / |k|, w k_y / |k| (xy weights)
    double N1, N2, N3;        // non-sort: g g2 e2, g g12 e2, g g1 (g = 1 / 2(g12 e2 - 1))
    double S1, S2;            // self: g g2 e2, g g1
    double hs;                // sorted-sum factor: 1/2 (0 without the sorted part)
};
static_assert(sizeof(KRec) == 16 * sizeof(double), "KRec layout");

__global__ void k_make_krec(LongConsts lc, const double4 *__restrict__ kv, int64_t Ptot,
                            KRec *__restrict__ kr)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Ptot) return;
    const double4 kk = kv[p];
    KRec r;
    r.kx = kk.x;
    r.ky = kk.y;
    r.k = kk.z;
    r.kxt = kk.x * kInv2Pi;
    r.kyt = kk.y * kInv2Pi;
    r.kl = kk.z * kLog2eD;
    r.w = kk.w;
    r.pad = 0.0;
    r.N1 = r.N2 = r.N3 = r.S1 = r.S2 = 0.0;
    const double e2 = exp(-2.0 * kk.z * lc.Lz);
    const double g = 0.5 / (lc.g12 * e2 - 1.0);
    r.wx = kk.w * (kk.x / kk.z);
    r.wy = kk.w * (kk.y / kk.z);
    if (lc.parts & QEM_LONG_NONSORT) {
        r.N1 = g * (lc.g2 * e2);
        r.N2 = g * (lc.g12 * e2);
        r.N3 = g * lc.g1;
    }
    if (lc.parts & QEM_LONG_SELF) {
        r.S1 = g * (lc.g2 * e2);
        r.S2 = g * lc.g1;
    }
    r.hs = (lc.parts & QEM_LONG_SORT) ? 0.5 : 0.0;
    kr[p] = r;
}

// one lane's row constants
struct LaneK {
    double kx, ky, k, w, wx, wy, N1, N2, N3, S1, S2, hs;
};

__device__ __forceinline__ LaneK load_lane(const KRec &r, bool mixed)
{
    LaneK L;
    L.kx = mixed ? r.kxt : r.kx;
    L.ky = mixed ? r.kyt : r.ky;
    L.k = mixed ? r.kl : r.k;
    L.w = r.w;
    L.wx = r.wx;
    L.wy = r.wy;
    L.N1 = r.N1;
    L.N2 = r.N2;
    L.N3 = r.N3;
    L.S1 = r.S1;
    L.S2 = r.S2;
    L.hs = r.hs;
    return L;
}

// cos, sin of k.rho and e^{+kz}, e^{-kz} from the lane's row constants
// (mixed: phase in turns reduced to [-1/2, 1/2] in FP64, then the hardware
// sine and cosine, which take their argument in turns (v_sin_f32,
// v_cos_f32; 1.2e-7 absolute over that range, measured on gfx950);
// e^{+-kz} = 2^{+-(n+f)} with the fraction f in FP32 and the integer n in
// the FP64 exponent, so no overflow)
template <int PREC>
__device__ __forceinline__ void lane_terms(const LaneK &L, double x, double y, double z,
                                           double &c, double &s, double &Ep, double &Em)
{
    if (PREC == QEM_PREC_FP64) {
        sincos(L.kx * x + L.ky * y, &s, &c);
        Ep = exp(L.k * z);
        Em = exp(-L.k * z);
    } else {
        double r = fma(L.kx, x, L.ky * y);
        r -= rint(r);
        const float t = (float)r;
        s = __builtin_amdgcn_sinf(t);
        c = __builtin_amdgcn_cosf(t);
        const double tz = L.k * z;
        const double nf = floor(tz);
        const float f = (float)(tz - nf);
        const int ni = (int)nf;
        Ep = ldexp((double)__builtin_amdgcn_exp2f(f), ni);
        Em = ldexp((double)__builtin_amdgcn_exp2f(-f), -ni);
    }
}

// Sum over the 64 lanes of a wave in a fixed order; lane 63 holds the sum.
// (pairs, quads, half rows, rows of 16, then rows 0+1 into 1, 2+3 into 3,
// and row 1 into rows 2, 3)
__device__ __forceinline__ double wave_sum63(double v)
{
    v = dev::row_sum16(v);
    v += dev::dpp_bcast_f64<0x142, 0xA>(v); // row_bcast:15 (rows 1, 3)
    v += dev::dpp_bcast_f64<0x143, 0xC>(v); // row_bcast:31 (rows 2, 3)
    return v;
}

// a + b without contraction into a fused multiply-add, so that passes A and
// B round their running sums identically
__device__ __forceinline__ double add_nc(double a, double b)
{
#pragma clang fp contract(off)
    return a + b;
}

constexpr int kLaneWaves = 4;   // waves (= chunks) per block in the k-lane passes

// Lane sums of three per-lane values for up to kRedBatch atoms at a time,
// through LDS: each lane stores its values of one atom in a row slot, and
// at a flush the 3 kRedBatch rows (slot j, component c: row 3 j + c) of 64
// values are summed by four lanes each (lane 4 r + q adds values 16 q ..
// 16 q + 15 of row r in order), then across the quad with two DPP steps, a
// fixed order.  Rows are kRedPitch doubles apart, so the stores of a slot
// are conflict-free and the 48 summing lanes share each bank pair at most
// twice.  A wave only reads what it wrote itself, and the LDS executes a
// wave's accesses in issue order; the wavefront fences keep the compiler
// from moving them across a flush.  Four atoms per flush keep the LDS at
// 25 KB per block of four waves, which the VGPR budget, not the LDS, then
// limits to five waves per SIMD.
constexpr int kRedBatch = 4;   // a power of two (slots are step & (kRedBatch - 1))
constexpr int kRedPitch = 65;

struct LaneReducer {
    double *buf;   // this wave's [kRedBatch][NC][kRedPitch], NC = 3 or 4 values per atom

    __device__ __forceinline__ void put(int slot, int lane, double x, double y, double z) const
    {
        double *r = buf + slot * 3 * kRedPitch;
        r[lane] = x;
        r[kRedPitch + lane] = y;
        r[2 * kRedPitch + lane] = z;
    }
    // four values per atom: the force and the potential
    __device__ __forceinline__ void put4(int slot, int lane, double x, double y, double z,
                                         double e) const
    {
        double *r = buf + slot * 4 * kRedPitch;
        r[lane] = x;
        r[kRedPitch + lane] = y;
        r[2 * kRedPitch + lane] = z;
        r[3 * kRedPitch + lane] = e;
    }

    // out[3 base + 3 j + c] (=, or += with add) for the first nb slots
    __device__ __forceinline__ void flush(int nb, int lane, double *__restrict__ out,
                                          int64_t base, bool add) const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int r = lane >> 2, q = lane & 3;
        const bool live = r < 3 * nb;
        double sum = 0.0;
        if (live) {
            const double *v = buf + r * kRedPitch + 16 * q;
#pragma unroll
            for (int m = 0; m < 16; ++m) sum += v[m];
        }
        sum += dev::dpp_perm_f64<0xB1>(sum);   // quad_perm [1,0,3,2]
        sum += dev::dpp_perm_f64<0x4E>(sum);   // quad_perm [2,3,0,1]
        if (live && q == 0) {
            double *o = out + 3 * base + r;
            *o = add ? *o + sum : sum;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};

// The last group of k rows holds Ptot mod 64 rows; when that is at most 32
// (or 16), pass B's waves there take 32 rows x 2 consecutive chunks
// (16 x 4) instead of 64 rows x 1 chunk, so that no lane idles: SPLIT
// sub-chunks per wave, lane = sub * (64 / SPLIT) + row.  The other groups,
// and pass A everywhere, use SPLIT = 1: pass A's short loop does not hide
// the per-lane atom loads of a split wave (measured: 58 us without the
// split, 61 with it, at N = 100k), pass B's does (125 -> 114 us).
__host__ __device__ constexpr int lane_split_of(int64_t rem)
{
    return rem == 0 ? 1 : rem <= 16 ? 4 : rem <= 32 ? 2 : 1;
}

template <int SPLIT>
__device__ constexpr int RW_of() { return 64 / SPLIT; }

// One wave's place in the k-lane grid: its rows, its sub-chunk, the atoms.
template <int SPLIT>
struct LaneSite {
    static constexpr int RW = 64 / SPLIT;   // rows per sub-chunk
    int64_t wfc;      // the wave's first chunk
    int64_t chunk;    // this lane's chunk (>= nchunk: none)
    int64_t p;        // this lane's k row (>= Ptot: none)
    int64_t a0, a1;   // this lane's atoms [a0, a1) (empty without a chunk)
    int64_t lmax;     // the longest sub-chunk of the wave (its first)
    __device__ __forceinline__ LaneSite(int wv, int lane, int64_t n, int chk, int nchunk)
    {
        const int sub = SPLIT == 1 ? 0 : lane / RW, rl = SPLIT == 1 ? lane : lane % RW;
        wfc = ((int64_t)blockIdx.x * kLaneWaves + wv) * SPLIT;
        chunk = wfc + sub;
        p = (int64_t)blockIdx.y * 64 + rl;
        a0 = chunk * chk;
        a1 = chunk < nchunk ? min(n, a0 + chk) : a0;
        lmax = min((int64_t)chk, n - wfc * chk);
    }
    // the atom of step t (clamped into the array for a lane past its end)
    __device__ __forceinline__ int64_t at(int64_t t, int64_t n) const
    {
        return min(a0 + t, n - 1);
    }
};

// Lane sums for a SPLIT-wave: the flush of LaneReducer with the 16-value
// row shares of lane 4 r + q summed over the quad (SPLIT 1), over pairs of
// the quad (SPLIT 2) or not at all (SPLIT 4), since the values 16 q ..
// 16 q + 15 of a row are then sub-chunk q's rows (or half of them).  Slot j
// of the batch that starts at step t0 is the atom a0 + t0 + j of each
// sub-chunk (up), or a1 - 1 - (t0 + j) (down); atoms past a sub-chunk's
// end are not written.
// With NC = 4 the rows are (atom j, component c) = row 4 j + c, and the
// fourth component (the potential) goes to oute[a].
template <int SPLIT, int NC = 3>
__device__ __forceinline__ void lane_flush(const LaneReducer &red, int nb, int lane,
                                           double *__restrict__ out, int64_t wfc, int chk,
                                           int64_t n, int nchunk, int64_t t0, bool down,
                                           bool add, double *__restrict__ oute = nullptr)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int r = lane >> 2, q = lane & 3;
    const bool live = r < NC * nb;
    double sum = 0.0;
    if (live) {
        const double *v = red.buf + r * kRedPitch + 16 * q;
        sum = v[0];
#pragma unroll
        for (int m = 1; m < 16; ++m) sum += v[m];
    }
    if (SPLIT <= 2) sum += dev::dpp_perm_f64<0xB1>(sum);   // quad_perm [1,0,3,2]
    if (SPLIT == 1) sum += dev::dpp_perm_f64<0x4E>(sum);   // quad_perm [2,3,0,1]
    constexpr int QPS = 4 / SPLIT;   // quad lanes per sub-chunk
    if (live && q % QPS == 0) {
        const int64_t ch = wfc + q / QPS;
        if (ch < nchunk) {
            const int64_t a0 = ch * chk, a1 = min(n, a0 + chk);
            const int j = r / NC, c = r - NC * j;
            const int64_t a = down ? a1 - 1 - (t0 + j) : a0 + t0 + j;
            if (a >= a0 && a < a1) {
                double *o = out + 3 * a + c;
                if constexpr (NC == 4)
                    if (c == 3) o = oute + a;
                *o = add ? *o + sum : sum;
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The lane sums of LaneReducer for one value per atom: up to kRedBatch1
// atoms per flush, one row of 64 values per atom, summed the same way (four
// lanes per row, 16 values each in order, then the quad).
constexpr int kRedBatch1 = 12;

struct LaneReducer1 {
    double *buf;   // this wave's [kRedBatch1][kRedPitch]

    __device__ __forceinline__ void put(int slot, int lane, double v) const
    {
        buf[slot * kRedPitch + lane] = v;
    }

    // out[base + j] (=, or += with add) for the first nb slots
    __device__ __forceinline__ void flush(int nb, int lane, double *__restrict__ out,
                                          int64_t base, bool add) const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int r = lane >> 2, q = lane & 3;
        const bool live = r < nb;
        double sum = 0.0;
        if (live) {
            const double *v = buf + r * kRedPitch + 16 * q;
            sum = v[0];
#pragma unroll
            for (int m = 1; m < 16; ++m) sum += v[m];
        }
        sum += dev::dpp_perm_f64<0xB1>(sum);   // quad_perm [1,0,3,2]
        sum += dev::dpp_perm_f64<0x4E>(sum);   // quad_perm [2,3,0,1]
        if (live && q == 0) {
            double *o = out + base + r;
            *o = add ? *o + sum : sum;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};
static_assert(kRedBatch1 * kRedPitch <= kRedBatch * 3 * kRedPitch, "LDS of the lane sums");

// Pass A (k lanes): per (k row, chunk) the four sums q cos e^{+kz},
// q sin e^{+kz}, q cos e^{-kz}, q sin e^{-kz}, accumulated in ascending z.
template <int PREC>
__global__ __launch_bounds__(64 * kLaneWaves) void k_lane_A(const KRec *__restrict__ kr,
                                                             int64_t Ptot,
                                                             const double4 *__restrict__ satom,
                                                             int64_t n, int chk, int nchunk,
                                                             double4 *__restrict__ partA)
{
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const LaneSite<1> w(wv, lane, n, chk, nchunk);
    if (w.wfc >= nchunk) return;
    const bool live = w.p < Ptot;
    const LaneK L = load_lane(kr[live ? w.p : Ptot - 1], PREC == QEM_PREC_MIXED);
    double sx = 0.0, sy = 0.0, sz = 0.0, sw = 0.0;
    // the next atom's (uniform, scalar) record is loaded one iteration
    // ahead, so it is in flight while this atom's terms are computed
    double4 nx = satom[w.a0];
    for (int64_t a = w.a0; a < w.a1; ++a) {
        const double4 at = nx;
        if (a + 1 < w.a1) nx = satom[a + 1];
        double c, s, Ep, Em;
        lane_terms<PREC>(L, at.x, at.y, at.z, c, s, Ep, Em);
        // the same rounded values and adds as pass B's running sums
        sx = add_nc(sx, at.w * (c * Ep));
        sy = add_nc(sy, at.w * (s * Ep));
        sz = add_nc(sz, at.w * (c * Em));
        sw = add_nc(sw, at.w * (s * Em));
    }
    if (live) partA[w.p * nchunk + w.chunk] = make_double4(sx, sy, sz, sw);
}

// Pass B (k lanes): forces of one chunk from one group of 64 k rows into
// partB[group][atom][3].  Prefix sums (below) run with the walk; the suffix
// sums (above) are the chunk offset plus the chunk total minus the running
// inclusive sum when the chunk is flat in e^{-kz} (k_max dz <= 1 over the
// chunk, so the subtraction loses at most a factor e in relative accuracy);
// otherwise a second walk in descending z accumulates them from the top.
// Every product that feeds a sum is an explicit fused multiply-add or a
// plain product, and the remaining sums are not contracted, so that the
// forces round the same with or without the potential (ENERGY).
//
// ENERGY: also the potential phi_i = dE/dq_i of the wave's 64 k rows into
// partE[group][atom] (the k = 0 potential is k_k0_apply's).  For k != 0,
//   phi_k(i) = w/(2k) [c e^{-kz} (A + P1) + s e^{-kz} (B + P2)
//                      + c e^{+kz} (C + P3) + s e^{+kz} (D + P4)]
// with the prefix/suffix sums A..D (the direct pair term) and P1..P4 the
// per-k derivative of the image term I_k = [g1 |S-|^2 + g2 e2 |S+|^2
// + 2 g1 g2 e2 Re(S+ conj S-)] / (2 (1 - g1 g2 e2)) with respect to q_i,
// divided by the atom's own terms.  With the non-sort row constants N1..N3
// (g = 1 / 2 (g12 e2 - 1), so 1 / (1 - g12 e2) = -2 g) these are
// P1 = -2 v2, P2 = -2 v1, P3 = -2 u2, P4 = -2 u1 for the u, v of the
// non-sort force, so the potential is
//   w/(2k) [(c e^{-kz} A + s e^{-kz} B) + (c e^{+kz} C + s e^{+kz} D) - 2 (X + Y)],
// X = c e^{+kz} u2 + s e^{+kz} u1, Y = c e^{-kz} v2 + s e^{-kz} v1, the
// first two brackets the sorted-sum force's own products.  It needs the
// non-sort constants (launch with the non-sort part on).
template <int PREC, int SPLIT, bool ENERGY = false>
__device__ __forceinline__ void lane_B_body(const KRec *__restrict__ kr, int64_t Ptot,
                                            const double4 *__restrict__ satom, int64_t n,
                                            int chk, int nchunk,
                                            const double4 *__restrict__ partA,
                                            const double4 *__restrict__ offA,
                                            const double4 *__restrict__ totA, int own_term,
                                            double *__restrict__ partB,
                                            double *__restrict__ partE, double *red_buf, int wv,
                                            int lane)
{
    constexpr int NC = ENERGY ? 4 : 3;
    const LaneSite<SPLIT> w(wv, lane, n, chk, nchunk);
    if (w.wfc >= nchunk) return;
    const bool live = w.p < Ptot;
    const int64_t pl = live ? w.p : Ptot - 1;
    const int64_t cl = w.chunk < nchunk ? w.chunk : nchunk - 1;
    LaneK L = load_lane(kr[pl], PREC == QEM_PREC_MIXED);
    const double kabs = live ? kr[pl].k : 0.0;
    if (!live) L.w = L.wx = L.wy = 0.0;   // a lane past the last row adds zeros
    // w/(2k), zero on a lane past the last row
    const double we = ENERGY ? L.w / (2.0 * kr[pl].k) : 0.0;
    const double4 T = totA[pl];
    // Fl_non_sort over all atoms, the atom's own term included: that term
    // is zero in xy (sin 0 = 0) and q e^{+-kz} (c^2 + s^2) times the z
    // factors, so the non-sort force is e^{+kz} (c u1 - s u2) + e^{-kz}
    // (c v1 - s v2) in xy and e^{+kz} (c u2 + s u1) - e^{-kz} (c v2 + s v1)
    // in z, with per-row constants from the totals; the own term,
    // -q (N1 e^{2kz} - N3 e^{-2kz}), and Flz_self, q (S1 e^{2kz} - S2
    // e^{-2kz}), combine into q (E1 e^{2kz} - E2 e^{-2kz}), which is zero
    // when both parts are on (S1 = N1, S2 = N3) and skipped then (own_term 0)
    const double u1 = fma(L.N1, T.y, L.N2 * T.w), u2 = fma(L.N1, T.x, L.N2 * T.z);
    const double v1 = fma(L.N2, T.y, L.N3 * T.w), v2 = fma(L.N2, T.x, L.N3 * T.z);
    const double E1 = L.S1 - L.N1, E2 = L.S2 - L.N3;
    const double4 off = offA[pl * nchunk + cl];
    const double4 own = partA[pl * nchunk + cl];
    const int64_t len = w.a1 - w.a0;
    // flat: k_max dz <= 1 over every sub-chunk of the wave (rows repeat
    // across sub-chunks, so the lane maximum of k is the rows' maximum)
    double kmax = kabs;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) kmax = fmax(kmax, __shfl_xor(kmax, d));
    double dz = len > 0 ? satom[w.a1 - 1].z - satom[w.a0].z : 0.0;
#pragma unroll
    for (int d = 32; d >= RW_of<SPLIT>(); d >>= 1) dz = fmax(dz, __shfl_xor(dz, d));
    const bool flat = __builtin_amdgcn_readfirstlane((int)(kmax * dz <= 1.0)) != 0;
    double *out = partB + ((int64_t)blockIdx.y * n) * 3;
    double *oute = ENERGY ? partE + (int64_t)blockIdx.y * n : nullptr;
    const LaneReducer red{red_buf};

    double A = off.x, B = off.y;   // exclusive prefix below the atom
    double PZ = 0.0, PW = 0.0;     // inclusive in-chunk sums of the e^{-kz} pair
    // the walk up the chunk, with the chunk's flatness and own_term (both
    // wave-uniform) as compile-time flags, so neither costs a select per atom
    const auto walk = [&](auto flat_c, auto own_c) {
        constexpr bool FLAT = decltype(flat_c)::value, OWN = decltype(own_c)::value;
        double4 nx = satom[w.at(0, n)];   // the next atom, loaded one iteration ahead
        for (int64_t t = 0; t < w.lmax; ++t) {
            const double4 at = nx;
            if (t + 1 < w.lmax) nx = satom[w.at(t + 1, n)];
            double fx = 0.0, fy = 0.0, fz = 0.0, fe = 0.0;
            if (SPLIT == 1 || t < len) {
                const double q = at.w;
                double c, s, Ep, Em;
                lane_terms<PREC>(L, at.x, at.y, at.z, c, s, Ep, Em);
                const double cp = c * Ep, sp = s * Ep, cm = c * Em, sm = s * Em;
                // Fl_non_sort (src/Force_long.jl:5-70), own term included
                // in the products cp, sp, cm, sm: four FMAs each
                const double Y = fma(cm, v2, sm * v1);
                const double gxn = fma(cp, u1, fma(-sp, u2, fma(cm, v1, -(sm * v2))));
                const double gzn = fma(cp, u2, fma(sp, u1, -Y));
                // Fl_sort (src/Force_long.jl:72-127): A, B below the atom; C, D
                // above it here for a flat chunk, else from the second walk
                double sxy = fma(sm, A, -(cm * B));
                const double z1 = fma(cm, A, sm * B);
                double szz = z1;
                double te = 0.0;
                if constexpr (ENERGY)
                    te = fma(-2.0, add_nc(fma(cp, u2, sp * u1), Y), z1);
                if constexpr (FLAT) {
                    PZ = add_nc(PZ, q * cm);
                    PW = add_nc(PW, q * sm);
                    const double C = off.z + (own.z - PZ), D = off.w + (own.w - PW);
                    const double z2 = fma(cp, C, sp * D);
                    sxy = add_nc(sxy, fma(sp, C, -(cp * D)));
                    szz = add_nc(szz, -z2);
                    if constexpr (ENERGY) te = add_nc(te, z2);
                }
                double gzs = fma(L.hs, szz, gzn);
                // Flz_self (src/Force_long.jl:164-184) minus the own term above
                if constexpr (OWN) gzs = fma(q, fma(E1, Ep * Ep, -(E2 * (Em * Em))), gzs);
                const double gx = q * fma(L.hs, sxy, gxn);
                const double gz = q * gzs;
                fx = L.wx * gx;
                fy = L.wy * gx;
                fz = L.w * gz;
                fe = we * te;
                A = fma(q, cp, A);
                B = fma(q, sp, B);
            }
            const int slot = (int)t & (kRedBatch - 1);
            if constexpr (ENERGY) red.put4(slot, lane, fx, fy, fz, fe);
            else red.put(slot, lane, fx, fy, fz);
            if (slot == kRedBatch - 1 || t + 1 == w.lmax)
                lane_flush<SPLIT, NC>(red, slot + 1, lane, out, w.wfc, chk, n, nchunk,
                                      t - slot, false, false, oute);
        }
    };
    if (flat) {
        if (own_term) walk(std::true_type{}, std::true_type{});
        else walk(std::true_type{}, std::false_type{});
    } else {
        if (own_term) walk(std::false_type{}, std::true_type{});
        else walk(std::false_type{}, std::false_type{});
    }
    if (flat) return;
    // exact suffix walk: the sorted-sum terms with C, D (above the atom),
    // from the top, added to the first walk's values
    double C = off.z, D = off.w;
    for (int64_t t = 0; t < w.lmax; ++t) {
        double fx = 0.0, fy = 0.0, fz = 0.0, fe = 0.0;
        if (SPLIT == 1 || t < len) {
            const double4 at = satom[w.a1 - 1 - t];
            const double q = at.w;
            double c, s, Ep, Em;
            lane_terms<PREC>(L, at.x, at.y, at.z, c, s, Ep, Em);
            const double cp = c * Ep, sp = s * Ep, cm = c * Em, sm = s * Em;
            const double z2 = fma(cp, C, sp * D);
            const double gx = q * (L.hs * fma(sp, C, -(cp * D)));
            const double gz = q * (L.hs * -z2);
            fx = L.wx * gx;
            fy = L.wy * gx;
            fz = L.w * gz;
            fe = we * z2;
            C = fma(q, cm, C);
            D = fma(q, sm, D);
        }
        const int slot = (int)t & (kRedBatch - 1);
        if constexpr (ENERGY) red.put4(slot, lane, fx, fy, fz, fe);
        else red.put(slot, lane, fx, fy, fz);
        if (slot == kRedBatch - 1 || t + 1 == w.lmax)
            lane_flush<SPLIT, NC>(red, slot + 1, lane, out, w.wfc, chk, n, nchunk, t - slot,
                                  true, true, oute);
    }
}

// Pass B over the k-lane grid; the last group of rows takes SPLIT
// sub-chunks per wave (lane_split_of).  ENERGY: the potential as well.
template <int PREC, bool ENERGY = false>
__global__ __launch_bounds__(64 * kLaneWaves) void k_lane_B(const KRec *__restrict__ kr,
                                                             int64_t Ptot,
                                                             const double4 *__restrict__ satom,
                                                             int64_t n, int chk, int nchunk,
                                                             int split_last,
                                                             const double4 *__restrict__ partA,
                                                             const double4 *__restrict__ offA,
                                                             const double4 *__restrict__ totA,
                                                             int own_term,
                                                             double *__restrict__ partB,
                                                             double *__restrict__ partE)
{
    constexpr int NC = ENERGY ? 4 : 3;
    __shared__ double red_lds[kLaneWaves][kRedBatch * NC * kRedPitch];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int split = blockIdx.y + 1 == gridDim.y ? split_last : 1;
    if (split == 4)
        lane_B_body<PREC, 4, ENERGY>(kr, Ptot, satom, n, chk, nchunk, partA, offA, totA,
                                     own_term, partB, partE, red_lds[wv], wv, lane);
    else if (split == 2)
        lane_B_body<PREC, 2, ENERGY>(kr, Ptot, satom, n, chk, nchunk, partA, offA, totA,
                                     own_term, partB, partE, red_lds[wv], wv, lane);
    else
        lane_B_body<PREC, 1, ENERGY>(kr, Ptot, satom, n, chk, nchunk, partA, offA, totA,
                                     own_term, partB, partE, red_lds[wv], wv, lane);
}

// ---------------------------------------------------------------------------
// Batch variance (k lanes): per atom and group of 64 k rows, the sum over the
// k != 0 rows of the row's force X = w F(k) and of |X|^2, into
// partV[group][atom] = (X_x, X_y, X_z, |X|^2).  Each lane needs its whole
// per-k force, so the suffix sums above each atom come from a descending
// pre-walk kept in LDS (chunks of at most kVarChunk atoms) instead of the
// subtractive form or a second walk.
constexpr int kVarChunk = 16;

template <int PREC>
__global__ __launch_bounds__(64 * kLaneWaves) void k_lane_V(const KRec *__restrict__ kr,
                                                             int64_t Ptot,
                                                             const double4 *__restrict__ satom,
                                                             int64_t n, int chk, int nchunk,
                                                             const double4 *__restrict__ offA,
                                                             const double4 *__restrict__ totA,
                                                             double4 *__restrict__ partV)
{
    __shared__ double2 sCD[kLaneWaves][kVarChunk][64];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t chunk = (int64_t)blockIdx.x * kLaneWaves + wv;
    if (chunk >= nchunk) return;
    const int64_t p = (int64_t)blockIdx.y * 64 + lane;
    const bool live = p < Ptot;
    const int64_t pl = live ? p : Ptot - 1;
    const LaneK L = load_lane(kr[pl], PREC == QEM_PREC_MIXED);
    const bool real = live;
    const double4 T = totA[pl];
    const double4 off = offA[pl * nchunk + chunk];
    const int64_t a0 = chunk * chk, a1 = min(n, a0 + chk);
    // each lane reads back only what it wrote: no barrier
    double C = off.z, D = off.w;
    for (int64_t a = a1 - 1; a >= a0; --a) {
        const double4 at = satom[a];
        double c, s, Ep, Em;
        lane_terms<PREC>(L, at.x, at.y, at.z, c, s, Ep, Em);
        sCD[wv][a - a0][lane] = make_double2(C, D);
        C += at.w * (c * Em);
        D += at.w * (s * Em);
    }
    double4 *out = partV + (int64_t)blockIdx.y * n;
    double A = off.x, B = off.y;
    for (int64_t a = a0; a < a1; ++a) {
        const double4 at = satom[a];
        const double q = at.w;
        double c, s, Ep, Em;
        lane_terms<PREC>(L, at.x, at.y, at.z, c, s, Ep, Em);
        const double cp = c * Ep, sp = s * Ep, cm = c * Em, sm = s * Em;
        const double vx = q * cp, vy = q * sp, vz = q * cm, vw = q * sm;
        const double2 cd = sCD[wv][a - a0][lane];
        // the force terms of k_lane_B with exact suffix sums
        const double f1 = T.x - vx, f2 = T.y - vy, f3 = T.z - vz, f4 = T.w - vw;
        const double X1 = c * f2 - s * f1, X2 = c * f4 - s * f3;
        const double Z1 = c * f1 + s * f2, Z2 = c * f3 + s * f4;
        const double gu = L.N1 * Ep + L.N2 * Em, gw = L.N3 * Em + L.N2 * Ep;
        const double guz = L.N1 * Ep - L.N2 * Em, gwz = L.N2 * Ep - L.N3 * Em;
        const double gxn = gu * X1 + gw * X2, gzn = guz * Z1 + gwz * Z2;
        const double sxy = sm * A - cm * B + sp * cd.x - cp * cd.y;
        const double szz = cm * A + sm * B - cp * cd.x - sp * cd.y;
        const double self = L.S1 * (Ep * Ep) - L.S2 * (Em * Em);
        const double gx = q * (gxn + L.hs * sxy);
        const double gz = q * (gzn + L.hs * szz + q * self);
        double fx = real ? L.wx * gx : 0.0;
        double fy = real ? L.wy * gx : 0.0;
        double fz = real ? L.w * gz : 0.0;
        double f2s = fx * fx + fy * fy + fz * fz;
        A += vx;
        B += vy;
        fx = wave_sum63(fx);
        fy = wave_sum63(fy);
        fz = wave_sum63(fz);
        f2s = wave_sum63(f2s);
        if (lane == 63) out[a] = make_double4(fx, fy, fz, f2s);
    }
}

// var[atom] = (P sum|X|^2 - |sum X|^2) / (P - 1) over the groups' partials
__global__ void k_var_R(const double4 *__restrict__ partV, int ngroup, int64_t n,
                        const int32_t *__restrict__ perm, double P, double *__restrict__ var)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double fx = 0.0, fy = 0.0, fz = 0.0, f2 = 0.0;
    for (int g = 0; g < ngroup; ++g) {
        const double4 v = partV[(int64_t)g * n + r];
        fx += v.x;
        fy += v.y;
        fz += v.z;
        f2 += v.w;
    }
    var[perm[r]] = (P * f2 - (fx * fx + fy * fy + fz * fz)) / (P - 1.0);
}

// ---------------------------------------------------------------------------
// Tangent passes in the k-lane layout (same walk as k_lane_A / k_lane_B).
// Along the direction v the per-atom terms move with
//   d(c e^{+kz}) = -s e^{+kz} dth + c e^{+kz} k v_z,   dth = k_x v_x + k_y v_y,
//   d(s e^{+kz}) =  c e^{+kz} dth + s e^{+kz} k v_z,
//   d(c e^{-kz}) = -s e^{-kz} dth - c e^{-kz} k v_z,
//   d(s e^{-kz}) =  c e^{-kz} dth - s e^{-kz} k v_z,
// and every sum carries its tangent.  The k = 0 term has no tangent (its z
// force is piecewise constant in z).

// values and tangents of the four per-atom channels (without the charge)
struct LaneTerms {
    double c, s, Ep, Em, cp, sp, cm, sm, dcp, dsp, dcm, dsm;
};

template <int PREC>
__device__ __forceinline__ LaneTerms lane_terms_t(const LaneK &L, double kx, double ky,
                                                  double kabs, const double4 &at,
                                                  const double3 &dv)
{
    LaneTerms T;
    lane_terms<PREC>(L, at.x, at.y, at.z, T.c, T.s, T.Ep, T.Em);
    T.cp = T.c * T.Ep;
    T.sp = T.s * T.Ep;
    T.cm = T.c * T.Em;
    T.sm = T.s * T.Em;
    const double dth = kx * dv.x + ky * dv.y, kvz = kabs * dv.z;
    T.dcp = T.cp * kvz - T.sp * dth;
    T.dsp = T.sp * kvz + T.cp * dth;
    T.dcm = -T.cm * kvz - T.sm * dth;
    T.dsm = T.cm * dth - T.sm * kvz;
    return T;
}

// Pass A (tangent): chunk sums of the four channels and of their tangents.
template <int PREC>
__global__ __launch_bounds__(64 * kLaneWaves) void k_lane_A_t(
    const KRec *__restrict__ kr, int64_t Ptot, const double4 *__restrict__ satom,
    const double *__restrict__ svx, const double *__restrict__ svy,
    const double *__restrict__ svz, int64_t n, int chk, int nchunk,
    double4 *__restrict__ partA, double4 *__restrict__ partT)
{
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t chunk = (int64_t)blockIdx.x * kLaneWaves + wv;
    if (chunk >= nchunk) return;
    const int64_t p = (int64_t)blockIdx.y * 64 + lane;
    const bool live = p < Ptot;
    const KRec &R = kr[live ? p : Ptot - 1];
    const LaneK L = load_lane(R, PREC == QEM_PREC_MIXED);
    const double kx = R.kx, ky = R.ky, kabs = R.k;
    const int64_t a0 = chunk * chk, a1 = min(n, a0 + chk);
    double sx = 0.0, sy = 0.0, sz = 0.0, sw = 0.0, tx = 0.0, ty = 0.0, tz = 0.0, tw = 0.0;
    // the next atom and its direction, loaded one iteration ahead
    double4 nx = satom[a0];
    double3 nv = make_double3(svx[a0], svy[a0], svz[a0]);
    for (int64_t a = a0; a < a1; ++a) {
        const double4 at = nx;
        const double3 dv = nv;
        if (a + 1 < a1) {
            nx = satom[a + 1];
            nv = make_double3(svx[a + 1], svy[a + 1], svz[a + 1]);
        }
        const LaneTerms T = lane_terms_t<PREC>(L, kx, ky, kabs, at, dv);
        const double q = at.w;
        sx = add_nc(sx, q * T.cp);
        sy = add_nc(sy, q * T.sp);
        sz = add_nc(sz, q * T.cm);
        sw = add_nc(sw, q * T.sm);
        tx = add_nc(tx, q * T.dcp);
        ty = add_nc(ty, q * T.dsp);
        tz = add_nc(tz, q * T.dcm);
        tw = add_nc(tw, q * T.dsm);
    }
    if (live) {
        partA[p * nchunk + chunk] = make_double4(sx, sy, sz, sw);
        partT[p * nchunk + chunk] = make_double4(tx, ty, tz, tw);
    }
}

// Pass B (tangent): tangent of the per-k force of each atom, summed over
// the wave's 64 rows, into partB[group][atom][3].
template <int PREC>
__global__ __launch_bounds__(64 * kLaneWaves) void k_lane_B_t(
    const KRec *__restrict__ kr, int64_t Ptot, const double4 *__restrict__ satom,
    const double *__restrict__ svx, const double *__restrict__ svy,
    const double *__restrict__ svz, int64_t n, int chk, int nchunk,
    const double4 *__restrict__ partA, const double4 *__restrict__ offA,
    const double4 *__restrict__ totA, const double4 *__restrict__ partT,
    const double4 *__restrict__ offT, const double4 *__restrict__ totT, int own_term,
    double *__restrict__ partB)
{
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t chunk = (int64_t)blockIdx.x * kLaneWaves + wv;
    if (chunk >= nchunk) return;
    const int64_t p = (int64_t)blockIdx.y * 64 + lane;
    const bool live = p < Ptot;
    const int64_t pl = live ? p : Ptot - 1;
    const KRec &R = kr[pl];
    LaneK L = load_lane(R, PREC == QEM_PREC_MIXED);
    const double kx = R.kx, ky = R.ky, kabs = R.k;
    if (!live) L.w = L.wx = L.wy = 0.0;   // a lane past the last row adds zeros
    const double4 T = totA[pl], dT = totT[pl];
    // the non-sort term with the atom's own term kept, as in k_lane_B: per
    // row the constants from the totals and from their tangents
    const double u1 = fma(L.N1, T.y, L.N2 * T.w), u2 = fma(L.N1, T.x, L.N2 * T.z);
    const double v1 = fma(L.N2, T.y, L.N3 * T.w), v2 = fma(L.N2, T.x, L.N3 * T.z);
    const double du1 = fma(L.N1, dT.y, L.N2 * dT.w), du2 = fma(L.N1, dT.x, L.N2 * dT.z);
    const double dv1 = fma(L.N2, dT.y, L.N3 * dT.w), dv2 = fma(L.N2, dT.x, L.N3 * dT.z);
    const double E1 = L.S1 - L.N1, E2 = L.S2 - L.N3;
    const double4 off = offA[pl * nchunk + chunk], offt = offT[pl * nchunk + chunk];
    const double4 own = partA[pl * nchunk + chunk], ownt = partT[pl * nchunk + chunk];
    const int64_t a0 = chunk * chk, a1 = min(n, a0 + chk);
    double kmax = live ? kabs : 0.0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) kmax = fmax(kmax, __shfl_xor(kmax, d));
    const bool flat = kmax * (satom[a1 - 1].z - satom[a0].z) <= 1.0;
    double *out = partB + ((int64_t)blockIdx.y * n) * 3;
    __shared__ double red_lds[kLaneWaves][kRedBatch * 3 * kRedPitch];
    const LaneReducer red{red_lds[wv]};

    double A = off.x, B = off.y, dA = offt.x, dB = offt.y;
    double PZ = 0.0, PW = 0.0, dPZ = 0.0, dPW = 0.0;
    // the next atom and its direction, loaded one iteration ahead
    double4 nx = satom[a0];
    double3 nv = make_double3(svx[a0], svy[a0], svz[a0]);
    for (int64_t a = a0; a < a1; ++a) {
        const double4 at = nx;
        const double3 dv = nv;
        if (a + 1 < a1) {
            nx = satom[a + 1];
            nv = make_double3(svx[a + 1], svy[a + 1], svz[a + 1]);
        }
        const LaneTerms X = lane_terms_t<PREC>(L, kx, ky, kabs, at, dv);
        const double q = at.w;
        const double vx = q * X.cp, vy = q * X.sp, vz = q * X.cm, vw = q * X.sm;
        const double tvx = q * X.dcp, tvy = q * X.dsp, tvz = q * X.dcm, tvw = q * X.dsm;
        PZ = add_nc(PZ, vz);
        PW = add_nc(PW, vw);
        dPZ = add_nc(dPZ, tvz);
        dPW = add_nc(dPW, tvw);
        double C = 0.0, D = 0.0, dC = 0.0, dD = 0.0;
        if (flat) {
            C = off.z + (own.z - PZ);
            D = off.w + (own.w - PW);
            dC = offt.z + (ownt.z - dPZ);
            dD = offt.w + (ownt.w - dPW);
        }
        const double c = X.c, s = X.s, Ep = X.Ep, Em = X.Em;
        const double dth = kx * dv.x + ky * dv.y, kvz = kabs * dv.z;
        // Fl_non_sort (src/Force_long.jl:5-70), own term included: the
        // force is Ep (c u1 - s u2) + Em (c v1 - s v2) in xy and
        // Ep (c u2 + s u1) - Em (c v2 + s v1) in z; along v, dc = -s dth,
        // ds = c dth, dEp = kvz Ep, dEm = -kvz Em and the per-row constants
        // move with the totals' tangents
        const double xu = c * u1 - s * u2, xv = c * v1 - s * v2;
        const double zu = c * u2 + s * u1, zv = c * v2 + s * v1;
        const double pu = du1 - dth * u2, qu = du2 + dth * u1;
        const double pv = dv1 - dth * v2, qv = dv2 + dth * v1;
        const double dxu = c * pu - s * qu, dxv = c * pv - s * qv;
        const double dzu = c * qu + s * pu, dzv = c * qv + s * pv;
        const double dgxn = Ep * fma(kvz, xu, dxu) + Em * (dxv - kvz * xv);
        const double dgzn = Ep * fma(kvz, zu, dzu) + Em * (kvz * zv - dzv);
        // Fl_sort (src/Force_long.jl:72-127)
        const double dsxy = X.dsm * A + X.sm * dA - X.dcm * B - X.cm * dB + X.dsp * C +
                            X.sp * dC - X.dcp * D - X.cp * dD;
        const double dszz = X.dcm * A + X.cm * dA + X.dsm * B + X.sm * dB - X.dcp * C -
                            X.cp * dC - X.dsp * D - X.sp * dD;
        double gzs = dgzn + L.hs * dszz;
        // Flz_self (src/Force_long.jl:164-184) minus the own term above:
        // the tangent of q (E1 Ep^2 - E2 Em^2)
        if (own_term) gzs += q * (2.0 * kvz * (E1 * (Ep * Ep) + E2 * (Em * Em)));
        const double gx = q * (dgxn + L.hs * dsxy);
        const double gz = q * gzs;
        const double fx = L.wx * gx, fy = L.wy * gx, fz = L.w * gz;
        A += vx;
        B += vy;
        dA += tvx;
        dB += tvy;
        const int slot = (int)(a - a0) & (kRedBatch - 1);
        red.put(slot, lane, fx, fy, fz);
        if (slot == kRedBatch - 1 || a == a1 - 1) red.flush(slot + 1, lane, out, a - slot, false);
    }
    if (flat) return;
    // exact suffix walk: the sorted-sum tangent terms with C, D (above),
    // in batches [lo, hi] taken from the top
    double C = off.z, D = off.w, dC = offt.z, dD = offt.w;
    int64_t lo = a1, hi = a1 - 1;
    for (int64_t a = a1 - 1; a >= a0; --a) {
        if (a < lo) {
            hi = a;
            lo = a - kRedBatch + 1 < a0 ? a0 : a - kRedBatch + 1;
        }
        const double4 at = satom[a];
        const double3 dv = make_double3(svx[a], svy[a], svz[a]);
        const LaneTerms X = lane_terms_t<PREC>(L, kx, ky, kabs, at, dv);
        const double q = at.w;
        const double dsxy = X.dsp * C + X.sp * dC - X.dcp * D - X.cp * dD;
        const double dszz = -X.dcp * C - X.cp * dC - X.dsp * D - X.sp * dD;
        const double gx = q * (L.hs * dsxy);
        const double gz = q * (L.hs * dszz);
        const double fx = L.wx * gx, fy = L.wy * gx, fz = L.w * gz;
        C += q * X.cm;
        D += q * X.sm;
        dC += q * X.dcm;
        dD += q * X.dsm;
        red.put((int)(a - lo), lane, fx, fy, fz);
        if (a == lo) red.flush((int)(hi - lo + 1), lane, out, lo, true);
    }
}

// Energy pass B (k lanes): per atom the potential phi_i = dE/dq_i of the
// wave's 64 k rows into partB[group][atom].  For k != 0,
//   phi_k(i) = w/(2k) [c e^{-kz} (A + P1) + s e^{-kz} (B + P2)
//                      + c e^{+kz} (C + P3) + s e^{+kz} (D + P4)]
// with the prefix/suffix sums A..D (the direct pair term) and P1..P4 the
// per-k derivative of the image term I_k = [g1 |S-|^2 + g2 e2 |S+|^2
// + 2 g1 g2 e2 Re(S+ conj S-)] / (2 (1 - g1 g2 e2)) with respect to q_i,
// divided by the atom's own terms.  (The k = 0 potential is k_k0_apply's.)
template <int PREC>
__global__ __launch_bounds__(64 * kLaneWaves) void k_lane_B_energy(
    LongConsts lc, const KRec *__restrict__ kr, int64_t Ptot, const double4 *__restrict__ satom,
    int64_t n, int chk, int nchunk, const double4 *__restrict__ partA,
    const double4 *__restrict__ offA, const double4 *__restrict__ totA,
    double *__restrict__ partB)
{
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t chunk = (int64_t)blockIdx.x * kLaneWaves + wv;
    if (chunk >= nchunk) return;
    const int64_t p = (int64_t)blockIdx.y * 64 + lane;
    const bool live = p < Ptot;
    const int64_t pl = live ? p : Ptot - 1;
    const KRec &R = kr[pl];
    const LaneK L = load_lane(R, PREC == QEM_PREC_MIXED);
    const double4 T = totA[pl];
    const double4 off = offA[pl * nchunk + chunk];
    const double4 own = partA[pl * nchunk + chunk];
    // per-row factors: w/(2k) (0 on an idle lane) and the image-term
    // derivative coefficients
    const double e2 = exp(-2.0 * R.k * lc.Lz);
    const double F = 1.0 / (1.0 - lc.g12 * e2);   // 2 / (2 (1 - g12 e2))
    const double we = live ? R.w / (2.0 * R.k) : 0.0;
    const double P1 = F * (lc.g1 * T.z + lc.g12 * e2 * T.x);
    const double P2 = F * (lc.g1 * T.w + lc.g12 * e2 * T.y);
    const double P3 = F * e2 * (lc.g2 * T.x + lc.g12 * T.z);
    const double P4 = F * e2 * (lc.g2 * T.y + lc.g12 * T.w);
    const int64_t a0 = chunk * chk, a1 = min(n, a0 + chk);
    double kmax = live ? R.k : 0.0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) kmax = fmax(kmax, __shfl_xor(kmax, d));
    const bool flat =
        __builtin_amdgcn_readfirstlane((int)(kmax * (satom[a1 - 1].z - satom[a0].z) <= 1.0)) != 0;
    double *out = partB + (int64_t)blockIdx.y * n;
    __shared__ double red_lds[kLaneWaves][kRedBatch1 * kRedPitch];
    const LaneReducer1 red{red_lds[wv]};

    double A = off.x, B = off.y, PZ = 0.0, PW = 0.0;
    // the walk up the chunk, the chunk's flatness a compile-time flag
    const auto walk = [&](auto flat_c) {
        constexpr bool FLAT = decltype(flat_c)::value;
        double4 nx = satom[a0];    // the next atom, loaded one iteration ahead
        int slot = 0;              // (a - a0) mod kRedBatch1
        for (int64_t a = a0; a < a1; ++a) {
            const double4 at = nx;
            if (a + 1 < a1) nx = satom[a + 1];
            const double q = at.w;
            double c, s, Ep, Em;
            lane_terms<PREC>(L, at.x, at.y, at.z, c, s, Ep, Em);
            const double cp = c * Ep, sp = s * Ep, cm = c * Em, sm = s * Em;
            double t = cm * (A + P1) + sm * (B + P2) + cp * P3 + sp * P4;
            if constexpr (FLAT) {
                PZ = add_nc(PZ, q * cm);
                PW = add_nc(PW, q * sm);
                t += cp * (off.z + (own.z - PZ)) + sp * (off.w + (own.w - PW));
            }
            A += q * cp;
            B += q * sp;
            red.put(slot, lane, we * t);
            if (slot == kRedBatch1 - 1 || a == a1 - 1) {
                red.flush(slot + 1, lane, out, a - slot, false);
                slot = 0;
            } else {
                ++slot;
            }
        }
    };
    if (flat) walk(std::true_type{});
    else walk(std::false_type{});
    if (flat) return;
    // exact suffix walk: the terms with C, D (above the atom), in batches
    // [lo, hi] taken from the top, added to the first walk's values
    double C = off.z, D = off.w;
    int64_t lo = a1, hi = a1 - 1;
    for (int64_t a = a1 - 1; a >= a0; --a) {
        if (a < lo) {
            hi = a;
            lo = a - kRedBatch1 + 1 < a0 ? a0 : a - kRedBatch1 + 1;
        }
        const double4 at = satom[a];
        const double q = at.w;
        double c, s, Ep, Em;
        lane_terms<PREC>(L, at.x, at.y, at.z, c, s, Ep, Em);
        const double cp = c * Ep, sp = s * Ep, cm = c * Em, sm = s * Em;
        red.put((int)(a - lo), lane, we * (cp * C + sp * D));
        C += q * cm;
        D += q * sm;
        if (a == lo) red.flush((int)(hi - lo + 1), lane, out, lo, true);
    }
}

// ---------------------------------------------------------------------------
// The k = 0 term (Flz_sort_k0, src/Force_long.jl:130-161) and its potential,
// off the k lanes.  For the atom of rank r in z order, with Q_b, QZ_b the
// sums of q and q z over the atoms before it and Q_a, QZ_a over those after,
//   F0_z(r) = w0 q (Q_b - Q_a),
//   phi0(r) = -w0 [z (Q_b - Q_a) - (QZ_b - QZ_a)],   w0 = 1 / (2 Lx Ly eps_0),
// into k0[r] = (F0_z, phi0).  Tiles of kK0Tile ranks, kK0Items consecutive
// ranks per thread: the tile sums, their exclusive scan (one block), then
// the per-atom pass, where the sums after an atom are the total minus those
// up to it (charges of either sign and no exponentials: no cancellation
// beyond the rounding of the totals).
constexpr int kK0Items = 8, kK0Tile = CH * kK0Items;

__global__ __launch_bounds__(CH) void k_k0_tiles(const double4 *__restrict__ satom, int64_t n,
                                                 double2 *__restrict__ tiles)
{
    __shared__ double2 sh[CH / 64];
    const int64_t base = (int64_t)blockIdx.x * kK0Tile + (int64_t)threadIdx.x * kK0Items;
    double2 v = make_double2(0.0, 0.0);
    for (int i = 0; i < kK0Items; ++i) {
        if (base + i < n) {
            const double4 at = satom[base + i];
            v.x += at.w;
            v.y += at.w * at.z;
        }
    }
    double2 tot;
    (void)block_excl_scan(v, &tot, sh);
    if (threadIdx.x == 0) tiles[blockIdx.x] = tot;
}

// in place: tiles[t] becomes the sum of the tiles before t, tiles[ntile]
// the total
__global__ __launch_bounds__(CH) void k_k0_scan(double2 *__restrict__ tiles, int ntile)
{
    __shared__ double2 sh[CH / 64];
    double2 run = make_double2(0.0, 0.0);
    for (int b = 0; b < ntile; b += CH) {
        const int m = b + (int)threadIdx.x;
        const double2 v = m < ntile ? tiles[m] : make_double2(0.0, 0.0);
        double2 tot;
        const double2 ex = block_excl_scan(v, &tot, sh);
        if (m < ntile) tiles[m] = make_double2(run.x + ex.x, run.y + ex.y);
        run.x += tot.x;
        run.y += tot.y;
    }
    if (threadIdx.x == 0) tiles[ntile] = run;
}

__global__ __launch_bounds__(CH) void k_k0_apply(const double4 *__restrict__ satom, int64_t n,
                                                 const double2 *__restrict__ tiles, int ntile,
                                                 double w0, double2 *__restrict__ k0)
{
    __shared__ double2 sh[CH / 64];
    const int64_t base = (int64_t)blockIdx.x * kK0Tile + (int64_t)threadIdx.x * kK0Items;
    double q[kK0Items], z[kK0Items];
    double2 v = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < kK0Items; ++i) {
        q[i] = z[i] = 0.0;
        if (base + i < n) {
            const double4 at = satom[base + i];
            q[i] = at.w;
            z[i] = at.z;
        }
        v.x += q[i];
        v.y += q[i] * z[i];
    }
    double2 tot;
    const double2 ex = block_excl_scan(v, &tot, sh);
    const double2 T = tiles[ntile], t0 = tiles[blockIdx.x];
    double Qb = t0.x + ex.x, QZb = t0.y + ex.y;
#pragma unroll
    for (int i = 0; i < kK0Items; ++i) {
        const double qz = q[i] * z[i];
        const double Qa = (T.x - Qb) - q[i], QZa = (T.y - QZb) - qz;
        const double dQ = Qb - Qa;
        if (base + i < n)
            k0[base + i] = make_double2(w0 * (q[i] * dQ), -w0 * (z[i] * dQ - (QZb - QZa)));
        Qb += q[i];
        QZb += qz;
    }
}

// Pass R: sum the k tiles (plus the k = 0 term when k0 is given) and
// scatter to input order.
__global__ void k_long_R(const double *__restrict__ partB, int ntile, int64_t n,
                         const int32_t *__restrict__ perm, const double2 *__restrict__ k0,
                         int accumulate, double *__restrict__ f)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int t = 0; t < ntile; ++t) {
        const double *o = partB + ((int64_t)t * n + r) * 3;
        fx += o[0];
        fy += o[1];
        fz += o[2];
    }
    if (k0) fz += k0[r].x;
    const int64_t a = perm[r];
    if (accumulate) {
        f[3 * a] += fx;
        f[3 * a + 1] += fy;
        f[3 * a + 2] += fz;
    } else {
        f[3 * a] = fx;
        f[3 * a + 1] = fy;
        f[3 * a + 2] = fz;
    }
}

// k_long_R fused with the short range's output pass: for the atom a of
// rank r, f[a] = (the slot-ordered short-range force of a, plus addz[a] on
// z) + (the long range of rank r), in that order, as the two passes would
// leave it.  slot_of: the slot of each atom.
__global__ void k_long_R_slots(const double *__restrict__ partB, int ntile, int64_t n,
                               const int32_t *__restrict__ perm,
                               const double2 *__restrict__ k0,
                               const int32_t *__restrict__ slot_of,
                               const double *__restrict__ sout, const double *__restrict__ addz,
                               double *__restrict__ f)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int t = 0; t < ntile; ++t) {
        const double *o = partB + ((int64_t)t * n + r) * 3;
        fx += o[0];
        fy += o[1];
        fz += o[2];
    }
    if (k0) fz += k0[r].x;
    const int64_t a = perm[r];
    const double *sv = sout + 3 * (int64_t)slot_of[a];
    double z = sv[2];
    if (addz) z = __dadd_rn(z, addz[a]);
    f[3 * a] = sv[0] + fx;
    f[3 * a + 1] = sv[1] + fy;
    f[3 * a + 2] = z + fz;
}

// ---------------------------------------------------------------------------
// Adjoint helpers (no counterpart in the reference).

// per k row: w_k/(2k) dI_k/dgamma_1,2 from the structure factors
__global__ void k_long_dgamma(LongConsts lc, const double4 *__restrict__ kv, int64_t Ptot,
                              const double4 *__restrict__ totA, double2 *__restrict__ out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Ptot) return;
    const double4 kk = kv[p];
    const double k = kk.z;
    const double4 S = totA[p];
    const double e2 = exp(-2.0 * k * lc.Lz);
    const double om = 1.0 - lc.g12 * e2;
    const double A3 = S.z * S.z + S.w * S.w;
    const double A1e = (e2 * S.x) * S.x + (e2 * S.y) * S.y;
    const double Xe = (e2 * S.x) * S.z + (e2 * S.y) * S.w;
    const double N = lc.g1 * A3 + lc.g2 * A1e + 2.0 * lc.g12 * Xe;
    const double d1 = (A3 + 2.0 * lc.g2 * Xe) / (2.0 * om) + N * lc.g2 * e2 / (2.0 * om * om);
    const double d2 = (A1e + 2.0 * lc.g1 * Xe) / (2.0 * om) + N * lc.g1 * e2 / (2.0 * om * om);
    const double f = kk.w / (2.0 * k);
    out[p] = make_double2(f * d1, f * d2);
}

__global__ void k_gather_sorted3(const double *__restrict__ v, const int32_t *__restrict__ perm,
                                 int64_t n, double *__restrict__ svx, double *__restrict__ svy,
                                 double *__restrict__ svz)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t a = perm[r];
    svx[r] = v[3 * a];
    svy[r] = v[3 * a + 1];
    svz[r] = v[3 * a + 2];
}

// sum the k tiles of a width-w per-atom quantity (for the potential,
// width 1, plus the k = 0 term's when k0 is given), scale, scatter to
// input order
__global__ void k_long_Rw(const double *__restrict__ partB, int ntile, int64_t n, int width,
                          const int32_t *__restrict__ perm, const double2 *__restrict__ k0,
                          int accumulate, double scale, double *__restrict__ f)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t a = perm[r];
    for (int c = 0; c < width; ++c) {
        double acc = 0.0;
        for (int t = 0; t < ntile; ++t) acc += partB[((int64_t)t * n + r) * width + c];
        if (k0 && width == 1) acc += k0[r].y;
        acc *= scale;
        if (accumulate) f[width * a + c] += acc;
        else f[width * a + c] = acc;
    }
}

#define LR_CHECK(expr)                                                        \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            err = std::string("long range: ") + #expr + ": " +                \
                  hipGetErrorString(e_);                                      \
            return e_;                                                        \
        }                                                                     \
    } while (0)

}  // namespace

hipError_t launch_zsort(LongWork &lw, int64_t n, const double *d_xyz, const double *d_q,
                        hipStream_t s, std::string &err)
{
    if (n <= 0) return hipSuccess;
    if (n > INT32_MAX) {
        err = "long range: more than 2^31-1 atoms";
        return hipErrorInvalidValue;
    }
    LR_CHECK(ensure(lw.perm, sizeof(int32_t) * (size_t)n));
    LR_CHECK(ensure(lw.satom, sizeof(double4) * (size_t)n));
    LR_CHECK(ensure_sort_scratch(lw.tmp, radix_sort_scratch_bytes(n, 8), s));
    LR_CHECK(radix_sort_z(d_xyz, d_q, n, lw.perm.as<int32_t>(), lw.satom.as<double4>(), lw.tmp.p,
                          s));
    return hipSuccess;
}

hipError_t launch_long_k0(const LongConsts &c, LongWork &lw, int64_t n, hipStream_t s,
                          std::string &err)
{
    if (n <= 0) return hipSuccess;
    const int ntile = (int)((n + kK0Tile - 1) / kK0Tile);
    LR_CHECK(ensure(lw.k0tile, sizeof(double2) * ((size_t)ntile + 1)));
    LR_CHECK(ensure(lw.k0out, sizeof(double2) * (size_t)n));
    double2 *tiles = lw.k0tile.as<double2>();
    const double4 *sa = lw.satom.as<double4>();
    k_k0_tiles<<<(unsigned)ntile, CH, 0, s>>>(sa, n, tiles);
    LR_CHECK(hipGetLastError());
    k_k0_scan<<<1, CH, 0, s>>>(tiles, ntile);
    LR_CHECK(hipGetLastError());
    k_k0_apply<<<(unsigned)ntile, CH, 0, s>>>(sa, n, tiles, ntile, 0.5 * c.inv_area_eps,
                                               lw.k0out.as<double2>());
    LR_CHECK(hipGetLastError());
    return hipSuccess;
}

namespace {
// k-lane passes: k rows in groups of 64 (one per lane); chunks of chk atoms
// sized so that the grid holds about 8192 waves
// (the last group's waves take split_last chunks each, lane_split_of)
struct LaneGrid {
    int ngroup, chk, nchunk, split_last;
};
LaneGrid lane_grid(int64_t n, int64_t Ptot)
{
    LaneGrid g;
    g.ngroup = (int)((Ptot + 63) / 64);
    g.split_last = lane_split_of(Ptot % 64);
    // chunks per group-wave row: (ngroup - 1) full groups plus 1 / split_last
    const int64_t units = (int64_t)(g.ngroup - 1) * g.split_last + 1;
    const int64_t want = (n * units + 8192LL * g.split_last - 1) / (8192LL * g.split_last);
    g.chk = (int)std::min<int64_t>(256, std::max<int64_t>(16, want));
    g.nchunk = (int)((n + g.chk - 1) / g.chk);
    return g;
}
}  // namespace

LongSizes long_force_sizes(int64_t n, int64_t Ptot)
{
    LongSizes z;
    const LaneGrid g = lane_grid(std::max<int64_t>(n, 1), std::max<int64_t>(Ptot, 1));
    const size_t P = (size_t)std::max<int64_t>(Ptot, 1), N = (size_t)std::max<int64_t>(n, 1);
    z.kv = sizeof(double4) * P;
    z.kr = sizeof(double) * 16 * P;
    z.perm = sizeof(int32_t) * N;
    z.satom = sizeof(double4) * N;
    z.sort = radix_sort_scratch_bytes((int64_t)N, 8);
    z.partA = sizeof(double4) * P * (size_t)g.nchunk;
    z.totA = sizeof(double4) * P;
    z.partB = sizeof(double) * 3 * (size_t)g.ngroup * N;
    z.k0tile = sizeof(double2) * ((N + kK0Tile - 1) / kK0Tile + 1);
    z.k0out = sizeof(double2) * N;
    return z;
}

hipError_t launch_long_krec(const LongConsts &c, LongWork &lw, int64_t Ptot, hipStream_t s,
                            std::string &err)
{
    if (Ptot <= 0) return hipSuccess;
    LR_CHECK(ensure(lw.kr, sizeof(double) * 16 * (size_t)Ptot));
    k_make_krec<<<nblk(Ptot, 256), 256, 0, s>>>(c, lw.kv.as<double4>(), Ptot, lw.kr.as<KRec>());
    LR_CHECK(hipGetLastError());
    return hipSuccess;
}

hipError_t launch_long_passes(const LongConsts &c, LongWork &lw, int64_t n, int64_t Ptot,
                              hipStream_t s, std::string &err, bool krec, int which)
{
    if (n <= 0 || Ptot <= 0) return hipSuccess;
    const LaneGrid lg = lane_grid(n, Ptot);
    const int ngroup = lg.ngroup, chk = lg.chk, nchunk = lg.nchunk, sl = lg.split_last;
    LR_CHECK(ensure(lw.partA, sizeof(double4) * (size_t)Ptot * nchunk));
    LR_CHECK(ensure(lw.offA, sizeof(double4) * (size_t)Ptot * nchunk));
    LR_CHECK(ensure(lw.totA, sizeof(double4) * (size_t)Ptot));
    LR_CHECK(ensure(lw.partB, sizeof(double) * 3 * (size_t)ngroup * (size_t)n));
    if (krec) {
        const hipError_t e = launch_long_krec(c, lw, Ptot, s, err);
        if (e != hipSuccess) return e;
    }
    const KRec *kr = lw.kr.as<KRec>();
    dim3 g((unsigned)((nchunk + kLaneWaves - 1) / kLaneWaves), (unsigned)ngroup);
    if (which & kLongPassAS) {
        if (c.precision == QEM_PREC_MIXED)
            k_lane_A<QEM_PREC_MIXED><<<g, 64 * kLaneWaves, 0, s>>>(
                kr, Ptot, lw.satom.as<double4>(), n, chk, nchunk, lw.partA.as<double4>());
        else
            k_lane_A<QEM_PREC_FP64><<<g, 64 * kLaneWaves, 0, s>>>(
                kr, Ptot, lw.satom.as<double4>(), n, chk, nchunk, lw.partA.as<double4>());
        LR_CHECK(hipGetLastError());
        k_long_S<<<(unsigned)Ptot, CH, 0, s>>>(lw.partA.as<double4>(), Ptot, nchunk,
                                               lw.offA.as<double4>(), lw.totA.as<double4>());
        LR_CHECK(hipGetLastError());
    }
    if (!(which & kLongPassB)) return hipSuccess;
    // the self term differs from the non-sort sum's own term unless both
    // parts are on or both off
    const int own_term = ((c.parts & QEM_LONG_NONSORT) != 0) != ((c.parts & QEM_LONG_SELF) != 0);
    if (c.precision == QEM_PREC_MIXED)
        k_lane_B<QEM_PREC_MIXED><<<g, 64 * kLaneWaves, 0, s>>>(
            kr, Ptot, lw.satom.as<double4>(), n, chk, nchunk, sl, lw.partA.as<double4>(),
            lw.offA.as<double4>(), lw.totA.as<double4>(), own_term, lw.partB.as<double>(),
            nullptr);
    else
        k_lane_B<QEM_PREC_FP64><<<g, 64 * kLaneWaves, 0, s>>>(
            kr, Ptot, lw.satom.as<double4>(), n, chk, nchunk, sl, lw.partA.as<double4>(),
            lw.offA.as<double4>(), lw.totA.as<double4>(), own_term, lw.partB.as<double>(),
            nullptr);
    LR_CHECK(hipGetLastError());
    return hipSuccess;
}

hipError_t launch_long_forces_energy_B(const LongConsts &c, LongWork &lw, int64_t n,
                                       int64_t Ptot, hipStream_t s, std::string &err)
{
    if (n <= 0 || Ptot <= 0) return hipSuccess;
    // the potential from the force pass needs the non-sort row constants
    if (!(c.parts & QEM_LONG_NONSORT)) {
        const hipError_t e = launch_long_passes(c, lw, n, Ptot, s, err, false, kLongPassB);
        return e != hipSuccess ? e : launch_long_energy_B(c, lw, n, Ptot, s, err);
    }
    const LaneGrid lg = lane_grid(n, Ptot);
    LR_CHECK(ensure(lw.partB, sizeof(double) * 3 * (size_t)lg.ngroup * (size_t)n));
    LR_CHECK(ensure(lw.partE, sizeof(double) * (size_t)lg.ngroup * (size_t)n));
    LR_CHECK(ensure(lw.q0part, sizeof(double2) * (size_t)Ptot));
    const KRec *kr = lw.kr.as<KRec>();
    const dim3 g((unsigned)((lg.nchunk + kLaneWaves - 1) / kLaneWaves), (unsigned)lg.ngroup);
    const int own_term = ((c.parts & QEM_LONG_NONSORT) != 0) != ((c.parts & QEM_LONG_SELF) != 0);
    if (c.precision == QEM_PREC_MIXED)
        k_lane_B<QEM_PREC_MIXED, true><<<g, 64 * kLaneWaves, 0, s>>>(
            kr, Ptot, lw.satom.as<double4>(), n, lg.chk, lg.nchunk, lg.split_last,
            lw.partA.as<double4>(), lw.offA.as<double4>(), lw.totA.as<double4>(), own_term,
            lw.partB.as<double>(), lw.partE.as<double>());
    else
        k_lane_B<QEM_PREC_FP64, true><<<g, 64 * kLaneWaves, 0, s>>>(
            kr, Ptot, lw.satom.as<double4>(), n, lg.chk, lg.nchunk, lg.split_last,
            lw.partA.as<double4>(), lw.offA.as<double4>(), lw.totA.as<double4>(), own_term,
            lw.partB.as<double>(), lw.partE.as<double>());
    LR_CHECK(hipGetLastError());
    k_long_dgamma<<<nblk(Ptot, 256), 256, 0, s>>>(c, lw.kv.as<double4>(), Ptot,
                                                  lw.totA.as<double4>(), lw.q0part.as<double2>());
    LR_CHECK(hipGetLastError());
    return hipSuccess;
}

hipError_t launch_long_sum_slots(LongWork &lw, int64_t n, int64_t Ptot, const int32_t *slot_of,
                                 const double *sout, const double *addz, double *d_f,
                                 hipStream_t s, std::string &err)
{
    if (n <= 0) return hipSuccess;
    const int ngroup = Ptot > 0 ? lane_grid(n, Ptot).ngroup : 0;
    k_long_R_slots<<<nblk(n, 256), 256, 0, s>>>(
        lw.partB.as<double>(), ngroup, n, lw.perm.as<int32_t>(),
        lw.k0_on ? lw.k0out.as<double2>() : nullptr, slot_of, sout, addz, d_f);
    LR_CHECK(hipGetLastError());
    return hipSuccess;
}

hipError_t launch_long_sum(LongWork &lw, int64_t n, int64_t Ptot, double *d_f, int accumulate,
                           hipStream_t s, std::string &err)
{
    if (n <= 0) return hipSuccess;
    if (Ptot <= 0 && !lw.k0_on) {
        // nothing to add; an overwrite still has to clear the output
        if (!accumulate) {
            k_fill_f64<<<nblk(3 * n, 256), 256, 0, s>>>(d_f, 3 * n, 0.0);
            LR_CHECK(hipGetLastError());
        }
        return hipSuccess;
    }
    const int ngroup = Ptot > 0 ? lane_grid(n, Ptot).ngroup : 0;
    k_long_R<<<nblk(n, 256), 256, 0, s>>>(lw.partB.as<double>(), ngroup, n, lw.perm.as<int32_t>(),
                                          lw.k0_on ? lw.k0out.as<double2>() : nullptr,
                                          accumulate, d_f);
    LR_CHECK(hipGetLastError());
    return hipSuccess;
}

hipError_t launch_long(const LongConsts &c, LongWork &lw, int64_t n, int64_t Ptot,
                       double *d_f, int accumulate, hipStream_t s, std::string &err)
{
    hipError_t e = launch_long_passes(c, lw, n, Ptot, s, err);
    if (e != hipSuccess) return e;
    return launch_long_sum(lw, n, Ptot, d_f, accumulate, s, err);
}

hipError_t launch_long_variance(const LongConsts &c, LongWork &lw, int64_t n, int64_t Ptot,
                               int64_t P, double *d_var, hipStream_t s, std::string &err)
{
    if (n <= 0) return hipSuccess;
    if (Ptot <= 0 || P < 2) {
        err = "the batch variance needs at least two k vectors";
        return hipErrorInvalidValue;
    }
    const int ngroup = (int)((Ptot + 63) / 64), chk = kVarChunk;
    const int nchunk = (int)((n + chk - 1) / chk);
    LR_CHECK(ensure(lw.kr, sizeof(double) * 16 * (size_t)Ptot));
    LR_CHECK(ensure(lw.partA, sizeof(double4) * (size_t)Ptot * nchunk));
    LR_CHECK(ensure(lw.offA, sizeof(double4) * (size_t)Ptot * nchunk));
    LR_CHECK(ensure(lw.totA, sizeof(double4) * (size_t)Ptot));
    LR_CHECK(ensure(lw.partB, sizeof(double4) * (size_t)ngroup * (size_t)n));
    KRec *kr = lw.kr.as<KRec>();
    k_make_krec<<<nblk(Ptot, 256), 256, 0, s>>>(c, lw.kv.as<double4>(), Ptot, kr);
    LR_CHECK(hipGetLastError());
    const dim3 g((unsigned)((nchunk + kLaneWaves - 1) / kLaneWaves), (unsigned)ngroup);
    const double4 *sa = lw.satom.as<double4>();
    if (c.precision == QEM_PREC_MIXED)
        k_lane_A<QEM_PREC_MIXED><<<g, 64 * kLaneWaves, 0, s>>>(kr, Ptot, sa, n, chk, nchunk,
                                                               lw.partA.as<double4>());
    else
        k_lane_A<QEM_PREC_FP64><<<g, 64 * kLaneWaves, 0, s>>>(kr, Ptot, sa, n, chk, nchunk,
                                                              lw.partA.as<double4>());
    LR_CHECK(hipGetLastError());
    k_long_S<<<(unsigned)Ptot, CH, 0, s>>>(lw.partA.as<double4>(), Ptot, nchunk,
                                           lw.offA.as<double4>(), lw.totA.as<double4>());
    LR_CHECK(hipGetLastError());
    double4 *pv = lw.partB.as<double4>();
    if (c.precision == QEM_PREC_MIXED)
        k_lane_V<QEM_PREC_MIXED><<<g, 64 * kLaneWaves, 0, s>>>(
            kr, Ptot, sa, n, chk, nchunk, lw.offA.as<double4>(), lw.totA.as<double4>(), pv);
    else
        k_lane_V<QEM_PREC_FP64><<<g, 64 * kLaneWaves, 0, s>>>(
            kr, Ptot, sa, n, chk, nchunk, lw.offA.as<double4>(), lw.totA.as<double4>(), pv);
    LR_CHECK(hipGetLastError());
    k_var_R<<<nblk(n, 256), 256, 0, s>>>(pv, ngroup, n, lw.perm.as<int32_t>(), (double)P, d_var);
    LR_CHECK(hipGetLastError());
    return hipSuccess;
}

hipError_t launch_long_energy_passes(const LongConsts &c, LongWork &lw, int64_t n,
                                     int64_t Ptot, hipStream_t s, std::string &err,
                                     const std::function<hipError_t()> *between)
{
    if (n <= 0 || Ptot <= 0) return hipSuccess;
    const LaneGrid lg = lane_grid(n, Ptot);
    const size_t pc = sizeof(double4) * (size_t)Ptot * lg.nchunk;
    LR_CHECK(ensure(lw.kr, sizeof(double) * 16 * (size_t)Ptot));
    LR_CHECK(ensure(lw.partA, pc));
    LR_CHECK(ensure(lw.offA, pc));
    LR_CHECK(ensure(lw.totA, sizeof(double4) * (size_t)Ptot));
    LR_CHECK(ensure(lw.partE, sizeof(double) * (size_t)lg.ngroup * (size_t)n));
    LR_CHECK(ensure(lw.q0part, sizeof(double2) * (size_t)Ptot));
    KRec *kr = lw.kr.as<KRec>();
    k_make_krec<<<nblk(Ptot, 256), 256, 0, s>>>(c, lw.kv.as<double4>(), Ptot, kr);
    LR_CHECK(hipGetLastError());
    const dim3 g((unsigned)((lg.nchunk + kLaneWaves - 1) / kLaneWaves), (unsigned)lg.ngroup);
    const double4 *sa = lw.satom.as<double4>();
    if (c.precision == QEM_PREC_MIXED)
        k_lane_A<QEM_PREC_MIXED><<<g, 64 * kLaneWaves, 0, s>>>(kr, Ptot, sa, n, lg.chk,
                                                               lg.nchunk, lw.partA.as<double4>());
    else
        k_lane_A<QEM_PREC_FP64><<<g, 64 * kLaneWaves, 0, s>>>(kr, Ptot, sa, n, lg.chk,
                                                              lg.nchunk, lw.partA.as<double4>());
    LR_CHECK(hipGetLastError());
    k_long_S<<<(unsigned)Ptot, CH, 0, s>>>(lw.partA.as<double4>(), Ptot, lg.nchunk,
                                           lw.offA.as<double4>(), lw.totA.as<double4>());
    LR_CHECK(hipGetLastError());
    if (between) LR_CHECK((*between)());
    return launch_long_energy_B(c, lw, n, Ptot, s, err);
}

hipError_t launch_long_energy_B(const LongConsts &c, LongWork &lw, int64_t n, int64_t Ptot,
                                hipStream_t s, std::string &err)
{
    if (n <= 0 || Ptot <= 0) return hipSuccess;
    const LaneGrid lg = lane_grid(n, Ptot);
    LR_CHECK(ensure(lw.partE, sizeof(double) * (size_t)lg.ngroup * (size_t)n));
    LR_CHECK(ensure(lw.q0part, sizeof(double2) * (size_t)Ptot));
    const KRec *kr = lw.kr.as<KRec>();
    const dim3 g((unsigned)((lg.nchunk + kLaneWaves - 1) / kLaneWaves), (unsigned)lg.ngroup);
    const double4 *sa = lw.satom.as<double4>();
    if (c.precision == QEM_PREC_MIXED)
        k_lane_B_energy<QEM_PREC_MIXED><<<g, 64 * kLaneWaves, 0, s>>>(
            c, kr, Ptot, sa, n, lg.chk, lg.nchunk, lw.partA.as<double4>(), lw.offA.as<double4>(),
            lw.totA.as<double4>(), lw.partE.as<double>());
    else
        k_lane_B_energy<QEM_PREC_FP64><<<g, 64 * kLaneWaves, 0, s>>>(
            c, kr, Ptot, sa, n, lg.chk, lg.nchunk, lw.partA.as<double4>(), lw.offA.as<double4>(),
            lw.totA.as<double4>(), lw.partE.as<double>());
    LR_CHECK(hipGetLastError());
    k_long_dgamma<<<nblk(Ptot, 256), 256, 0, s>>>(c, lw.kv.as<double4>(), Ptot,
                                                  lw.totA.as<double4>(), lw.q0part.as<double2>());
    LR_CHECK(hipGetLastError());
    return hipSuccess;
}

hipError_t launch_long_energy_sum(LongWork &lw, int64_t n, int64_t Ptot, double *d_phi,
                                  int accumulate, double *h_q0, hipStream_t s,
                                  std::string &err)
{
    if (n <= 0 || (Ptot <= 0 && !lw.k0_on)) return hipSuccess;
    const int ngroup = Ptot > 0 ? lane_grid(n, Ptot).ngroup : 0;
    k_long_Rw<<<nblk(n, 256), 256, 0, s>>>(lw.partE.as<double>(), ngroup, n, 1,
                                           lw.perm.as<int32_t>(),
                                           lw.k0_on ? lw.k0out.as<double2>() : nullptr,
                                           accumulate, 1.0, d_phi);
    LR_CHECK(hipGetLastError());
    if (Ptot > 0)
        LR_CHECK(hipMemcpyAsync(h_q0, lw.q0part.p, sizeof(double2) * (size_t)Ptot,
                                hipMemcpyDeviceToHost, s));
    return hipSuccess;
}

hipError_t launch_long_energy(const LongConsts &c, LongWork &lw, int64_t n, int64_t Ptot,
                              double *d_phi, int accumulate, double *h_dgamma, hipStream_t s,
                              std::string &err)
{
    h_dgamma[0] = h_dgamma[1] = 0.0;
    if (n <= 0 || (Ptot <= 0 && !lw.k0_on)) return hipSuccess;
    std::vector<double> h(2 * (size_t)std::max<int64_t>(Ptot, 0));
    hipError_t e = lw.k0_on ? launch_long_k0(c, lw, n, s, err) : hipSuccess;
    if (e == hipSuccess) e = launch_long_energy_passes(c, lw, n, Ptot, s, err);
    if (e == hipSuccess)
        e = launch_long_energy_sum(lw, n, Ptot, d_phi, accumulate, h.data(), s, err);
    if (e != hipSuccess) return e;
    LR_CHECK(hipStreamSynchronize(s));
    for (int64_t p = 0; p < Ptot; ++p) {
        h_dgamma[0] += h[2 * (size_t)p];
        h_dgamma[1] += h[2 * (size_t)p + 1];
    }
    return hipSuccess;
}

hipError_t launch_long_tangent_sum(LongWork &lw, int64_t n, int64_t Ptot, double scale,
                                   double *d_out, int accumulate, hipStream_t s,
                                   std::string &err)
{
    if (n <= 0 || Ptot <= 0) return hipSuccess;
    const LaneGrid lg = lane_grid(n, Ptot);
    k_long_Rw<<<nblk(n, 256), 256, 0, s>>>(lw.partB.as<double>(), lg.ngroup, n, 3,
                                           lw.perm.as<int32_t>(), nullptr, accumulate, scale,
                                           d_out);
    LR_CHECK(hipGetLastError());
    return hipSuccess;
}

hipError_t launch_long_tangent(const LongConsts &c, LongWork &lw, int64_t n, int64_t Ptot,
                               const double *d_v, double scale, double *d_out, int accumulate,
                               hipStream_t s, std::string &err, bool with_sum,
                               const std::function<hipError_t()> *between)
{
    if (n <= 0 || Ptot <= 0) return hipSuccess;
    const LaneGrid lg = lane_grid(n, Ptot);
    const size_t nd = sizeof(double) * (size_t)n;
    LR_CHECK(ensure(lw.svx, nd));
    LR_CHECK(ensure(lw.svy, nd));
    LR_CHECK(ensure(lw.svz, nd));
    LR_CHECK(ensure(lw.kr, sizeof(double) * 16 * (size_t)Ptot));
    const size_t pc = sizeof(double4) * (size_t)Ptot * lg.nchunk;
    LR_CHECK(ensure(lw.partA, pc));
    LR_CHECK(ensure(lw.offA, pc));
    LR_CHECK(ensure(lw.totA, sizeof(double4) * (size_t)Ptot));
    LR_CHECK(ensure(lw.partT, pc));
    LR_CHECK(ensure(lw.offT, pc));
    LR_CHECK(ensure(lw.totT, sizeof(double4) * (size_t)Ptot));
    LR_CHECK(ensure(lw.partB, sizeof(double) * 3 * (size_t)lg.ngroup * (size_t)n));
    k_gather_sorted3<<<nblk(n, 256), 256, 0, s>>>(d_v, lw.perm.as<int32_t>(), n, lw.svx.as<double>(),
                                                  lw.svy.as<double>(), lw.svz.as<double>());
    KRec *kr = lw.kr.as<KRec>();
    k_make_krec<<<nblk(Ptot, 256), 256, 0, s>>>(c, lw.kv.as<double4>(), Ptot, kr);
    LR_CHECK(hipGetLastError());
    const dim3 g((unsigned)((lg.nchunk + kLaneWaves - 1) / kLaneWaves), (unsigned)lg.ngroup);
    const double4 *sa = lw.satom.as<double4>();
    if (c.precision == QEM_PREC_MIXED)
        k_lane_A_t<QEM_PREC_MIXED><<<g, 64 * kLaneWaves, 0, s>>>(
            kr, Ptot, sa, lw.svx.as<double>(), lw.svy.as<double>(), lw.svz.as<double>(), n,
            lg.chk, lg.nchunk, lw.partA.as<double4>(), lw.partT.as<double4>());
    else
        k_lane_A_t<QEM_PREC_FP64><<<g, 64 * kLaneWaves, 0, s>>>(
            kr, Ptot, sa, lw.svx.as<double>(), lw.svy.as<double>(), lw.svz.as<double>(), n,
            lg.chk, lg.nchunk, lw.partA.as<double4>(), lw.partT.as<double4>());
    LR_CHECK(hipGetLastError());
    k_long_S<<<(unsigned)Ptot, CH, 0, s>>>(lw.partA.as<double4>(), Ptot, lg.nchunk,
                                           lw.offA.as<double4>(), lw.totA.as<double4>());
    k_long_S<<<(unsigned)Ptot, CH, 0, s>>>(lw.partT.as<double4>(), Ptot, lg.nchunk,
                                           lw.offT.as<double4>(), lw.totT.as<double4>());
    LR_CHECK(hipGetLastError());
    if (between) LR_CHECK((*between)());
    // the self term differs from the non-sort sum's own term unless both
    // parts are on or both off
    const int own_term = ((c.parts & QEM_LONG_NONSORT) != 0) != ((c.parts & QEM_LONG_SELF) != 0);
    if (c.precision == QEM_PREC_MIXED)
        k_lane_B_t<QEM_PREC_MIXED><<<g, 64 * kLaneWaves, 0, s>>>(
            kr, Ptot, sa, lw.svx.as<double>(), lw.svy.as<double>(), lw.svz.as<double>(), n,
            lg.chk, lg.nchunk, lw.partA.as<double4>(), lw.offA.as<double4>(),
            lw.totA.as<double4>(), lw.partT.as<double4>(), lw.offT.as<double4>(),
            lw.totT.as<double4>(), own_term, lw.partB.as<double>());
    else
        k_lane_B_t<QEM_PREC_FP64><<<g, 64 * kLaneWaves, 0, s>>>(
            kr, Ptot, sa, lw.svx.as<double>(), lw.svy.as<double>(), lw.svz.as<double>(), n,
            lg.chk, lg.nchunk, lw.partA.as<double4>(), lw.offA.as<double4>(),
            lw.totA.as<double4>(), lw.partT.as<double4>(), lw.offT.as<double4>(),
            lw.totT.as<double4>(), own_term, lw.partB.as<double>());
    LR_CHECK(hipGetLastError());
    if (!with_sum) return hipSuccess;
    return launch_long_tangent_sum(lw, n, Ptot, scale, d_out, accumulate, s, err);
}

}  // namespace qem
