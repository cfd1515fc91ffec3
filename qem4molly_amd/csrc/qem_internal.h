// Internal declarations shared by the libqem_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/qem.h"

namespace qem {

// Grow-only device buffer owned by a context.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    template <class T> T *as() const { return static_cast<T *>(p); }
};
hipError_t ensure(DevBuf &b, size_t bytes);
void release(DevBuf &b);

// Kernel template precision for FP64 contexts whose Bessel argument range
// exceeds the FP64 table: FP64 arithmetic with the device library's j0/j1.
constexpr int kPrecFP64Lib = 2;

// Constants of one short-range launch (kernel argument, passed by value).
struct ShortConsts {
    double Lx, Ly, Lz;
    double rc2;          // r_cutoff^2
    double alpha;
    double g1, g2, g12;  // gamma_1, gamma_2, gamma_1*gamma_2
    double ln_acc;       // log(accuracy)
    double inv2pieps;    // 1 / (2 pi eps_0)
    uint32_t sem;        // QEM_SEM_* bits
    int32_t Nt;
    int32_t single;      // point-charge part only (F_short_ij single=true)
    int32_t jt_n;        // Bessel table intervals (mixed precision)
    float jt_dx, jt_inv_dx;  // interval width and its inverse
    const float4 *ncon;  // device node constants (mixed precision), or null
    const float *nconi;  // per node 1 / (1 + x_n) (the fused energy sums), or null
    const double2 *ncon64;   // FP64 node constants: (1 + x_n, log2(w_n (1 + x_n)))
    double d_c2, d_rG, d_hinvLz;  // FP64: -2 Lz log2(e), cg / d_c2^2 with
                                  // cg = -log2(e) / (4 alpha), 1 / (2 Lz)
    int32_t jd_n;        // FP64 Bessel table intervals (0: device-library j0/j1)
    double jd_inv_dx;    // its inverse interval width
    const double *jdtab; // FP64 table in device memory (staged in LDS by the kernels)
    double inv_Lx, inv_Ly;   // 1/Lx, 1/Ly
    float f_2Lz, f_nlnacc;   // FP32: 2 Lz, -log(accuracy)
    float f_2alpha, f_m4alnacc;  // FP32: 2 alpha, -4 alpha log(accuracy)
    float f_c2, f_rG, f_hinvLz;  // FP32 node-loop constants: -2 Lz log2(e),
                                 // cg / f_c2^2 with cg = -log2(e) / (4 alpha), 1/(2 Lz)
    // pole splitting (gamma_1 gamma_2 > 1 with QEM_SEM_POLE_SPLIT)
    int32_t pole;             // 1: the pair and self kernels use the pole rule
    int32_t pm[2];            // nodes of the point and the gaussian set
    const double2 *pnode;     // (k_n, c_n) per node, point set then gaussian set
    const float2 *pnodef;     // the same in FP32
    double pk0, pdel;         // k0 and the central-difference step 0.01 eps
    double pWx[2], pWm[2];    // per set: sum of w/(2 Lz (k - k0)) over the side
                              // segments, sum of w over the middle segment
    double ppf[2][3];         // per set: its factor at k0 - del, k0, k0 + del
    double pdpf[2];           // per set: the factor's k-derivative at k0
};

// The pole rule of one set: the node layout and its constants (host side).
struct PoleSet {
    std::vector<double> k, c;  // nodes and node constants
    double Wx = 0.0, Wm = 0.0;
    int n1 = 0, n3 = 0;
    double kf = 0.0;
};
// Builds the rule for one set (gauss = the gaussian set): k0 = ln(g12)/(2 Lz),
// kf = max(kf_base, 2.1 k0), N1 = floor(N_t k0 / kf) nodes on [0, k0 - eps]
// and [k0 + eps, 2 k0], 10 on [k0 - eps, k0 + eps], N_t - 2 N1 on [2 k0, kf].
// Returns false when the rule does not exist.
bool make_pole_set(const qem_params &p, double eps, bool gauss, PoleSet &out);
// The half-width used for p (p.pole_eps, or k0 / 20 when that is <= 0).
double pole_eps_of(const qem_params &p);
// Whether p's short range uses the pole rule.
inline bool pole_active(const qem_params &p)
{
    return (p.semantics & QEM_SEM_POLE_SPLIT) && p.gamma_1 * p.gamma_2 > 1.0;
}

// Gauss-Legendre rule on [-1,1] (ascending) plus its FP32 copy.
struct GLRule {
    int n = 0;
    std::vector<double> x, w;
};
GLRule make_gl_rule(int n);

// Per interval [m dx, (m+1) dx): coefficients c0..cD of a polynomial in
// t in [0, 1) fitted to J_order(m dx + t dx) (order 0 or 1) at Chebyshev
// points (least squares in long double), D + 1 values per interval; two
// padding intervals past x_max.  Both paths tabulate J0 and take J1 from the
// derivative: the mixed path with D = 3 (one float4), the FP64 path D = 7.
std::vector<long double> fit_bessel_table(double x_max, double dx, int D, int order,
                                          int *n_intervals);
// The FP64 table: J0 interpolated at D + 1 Chebyshev points per interval.
std::vector<long double> interp_j0_table(double x_max, double dx, int D, int *n_intervals);

// ---- sorting and scanning (radix_sort.hip) --------------------------------
// Stable LSD radix sorts (8-bit digits) and an exclusive scan, hand-written.
// Each takes a scratch buffer of the given size whose first 256 bytes hold
// the launches' completion counters: zero them (radix_sort_init) whenever
// the buffer is new; every sort and scan leaves them at zero.
size_t radix_sort_scratch_bytes(int64_t n, int key_bytes);
hipError_t radix_sort_init(void *scratch, hipStream_t s);
// ensure() of a sort/scan scratch buffer, with the counters zeroed when the
// buffer is (re)allocated
hipError_t ensure_sort_scratch(DevBuf &b, size_t bytes, hipStream_t s);
// sortperm of the z coordinates (ascending, ties by atom index): perm[r] is
// the atom of rank r, satom[r] its (x, y, z, q).  Scratch:
// radix_sort_scratch_bytes(n, 8).
hipError_t radix_sort_z(const double *d_xyz, const double *d_q, int64_t n, int32_t *perm,
                        double4 *satom, void *scratch, hipStream_t s);
// keys (< 2^bits) sorted stably: keys_out the sorted keys, idx_out their
// input positions.  Scratch: radix_sort_scratch_bytes(n, 4).
hipError_t radix_sort_u32(const uint32_t *keys, int64_t n, int bits, uint32_t *keys_out,
                          int32_t *idx_out, void *scratch, hipStream_t s);
// out[i] = sum of in[0 .. i-1] for m int64 values.  Scratch:
// scan_scratch_bytes(m).
size_t scan_scratch_bytes(int64_t m);
hipError_t scan_exclusive_i64(const int64_t *in, int64_t *out, int64_t m, void *scratch,
                              hipStream_t s);

// ---- neighbour list (neighbor.hip) ------------------------------------
struct NeighborList {
    int64_t n = -1;       // atom count the list was built for
    int64_t npairs = 0;
    int64_t builds = 0;
    DevBuf pi, pj;        // int32 pair endpoints
    // per atom a: pairs with a as first endpoint are ids
    // i_idx[i_start[a] .. i_end[a]) (or the ids themselves when !has_i_idx),
    // pairs with a as second endpoint are j_idx[j_start[a] .. j_end[a])
    bool has_i_idx = false;
    DevBuf i_start, i_end, i_idx;
    DevBuf j_start, j_end, j_idx;
    // per pair p: its place in the second-endpoint grouping (jpos[j_idx[u]]
    // = u), where the pair kernels write the record of its second atom
    DevBuf jpos;
    // Endpoints, i_*/j_* and per-atom kernel rows are in slot numbering:
    // for a built list the slots are the atoms in cell order (slot s is
    // atom atom_idx_sorted[s]), so pair neighbours are near in memory; for
    // an installed list slot = atom (slot_order false).
    bool slot_order = false;
    DevBuf slot_of;       // built lists: the slot of each atom (inverse of atom_idx_sorted)
    DevBuf sxyz, sq, sv;  // per call: positions, charges, directions in slot order
    DevBuf sout;          // per call: per-slot results before the scatter
    DevBuf sout2;         // per call: per-slot potentials (phi, psi) of the fused call
    // scratch
    DevBuf cell_key, cell_key_sorted, atom_idx_sorted;
    DevBuf cell_start, cell_end, counts, row_ptr, tmp;
};
// Slot-ordered copies: rows of w doubles, dst[s] = src[atom_idx_sorted[s]]
// (to slots) and dst[atom_idx_sorted[s]] = src[s] (from slots).  With
// addz (w == 3, one double per atom, atom order) the from-slots copy writes
// src[s].z + addz[atom] as the z component.
hipError_t launch_rows_to_slots(const NeighborList &nl, int64_t n, int w, const double *src,
                                double *dst, hipStream_t s);
hipError_t launch_rows_from_slots(const NeighborList &nl, int64_t n, int w, const double *src,
                                  double *dst, hipStream_t s, const double *addz = nullptr);
// Positions (N x 3) and charges (N) into slot order in one pass.
hipError_t launch_atoms_to_slots(const NeighborList &nl, int64_t n, const double *d_xyz,
                                 const double *d_q, double *sxyz, double *sq, hipStream_t s);
// Builds the list for device coordinates; synchronises the stream once
// (to learn the pair count).  Returns hipSuccess or the first error.
hipError_t build_neighbor_list(NeighborList &nl, int64_t n, const double *d_xyz,
                               double Lx, double Ly, double rc, hipStream_t s,
                               std::string &err);
// Installs a caller-given pair list (device copies of pairs, i < j not required).
hipError_t install_pairs(NeighborList &nl, int64_t n, const int32_t *h_pairs,
                         int64_t npairs, hipStream_t s, std::string &err);

// ---- short range (short_range.hip) ------------------------------------
struct ShortWork {
    DevBuf glx, glw;      // double nodes / weights
    DevBuf glxf, glwf;    // float nodes / weights
    DevBuf ncon;          // float4 per node: (1 + x_n, (1 + x_n)^2, log2 w_n,
                          //                   log2(w_n (1 + x_n)))
    DevBuf nconi;         // float per node: 1 / (1 + x_n)
    DevBuf ncon64;        // double2 per node: (1 + x_n, log2(w_n (1 + x_n)))
    DevBuf jtab;          // one float4 (cubic coefficients) per interval
    int jt_n = 0;
    float jt_dx = 0.f;
    DevBuf jdtab;         // FP64: 9 doubles per interval (degree-7 J0 coefficients, pad)
    int jd_n = 0;
    double jd_dx = 0.0;
    DevBuf fpair;         // endpoint records, pair_rec_bytes(npairs)
    DevBuf pnode, pnodef; // pole-splitting nodes (double2, float2)
};
// Endpoint records.  A pair kernel over a list writes two records per pair
// p, one per endpoint, each holding what that atom adds up: rec[p] for the
// first atom and rec[npairs + jpos[p]] for the second.  An atom's records
// are then two contiguous runs, [i_start, i_end) of the first half (through
// i_idx for an installed list) and [j_start, j_end) of the second, which
// the gathers read in order.  Bytes for the records of npairs pairs, rec
// bytes each (12 for float3, 16 for float4, 32 for double4): two per pair.
inline size_t pair_rec_bytes(int64_t npairs, size_t rec)
{
    return 2 * rec * (size_t)(npairs > 0 ? npairs : 1);
}
// Evaluates every pair of the list into endpoint records in d_fpair: first
// atom (fx, fy, fz_i), second atom (-fx, -fy, fz_j) (float3 in mixed
// precision, double4 with w = 0 otherwise).
hipError_t launch_pair_forces(const ShortConsts &c, int precision,
                              const ShortWork &ws, const int32_t *d_pi,
                              const int32_t *d_pj, const int32_t *d_jpos, int64_t npairs,
                              const double *d_xyz, const double *d_q,
                              double *d_fpair, hipStream_t s);
// Mixed precision: the force records of launch_pair_forces (float4), whose
// w carries the pair potential times the other endpoint's charge (first
// atom q_j G, second atom q_i G), from the same node loops; and per block of
// the launch (pair_fe_blocks of them) the pairs' sum of
// (q_i q_j dG/dgamma_1, q_i q_j dG/dgamma_2) in d_gpart (2 doubles each).
int pair_fe_blocks(int64_t npairs);
// With d_sums (tan_sums_bytes(npairs)) the kernel also keeps every pair's
// force-tangent node sums there, for launch_short_tangent's saved path.
hipError_t launch_pair_forces_energy(const ShortConsts &c, const ShortWork &ws,
                                     const int32_t *d_pi, const int32_t *d_pj,
                                     const int32_t *d_jpos, int64_t npairs,
                                     const double *d_xyz, const double *d_q, double *d_fpair,
                                     double *d_gpart, hipStream_t s, float *d_sums = nullptr);
// Bytes of the saved tangent node sums of npairs pairs (224 per pair:
// twelve float4 and four float2 arrays).
constexpr int kTanSumBytesPerPair = 12 * 16 + 4 * 8;
inline size_t tan_sums_bytes(int64_t npairs)
{
    return (size_t)kTanSumBytesPerPair * (size_t)(npairs > 0 ? npairs : 1);
}
// Pair records of the force path (launch_pair_forces -> launch_gather_self):
// float3 (12 bytes) for the mixed-precision list kernel, double4 (32 bytes)
// otherwise (FP64, and the pole-splitting kernels).
inline bool pair_records_f32(const ShortConsts &c, int precision)
{
    return precision == QEM_PREC_MIXED && !c.pole;
}
inline int force_rec_bytes(const ShortConsts &c, int precision)
{
    return pair_records_f32(c, precision) ? 12 : 32;
}
// Per atom: gather pair forces (i side and j side) and add the self term.
hipError_t launch_gather_self(const ShortConsts &c, int precision,
                              const ShortWork &ws, const NeighborList &nl,
                              int64_t n, const double *d_xyz,
                              const double *d_q, const double *d_fpair,
                              double *d_f, hipStream_t s);
// The pole-splitting variants (pole.hip): same outputs as the kernels they
// replace when ShortConsts::pole is set.
hipError_t launch_pair_pole(const ShortConsts &c, int precision, const ShortWork &ws,
                            const int32_t *d_pi, const int32_t *d_pj, const int32_t *d_jpos,
                            int64_t npairs,
                            const double *d_xyz, const double *d_q, double *d_fpair,
                            hipStream_t s);
hipError_t launch_pair_direct_pole(const ShortConsts &c, int precision, const ShortWork &ws,
                                   int64_t m, const double *d_ri, const double *d_rj,
                                   const double *d_qi, const double *d_qj, double *d_fpair,
                                   hipStream_t s);
hipError_t launch_self_pole(const ShortConsts &c, int precision, int64_t n, const double *d_xyz,
                            const double *d_q, int zero_xy, double *d_f, hipStream_t s);
// Self term only (F_short_i), n atoms.
hipError_t launch_self_only(const ShortConsts &c, int precision,
                            const ShortWork &ws, int64_t n,
                            const double *d_xyz, const double *d_q,
                            double *d_f, hipStream_t s);
// Independent pairs (F_short_ij): positions/charges given per pair.
hipError_t launch_pair_direct(const ShortConsts &c, int precision,
                              const ShortWork &ws, int64_t m,
                              const double *d_ri, const double *d_rj,
                              const double *d_qi, const double *d_qj,
                              double *d_fpair, hipStream_t s);

// ---- long range (long_range.hip) ---------------------------------------
struct LongConsts {
    double Lz, g1, g2, g12;
    double inv_area_eps;  // 1 / (Lx Ly eps_0)
    double weight;        // batch weight (sum_K / P), or < 0 for exp(-k^2/4a)
    double inv4alpha;
    uint32_t parts;       // QEM_LONG_* bits
    int32_t precision;
};
struct LongWork {
    DevBuf perm, tmp;        // z order (rank -> atom); the sort's scratch
    DevBuf satom;            // atoms in z order, (x, y, z, q) per atom
    DevBuf kr;               // per k row derived constants of the force passes
    DevBuf kv;               // P x 4 doubles (kx, ky, k, weight)
    DevBuf partA;            // [P][nchunk][4] chunk sums
    DevBuf offA;             // [P][nchunk][4] chunk offsets
    DevBuf totA;             // [P][4] totals
    DevBuf q0part;           // per k row: dE_k/dgamma_1,2 (energy)
    DevBuf partB;            // [ntile][n][3]
    DevBuf partE;            // [ntile][n] (energy: the potential of each atom)
    DevBuf svx, svy, svz;    // tangent directions in z order
    DevBuf partT, offT, totT;  // tangent chunk sums / offsets / totals
    DevBuf k0tile;           // k = 0 term: tile sums of (q, q z), then their prefix
    DevBuf k0out;            // [n] (F0_z, phi0) of each atom in z order
    bool k0_on = false;      // the staged batch carries the k = 0 term
    static constexpr int kSlots = 4;   // pinned staging ring for k vectors
    void *h_kv[kSlots] = {nullptr, nullptr, nullptr, nullptr};
    size_t h_kv_cap[kSlots] = {0, 0, 0, 0};
    hipEvent_t kv_ev[kSlots] = {nullptr, nullptr, nullptr, nullptr};
    int kv_slot = 0;
};
// Bytes each LongWork buffer needs for a force call over n atoms and Ptot k
// rows (k != 0): lw.kv, lw.kr, lw.perm, lw.satom, lw.tmp (the z sort),
// lw.partA and lw.offA (each), lw.totA, lw.partB, and the k = 0 term's
// lw.k0tile and lw.k0out.
struct LongSizes {
    size_t kv, kr, perm, satom, sort, partA, totA, partB, k0tile, k0out;
};
LongSizes long_force_sizes(int64_t n, int64_t Ptot);
// z-sort; fills lw.perm (int32, rank -> atom) and lw.satom (atoms in z order).
hipError_t launch_zsort(LongWork &lw, int64_t n, const double *d_xyz,
                        const double *d_q, hipStream_t s, std::string &err);
// The k = 0 term of the atoms in lw.satom: lw.k0out[r] = (F0_z, phi0) of
// the atom of rank r (force and potential, weight 1 / (2 Lx Ly eps_0)).  The
// sums (launch_long_sum*, launch_long_energy_sum) add it when lw.k0_on.
hipError_t launch_long_k0(const LongConsts &c, LongWork &lw, int64_t n, hipStream_t s,
                          std::string &err);
// Full long-range force for the k vectors already in lw.kv (P rows).
// accumulate != 0 adds into d_f instead of overwriting.
hipError_t launch_long(const LongConsts &c, LongWork &lw, int64_t n,
                       int64_t P, double *d_f, int accumulate, hipStream_t s,
                       std::string &err);
// launch_long in parts: the per-k constants (lw.kr) of the rows in lw.kv;
// the passes up to the per-group partial forces (lw.partB), with the per-k
// constants first unless krec is false (they are then already in lw.kr for
// these rows); and the sum into d_f, which may run on another stream once
// the passes are done.
hipError_t launch_long_krec(const LongConsts &c, LongWork &lw, int64_t P, hipStream_t s,
                            std::string &err);
// which: the chunk sums and offsets (passes A and S, kLongPassAS), the
// per-group partial forces from them (pass B, kLongPassB), or both.
constexpr int kLongPassAS = 1, kLongPassB = 2, kLongPassAll = 3;
hipError_t launch_long_passes(const LongConsts &c, LongWork &lw, int64_t n, int64_t P,
                              hipStream_t s, std::string &err, bool krec = true,
                              int which = kLongPassAll);
// The sum into d_f fused with the short range's output pass: d_f[a] =
// sout[slot_of[a]] (+ addz[a] on z) + the long range of atom a.
hipError_t launch_long_sum_slots(LongWork &lw, int64_t n, int64_t P, const int32_t *slot_of,
                                 const double *sout, const double *addz, double *d_f,
                                 hipStream_t s, std::string &err);
hipError_t launch_long_sum(LongWork &lw, int64_t n, int64_t P, double *d_f, int accumulate,
                           hipStream_t s, std::string &err);

// Batch variance of the random-batch long-range force for the k rows in
// lw.kv (P sampled rows; the k = 0 term is not sampled and is left out):
// d_var[atom] = (P sum_p |X_p|^2 - |sum_p X_p|^2) / (P - 1), X_p the row's
// weighted force on the atom.  P >= 2.
hipError_t launch_long_variance(const LongConsts &c, LongWork &lw, int64_t n, int64_t Ptot,
                               int64_t P, double *d_var, hipStream_t s, std::string &err);
// Energy: phi (dE/dq) of the k rows in lw.kv into d_phi (overwrite or add);
// sum over k of dE_k/dgamma_1,2 into h_dgamma (synchronises the stream).
hipError_t launch_long_energy(const LongConsts &c, LongWork &lw, int64_t n, int64_t Ptot,
                              double *d_phi, int accumulate, double *h_dgamma, hipStream_t s,
                              std::string &err);
// launch_long_energy in two parts that do not synchronise: the passes (with
// the per-k dE_k/dgamma on the device), then, on the same or a joined stream,
// the phi sum into d_phi and the copy of the per-k dE_k/dgamma into h_q0
// (2 Ptot doubles, summed by the caller once s is done).
hipError_t launch_long_energy_passes(const LongConsts &c, LongWork &lw, int64_t n,
                                     int64_t Ptot, hipStream_t s, std::string &err,
                                     const std::function<hipError_t()> *between = nullptr);
hipError_t launch_long_energy_sum(LongWork &lw, int64_t n, int64_t Ptot, double *d_phi,
                                  int accumulate, double *h_q0, hipStream_t s,
                                  std::string &err);
// The potential pass (and the per-k dE_k/dgamma) alone, for k rows whose
// chunk sums, offsets and totals the force passes (launch_long_passes)
// have just formed: the two share passes A and S.
hipError_t launch_long_energy_B(const LongConsts &c, LongWork &lw, int64_t n, int64_t Ptot,
                                hipStream_t s, std::string &err);
// Pass B of launch_long_passes and launch_long_energy_B in one: the partial
// forces (lw.partB), the potential (lw.partE) and the per-k dE_k/dgamma, for
// k rows whose chunk sums, offsets and totals passes A and S have formed.
hipError_t launch_long_forces_energy_B(const LongConsts &c, LongWork &lw, int64_t n,
                                       int64_t Ptot, hipStream_t s, std::string &err);
// Self energies per atom: d_phi = 2 q s_i (N) and d_psi = 2 q ds_i/dgamma_1,2
// (N x 2), overwritten.
hipError_t launch_self_energy(const ShortConsts &c, int precision, const ShortWork &ws,
                              int64_t n, const double *d_xyz, const double *d_q, double *d_phi,
                              double *d_psi, hipStream_t s);
// Per atom from the records of launch_pair_forces_energy (mixed precision)
// and the self force: the short-range force into d_f (N x 3, overwritten);
// the pairs' dE/dq added to d_phi (N), which holds the self energies of
// launch_self_energy.
hipError_t launch_gather_forces_energy(const ShortConsts &c, const ShortWork &ws,
                                       const NeighborList &nl, int64_t n, const double *d_xyz,
                                       const double *d_q, const double *d_rec, double *d_f,
                                       double *d_phi, hipStream_t s);
// Tangent of the long-range force along d_v (N x 3), times scale; between,
// when given, is called once the chunk sums and offsets are queued, before
// pass B.  Without with_sum the call stops at the per-group partial
// tangents (lw.partB);
// launch_long_tangent_sum then adds (or writes) them into d_out, possibly
// on another stream once the passes are done.
hipError_t launch_long_tangent(const LongConsts &c, LongWork &lw, int64_t n, int64_t Ptot,
                               const double *d_v, double scale, double *d_out, int accumulate,
                               hipStream_t s, std::string &err, bool with_sum = true,
                               const std::function<hipError_t()> *between = nullptr);
hipError_t launch_long_tangent_sum(LongWork &lw, int64_t n, int64_t Ptot, double scale,
                                   double *d_out, int accumulate, hipStream_t s,
                                   std::string &err);
// Short-range energy: potential records in d_pot (pair_rec_bytes: per
// endpoint the other atom's charge times (G, dG/dgamma_1, dG/dgamma_2),
// float3 in mixed precision, double4 in FP64), per atom phi (N) and psi
// (N x 2: d phi / d gamma_1,2), self terms included.
// The self terms run on self_s (default s) and are queued right after the
// pair kernel; after_pair (optional) is called after them.  With self_s
// another stream, the caller orders it: self_s must see the inputs, and
// the gather (on s) must wait for the self terms.
hipError_t launch_short_energy(const ShortConsts &c, int precision, const ShortWork &ws,
                               const NeighborList &nl, int64_t n, const double *d_xyz,
                               const double *d_q, double *d_pot, double *d_phi,
                               double *d_psi, hipStream_t s,
                               const std::function<void()> *after_pair = nullptr,
                               hipStream_t self_s = nullptr);
// Short-range force tangent along d_v, times scale, into d_out (N x 3);
// d_tp holds the tangent records (pair_rec_bytes: first atom (dfx, dfy,
// dfz_i), second atom (-dfx, -dfy, dfz_j); float3 in mixed precision,
// double4 in FP64).
// The self term runs on self_s (default s) and is queued right after the
// pair kernel, then after_pair (optional) is called, then the gather is
// queued on s.  With self_s another stream, the caller orders it: self_s
// must see the inputs, and the gather must wait for the self term.
// d_sums (mixed precision): the pairs' node sums that
// launch_pair_forces_energy kept at these positions, used instead of the
// node loops.
hipError_t launch_short_tangent(const ShortConsts &c, int precision, const ShortWork &ws,
                                const NeighborList &nl, int64_t n, const double *d_xyz,
                                const double *d_q, const double *d_v, double *d_tp,
                                double scale, double *d_out, hipStream_t s,
                                const std::function<void()> *after_pair = nullptr,
                                hipStream_t self_s = nullptr,
                                const float *d_sums = nullptr);
// out3 = (1/2 q.phi, 1/2 q.psi_1, 1/2 q.psi_2); psi may be null; d_part is
// scratch of half_dots_scratch() doubles.  With d_gpart, the sums of its
// ngpart pairs of doubles are added to out3[1] and out3[2].
hipError_t launch_half_dots(int64_t n, const double *d_q, const double *d_phi,
                            const double *d_psi, double *d_out3, double *d_part, hipStream_t s,
                            const double *d_gpart = nullptr, int ngpart = 0);
int half_dots_scratch();

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);

// ---- Substrate_LJ walls (wall.hip) --------------------------------------
// Up to kMaxWallsPerLaunch walls in one pass (kernel argument, by value).
constexpr int kMaxWallsPerLaunch = 8;
struct WallSet {
    qem_wall w[kMaxWallsPerLaunch];
    int32_t nw;
    uint32_t epoch;   // nonzero; written to the flag when an atom is on the wrong side
};
// The walls' z forces into d_f (overwrite or add, walls in order): rows of
// fstride doubles with z at column fstride - 1 (3: an n x 3 force array,
// whose x and y an overwrite zeroes; 1: z only).  *d_flag becomes ws.epoch
// if some atom is on the wrong side of some wall; the launch never clears
// the flag: each call brings a new epoch.
hipError_t launch_substrate_lj(const WallSet &ws, int64_t n, const double *d_xyz, double *d_f,
                               int fstride, int accumulate, uint32_t *d_flag, hipStream_t s);
hipError_t launch_zero_flag(uint32_t *d_flag, hipStream_t s);
// d_f[3 i + 2] += d_z[i]
hipError_t launch_add_z(int64_t n, const double *d_z, double *d_f, hipStream_t s);

}  // namespace qem

// The context behind the opaque qem_ctx handle.
struct qem_ctx {
    qem_params p{};
    std::string err;
    hipStream_t stream = nullptr;
    // the force call's long range runs on `side` alongside the short range
    // on the call's stream: fork_ev orders it after the inputs, join_ev the
    // final sum after it, and sums_ev (adjoint calls) the gather after the
    // side stream's self terms and chunk sums
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr, sums_ev = nullptr;
    // pinned host memory for the energy calls' small device-to-host copies
    // (the three half-dot sums at [0, 3), the per-k dE_k/dgamma from [4)): a
    // copy into pageable memory holds the host until it is done, before it
    // can queue the next launch.  Kept here, not on the stack, so an early
    // error return leaves no copy pending into freed memory
    double *h_out = nullptr;
    size_t h_out_cap = 0;   // doubles
    // the pairs' force-tangent node sums kept by the last saving forces +
    // energy call (tan_sums_bytes), valid for the atom count, pair count and
    // neighbour-list build it was made with
    qem::DevBuf tsums;
    bool tsums_ok = false;
    int64_t tsums_n = -1, tsums_np = -1, tsums_builds = -1;
    // K set (host)
    std::vector<double> kset;   // |K| x 3
    std::vector<float> kprob;   // |K|
    double sum_K = 0.0;
    int64_t kset_mxm = 0, kset_mym = 0;   // the enumeration's |m_x|, |m_y| bounds
    std::vector<double> alias_prob;
    std::vector<int64_t> alias_idx;
    // quadrature
    qem::GLRule gl;
    qem::ShortWork sw;
    qem::NeighborList nl;
    int64_t n_steps = 0;
    qem::LongWork lw;
    // host-pointer staging
    qem::DevBuf d_xyz, d_q, d_f, d_f2, d_aux, d_adj;
    qem::DevBuf d_flag;   // device status word (wall domain check)
    qem::DevBuf d_wz;     // wall z forces of a general-interaction step (atom order)
    // pole splitting: node counts and weights of the point and gaussian sets
    int pole_m[2] = {0, 0};
    double pole_wx[2] = {0.0, 0.0}, pole_wm[2] = {0.0, 0.0};
    uint32_t wall_epoch = 0;
    uint32_t *h_flag = nullptr;     // pinned copy of the status word
    hipEvent_t wall_ev = nullptr;   // recorded after that copy
    // optional per-phase timing (HIP events on the launch stream)
    bool timing = false;
    struct Mark {
        int phase;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;           // recorded, not yet read
    std::vector<hipEvent_t> ev_pool;   // free events
    double phase_ms[QEM_NPHASES] = {0, 0, 0, 0};
    int64_t phase_calls[QEM_NPHASES] = {0, 0, 0, 0};
};
