// 2-D periodic cell list and half neighbour list for the QEM short range.
//
// Stands in for CellListMap.neighborlist(xy, r_cutoff; unitcell=[Lx, Ly])
// as called at src/QEM_short.jl:34-35: all unordered pairs whose minimum-
// image xy distance is <= r_cutoff, reported once with i < j.
//
// Layout: atoms are bucketed into an ncx x ncy grid of cells at least
// r_cutoff wide (stable radix sort of cell ids, radix_sort.hip, so atoms
// inside a cell stay in index order).  Cell ids follow a Morton (Z-order) curve, so that the
// cells around a cell are mostly near it in the id order (a row-major order
// puts the cells above and below a whole row of cells away).  The sorted order numbers the atoms by slot: slot s is
// atom atom_idx_sorted[s], and a cell's atoms are a run of slots.  One
// thread per slot walks the 3 x 3 block of cells around its own, first
// counting and then writing its partners with a larger atom index (the
// pair keeps the orientation i < j of the atom indices).  Pairs are stored
// as slot pairs grouped by first slot, in a deterministic order.  A second
// stable sort of the pair list by second slot gives, per slot, the pairs in
// which it is the second endpoint; the force gather uses both groupings so
// that the per-atom reduction order is fixed (no atomics).  Since pair
// neighbours are near each other in slot order, the per-pair position reads
// and the gather's reads of per-pair results stay within nearby cache lines.
#include "qem_internal.h"

namespace qem {
namespace {

struct CellGrid {
    double Lx, Ly;
    int ncx, ncy;
};

__device__ __forceinline__ int cell_of(double x, double L, int nc)
{
    double xw = x - L * floor(x / L);
    int c = (int)(xw / L * (double)nc);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

// Minimum-image squared xy distance.  Contraction is disabled so the value
// (and hence the pair set) is bit-identical to a plain IEEE evaluation of
// dx - L*rint(dx/L), dx*dx + dy*dy.
__device__ __forceinline__ double pair_d2(double xi, double yi, double xj,
                                          double yj, double Lx, double Ly)
{
#pragma clang fp contract(off)
    double dx = xi - xj, dy = yi - yj;
    double mx = dx - Lx * rint(dx / Lx);
    double my = dy - Ly * rint(dy / Ly);
    return mx * mx + my * my;
}

// Morton id of cell (cx, cy): the bits of cx and cy interleaved (cx, cy < 2^16)
__host__ __device__ __forceinline__ uint32_t spread_bits(uint32_t v)
{
    v &= 0x0000ffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__host__ __device__ __forceinline__ uint32_t cell_id(int cx, int cy)
{
    return spread_bits((uint32_t)cx) | (spread_bits((uint32_t)cy) << 1);
}

// inverse of spread_bits: the even bits of v packed together
__device__ __forceinline__ int gather_bits(uint32_t v)
{
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0f0f0f0fu;
    v = (v | (v >> 4)) & 0x00ff00ffu;
    v = (v | (v >> 8)) & 0x0000ffffu;
    return (int)v;
}

__global__ void k_cell_keys(const double *__restrict__ xyz, int64_t n, CellGrid g,
                            uint32_t *__restrict__ key)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cx = cell_of(xyz[3 * i], g.Lx, g.ncx);
    int cy = cell_of(xyz[3 * i + 1], g.Ly, g.ncy);
    key[i] = cell_id(cx, cy);
}

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i64(int64_t *__restrict__ p, int64_t n, int64_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// start/end of each run of equal keys in a sorted key array
__global__ void k_run_bounds(const uint32_t *__restrict__ skey, int64_t n,
                             int32_t *__restrict__ start, int32_t *__restrict__ end)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t k = skey[t];
    if (t == 0 || skey[t - 1] != k) start[k] = (int32_t)t;
    if (t == n - 1 || skey[t + 1] != k) end[k] = (int32_t)(t + 1);
}

template <bool FILL>
__global__ void k_nl_pass(const double *__restrict__ sxyz, int64_t n, CellGrid g, double rc2,
                          const uint32_t *__restrict__ skey, const int32_t *__restrict__ cstart,
                          const int32_t *__restrict__ cend, const int32_t *__restrict__ sidx,
                          int64_t *__restrict__ counts, const int64_t *__restrict__ row_ptr,
                          int32_t *__restrict__ pi, int32_t *__restrict__ pj)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double xi = sxyz[3 * s], yi = sxyz[3 * s + 1];
    const int32_t ai = sidx[s];
    const uint32_t key = skey[s];
    const int cx = gather_bits(key), cy = gather_bits(key >> 1);
    const int nxi = g.ncx >= 3 ? 3 : g.ncx, nyi = g.ncy >= 3 ? 3 : g.ncy;
    int64_t cnt = 0;
    const int64_t base = FILL ? row_ptr[s] : 0;
    for (int oy = 0; oy < nyi; ++oy) {
        const int cyy = g.ncy >= 3 ? (cy + oy - 1 + g.ncy) % g.ncy : oy;
        for (int ox = 0; ox < nxi; ++ox) {
            const int cxx = g.ncx >= 3 ? (cx + ox - 1 + g.ncx) % g.ncx : ox;
            const int c = (int)cell_id(cxx, cyy);
            const int t1 = cend[c];
            for (int t = cstart[c]; t < t1; ++t) {
                if (sidx[t] <= ai) continue;
                const double d2 = pair_d2(xi, yi, sxyz[3 * (int64_t)t], sxyz[3 * (int64_t)t + 1],
                                          g.Lx, g.Ly);
                if (d2 <= rc2) {
                    if (FILL) {
                        pi[base + cnt] = (int32_t)s;
                        pj[base + cnt] = t;
                    }
                    ++cnt;
                }
            }
        }
    }
    if (!FILL) counts[s] = cnt;
}

__global__ void k_rows_to_slots(const double *__restrict__ src, const int32_t *__restrict__ perm,
                                int64_t n, int w, double *__restrict__ dst)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * w) return;
    const int64_t s = e / w, c = e - s * w;
    dst[e] = src[(int64_t)perm[s] * w + c];
}

// positions and charges of slot s from atom perm[s], one thread per slot
__global__ void k_atoms_to_slots(const double *__restrict__ xyz, const double *__restrict__ q,
                                 const int32_t *__restrict__ perm, int64_t n,
                                 double *__restrict__ sxyz, double *__restrict__ sq)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t a = perm[s];
    sxyz[3 * s] = xyz[3 * a];
    sxyz[3 * s + 1] = xyz[3 * a + 1];
    sxyz[3 * s + 2] = xyz[3 * a + 2];
    sq[s] = q[a];
}

__global__ void k_rows_from_slots(const double *__restrict__ src,
                                  const int32_t *__restrict__ perm, int64_t n, int w,
                                  double *__restrict__ dst, const double *__restrict__ addz)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * w) return;
    const int64_t s = e / w, c = e - s * w;
    const int64_t r = perm[s];
    double v = src[e];
    if (addz && c == 2) v = __dadd_rn(v, addz[r]);
    dst[r * w + c] = v;
}

__global__ void k_row_bounds(const int64_t *__restrict__ row_ptr, int64_t n,
                             int32_t *__restrict__ start, int32_t *__restrict__ end)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    start[i] = (int32_t)row_ptr[i];
    end[i] = (int32_t)row_ptr[i + 1];
}

// inv[perm[t]] = t for a permutation perm of [0, m)
__global__ void k_inverse_perm(const int32_t *__restrict__ perm, int64_t m,
                               int32_t *__restrict__ inv)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    inv[perm[t]] = (int32_t)t;
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

inline int bits_for(uint64_t maxval)
{
    int b = 1;
    while (b < 32 && (1ull << b) <= maxval) ++b;
    return b;
}

#define NL_CHECK(expr)                                                        \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            err = std::string("neighbour list: ") + #expr + ": " +            \
                  hipGetErrorString(e_);                                      \
            return e_;                                                        \
        }                                                                     \
    } while (0)

// Sorts m int32 endpoint values and records, for every atom, the run of
// pair ids having that endpoint (stable: pair ids ascending inside a run).
hipError_t group_by_endpoint(NeighborList &nl, const int32_t *d_end, int64_t m, int64_t n,
                             DevBuf &idx_out, DevBuf &start, DevBuf &end, hipStream_t s,
                             std::string &err)
{
    NL_CHECK(ensure(start, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)));
    NL_CHECK(ensure(end, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)));
    if (n > 0) {
        k_fill_i32<<<nblk(n, 256), 256, 0, s>>>(start.as<int32_t>(), n, 0);
        k_fill_i32<<<nblk(n, 256), 256, 0, s>>>(end.as<int32_t>(), n, 0);
    }
    if (m == 0) return hipSuccess;
    NL_CHECK(ensure(nl.cell_key_sorted, sizeof(uint32_t) * (size_t)m));
    NL_CHECK(ensure(idx_out, sizeof(int32_t) * (size_t)m));
    const int ebits = bits_for((uint64_t)(n > 0 ? n - 1 : 0));
    NL_CHECK(ensure_sort_scratch(nl.tmp, radix_sort_scratch_bytes(m, 4), s));
    NL_CHECK(radix_sort_u32(reinterpret_cast<const uint32_t *>(d_end), m, ebits,
                            nl.cell_key_sorted.as<uint32_t>(), idx_out.as<int32_t>(), nl.tmp.p,
                            s));
    k_run_bounds<<<nblk(m, 256), 256, 0, s>>>(nl.cell_key_sorted.as<uint32_t>(), m,
                                              start.as<int32_t>(), end.as<int32_t>());
    return hipGetLastError();
}

// jpos[p]: pair p's place in the second-endpoint grouping, where the pair
// kernels write the record of its second atom (the inverse of j_idx).
hipError_t second_endpoint_places(NeighborList &nl, int64_t m, hipStream_t s, std::string &err)
{
    NL_CHECK(ensure(nl.jpos, sizeof(int32_t) * (size_t)(m > 0 ? m : 1)));
    if (m == 0) return hipSuccess;
    k_inverse_perm<<<nblk(m, 256), 256, 0, s>>>(nl.j_idx.as<int32_t>(), m, nl.jpos.as<int32_t>());
    return hipGetLastError();
}

}  // namespace

hipError_t build_neighbor_list(NeighborList &nl, int64_t n, const double *d_xyz, double Lx,
                               double Ly, double rc, hipStream_t s, std::string &err)
{
    nl.n = n;
    nl.npairs = 0;
    nl.has_i_idx = false;
    nl.slot_order = true;
    ++nl.builds;
    NL_CHECK(ensure(nl.i_start, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)));
    NL_CHECK(ensure(nl.i_end, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)));
    NL_CHECK(ensure(nl.j_start, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)));
    NL_CHECK(ensure(nl.j_end, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)));
    if (n == 0) return hipSuccess;
    if (n > INT32_MAX) {
        err = "neighbour list: more than 2^31-1 atoms";
        return hipErrorInvalidValue;
    }
    // cells at least r_cutoff wide (a hair wider so rounding in the cell
    // index cannot separate a pair by two cells)
    CellGrid g;
    g.Lx = Lx;
    g.Ly = Ly;
    double wx = rc * (1.0 + 1e-9), wy = rc * (1.0 + 1e-9);
    g.ncx = (int)std::min(1024.0, std::max(1.0, std::floor(Lx / wx)));
    g.ncy = (int)std::min(1024.0, std::max(1.0, std::floor(Ly / wy)));
    // Morton ids run up to that of the last cell
    const int64_t ncell = (int64_t)cell_id(g.ncx - 1, g.ncy - 1) + 1;

    NL_CHECK(ensure(nl.cell_key, sizeof(uint32_t) * (size_t)n));
    NL_CHECK(ensure(nl.cell_key_sorted, sizeof(uint32_t) * (size_t)n));
    NL_CHECK(ensure(nl.atom_idx_sorted, sizeof(int32_t) * (size_t)n));
    NL_CHECK(ensure(nl.cell_start, sizeof(int32_t) * (size_t)ncell));
    NL_CHECK(ensure(nl.cell_end, sizeof(int32_t) * (size_t)ncell));
    NL_CHECK(ensure(nl.counts, sizeof(int64_t) * (size_t)(n + 1)));
    NL_CHECK(ensure(nl.row_ptr, sizeof(int64_t) * (size_t)(n + 1)));

    k_cell_keys<<<nblk(n, 256), 256, 0, s>>>(d_xyz, n, g, nl.cell_key.as<uint32_t>());
    const int kbits = bits_for((uint64_t)(ncell - 1));
    NL_CHECK(ensure_sort_scratch(nl.tmp, std::max(radix_sort_scratch_bytes(n, 4),
                                                  scan_scratch_bytes(n + 1)), s));
    NL_CHECK(radix_sort_u32(nl.cell_key.as<uint32_t>(), n, kbits,
                            nl.cell_key_sorted.as<uint32_t>(), nl.atom_idx_sorted.as<int32_t>(),
                            nl.tmp.p, s));
    k_fill_i32<<<nblk(ncell, 256), 256, 0, s>>>(nl.cell_start.as<int32_t>(), ncell, 0);
    k_fill_i32<<<nblk(ncell, 256), 256, 0, s>>>(nl.cell_end.as<int32_t>(), ncell, 0);
    k_run_bounds<<<nblk(n, 256), 256, 0, s>>>(nl.cell_key_sorted.as<uint32_t>(), n,
                                              nl.cell_start.as<int32_t>(),
                                              nl.cell_end.as<int32_t>());
    const double rc2 = rc * rc;
    NL_CHECK(ensure(nl.sxyz, sizeof(double) * 3 * (size_t)n));
    k_rows_to_slots<<<nblk(3 * n, 256), 256, 0, s>>>(d_xyz, nl.atom_idx_sorted.as<int32_t>(), n, 3,
                                                      nl.sxyz.as<double>());
    NL_CHECK(ensure(nl.slot_of, sizeof(int32_t) * (size_t)n));
    k_inverse_perm<<<nblk(n, 256), 256, 0, s>>>(nl.atom_idx_sorted.as<int32_t>(), n,
                                                nl.slot_of.as<int32_t>());
    // counts[n] = 0 so the exclusive scan's last entry is the total
    k_fill_i64<<<1, 64, 0, s>>>(nl.counts.as<int64_t>() + n, 1, 0);
    k_nl_pass<false><<<nblk(n, 128), 128, 0, s>>>(
        nl.sxyz.as<double>(), n, g, rc2, nl.cell_key_sorted.as<uint32_t>(),
        nl.cell_start.as<int32_t>(), nl.cell_end.as<int32_t>(), nl.atom_idx_sorted.as<int32_t>(),
        nl.counts.as<int64_t>(), nullptr, nullptr, nullptr);
    NL_CHECK(hipGetLastError());
    NL_CHECK(scan_exclusive_i64(nl.counts.as<int64_t>(), nl.row_ptr.as<int64_t>(), n + 1,
                                nl.tmp.p, s));
    int64_t total = 0;
    NL_CHECK(hipMemcpyAsync(&total, nl.row_ptr.as<int64_t>() + n, sizeof(int64_t),
                            hipMemcpyDeviceToHost, s));
    NL_CHECK(hipStreamSynchronize(s));
    if (total > INT32_MAX) {
        err = "neighbour list: more than 2^31-1 pairs";
        return hipErrorInvalidValue;
    }
    nl.npairs = total;
    NL_CHECK(ensure(nl.pi, sizeof(int32_t) * (size_t)(total > 0 ? total : 1)));
    NL_CHECK(ensure(nl.pj, sizeof(int32_t) * (size_t)(total > 0 ? total : 1)));
    if (total > 0) {
        k_nl_pass<true><<<nblk(n, 128), 128, 0, s>>>(
            nl.sxyz.as<double>(), n, g, rc2, nl.cell_key_sorted.as<uint32_t>(),
            nl.cell_start.as<int32_t>(), nl.cell_end.as<int32_t>(),
            nl.atom_idx_sorted.as<int32_t>(), nullptr, nl.row_ptr.as<int64_t>(),
            nl.pi.as<int32_t>(), nl.pj.as<int32_t>());
        NL_CHECK(hipGetLastError());
    }
    k_row_bounds<<<nblk(n, 256), 256, 0, s>>>(nl.row_ptr.as<int64_t>(), n,
                                              nl.i_start.as<int32_t>(), nl.i_end.as<int32_t>());
    NL_CHECK(hipGetLastError());
    NL_CHECK(group_by_endpoint(nl, nl.pj.as<int32_t>(), total, n, nl.j_idx, nl.j_start,
                               nl.j_end, s, err));
    return second_endpoint_places(nl, total, s, err);
}

hipError_t install_pairs(NeighborList &nl, int64_t n, const int32_t *h_pairs, int64_t npairs,
                         hipStream_t s, std::string &err)
{
    nl.n = n;
    nl.npairs = npairs;
    nl.has_i_idx = true;
    nl.slot_order = false;
    ++nl.builds;
    std::vector<int32_t> a((size_t)npairs), b((size_t)npairs);
    for (int64_t t = 0; t < npairs; ++t) {
        a[(size_t)t] = h_pairs[2 * t];
        b[(size_t)t] = h_pairs[2 * t + 1];
    }
    NL_CHECK(ensure(nl.pi, sizeof(int32_t) * (size_t)(npairs > 0 ? npairs : 1)));
    NL_CHECK(ensure(nl.pj, sizeof(int32_t) * (size_t)(npairs > 0 ? npairs : 1)));
    if (npairs > 0) {
        NL_CHECK(hipMemcpyAsync(nl.pi.p, a.data(), sizeof(int32_t) * (size_t)npairs,
                                hipMemcpyHostToDevice, s));
        NL_CHECK(hipMemcpyAsync(nl.pj.p, b.data(), sizeof(int32_t) * (size_t)npairs,
                                hipMemcpyHostToDevice, s));
    }
    NL_CHECK(group_by_endpoint(nl, nl.pi.as<int32_t>(), npairs, n, nl.i_idx, nl.i_start,
                               nl.i_end, s, err));
    NL_CHECK(group_by_endpoint(nl, nl.pj.as<int32_t>(), npairs, n, nl.j_idx, nl.j_start,
                               nl.j_end, s, err));
    NL_CHECK(second_endpoint_places(nl, npairs, s, err));
    // the host vectors must outlive the asynchronous copies
    NL_CHECK(hipStreamSynchronize(s));
    return hipSuccess;
}

hipError_t launch_rows_to_slots(const NeighborList &nl, int64_t n, int w, const double *src,
                                double *dst, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_rows_to_slots<<<nblk(n * w, 256), 256, 0, s>>>(src, nl.atom_idx_sorted.as<int32_t>(), n, w,
                                                      dst);
    return hipGetLastError();
}

hipError_t launch_atoms_to_slots(const NeighborList &nl, int64_t n, const double *d_xyz,
                                 const double *d_q, double *sxyz, double *sq, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_atoms_to_slots<<<nblk(n, 256), 256, 0, s>>>(d_xyz, d_q, nl.atom_idx_sorted.as<int32_t>(), n,
                                                  sxyz, sq);
    return hipGetLastError();
}

hipError_t launch_rows_from_slots(const NeighborList &nl, int64_t n, int w, const double *src,
                                  double *dst, hipStream_t s, const double *addz)
{
    if (n <= 0) return hipSuccess;
    if (addz && w != 3) return hipErrorInvalidValue;
    k_rows_from_slots<<<nblk(n * w, 256), 256, 0, s>>>(src, nl.atom_idx_sorted.as<int32_t>(), n,
                                                        w, dst, addz);
    return hipGetLastError();
}

}  // namespace qem
