// qem_long_forces_noalloc_dev: the long-range force call (qem_long_forces_dev)
// without any device allocation, for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  No counterpart in the reference.
//
// Its work buffers -- the staged k rows, the z sort (keys, values,
// histograms, counters), the z-ordered atoms, the per-k constants, the chunk
// sums and offsets, the per-group partial forces, the k = 0 term's tile sums
// and per-atom values -- are carved from one
// file-scope device array, `scratch`, of QEM_NOALLOC_SCRATCH_BYTES.  The
// array exists once per GPU, from the time the library's code object is
// loaded on it until the process ends, and every call on that GPU shares it:
// calls that overlap on one GPU race.  Its device address is taken once per
// device with hipGetSymbolAddress (the array's name is only a host-side
// stand-in); every launch gets that pointer.
//
// The k rows are built on the device from the k indices, which travel as
// kernel arguments (no host staging, no copy): index -> (m_x, m_y) by the
// K-set enumeration (src/K_set_generator.jl:32-42: m_x outer, m_y inner,
// (0, 0) skipped), k_x = m_x * 2pi / L_x, |k| = sqrt(k_x^2 + k_y^2), in the
// same operations as the host's table, so the rows are the same bits.
// Everything after the rows is the kernels of qem_long_forces_dev.
#include <mutex>

#include "qem_internal.h"

namespace qem {
namespace {

constexpr size_t kScratchBytes = QEM_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
constexpr double kTwoPi = 6.283185307179586;  // Float64(2pi), as the host K set

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the
// first call on that device).
hipError_t scratch_ptr(int dev, void **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return hipSuccess;
}

// k indices of one launch (kernel argument)
constexpr int kIdxPerLaunch = 256;
struct KIdxChunk {
    int32_t idx[kIdxPerLaunch];
    int32_t count, off;
};
struct KSetGeom {
    double Lx, Ly;
    double wrow;   // batch weight / (Lx Ly eps_0), as the host's staged rows
    int32_t mxm, mym;
};

__global__ void k_rows_from_idx(KIdxChunk ch, KSetGeom g, double4 *__restrict__ kv)
{
#pragma clang fp contract(off)
    const int t = threadIdx.x;
    if (t < ch.count) {
        const int64_t ny = 2 * (int64_t)g.mym + 1;
        const int64_t center = (int64_t)g.mxm * ny + g.mym;   // (0, 0), skipped
        int64_t j = ch.idx[t];
        if (j >= center) ++j;
        const int64_t mx = j / ny - g.mxm, my = j % ny - g.mym;
        const double kx = (double)mx * kTwoPi / g.Lx;
        const double ky = (double)my * kTwoPi / g.Ly;
        const double k = sqrt(kx * kx + ky * ky);
        kv[ch.off + t] = make_double4(kx, ky, k, g.wrow);
    }
}

inline size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }

int fail(qem_ctx *ctx, int code, const char *msg)
{
    ctx->err.assign(msg);   // capacity reserved at qem_create: no allocation
    return code;
}

}  // namespace
}  // namespace qem

using namespace qem;

extern "C" int qem_long_forces_noalloc_dev(qem_ctx *ctx, int64_t n, const double *d_xyz,
                                           const double *d_q, const int64_t *k_idx, int64_t P,
                                           double *d_f, void *stream)
{
    if (!ctx) return QEM_ERR_INVALID_ARG;
    ctx->err.clear();
    if (hipSetDevice(ctx->p.device) != hipSuccess) return fail(ctx, QEM_ERR_HIP, "hipSetDevice");
    if (n < 0 || P < 0 || (P > 0 && !k_idx) || (n > 0 && (!d_xyz || !d_q || !d_f)))
        return fail(ctx, QEM_ERR_INVALID_ARG, "bad arguments");
    if (n > QEM_NOALLOC_MAX_N)
        return fail(ctx, QEM_ERR_CAPACITY, "qem_long_forces_noalloc_dev: n above QEM_NOALLOC_MAX_N");
    if (P > QEM_NOALLOC_MAX_P)
        return fail(ctx, QEM_ERR_CAPACITY, "qem_long_forces_noalloc_dev: P above QEM_NOALLOC_MAX_P");
    if (ctx->kprob.empty())
        return fail(ctx, QEM_ERR_UNSUPPORTED,
                    "qem_long_forces_noalloc_dev: the K set is not built (rbe_p > 0 at "
                    "qem_create, or call qem_kset_size once)");
    const int64_t nk = (int64_t)ctx->kprob.size();
    for (int64_t t = 0; t < P; ++t)
        if (k_idx[t] < 0 || k_idx[t] >= nk)
            return fail(ctx, QEM_ERR_INVALID_ARG, "k index out of range");
    if (n == 0) return QEM_OK;
    hipStream_t s = stream == QEM_STREAM_NULL ? nullptr
                                              : (stream ? static_cast<hipStream_t>(stream)
                                                        : ctx->stream);
    void *base = nullptr;
    if (hipError_t e = scratch_ptr(ctx->p.device, &base); e != hipSuccess)
        return fail(ctx, QEM_ERR_HIP, "hipGetSymbolAddress(scratch)");

    // carve the scratch: the sort's counters first, at a fixed offset
    const int64_t Ptot = P;   // the sampled rows (the k = 0 term has no row)
    const LongSizes z = long_force_sizes(n, Ptot);
    char *b = static_cast<char *>(base);
    size_t off = 0;
    LongWork lw;
    auto take = [&](DevBuf &d, size_t bytes) {
        d.p = b + off;
        d.cap = bytes;
        off += a256(bytes);
    };
    take(lw.tmp, z.sort);
    take(lw.kv, z.kv);
    take(lw.kr, z.kr);
    take(lw.perm, z.perm);
    take(lw.satom, z.satom);
    take(lw.partA, z.partA);
    take(lw.offA, z.partA);
    take(lw.totA, z.totA);
    take(lw.partB, z.partB);
    take(lw.k0tile, z.k0tile);
    take(lw.k0out, z.k0out);
    lw.k0_on = true;
    if (off > kScratchBytes)
        return fail(ctx, QEM_ERR_CAPACITY, "qem_long_forces_noalloc_dev: scratch too small");

    // the sort's completion counters start at zero (a fill kernel, not a
    // memset, so that the call can be captured in a graph)
    if (radix_sort_init(lw.tmp.p, s) != hipSuccess)
        return fail(ctx, QEM_ERR_HIP, "radix_sort_init");

    // k rows from the indices
    const qem_params &p = ctx->p;
    KSetGeom g;
    g.Lx = p.Lx;
    g.Ly = p.Ly;
    const double inv_ae = 1.0 / (p.Lx * p.Ly * p.eps_0);
    g.wrow = (P > 0 ? ctx->sum_K / (double)P : 0.0) * inv_ae;
    g.mxm = (int32_t)ctx->kset_mxm;
    g.mym = (int32_t)ctx->kset_mym;
    double4 *kv = lw.kv.as<double4>();
    for (int64_t o = 0; o < P; o += kIdxPerLaunch) {
        KIdxChunk ch;
        ch.off = (int32_t)o;
        ch.count = (int32_t)std::min<int64_t>(kIdxPerLaunch, P - o);
        for (int t = 0; t < ch.count; ++t) ch.idx[t] = (int32_t)k_idx[o + t];
        k_rows_from_idx<<<1, kIdxPerLaunch, 0, s>>>(ch, g, kv);
    }
    if (hipGetLastError() != hipSuccess) return fail(ctx, QEM_ERR_HIP, "k rows");

    LongConsts c;
    c.Lz = p.Lz;
    c.g1 = p.gamma_1;
    c.g2 = p.gamma_2;
    c.g12 = p.gamma_1 * p.gamma_2;
    c.inv_area_eps = inv_ae;
    c.weight = P > 0 ? ctx->sum_K / (double)P : 0.0;
    c.inv4alpha = 0.25 / p.alpha;
    c.parts = QEM_LONG_ALL;
    c.precision = p.precision;
    // the same launches as qem_long_forces_dev; every buffer above already
    // has the capacity these ask for, so nothing is allocated
    if (launch_zsort(lw, n, d_xyz, d_q, s, ctx->err) != hipSuccess ||
        launch_long_k0(c, lw, n, s, ctx->err) != hipSuccess ||
        launch_long_krec(c, lw, Ptot, s, ctx->err) != hipSuccess ||
        launch_long_passes(c, lw, n, Ptot, s, ctx->err, false) != hipSuccess ||
        launch_long_sum(lw, n, Ptot, d_f, 0, s, ctx->err) != hipSuccess)
        return QEM_ERR_HIP;
    return QEM_OK;
}
