// Stable LSD radix sort of (key, int32 value) pairs and an exclusive scan,
// hand-written for gfx950.  Used for the z order of the long range
// (sortperm, src/QEM_long.jl:39-40; the O(N) bucketing the reference meant
// to use is src/bulket_sorting.jl:4-9) and for the neighbour list's cell and
// pair groupings.
//
// Sort layout.  The input is cut into tiles of T = 256 * items keys, one
// block of 256 threads per tile, 8-bit digits (256 bins, one per thread).
//   load   (1 launch)  keys and values from a source (the z coordinates, an
//                      array) and each tile's counts of digit 0;
//   pass d (1 launch per digit)  each block forms its tile's output offsets
//                      from every tile's counts of digit d (the digit's
//                      total over all tiles below it plus its count in the
//                      tiles before this one, an L2-resident read), scatters
//                      its keys stably to offset + rank, and counts digit
//                      d + 1 at the keys' new tiles (global atomics: integer
//                      counts, so their order does not matter); the last
//                      pass writes through an output functor (e.g. the
//                      permutation and the z-ordered atoms).
// The counts rotate through three buffers: pass d reads buffer d % 3 (filled
// by the launch before it), adds into (d + 1) % 3 (cleared by the launch
// before it) and clears (d + 2) % 3 (read by the launch before it), so the
// kernel boundaries order every access: no block waits for another and no
// device-scope fence is needed.
//
// Stable rank inside a tile: the tile is walked in rounds of 256 keys in
// position order; in each round a wave finds the lanes holding its digit
// with eight ballots, the waves' per-digit counts go through LDS, and each
// digit's thread turns them into offsets (four waves in order, plus the
// digit's running count): equal keys keep their input order.
#include "qem_internal.h"

namespace qem {
namespace {

constexpr int RS = 256;   // threads per block = digit bins

// per-block LDS of the ranking rounds
struct RankLds {
    uint32_t cnt[2][4][RS];   // per round (double-buffered): wave counts, then offsets
    uint32_t run[RS];         // next output position of each digit
    int last;                 // this block finished last
};

// Barrier that orders LDS accesses only: global loads and stores issued
// before it stay in flight (a full __syncthreads waits for them).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, int shift)
{
    return (uint32_t)(key >> shift) & 0xFFu;
}

// One round: the 256 keys at positions base + threadIdx.x (valid ones)
// ranked stably by digit; returns the output position of this thread's key.
// lds.cnt[buf] holds zeros on entry; lds.cnt[buf ^ 1] is zero on return.
__device__ __forceinline__ uint32_t rank_round(RankLds &lds, int buf, bool valid, uint32_t d)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t peer = (uint32_t)__popcll(m & lt);
    if (valid && (m & lt) == 0) lds.cnt[buf][wave][d] = (uint32_t)__popcll(m);
    lds_barrier();
    {
        const int t = threadIdx.x;
        const uint32_t c0 = lds.cnt[buf][0][t], c1 = lds.cnt[buf][1][t];
        const uint32_t c2 = lds.cnt[buf][2][t], c3 = lds.cnt[buf][3][t];
        const uint32_t b0 = lds.run[t];
        lds.cnt[buf][0][t] = b0;
        lds.cnt[buf][1][t] = b0 + c0;
        lds.cnt[buf][2][t] = b0 + c0 + c1;
        lds.cnt[buf][3][t] = b0 + c0 + c1 + c2;
        lds.run[t] = b0 + c0 + c1 + c2 + c3;
#pragma unroll
        for (int w = 0; w < 4; ++w) lds.cnt[buf ^ 1][w][t] = 0;
    }
    lds_barrier();
    return valid ? lds.cnt[buf][wave][d] + peer : 0u;
}

// Exclusive scan over the 256 threads of the block (one value each).
__device__ __forceinline__ uint32_t block_excl_u32(uint32_t v, uint32_t *sh /* [4] */)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(inc, o);
        if (lane >= o) inc += u;
    }
    if (lane == 63) sh[wave] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (int w = 0; w < wave; ++w) base += sh[w];
    __syncthreads();
    return base + inc - v;
}

// Counts this block in; true (in every thread) for the block that finishes
// last, after which it may read what every other block wrote.
__device__ __forceinline__ bool finished_last(unsigned *counter, int *flag)
{
    __threadfence();            // release this thread's writes device-wide
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned prev = atomicInc(counter, gridDim.x - 1);   // wraps to 0
        *flag = prev == gridDim.x - 1;
    }
    __syncthreads();
    const bool last = *flag != 0;
    if (last) __threadfence();  // acquire: see the other blocks' writes
    return last;
}

// ---- sources and outputs --------------------------------------------------

// z of an N x 3 coordinate array as an order-preserving unsigned key: the
// IEEE bits with the sign bit flipped for non-negative values and all bits
// flipped for negative ones (ascending keys = ascending z, -0 before +0);
// the value is the atom index, so equal z keep index order.
struct ZSource {
    const double *xyz;
    __device__ __forceinline__ void get(int64_t i, uint64_t &k, int32_t &v) const
    {
        const uint64_t b = (uint64_t)__double_as_longlong(xyz[3 * i + 2]);
        k = (b >> 63) ? ~b : (b | 0x8000000000000000ull);
        v = (int32_t)i;
    }
};

// keys from an array (read as unsigned), values the positions
struct KeyIotaSource {
    const uint32_t *key;
    __device__ __forceinline__ void get(int64_t i, uint32_t &k, int32_t &v) const
    {
        k = key[i];
        v = (int32_t)i;
    }
};

// the last pass of the z sort: the permutation (rank -> atom) and the atoms
// in z order as (x, y, z, q)
struct ZOutput {
    const double *xyz, *q;
    int32_t *perm;
    double4 *satom;
    __device__ __forceinline__ void put(uint32_t pos, uint64_t, int32_t v) const
    {
        perm[pos] = v;
        const int64_t a = v;
        satom[pos] = make_double4(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2], q[a]);
    }
};

struct PairOutput {
    uint32_t *keys;
    int32_t *vals;
    __device__ __forceinline__ void put(uint32_t pos, uint32_t k, int32_t v) const
    {
        keys[pos] = k;
        vals[pos] = v;
    }
};

// ---- kernels ----------------------------------------------------------------

// Counts are tile-major: C[t * RS + d] for tile t and digit d.
template <typename K, class Src>
__global__ __launch_bounds__(RS) void k_rs_load(Src src, int64_t n, int items, K *__restrict__ ka,
                                                int32_t *__restrict__ va, uint32_t *__restrict__ C0,
                                                uint32_t *__restrict__ C1)
{
    __shared__ uint32_t hist[RS];
    const int64_t T = (int64_t)RS * items;
    hist[threadIdx.x] = 0;
    C1[(size_t)blockIdx.x * RS + threadIdx.x] = 0;   // for the first pass's atomics
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * T;
    for (int r = 0; r < items; ++r) {
        const int64_t i = t0 + (int64_t)r * RS + threadIdx.x;
        if (i < n) {
            K k;
            int32_t v;
            src.get(i, k, v);
            ka[i] = k;
            va[i] = v;
            atomicAdd(&hist[digit_of(k, 0)], 1u);
        }
    }
    __syncthreads();
    C0[(size_t)blockIdx.x * RS + threadIdx.x] = hist[threadIdx.x];
}

template <typename K, bool LAST, class Out>
__global__ __launch_bounds__(RS) void k_rs_pass(const K *__restrict__ kin,
                                                const int32_t *__restrict__ vin, K *__restrict__ kout,
                                                int32_t *__restrict__ vout, Out out, int64_t n,
                                                int items, int shift, int key_bits,
                                                const uint32_t *__restrict__ Ccur,
                                                uint32_t *__restrict__ Cnxt,
                                                uint32_t *__restrict__ Cclr)
{
    __shared__ RankLds lds;
    __shared__ uint32_t sh[4];
    const int ntiles = gridDim.x;
    const int64_t T = (int64_t)RS * items;
    const int64_t t0 = (int64_t)blockIdx.x * T;
    if (!LAST) Cclr[(size_t)blockIdx.x * RS + threadIdx.x] = 0;   // for the pass after next
    {
        // this tile's offset of digit d: the counts of the smaller digits in
        // every tile plus those of digit d in the tiles before this one
        const int d = threadIdx.x;
        uint32_t tot = 0, pre = 0;
#pragma unroll 16
        for (int t = 0; t < ntiles; ++t) {
            const uint32_t h = Ccur[(size_t)t * RS + d];
            tot += h;
            pre += t < (int)blockIdx.x ? h : 0u;
        }
        lds.run[d] = block_excl_u32(tot, sh) + pre;
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) lds.cnt[0][w][threadIdx.x] = 0;
    // the first round's key, then each round's load issued one round ahead
    int64_t i = t0 + threadIdx.x;
    K kn = 0;
    int32_t vn = 0;
    if (i < n) {
        kn = kin[i];
        vn = vin[i];
    }
    lds_barrier();
    int buf = 0;
    for (int r = 0; r < items; ++r) {
        const bool valid = i < n;
        const K k = kn;
        const int32_t v = vn;
        i += RS;
        if (r + 1 < items && i < n) {
            kn = kin[i];
            vn = vin[i];
        }
        const uint32_t pos = rank_round(lds, buf, valid, valid ? digit_of(k, shift) : 0u);
        buf ^= 1;
        if (LAST) {
            if (valid) out.put(pos, k, v);
        } else {
            if (valid) {
                kout[pos] = k;
                vout[pos] = v;
            }
            // the next digit's count at the key's new tile, one atomic per
            // group of lanes with the same (tile, digit): keys whose next
            // digits repeat (e.g. neighbouring atoms' indices) would
            // otherwise pile onto a few counters
            const uint32_t key2 = valid ? (uint32_t)(pos / T) * RS + digit_of(k, shift + 8) : 0u;
            uint64_t m = __ballot(valid);
            for (int b = 0; b < key_bits; ++b) {
                const bool bit = (key2 >> b) & 1u;
                const uint64_t bal = __ballot(bit);
                m &= bit ? bal : ~bal;
            }
            const int lane = threadIdx.x & 63;
            const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
            if (valid && (m & lt) == 0) atomicAdd(&Cnxt[key2], (uint32_t)__popcll(m));
        }
    }
}

__global__ void k_zero_u32(uint32_t *p, int64_t m)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[i] = 0;
}

// ---- the scan -----------------------------------------------------------------

// Block sums of tiles of RS * items int64 values; the last block turns them
// into exclusive block offsets (in place).
__global__ __launch_bounds__(RS) void k_scan_up(const int64_t *__restrict__ in, int64_t m,
                                                int items, int64_t *__restrict__ part,
                                                unsigned *counter)
{
    __shared__ int64_t red[RS];
    __shared__ int flag;
    const int64_t t0 = (int64_t)blockIdx.x * RS * items;
    int64_t s = 0;
    for (int r = 0; r < items; ++r) {
        const int64_t i = t0 + (int64_t)r * RS + threadIdx.x;
        if (i < m) s += in[i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = RS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
    if (!finished_last(counter, &flag)) return;
    if (threadIdx.x == 0) {
        int64_t run = 0;
        for (unsigned b = 0; b < gridDim.x; ++b) {
            const int64_t v = __hip_atomic_load(part + b, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            part[b] = run;
            run += v;
        }
    }
}

// Each tile's exclusive scan plus its block offset.
__global__ __launch_bounds__(RS) void k_scan_down(const int64_t *__restrict__ in, int64_t m,
                                                  int items, const int64_t *__restrict__ part,
                                                  int64_t *__restrict__ out)
{
    __shared__ int64_t sh[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t t0 = (int64_t)blockIdx.x * RS * items;
    int64_t carry = part[blockIdx.x];
    for (int r = 0; r < items; ++r) {
        const int64_t i = t0 + (int64_t)r * RS + threadIdx.x;
        const int64_t v = i < m ? in[i] : 0;
        int64_t inc = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t u = __shfl_up(inc, o);
            if (lane >= o) inc += u;
        }
        if (lane == 63) sh[wave] = inc;
        __syncthreads();
        int64_t base = carry, tot = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) base += sh[w];
            tot += sh[w];
        }
        if (i < m) out[i] = base + inc - v;
        carry += tot;
        __syncthreads();
    }
}

// ---- host side ------------------------------------------------------------------

struct SortPlan {
    int64_t n;
    int items, ntiles;
};

SortPlan plan_of(int64_t n)
{
    SortPlan p;
    p.n = n;
    // enough items per thread that the offset read (256 x ntiles counts per
    // block) stays short: at most ~1024 tiles
    const int64_t per = (n + (int64_t)RS * 1024 - 1) / ((int64_t)RS * 1024);
    p.items = (int)std::max<int64_t>(4, per);
    p.ntiles = (int)std::max<int64_t>(1, (n + (int64_t)RS * p.items - 1) / ((int64_t)RS * p.items));
    return p;
}

// scratch layout: counters (the scan's), C[3], keys A/B, values A/B
struct SortMem {
    unsigned *counter;
    uint32_t *C[3];
    void *ka, *kb;
    int32_t *va, *vb;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t sort_bytes_of(int64_t n, int key_bytes)
{
    const SortPlan p = plan_of(n);
    const size_t cb = align256(sizeof(uint32_t) * RS * (size_t)p.ntiles);
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    return 256 + 3 * cb + 2 * align256((size_t)key_bytes * nn) + 2 * align256(4 * nn);
}

SortMem carve(void *scratch, int64_t n, int key_bytes)
{
    const SortPlan p = plan_of(n);
    const size_t cb = align256(sizeof(uint32_t) * RS * (size_t)p.ntiles);
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    char *b = static_cast<char *>(scratch);
    SortMem m;
    m.counter = reinterpret_cast<unsigned *>(b);
    b += 256;
    for (int i = 0; i < 3; ++i) {
        m.C[i] = reinterpret_cast<uint32_t *>(b);
        b += cb;
    }
    m.ka = b;
    b += align256((size_t)key_bytes * nn);
    m.kb = b;
    b += align256((size_t)key_bytes * nn);
    m.va = reinterpret_cast<int32_t *>(b);
    b += align256(4 * nn);
    m.vb = reinterpret_cast<int32_t *>(b);
    return m;
}

template <typename K, class Src, class Out>
hipError_t run_sort(Src src, Out out, int64_t n, int bits, void *scratch, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (n > (int64_t)INT32_MAX) return hipErrorInvalidValue;
    const SortPlan p = plan_of(n);
    const SortMem m = carve(scratch, n, (int)sizeof(K));
    const int passes = std::max(1, (bits + 7) / 8);
    K *ka = static_cast<K *>(m.ka), *kb = static_cast<K *>(m.kb);
    int32_t *va = m.va, *vb = m.vb;
    k_rs_load<K, Src><<<p.ntiles, RS, 0, s>>>(src, n, p.items, ka, va, m.C[0], m.C[1]);
    // bits of a (tile, next digit) counter index
    int key_bits = 8;
    while ((1ll << (key_bits - 8)) < p.ntiles) ++key_bits;
    for (int d = 0; d < passes; ++d) {
        uint32_t *cc = m.C[d % 3], *cn = m.C[(d + 1) % 3], *cz = m.C[(d + 2) % 3];
        if (d == passes - 1)
            k_rs_pass<K, true, Out><<<p.ntiles, RS, 0, s>>>(ka, va, kb, vb, out, n, p.items,
                                                            8 * d, key_bits, cc, cn, cz);
        else
            k_rs_pass<K, false, Out><<<p.ntiles, RS, 0, s>>>(ka, va, kb, vb, out, n, p.items,
                                                             8 * d, key_bits, cc, cn, cz);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    return hipGetLastError();
}

}  // namespace

size_t radix_sort_scratch_bytes(int64_t n, int key_bytes) { return sort_bytes_of(n, key_bytes); }

hipError_t radix_sort_init(void *scratch, hipStream_t s)
{
    k_zero_u32<<<1, 64, 0, s>>>(static_cast<uint32_t *>(scratch), 64);
    return hipGetLastError();
}

hipError_t ensure_sort_scratch(DevBuf &b, size_t bytes, hipStream_t s)
{
    const size_t cap0 = b.cap;
    const void *p0 = b.p;
    hipError_t e = ensure(b, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) return e;
    if (b.cap != cap0 || b.p != p0) return radix_sort_init(b.p, s);
    return hipSuccess;
}

hipError_t radix_sort_z(const double *d_xyz, const double *d_q, int64_t n, int32_t *perm,
                        double4 *satom, void *scratch, hipStream_t s)
{
    return run_sort<uint64_t>(ZSource{d_xyz}, ZOutput{d_xyz, d_q, perm, satom}, n, 64, scratch,
                              s);
}

hipError_t radix_sort_u32(const uint32_t *keys, int64_t n, int bits, uint32_t *keys_out,
                          int32_t *idx_out, void *scratch, hipStream_t s)
{
    return run_sort<uint32_t>(KeyIotaSource{keys}, PairOutput{keys_out, idx_out}, n, bits,
                              scratch, s);
}

size_t scan_scratch_bytes(int64_t m)
{
    const int items = (int)std::max<int64_t>(4, (m + (int64_t)RS * 1024 - 1) / ((int64_t)RS * 1024));
    const int64_t nb = std::max<int64_t>(1, (m + (int64_t)RS * items - 1) / ((int64_t)RS * items));
    return 256 + align256(sizeof(int64_t) * (size_t)nb);
}

hipError_t scan_exclusive_i64(const int64_t *in, int64_t *out, int64_t m, void *scratch,
                              hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    const int items = (int)std::max<int64_t>(4, (m + (int64_t)RS * 1024 - 1) / ((int64_t)RS * 1024));
    const int64_t nb = std::max<int64_t>(1, (m + (int64_t)RS * items - 1) / ((int64_t)RS * items));
    unsigned *counter = static_cast<unsigned *>(scratch);
    int64_t *part = reinterpret_cast<int64_t *>(static_cast<char *>(scratch) + 256);
    k_scan_up<<<(unsigned)nb, RS, 0, s>>>(in, m, items, part, counter);
    k_scan_down<<<(unsigned)nb, RS, 0, s>>>(in, m, items, part, out);
    return hipGetLastError();
}

}  // namespace qem
