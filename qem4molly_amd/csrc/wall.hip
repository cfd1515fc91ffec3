// Substrate_LJ walls on gfx950: a 12-6 Lennard-Jones wall force on each
// atom's z coordinate (src/Substrate_LJ.jl:7-46).  One thread per atom; the
// whole pass reads 24 B and writes 24 B per atom (HBM-bound, microseconds
// at N = 100,000).
#include "qem_internal.h"

namespace qem {
namespace {

__global__ void k_substrate_lj(qem_wall w, int64_t n, const double *__restrict__ xyz,
                               int accumulate, double *__restrict__ f, int *__restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dz = xyz[3 * i + 2] - w.sub_pos;
    // the reference's @assert inter.direction * dz > 0 (src/Substrate_LJ.jl:37):
    // every offending thread stores the same value
    if (!((double)w.direction * dz > 0.0)) *bad = 1;
    const double dcut = (w.semantics & QEM_WALL_SEM_ABS_CUTOFF) ? fabs(dz) : dz;
    double fz = 0.0;
    if (dcut < w.cutoff) {
        // substrate_LJ_force (src/Substrate_LJ.jl:19-22)
        const double invr2 = 1.0 / (dz * dz);
        const double six = (w.sigma * w.sigma * invr2) * (w.sigma * w.sigma * invr2) *
                           (w.sigma * w.sigma * invr2);
        fz = (24.0 * w.eps * invr2) * (2.0 * six * six - six);
        if (w.semantics & QEM_WALL_SEM_TRUE_FORCE) fz *= fabs(dz);
        fz *= (double)w.direction;
    }
    if (accumulate) {
        f[3 * i + 2] += fz;
    } else {
        f[3 * i] = 0.0;
        f[3 * i + 1] = 0.0;
        f[3 * i + 2] = fz;
    }
}

__global__ void k_fill_i32_one(int *__restrict__ p) { *p = 0; }

}  // namespace

hipError_t launch_substrate_lj(const qem_wall &w, int64_t n, const double *d_xyz, double *d_f,
                               int accumulate, int *d_bad, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_fill_i32_one<<<1, 1, 0, s>>>(d_bad);
    k_substrate_lj<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(w, n, d_xyz, accumulate, d_f,
                                                                d_bad);
    return hipGetLastError();
}

}  // namespace qem
