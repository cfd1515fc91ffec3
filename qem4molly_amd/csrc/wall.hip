// Substrate_LJ walls on gfx950: a 12-6 Lennard-Jones wall force on each
// atom's z coordinate (src/Substrate_LJ.jl:7-46).  One thread per atom and
// every wall of the call in that thread, so the pass reads z (and the z force
// when accumulating) once and writes it once: HBM-bound, microseconds at
// N = 100,000.
#include "qem_internal.h"

namespace qem {
namespace {

template <int FS>
__global__ void k_substrate_lj(WallSet ws, int64_t n, const double *__restrict__ xyz,
                               int accumulate, double *__restrict__ f,
                               uint32_t *__restrict__ flag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double z = xyz[3 * i + 2];
    double acc = accumulate ? f[FS * i + FS - 1] : 0.0;
    bool bad = false;
    for (int k = 0; k < ws.nw; ++k) {
        const qem_wall &w = ws.w[k];
        const double dz = z - w.sub_pos;
        // the reference's @assert inter.direction * dz > 0 (src/Substrate_LJ.jl:37)
        bad |= !((double)w.direction * dz > 0.0);
        const double dcut = (w.semantics & QEM_WALL_SEM_ABS_CUTOFF) ? fabs(dz) : dz;
        if (dcut < w.cutoff) {
            // substrate_LJ_force (src/Substrate_LJ.jl:19-22)
            const double invr2 = 1.0 / (dz * dz);
            const double s2 = w.sigma * w.sigma * invr2;
            const double six = s2 * s2 * s2;
            double fz = (24.0 * w.eps * invr2) * (2.0 * six * six - six);
            if (w.semantics & QEM_WALL_SEM_TRUE_FORCE) fz *= fabs(dz);
            fz *= (double)w.direction;
            // the walls add one after another, as separate interactions do
            acc = __dadd_rn(acc, fz);
        }
    }
    if (FS == 3 && !accumulate) {
        f[FS * i] = 0.0;
        f[FS * i + 1] = 0.0;
    }
    f[FS * i + FS - 1] = acc;
    // every offending thread stores the same value: this call's epoch
    if (bad) *flag = ws.epoch;
}

__global__ void k_zero_u32(uint32_t *__restrict__ p) { *p = 0u; }

__global__ void k_add_z(int64_t n, const double *__restrict__ z, double *__restrict__ f)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[3 * i + 2] = __dadd_rn(f[3 * i + 2], z[i]);
}

}  // namespace

hipError_t launch_zero_flag(uint32_t *d_flag, hipStream_t s)
{
    k_zero_u32<<<1, 1, 0, s>>>(d_flag);
    return hipGetLastError();
}

hipError_t launch_substrate_lj(const WallSet &ws, int64_t n, const double *d_xyz, double *d_f,
                               int fstride, int accumulate, uint32_t *d_flag, hipStream_t s)
{
    if (n <= 0 || ws.nw <= 0) return hipSuccess;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (fstride == 3)
        k_substrate_lj<3><<<g, 256, 0, s>>>(ws, n, d_xyz, accumulate, d_f, d_flag);
    else if (fstride == 1)
        k_substrate_lj<1><<<g, 256, 0, s>>>(ws, n, d_xyz, accumulate, d_f, d_flag);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_add_z(int64_t n, const double *d_z, double *d_f, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_add_z<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, d_z, d_f);
    return hipGetLastError();
}

}  // namespace qem
