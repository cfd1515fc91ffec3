/*
 * qem.h -- C ABI of libqem_hip.so, the MI355X (gfx950) implementation of the
 * QEM4Molly quasi-Ewald force path (short range: Gauss-Legendre quadrature of
 * the image-charge Sommerfeld integrals over a 2-D cell list; long range:
 * random-batch k-space sum with z-sorted prefix sums and the k = 0 term).
 *
 * Every entry point replaces a piece of the reference's Molly plugin
 * interface; the reference file:line it stands in for is given beside it
 * (paths relative to the QEM4Molly repository root).
 *
 * Conventions
 *   - Every function returns an int status (QEM_OK = 0); no exception or
 *     abort crosses the ABI.  qem_last_error(ctx) gives the message.
 *   - Indices are 0-based (a Julia host adds/subtracts 1).
 *   - Coordinates are N x 3 row-major doubles (x0,y0,z0,x1,...), the memory
 *     layout of Julia's Vector{SVector{3,Float64}}; charges are N doubles;
 *     forces are N x 3 doubles, overwritten.
 *   - Host-pointer entry points are synchronous and validate their inputs.
 *     *_dev entry points take device pointers and a hipStream_t passed as
 *     void* (NULL = the context's own non-blocking stream; QEM_STREAM_NULL =
 *     HIP's null stream, which a caller working on the default stream, such
 *     as torch's default stream, needs for its own work to be ordered with
 *     the library's); they enqueue work and return without synchronising,
 *     and do not validate array contents.
 *   - A context is not thread-safe: use one per host thread or stream.
 *     Device workspaces belong to the context and grow on demand.
 */
#ifndef QEM_H
#define QEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- streams ------------------------------------------------------------ */
#define QEM_STREAM_NULL ((void *)1)   /* HIP's null (default) stream */

/* ---- status codes ------------------------------------------------------ */
#define QEM_OK                 0
#define QEM_ERR_INVALID_ARG    1  /* null pointer, negative size, bad enum   */
#define QEM_ERR_DOMAIN         2  /* NaN/Inf input, z outside (0, Lz), ...   */
#define QEM_ERR_HIP            3  /* HIP runtime error                        */
#define QEM_ERR_NOMEM          4  /* device or host allocation failed         */
#define QEM_ERR_UNSUPPORTED    5  /* parameters outside what is implemented   */
#define QEM_ERR_CAPACITY       6  /* caller buffer too small                  */

/* ---- precision modes ------------------------------------------------------ */
#define QEM_PREC_FP64   0  /* every operation in double precision             */
#define QEM_PREC_MIXED  1  /* FP32 nodes, Bessel, exp, sincos; FP64 sums      */

/* ---- semantics switches (0 = the reference as written) --------------------
 * QEM_SEM_SWAP_LIMITS   gaussian integrals on k_f2, point-charge integrals on
 *                       k_f1 (src/greens_function.jl:26-27 assigns them the
 *                       other way round, src/Force_short.jl:48,51,64,67)
 * QEM_SEM_MINIMAGE_DIR  xy force direction from the minimum-image separation
 *                       (src/Force_short.jl:57-58 uses raw x_i - x_j)
 * QEM_SEM_NEWTON_IMAGES z force on j from d/dz_j: images 2 and 3 push i and j
 *                       the same way (src/QEM_short.jl:47 negates all of F_ij)
 * QEM_SEM_POLE_SPLIT    for gamma_1 gamma_2 > 1, where the image-series
 *                       denominator gamma_1 gamma_2 e^{-2 k Lz} - 1 vanishes
 *                       at k0 = ln(gamma_1 gamma_2)/(2 Lz): the short-range
 *                       integrals (all images together, limits
 *                       max(-ln(acc)/(2 Lz), 2.1 k0) for the point-charge and
 *                       max(sqrt(-4 alpha ln acc), 2.1 k0) for the gaussian
 *                       ones) as principal values by the segment rule of
 *                       src/python_lib/Force_short_norm_njit.py:11-414 and
 *                       green_norm_njit.py:13-32, middle half-width pole_eps.
 *                       Without it (as written, src/Force_short.jl) the
 *                       Gauss-Legendre nodes run across the pole; that is
 *                       accepted in FP64 only.  No effect for g1 g2 <= 1. */
#define QEM_SEM_SWAP_LIMITS    1u
#define QEM_SEM_MINIMAGE_DIR   2u
#define QEM_SEM_NEWTON_IMAGES  4u
#define QEM_SEM_CORRECTED      7u
#define QEM_SEM_POLE_SPLIT     8u

/* ---- long-range parts (qem_long_terms) ------------------------------------ */
#define QEM_LONG_NONSORT  1u  /* Fl_non_sort  src/Force_long.jl:5-70    */
#define QEM_LONG_SORT     2u  /* Fl_sort      src/Force_long.jl:72-127  */
#define QEM_LONG_SELF     4u  /* Flz_self     src/Force_long.jl:164-184 */
#define QEM_LONG_K0       8u  /* Flz_sort_k0  src/Force_long.jl:130-161 */
#define QEM_LONG_ALL     15u

/* Parameters: the union of the QEM_short and QEM_long keyword arguments
 * (src/QEM_short.jl:19, src/QEM_long.jl:19) plus precision and semantics. */
typedef struct qem_params {
    double Lx, Ly, Lz;      /* box; periodic in x, y; interfaces at z=0, Lz */
    double r_cutoff;        /* QEM_short r_cutoff                            */
    double alpha;           /* splitting parameter                           */
    double gamma_1;         /* dielectric contrast at z = 0  (gamma_down)    */
    double gamma_2;         /* dielectric contrast at z = Lz (gamma_up)      */
    double eps_0;
    double accuracy;        /* sets the quadrature limits and the K set      */
    int32_t N_t;            /* Gauss-Legendre order, 1..256                  */
    int32_t iter_period;    /* neighbour-list rebuild period (>= 1)          */
    int32_t rbe_p;          /* random-batch size P                           */
    int32_t precision;      /* QEM_PREC_FP64 | QEM_PREC_MIXED                */
    uint32_t semantics;     /* QEM_SEM_* bitmask                             */
    int32_t device;         /* HIP device ordinal                            */
    double pole_eps;        /* QEM_SEM_POLE_SPLIT middle half-width eps
                               (<= 0: k0 / 20); the reference's INCAR 'eps' */
} qem_params;

typedef struct qem_ctx qem_ctx;

/* Reference keyword defaults: iter_period=100, gamma=0, eps_0=1,
 * accuracy=1e-6, N_t=30 (src/QEM_short.jl:19), rbe_p=30, alpha=1
 * (src/QEM_long.jl:19); Lx=Ly=Lz=1, r_cutoff=1, FP64, as written,
 * pole_eps=0 (k0/20).
 * gamma_1 gamma_2 may take any finite value.  For gamma_1 gamma_2 > 1,
 * mixed precision needs QEM_SEM_POLE_SPLIT, and with it qem_create returns
 * QEM_ERR_UNSUPPORTED when the rule has no nodes for a segment
 * (floor(N_t k0 / k_f) < 1), as the reference's leggauss(0) raises.  The
 * energy and HVP entry points need |gamma_1 gamma_2| < 1. */
void qem_default_params(qem_params *p);

/* Stands in for the QEM_short / QEM_long constructors
 * (src/QEM_short.jl:19-25, src/QEM_long.jl:19-30).  The K set
 * (src/K_set_generator.jl:5-45) is built here when rbe_p > 0. */
int qem_create(const qem_params *p, qem_ctx **out);
int qem_destroy(qem_ctx *ctx);
const char *qem_last_error(const qem_ctx *ctx);
const char *qem_status_string(int status);
int qem_get_params(const qem_ctx *ctx, qem_params *p);

/* QEM_inter's derivation of alpha, r_cutoff and k_cutoff from n_atoms and s
 * (src/Quasi_Ewald_Method.jl:3-24). */
int qem_inter_params(int64_t n_atoms, double Lx, double Ly, double s,
                     int32_t rbe_mode, double *alpha, double *r_cutoff,
                     double *k_cutoff);

/* ---- K set / random batch (src/K_set_generator.jl:5-45,
 *      src/Force_long.jl:235) ------------------------------------------- */
int qem_kset_size(qem_ctx *ctx, int64_t *n);
/* kxkyk: n x 3 (k_x, k_y, |k|); prob: n Float32 weights; sum_K scalar */
int qem_kset_copy(qem_ctx *ctx, double *kxkyk, float *prob, double *sum_K);
/* Draws P indices with replacement, probability prob[i]/sum(prob), with a
 * seeded alias sampler (stands in for StatsBase.sample(K_set, Prob, rbe_p),
 * src/Force_long.jl:235, whose stream cannot be reproduced outside Julia). */
int qem_sample_k(qem_ctx *ctx, uint64_t seed, int64_t P, int64_t *k_idx);

/* ---- short range --------------------------------------------------------- */
/* The 2-D neighbour list (CellListMap.neighborlist at src/QEM_short.jl:34-35):
 * pairs i < j with minimum-image xy distance <= r_cutoff.  Writes
 * min(npairs, cap) pairs (2 x int32 each) and sets *npairs to the total. */
int qem_neighbor_list(qem_ctx *ctx, int64_t n, const double *xyz, int64_t cap,
                      int32_t *pairs, int64_t *npairs);

/* Molly.forces(::QEM_short, sys) (src/QEM_short.jl:27-59): advances the
 * context's step counter, rebuilds the neighbour list when
 * (n_steps - 1) % iter_period == 0 (or when n changes), evaluates every
 * listed pair still within r_cutoff, adds the self-image term. */
int qem_short_forces(qem_ctx *ctx, int64_t n, const double *xyz,
                     const double *q, double *f);
/* Same accumulation over a caller-given pair list (no step counter). */
int qem_short_forces_pairs(qem_ctx *ctx, int64_t n, const double *xyz,
                           const double *q, int64_t npairs,
                           const int32_t *pairs, double *f);
/* F_short_ij (src/Force_short.jl:25-78) for m independent pairs:
 * f_i is m x 3 (force on i), fz_j is m (z force on j; may be NULL). */
int qem_short_pair(qem_ctx *ctx, int64_t m, const double *qi, const double *qj,
                   const double *ri, const double *rj, int32_t single,
                   double *f_i, double *fz_j);
/* F_short_i (src/Force_short.jl:82-107) for n atoms: f is n x 3. */
int qem_short_self(qem_ctx *ctx, int64_t n, const double *q,
                   const double *xyz, int32_t single, double *f);

/* ---- long range ---------------------------------------------------------- */
/* sortperm of z (src/QEM_long.jl:39-40): ascending, ties by index, 0-based. */
int qem_sortperm_z(qem_ctx *ctx, int64_t n, const double *xyz, int64_t *zlist);
/* Molly.forces(::QEM_long, sys) (src/QEM_long.jl:32-44 -> F_long,
 * src/Force_long.jl:201-242, random-batch branch) with the batch given as
 * P indices into the K set: F = F_k0 + (sum_K / P) sum_p F_l_total(k_p). */
int qem_long_forces(qem_ctx *ctx, int64_t n, const double *xyz,
                    const double *q, const int64_t *k_idx, int64_t P,
                    double *f);
/* Explicit k vectors (P x 3: k_x, k_y, |k|): F = [K0] + weight *
 * sum_p (selected parts of F_l_total(k_p)); weight < 0 requests
 * exp(-k^2/(4 alpha)) per k (the full-sum weighting).  Covers Fl_non_sort,
 * Fl_sort, Flz_self, Flz_sort_k0 and F_l_total individually. */
int qem_long_terms(qem_ctx *ctx, int64_t n, const double *xyz,
                   const double *q, const double *kvecs, int64_t P,
                   double weight, uint32_t parts, double *f);

/* Random-batch diagnostics (no counterpart in the reference, whose batch
 * estimator is src/Force_long.jl:201-218,234-242).  For the batch k_idx
 * (P >= 2), with X_ip = (sum_K / P) F_l_total(k_p) on atom i, the k = 0 term
 * left out (it is not sampled):
 *   var_i = (P sum_p |X_ip|^2 - |sum_p X_ip|^2) / (P - 1),
 * the unbiased estimate, from the batch itself, of E|F_i - E F_i|^2 for the
 * batch force F_i (whose mean over batches is the full k sum weighted by
 * exp(-k^2/(4 alpha))).  f (may be NULL) receives qem_long_forces for the
 * same batch.  P <= 65536 (QEM_ERR_UNSUPPORTED beyond). */
int qem_long_variance(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                      const int64_t *k_idx, int64_t P, double *f, double *var);
int qem_long_variance_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                          const int64_t *k_idx, int64_t P, double *d_f, double *d_var,
                          void *stream);

/* ---- both halves (one MD step's electrostatics) ------------------------ */
int qem_forces(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
               const int64_t *k_idx, int64_t P, double *f);

/* ---- device-pointer variants -------------------------------------------- */
int qem_short_forces_dev(qem_ctx *ctx, int64_t n, const double *d_xyz,
                         const double *d_q, double *d_f, void *stream);
/* k_idx is a host array (P is small); it is staged through pinned memory. */
int qem_long_forces_dev(qem_ctx *ctx, int64_t n, const double *d_xyz,
                        const double *d_q, const int64_t *k_idx, int64_t P,
                        double *d_f, void *stream);
int qem_forces_dev(qem_ctx *ctx, int64_t n, const double *d_xyz,
                   const double *d_q, const int64_t *k_idx, int64_t P,
                   double *d_f, void *stream);
/* qem_long_terms on device data (kvecs stays a host array, P x 3); with
 * accumulate != 0 the terms are added to d_f.  With weight < 0 over the whole
 * K set this is the deterministic full k sum (rbe_mode = false). */
int qem_long_terms_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                       const double *kvecs, int64_t P, double weight, uint32_t parts,
                       double *d_f, int32_t accumulate, void *stream);

/* ---- Substrate_LJ walls (src/Substrate_LJ.jl:7-46) ----------------------
 * A 12-6 Lennard-Jones wall at z = sub_pos acting on every atom's z
 * coordinate, the other general interaction of the reference's MD setup
 * (test/QEM_test.jl:20-28).  As written (src/Substrate_LJ.jl:24-46), with
 * dz = z_i - sub_pos: atoms with dz < cutoff (signed: a wall with
 * direction -1 acts at every distance) get
 *   F_z = direction * 24 eps / dz^2 * (2 (sigma/dz)^12 - (sigma/dz)^6),
 * which is -dU/d(dz) / dz for U = 4 eps ((sigma/dz)^12 - (sigma/dz)^6);
 * direction * dz <= 0 for any atom fails the reference's @assert. */
#define QEM_WALL_SEM_ABS_CUTOFF 1u  /* cutoff on |dz| instead of signed dz       */
#define QEM_WALL_SEM_TRUE_FORCE 2u  /* F_z = -dU/dz (as written: -dU/dz / |dz|) */

typedef struct qem_wall {
    double cutoff, sigma, eps, sub_pos;
    int32_t direction;   /* +1: wall below the atoms, -1: above */
    uint32_t semantics;  /* QEM_WALL_SEM_* bits, 0 = as written  */
} qem_wall;

/* One wall's forces (host pointers, n x 3); accumulate != 0 adds them to f.
 * QEM_ERR_DOMAIN if direction * dz <= 0 for some atom (f is untouched). */
int qem_substrate_lj(qem_ctx *ctx, const qem_wall *w, int64_t n, const double *xyz,
                     double *f, int32_t accumulate);
/* Device pointers.  The domain check reads a device flag, so the call
 * synchronises the stream; on QEM_ERR_DOMAIN d_f holds partial results. */
int qem_substrate_lj_dev(qem_ctx *ctx, const qem_wall *w, int64_t n, const double *d_xyz,
                         double *d_f, int32_t accumulate, void *stream);
/* All general interactions of a Molly system in one device-resident call
 * (no counterpart in the reference, whose Molly loop moves coordinates and
 * forces across per interaction): d_f = sum of the walls + QEM_short (step
 * counter and list rebuilds as qem_short_forces) + QEM_long for the k batch,
 * summed in the reference's order (walls, QEM_short, QEM_long;
 * test/QEM_test.jl:28).  nwalls may be 0.  With walls, the call waits for
 * the wall pass's domain flag only, after the QEM kernels are queued behind
 * it, so consecutive calls keep the stream busy; on QEM_ERR_DOMAIN d_f holds
 * the full sum all the same. */
int qem_general_forces_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                           const int64_t *k_idx, int64_t P, const qem_wall *walls,
                           int32_t nwalls, double *d_f, void *stream);

/* ---- adjoint (no counterpart in the reference, which has forces only) ----
 * Energy E whose negative gradient at frozen quadrature limits is the force
 * (with QEM_SEM_CORRECTED the physical slab energy, split by
 * exp(-k^2/(4 alpha))): pair terms are the J0 transforms of Gamma_2 and
 * Gamma_1 e^{-k^2/4a} (src/greens_function.jl:68-88) over the nodes and
 * limits of src/Force_short.jl:48-67, self terms the images of each charge,
 * long range the structure-factor energy of each sampled k and the k = 0
 * term.  dE_dq (N, may be NULL) is the per-atom potential; dE_dgamma (2,
 * may be NULL) = (dE/dgamma_1, dE/dgamma_2).  The short range uses the
 * context's neighbour list (built if absent; the step counter is untouched). */
int qem_energy(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
               const int64_t *k_idx, int64_t P, double *E, double *dE_dq,
               double *dE_dgamma);
/* d_dE_dq is device memory (may be NULL); E and dE_dgamma are host outputs
 * (the call synchronises the stream). */
int qem_energy_dev(qem_ctx *ctx, int64_t n, const double *d_xyz,
                   const double *d_q, const int64_t *k_idx, int64_t P,
                   double *d_dE_dq, double *E, double *dE_dgamma, void *stream);
/* Forces, E, dE/dq and dE/dgamma of one step in one call: qem_forces_dev
 * and qem_energy_dev together (the step counter and neighbour list advance
 * as in qem_forces_dev).  In mixed precision the pair quadrature runs once
 * for both, its node loops also forming the energy sums, and the long
 * range's chunk sums serve both the force and the potential; the forces are
 * those of qem_forces_dev bit for bit.  d_f (n x 3) and d_dE_dq (n, may be
 * NULL) are device memory, E and dE_dgamma host outputs (the call
 * synchronises the stream).  Needs |gamma_1 gamma_2| < 1. */
int qem_forces_energy_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                          const int64_t *k_idx, int64_t P, double *d_f, double *d_dE_dq,
                          double *E, double *dE_dgamma, void *stream);
int qem_forces_energy(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                      const int64_t *k_idx, int64_t P, double *f, double *E, double *dE_dq,
                      double *dE_dgamma);
/* Hessian-vector product hv = -(dF/dx) v of the force as computed (short +
 * long range for the k batch; the k = 0 term is piecewise constant in z),
 * differentiating the quadrature through the z-dependent limits.  For a
 * conservative force this is H v; v and hv are N x 3. */
int qem_hvp(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
            const int64_t *k_idx, int64_t P, const double *v, double *hv);
int qem_hvp_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                const int64_t *k_idx, int64_t P, const double *d_v, double *d_hv,
                void *stream);
/* The adjoint step with one pair quadrature in the forward (no reference
 * counterpart; for callers that take forces, energy and then a
 * Hessian-vector product at the same positions, as an autograd backward
 * does).  qem_forces_energy_save(_dev) is qem_forces_energy(_dev) whose
 * pair kernel also keeps, in mixed precision, every pair's force-tangent
 * node sums in the context (224 bytes per pair of device memory, grown
 * once); qem_hvp_saved(_dev) is qem_hvp(_dev) that combines those sums
 * instead of running the tangent quadrature.  The caller passes
 * qem_hvp_saved(_dev) the positions, charges and k indices of the last save
 * call (the library cannot tell changed positions apart); it uses the sums
 * when they belong to the current atom count and neighbour list (no
 * rebuild since) and the context is mixed precision, and otherwise
 * computes the product as qem_hvp(_dev) does.  Any forces + energy call
 * without saving drops the sums.  The product equals qem_hvp's within the
 * mixed-precision tolerance (the sums are taken at the forward's nodes). */
int qem_forces_energy_save_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                               const int64_t *k_idx, int64_t P, double *d_f, double *d_dE_dq,
                               double *E, double *dE_dgamma, void *stream);
int qem_forces_energy_save(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                           const int64_t *k_idx, int64_t P, double *f, double *E,
                           double *dE_dq, double *dE_dgamma);
int qem_hvp_saved_dev(qem_ctx *ctx, int64_t n, const double *d_xyz, const double *d_q,
                      const int64_t *k_idx, int64_t P, const double *d_v, double *d_hv,
                      void *stream);
int qem_hvp_saved(qem_ctx *ctx, int64_t n, const double *xyz, const double *q,
                  const int64_t *k_idx, int64_t P, const double *v, double *hv);

/* ---- allocation-free long range ------------------------------------------
 * No counterpart in the reference.  qem_long_forces_dev for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The call allocates nothing (no
 * hipMalloc, hipFree, pinned staging or other allocating routine): its work
 * buffers (z sort, per-k constants, chunk sums, per-group partial forces)
 * live in one fixed device array of QEM_NOALLOC_SCRATCH_BYTES that the
 * library holds on each GPU it runs on, from the first call there until the
 * process ends, and the k indices travel to the device as kernel arguments.
 * Limits: n <= QEM_NOALLOC_MAX_N and P <= QEM_NOALLOC_MAX_P
 * (QEM_ERR_CAPACITY beyond); the context's K set must exist (rbe_p > 0 at
 * qem_create, or one qem_kset_size call; QEM_ERR_UNSUPPORTED otherwise).
 * Every call on one GPU uses that GPU's one array, whatever the context:
 * calls that overlap on one GPU race.  Same results as qem_long_forces_dev,
 * bit for bit. */
#define QEM_NOALLOC_MAX_N 131072
#define QEM_NOALLOC_MAX_P 512
#define QEM_NOALLOC_SCRATCH_BYTES ((size_t)72 << 20)
int qem_long_forces_noalloc_dev(qem_ctx *ctx, int64_t n, const double *d_xyz,
                                const double *d_q, const int64_t *k_idx, int64_t P,
                                double *d_f, void *stream);

/* ---- introspection ------------------------------------------------------ */
/* device bytes held by the context's workspaces (they only grow: the
 * current value is the peak) */
int qem_device_bytes(qem_ctx *ctx, int64_t *bytes);
/* pairs in the current neighbour list, its build count, and the step count */
int qem_short_state(qem_ctx *ctx, int64_t *npairs, int64_t *n_builds,
                    int64_t *n_steps);

/* Per-phase GPU time, measured with HIP events recorded on the stream each
 * phase is launched on (no counterpart in the reference; for benchmarks).
 * qem_timing_read synchronises, returns the summed milliseconds and launch
 * counts per phase since the last reset, and optionally resets them. */
#define QEM_PHASE_NEIGHBOR  0  /* neighbour-list build            */
#define QEM_PHASE_PAIR      1  /* short-range pair quadrature     */
#define QEM_PHASE_GATHER    2  /* per-atom gather + self term     */
#define QEM_PHASE_LONG      3  /* z sort + k-space passes         */
#define QEM_NPHASES         4
int qem_timing_enable(qem_ctx *ctx, int32_t on);
int qem_timing_read(qem_ctx *ctx, double *ms, int64_t *calls, int32_t reset);

/* Gauss-Legendre rule on [-1, 1], ascending nodes (stands in for
 * GaussQuadrature.legendre(n), used at src/Gaussian_int.jl:11); 1 <= n <= 256. */
int qem_gauss_legendre(int32_t n, double *x, double *w);
/* The tabulated J0/J1 pair of the context's kernels, evaluated on the GPU
 * for n arguments in [0, table range]: the FP32 cubic table in
 * QEM_PREC_MIXED contexts, the FP64 degree-7 table in QEM_PREC_FP64 ones
 * (no counterpart in the reference: it replaces SpecialFunctions.besselj0/
 * besselj1 at src/Force_short.jl:5-21).  QEM_ERR_UNSUPPORTED when the
 * context has no table. */
int qem_bessel01_table(qem_ctx *ctx, int64_t n, const double *x, double *j0,
                       double *j1);

/* ---- device memory utilities (no counterpart in the reference) ----------
 * For hosts without their own GPU allocator (ctypes, Julia's ccall). */
int qem_dev_malloc(void **ptr, size_t bytes);
int qem_dev_free(void *ptr);
int qem_memcpy_htod(void *dst, const void *src, size_t bytes);
int qem_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int qem_device_synchronize(void);
/* copies bytes with a plain load/store kernel (traffic calibration) */
int qem_dev_copy_kernel(void *dst, const void *src, size_t bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* QEM_H */
