// bx_gravity.hip -- gravity from a given potential: calc_gforce
// (utils.f90:549-571), the external point mass of external_grav
// (gravity.f90:1477-1549) with its mass-moment reduction, and the
// configuration calls.  The potential comes from the caller: the reference's
// self-gravity solver is not part of this library.
#include <cmath>
#include <vector>

#include "bx_internal.h"
#include "bx_math.h"

namespace bx {

constexpr int kGThreads = 256;

__device__ __forceinline__ double centre_of(double d, int i, int n) { return d * ((double)i - ((double)n + 1.0) / 2.0); }

// utils.f90:549-571 on interior cells (boundary 0); neighbours as ineigh
// (init_grid.f90): x+-1, y+-nx, z+-nx*ny, the cell itself along a flat z.
__global__ __launch_bounds__(kGThreads) void k_calc_gforce(int64_t n, int nx, int ny, int nz, bool flatz,
                                                          const double *__restrict__ phi, double *__restrict__ gf,
                                                          double dx, double dy, double dz)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int64_t pl = (int64_t)nx * ny;
    const int iz = (int)(c / pl) + 1, iy = (int)((c % pl) / nx) + 1, ix = (int)(c % nx) + 1;
    const bool b1 = ix == 1 || ix == nx || iy == 1 || iy == ny || (nz >= 6 && (iz == 1 || iz == nz));
    const bool b2 = ix == 2 || ix == nx - 1 || iy == 2 || iy == ny - 1 || (nz >= 6 && (iz == 2 || iz == nz - 1));
    if (b1 || b2) return;
    const int64_t sz = flatz ? 0 : pl;
    gf[c] = -(phi[c + 1] - phi[c - 1]) / (2.0 * dx);
    gf[n + c] = -(phi[c + nx] - phi[c - nx]) / (2.0 * dy);
    gf[2 * n + c] = -(phi[c + sz] - phi[c - sz]) / (2.0 * dz);
}

// gravity.f90:1492-1494 terms of cell c
__device__ __forceinline__ void moment_terms(int64_t c, int nx, int ny, int nz, const double *__restrict__ rho,
                                             double dx, double dy, double dz, double &tx, double &ty, double &tz)
{
    const int64_t pl = (int64_t)nx * ny;
    const int iz = (int)(c / pl) + 1, iy = (int)((c % pl) / nx) + 1, ix = (int)(c % nx) + 1;
    const double r = rho[c];
    tx = centre_of(dx, ix, nx) * r * dx * dy * dz;
    ty = centre_of(dy, iy, ny) * r * dx * dy * dz;
    tz = centre_of(dz, iz, nz) * r * dx * dy * dz;
}

// The mass moments summed in cell order, as the reference's reduction does on
// one OpenMP thread: one wave computes the terms of 64 cells at a time and
// folds them into the running sums lane by lane (readlane, no LDS round trip),
// while the next 64 cells load.  The private sum starts at +0 and is added to
// the shared zero at the end (OpenMP reduction semantics).
__device__ __forceinline__ double lane_value(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__global__ __launch_bounds__(64) void k_moments_seq(int64_t n, int nx, int ny, int nz, const double *__restrict__ rho,
                                                   double dx, double dy, double dz, double *__restrict__ sums)
{
    const int lane = threadIdx.x;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    double tx = 0.0, ty = 0.0, tz = 0.0;
    if (lane < n) moment_terms(lane, nx, ny, nz, rho, dx, dy, dz, tx, ty, tz);
    for (int64_t b = 0; b < n; b += 64) {
        double nxt[3] = {0.0, 0.0, 0.0};
        if (b + 64 + lane < n) moment_terms(b + 64 + lane, nx, ny, nz, rho, dx, dy, dz, nxt[0], nxt[1], nxt[2]);
        const int m = n - b < 64 ? (int)(n - b) : 64;
        for (int l = 0; l < m; l++) {
            sx = sx + lane_value(tx, l);
            sy = sy + lane_value(ty, l);
            sz = sz + lane_value(tz, l);
        }
        tx = nxt[0];
        ty = nxt[1];
        tz = nxt[2];
    }
    if (lane == 0) {
        sums[0] = 0.0 + sx;
        sums[1] = 0.0 + sy;
        sums[2] = 0.0 + sz;
    }
}

// Fixed-order tree: block partials (pairwise in shared memory), then one block
// folds the partials (strided per thread, then pairwise).  Deterministic, not
// the sequential order.
__device__ __forceinline__ void block_tree3(double v[3], double (*sh)[kGThreads])
{
    for (int k = 0; k < 3; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = kGThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] = sh[k][threadIdx.x] + sh[k][threadIdx.x + off];
        __syncthreads();
    }
    for (int k = 0; k < 3; k++) v[k] = sh[k][0];
}

__global__ __launch_bounds__(kGThreads) void k_moments_part(int64_t n, int nx, int ny, int nz,
                                                           const double *__restrict__ rho, double dx, double dy,
                                                           double dz, double *__restrict__ part)
{
    __shared__ double sh[3][kGThreads];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (c < n) moment_terms(c, nx, ny, nz, rho, dx, dy, dz, v[0], v[1], v[2]);
    block_tree3(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) part[(int64_t)k * gridDim.x + blockIdx.x] = v[k];
}

__global__ __launch_bounds__(kGThreads) void k_moments_final(int nb, const double *__restrict__ part,
                                                            double *__restrict__ sums)
{
    __shared__ double sh[3][kGThreads];
    // thread t folds partials t, t+256, ... in order, then one tree over the threads
    double v[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += kGThreads)
        for (int k = 0; k < 3; k++) v[k] = v[k] + part[(int64_t)k * nb + b];
    block_tree3(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) sums[k] = 0.0 + v[k];
}

// utils.f90:486-510: cubic-spline softening and its derivative
__device__ __forceinline__ double wspline3(double x)
{
    if (x < 1.0) return x * (1.5 * x - 0.75 * x * x);
    if (x < 2.0) return 1.0 - 0.25 * ((2.0 - x) * (2.0 - x) * (2.0 - x));
    return 1.0;
}
__device__ __forceinline__ double derivwspline3(double x)
{
    if (x < 1.0) return x * (3.0 - 2.25 * x);
    if (x < 2.0) return 0.75 * ((2.0 - x) * (2.0 - x));
    return 0.0;
}

// gravity.f90:1500-1541 on every cell (2-D: the point mass acts in x and y)
__global__ __launch_bounds__(kGThreads) void k_external_grav(int64_t n, int nx, int ny, double dx, double dy,
                                                            const double *__restrict__ sums, double mass,
                                                            double radius, double *__restrict__ gf)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double xobj = -sums[0] / mass, yobj = -sums[1] / mass;
    const int ix = (int)(c % nx) + 1, iy = (int)((c / nx) % ny) + 1;
    const double x = centre_of(dx, ix, nx) - xobj, y = centre_of(dy, iy, ny) - yobj;
    const double r = sqrt(x * x + y * y);
    const double h1 = r / radius;
    const double w = wspline3(h1), dw = derivwspline3(h1);
    const double r3 = (r * r) * r, r2 = r * r;
    gf[c] = gf[c] - mass * x / r3 * w + mass * dw * x / (r2 * radius);
    gf[n + c] = gf[n + c] - mass * y / r3 * w + mass * dw * y / (r2 * radius);
}

void gravity_free(bx_ctx *c)
{
    for (double *p : {c->phi, c->gforce, c->grav_sums, c->gfused})
        if (p) (void)hipFree(p);
    c->phi = c->gforce = c->grav_sums = c->gfused = nullptr;
    c->grav_flags = 0;
    c->gforce_gen = c->gfused_gen = -1;
}

// The ghosts of the force planes as the fused ghost fill sets the momenta's:
// each ghost takes its (interior) donor's force with the ghost's sign flips.
__global__ __launch_bounds__(kGThreads) void k_gf_ghosts(int64_t n, int64_t ng, const int32_t *__restrict__ dst,
                                                        const int32_t *__restrict__ src,
                                                        const int32_t *__restrict__ sign, double *__restrict__ gf)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ng) return;
    const int64_t d = dst[k], s = src[k];
    const int sg = sign[k];
    for (int f = 0; f < 3; f++) {
        const double v = gf[f * n + s];
        gf[f * n + d] = ((sg >> f) & 1) ? -v : v;
    }
}

bool grav_settled(const bx_ctx *c) { return !c->grav_flags || c->gforce_gen == c->phi_gen; }

// The force planes of the fused kernels from the settled gforce (allocates
// them on first use; a no-op while they are current).
int grav_fused_planes(bx_ctx *c)
{
    if (!c->grav_flags || c->gfused_gen == c->gforce_gen) return BX_OK;
    if (!c->gfused) BX_HIP(hipMalloc((void **)&c->gfused, sizeof(double) * 3 * (size_t)c->n));
    BX_HIP(hipMemcpyAsync(c->gfused, c->gforce, sizeof(double) * 3 * (size_t)c->n, hipMemcpyDeviceToDevice,
                          c->stream));
    if (c->nghost > 0) {
        hipLaunchKernelGGL(k_gf_ghosts, dim3(blocks_for(c->nghost, kGThreads)), dim3(kGThreads), 0, c->stream, c->n,
                           c->nghost, c->ghost_dst, c->nscratch ? c->fghost_src : c->ghost_src,
                           c->nscratch ? c->fghost_sign : c->ghost_sign, c->gfused);
        BX_HIP(hipGetLastError());
    }
    c->gfused_gen = c->gforce_gen;
    return BX_OK;
}

static int nblocks(const bx_ctx *c) { return (int)blocks_for(c->n, kGThreads); }

int launch_calc_gforce(bx_ctx *c)
{
    hipLaunchKernelGGL(k_calc_gforce, dim3(blocks_for(c->n, kGThreads)), dim3(kGThreads), 0, c->stream, c->n, c->nx,
                       c->ny, c->nz, c->flatz, c->phi, c->gforce, c->prm.dx, c->prm.dy, c->prm.dz);
    BX_HIP(hipGetLastError());
    // the interior from phi alone; the ghosts (never read by source()) keep
    // their values
    c->gforce_gen = c->phi_gen;
    return BX_OK;
}

int launch_external_grav(bx_ctx *c)
{
    c->gforce_gen = -1;  // the point mass adds to gforce: no longer phi's alone
    const double dx = c->prm.dx, dy = c->prm.dy, dz = c->prm.dz;
    if (!(c->grav_flags & BX_GRAV_COM_SEQUENTIAL)) {
        const int nb = nblocks(c);
        hipLaunchKernelGGL(k_moments_part, dim3(nb), dim3(kGThreads), 0, c->stream, c->n, c->nx, c->ny, c->nz, c->cons,
                           dx, dy, dz, c->grav_sums + 4);
        BX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(kGThreads), 0, c->stream, nb, c->grav_sums + 4, c->grav_sums);
    } else {
        hipLaunchKernelGGL(k_moments_seq, dim3(1), dim3(64), 0, c->stream, c->n, c->nx, c->ny, c->nz, c->cons, dx, dy,
                           dz, c->grav_sums);
    }
    BX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_external_grav, dim3(blocks_for(c->n, kGThreads)), dim3(kGThreads), 0, c->stream, c->n, c->nx,
                       c->ny, dx, dy, c->grav_sums, c->obj_mass, c->obj_radius, c->gforce);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

}  // namespace bx

using namespace bx;

extern "C" int bx_gravity_config(bx_ctx *c, int32_t flags, double object_mass, double object_radius)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_gravity_config: null context");
    if (flags & ~(BX_GRAV_POTENTIAL | BX_GRAV_EXTERNAL | BX_GRAV_COM_SEQUENTIAL))
        return set_error(BX_ERR_ARG, "bx_gravity_config: unknown flag");
    if (c->prm.precision != BX_PREC_FP64) return set_error(BX_ERR_UNSUPPORTED, "bx_gravity_config: FP64 contexts only");
    if (c->fused_ready || c->adj)
        return set_error(BX_ERR_STATE, "bx_gravity_config: call before the first fused step or adjoint run");
    BX_HIP(hipSetDevice(c->device));
    gravity_free(c);
    if (!flags) {
        c->fused_unsupported = false;  // fused_prepare decides again on the next fused step
        return BX_OK;
    }
    BX_HIP(hipMalloc((void **)&c->phi, sizeof(double) * c->n));
    BX_HIP(hipMalloc((void **)&c->gforce, sizeof(double) * 3 * c->n));
    BX_HIP(hipMalloc((void **)&c->grav_sums, sizeof(double) * (4 + 3 * (size_t)nblocks(c))));
    BX_HIP(hipMemsetAsync(c->phi, 0, sizeof(double) * c->n, c->stream));  // init_grid.f90:291, 299
    BX_HIP(hipMemsetAsync(c->gforce, 0, sizeof(double) * 3 * c->n, c->stream));
    BX_HIP(hipMemsetAsync(c->grav_sums, 0, sizeof(double) * 4, c->stream));
    c->grav_flags = flags;
    c->obj_mass = object_mass;
    c->obj_radius = object_radius;
    c->phi_gen++;  // gforce (zero) is not yet the force of phi
    c->fused_valid = false;
    if (c->fused_lz != 0 && c->fused_occ != 11) {
        // a tuned variant other than the 64 x 8 tiled kernel, which alone
        // has the gravity form: back to the default work split
        c->fused_ly = c->fused_lz = 0;
        c->fused_occ = 0;
    }
    // fused_prepare decides on the next fused step: a potential alone takes
    // the fused kernels, the point mass the exact sequence
    c->fused_unsupported = false;
    return BX_OK;
}

extern "C" int bx_set_phi(bx_ctx *c, const double *phi)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_set_phi: null context");
    if (!c->grav_flags) return set_error(BX_ERR_STATE, "bx_set_phi: configure gravity first (bx_gravity_config)");
    BX_HIP(hipSetDevice(c->device));
    c->phi_gen++;  // gforce follows at the next calc_gforce
    c->fused_valid = false;
    if (!phi) {
        BX_HIP(hipMemsetAsync(c->phi, 0, sizeof(double) * c->n, c->stream));
        return BX_OK;
    }
    BX_HIP(hipMemcpyAsync(c->phi, phi, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
    BX_HIP(hipStreamSynchronize(c->stream));
    return BX_OK;
}

extern "C" int bx_calc_gforce(bx_ctx *c)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_calc_gforce: null context");
    if (!c->grav_flags) return set_error(BX_ERR_STATE, "bx_calc_gforce: configure gravity first");
    BX_HIP(hipSetDevice(c->device));
    return launch_calc_gforce(c);
}

extern "C" int bx_external_grav(bx_ctx *c, double *com)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_external_grav: null context");
    if (!c->grav_flags) return set_error(BX_ERR_STATE, "bx_external_grav: configure gravity first");
    BX_HIP(hipSetDevice(c->device));
    int rc = launch_external_grav(c);
    if (rc || !com) return rc;
    BX_HIP(hipMemcpyAsync(com, c->grav_sums, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BX_HIP(hipStreamSynchronize(c->stream));
    return BX_OK;
}

extern "C" int bx_download_gforce(bx_ctx *c, double *gforce)
{
    if (!c || !gforce) return set_error(BX_ERR_ARG, "bx_download_gforce: null argument");
    if (!c->grav_flags) return set_error(BX_ERR_STATE, "bx_download_gforce: configure gravity first");
    BX_HIP(hipSetDevice(c->device));
    std::vector<double> planes(3 * (size_t)c->n);
    BX_HIP(hipMemcpyAsync(planes.data(), c->gforce, planes.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BX_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < c->n; i++)
        for (int k = 0; k < 3; k++) gforce[3 * i + k] = planes[(size_t)k * c->n + i];
    return BX_OK;
}
