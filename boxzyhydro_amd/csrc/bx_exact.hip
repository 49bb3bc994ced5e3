// bx_exact.hip -- the hydro step as the reference structures it: one kernel per
// reference subroutine, acting on device-resident SoA planes.  These kernels
// are the plain FP64 drop-in behind bx_state/bx_velocity/bx_courant/
// bx_source/bx_set_ghost_cells/bx_flux/bx_avisc, and the exact fallback of
// the fused step.  Arithmetic order follows the cited reference lines so the
// results are bit-identical to the reference's FP64 step.
#include "bx_cells.h"

namespace bx {

constexpr int kThreads = 256;

// max-reduce non-negative doubles of a block into *dst (as IEEE bits)
__device__ __forceinline__ void block_max_to(double v, unsigned long long *dst)
{
    __shared__ double wmax[kThreads / 64];
    for (int off = 32; off > 0; off >>= 1) v = vmax(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wmax[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmax[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) m = vmax(m, wmax[k]);
        // skip the atomic when the stored maximum is already larger
        const unsigned long long b = (unsigned long long)__double_as_longlong(m);
        if (b > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(dst, b);
    }
}

// ------------------------------------------------------------------ state
__global__ __launch_bounds__(kThreads) void k_state(int64_t n, double *__restrict__ cons,
                                                   const double *__restrict__ u, double *__restrict__ p,
                                                   double *__restrict__ ad, double gam, double seps,
                                                   double muc, double rgas)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) state_cell(n, c, cons, u, p, ad, gam, seps, muc, rgas);
}

// --------------------------------------------------------------- velocity
__global__ __launch_bounds__(kThreads) void k_velocity(int64_t n, double *__restrict__ cons,
                                                      double *__restrict__ u, double vlimit,
                                                      double seps, unsigned long long *maxv)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double vm = c < n ? velocity_cell(n, c, cons, u, vlimit, seps) : 0.0;
    if (maxv) block_max_to(vm, maxv);
}

// ---------------------------------------------------------------- courant
// courant.f90:17-29: max over all cells
__global__ __launch_bounds__(kThreads) void k_courant(int64_t n, const double *__restrict__ cons,
                                                     const double *__restrict__ u,
                                                     const double *__restrict__ p,
                                                     const double *__restrict__ ad, double dx,
                                                     double dy, double dz, unsigned long long *rate)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    block_max_to(c < n ? courant_cell(n, c, cons, u, p, ad, dx, dy, dz) : 0.0, rate);
}

// ----------------------------------------------------------------- source
__global__ __launch_bounds__(kThreads) void k_source(int64_t n, int nx, int ny, int nz,
                                                    double *__restrict__ cons,
                                                    const double *__restrict__ u, double dt,
                                                    double g0, double g1, double g2, double seps,
                                                    int ang, double dx, double dy,
                                                    const double *__restrict__ gf)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) source_cell(n, c, nx, ny, nz, cons, u, dt, g0, g1, g2, seps, ang, dx, dy, gf);
}

// ------------------------------------------------------------ ghost cells
__global__ __launch_bounds__(kThreads) void k_ghost_snapshot(int64_t n, int64_t ns,
                                                            const int32_t *__restrict__ ssrc,
                                                            const double *__restrict__ planes,
                                                            double *__restrict__ scratch)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ns) ghost_snapshot_one(n, ns, k, ssrc, planes, scratch);
}

__global__ __launch_bounds__(kThreads) void k_ghosts(int64_t n, int64_t ng, const int32_t *__restrict__ dst,
                                                    const int32_t *__restrict__ src,
                                                    const int32_t *__restrict__ sign,
                                                    double *__restrict__ planes,
                                                    const double *__restrict__ scratch, int64_t ns)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ng) ghost_fill_one(n, k, dst, src, sign, planes, scratch, ns);
}

// ------------------------------------------------------------ flux: states
// flux.f90:47-116: slopes (calculate_slopes, 253-295) and interface states
// (calculate_states, 312-482) of every cell; ghosts get zero slopes.
// Output: 42 planes, index (dir*14 + k), k = 0..6 minus (rho,u,v,w,p,eps,gamma), 7..13 plus.
// ANG (flux_ang.f90:301-482): 54 planes, index (dir*18 + k), k = 0..8 minus
// (rho,u,v,w,p,eps,gamma,rmom,amom), 9..17 plus; rmom and amom are the
// transformed cons(2), cons(3) and get no transverse term.
template <bool ANG>
__global__ __launch_bounds__(kThreads) void k_flux_states(int64_t n, int nx, int ny, int nz, bool flatz,
                                                         const double *__restrict__ cons,
                                                         const double *__restrict__ u,
                                                         const double *__restrict__ p,
                                                         const double *__restrict__ ad,
                                                         double *__restrict__ st, double dt, double dx,
                                                         double dy, double dz, double seps)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int ix, iy, iz;
    coords(c, nx, ny, ix, iy, iz);
    const bool ghost = boundary_of(ix, iy, iz, nx, ny, nz) > 0;
    const int64_t stride[3] = {1, (int64_t)nx, flatz ? 0 : (int64_t)nx * ny};
    const double ds[3] = {dx, dy, dz};

    // own primitives (u from the last velocity() call, p from the last state())
    const double rho = cons[c], E = cons[4 * n + c];
    const double v[3] = {u[c], u[n + c], u[2 * n + c]};
    const double pc = p[c], gc = ad[c];
    const double eps = vmax(E - 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), seps);
    const double sound = sqrt(pc * gc / rho);

    constexpr int NS = ANG ? 9 : 7;
    for (int d = 0; d < 3; d++) {
        double su[3] = {0.0, 0.0, 0.0}, sp = 0.0, sd = 0.0, se = 0.0, sgm = 0.0, srm = 0.0, sam = 0.0;
        if (!ghost) {
            const int64_t cm = c - stride[d], cp = c + stride[d];
            for (int j = 0; j < 3; j++) su[j] = slope(u[j * n + cm], v[j], u[j * n + cp]);
            double em, ep;
            {
                const double r = cons[cm];
                const double a0 = u[cm], a1 = u[n + cm], a2 = u[2 * n + cm];
                em = vmax(cons[4 * n + cm] - 0.5 * r * (a0 * a0 + a1 * a1 + a2 * a2), seps);
            }
            {
                const double r = cons[cp];
                const double a0 = u[cp], a1 = u[n + cp], a2 = u[2 * n + cp];
                ep = vmax(cons[4 * n + cp] - 0.5 * r * (a0 * a0 + a1 * a1 + a2 * a2), seps);
            }
            sp = slope(p[cm], pc, p[cp]);
            sd = slope(cons[cm], rho, cons[cp]);
            se = slope(em, eps, ep);
            sgm = slope(ad[cm], gc, ad[cp]);
            if (ANG) {  // flux_ang.f90:319-320
                srm = slope(cons[n + cm], cons[n + c], cons[n + cp]);
                sam = slope(cons[2 * n + cm], cons[2 * n + c], cons[2 * n + cp]);
            }
        }
        const double lp = vmax(v[d] + sound, 0.0) * dt / ds[d];
        const double lm = vmin(v[d] - sound, 0.0) * dt / ds[d];
        const double zc = 0.0 * (dt * 0.5 / ds[d]);  // transverse terms: flux.f90:391-479
        double mi[NS], pl[NS];
        muscl(rho, lm, lp, sd, mi[0], pl[0]);
        for (int j = 0; j < 3; j++) muscl(v[j], lm, lp, su[j], mi[1 + j], pl[1 + j]);
        muscl(pc, lm, lp, sp, mi[4], pl[4]);
        muscl(eps, lm, lp, se, mi[5], pl[5]);
        muscl(gc, lm, lp, sgm, mi[6], pl[6]);
        for (int k = 0; k < 7; k++) {
            st[(int64_t)(d * 2 * NS + k) * n + c] = mi[k] + zc;
            st[(int64_t)(d * 2 * NS + NS + k) * n + c] = pl[k] + zc;
        }
        if constexpr (ANG) {  // flux_ang.f90:377-380
            muscl(cons[n + c], lm, lp, srm, mi[7], pl[7]);
            muscl(cons[2 * n + c], lm, lp, sam, mi[8], pl[8]);
            for (int k = 7; k < 9; k++) {
                st[(int64_t)(d * 2 * NS + k) * n + c] = mi[k];
                st[(int64_t)(d * 2 * NS + NS + k) * n + c] = pl[k];
            }
        }
    }
}

__device__ __forceinline__ Side<double> load_side(const double *__restrict__ st, int64_t n, int d, int pm,
                                                 int64_t c)
{
    const double *b = st + (int64_t)(d * 14 + 7 * pm) * n + c;
    Side<double> s;
    s.d = b[0];
    s.v[0] = b[n];
    s.v[1] = b[2 * n];
    s.v[2] = b[3 * n];
    s.p = b[4 * n];
    s.e = b[5 * n];
    s.g = b[6 * n];
    return s;
}

template <int dir>
__device__ __forceinline__ void dir_fluxes(const double *__restrict__ st, int64_t n, int64_t c, int64_t s,
                                           double f0[5], double f1[5])
{
    face_flux<double, dir>(load_side(st, n, dir, 1, c - s), load_side(st, n, dir, 0, c), f0);
    face_flux<double, dir>(load_side(st, n, dir, 1, c), load_side(st, n, dir, 0, c + s), f1);
}

// flux_ang.f90:523-618 (get_fluxes_ang): the momentum slots carry the
// reconstructed r- and angular momenta, and the pressure force is projected
// with the centre (x, y, r) of the cell whose faces these are.
template <int dir>
__device__ __forceinline__ void face_flux_ang(const double *__restrict__ st, int64_t n, int64_t cl, int64_t cr,
                                              double x, double y, double r, double F[5])
{
    const double *bl = st + (int64_t)(dir * 18 + 9) * n + cl;  // plus state of the left cell
    const double *br = st + (int64_t)(dir * 18) * n + cr;      // minus state of the right cell
    const double dl = bl[0], vxl = bl[n], vyl = bl[2 * n], vzl = bl[3 * n], pl = bl[4 * n], el = bl[5 * n],
                 gl = bl[6 * n];
    const double dr = br[0], vxr = br[n], vyr = br[2 * n], vzr = br[3 * n], pr = br[4 * n], er = br[5 * n],
                 gr = br[6 * n];
    double sl[5], sr[5], fl[5], fr[5];
    sl[0] = dl;
    sl[1] = bl[7 * n];
    sl[2] = bl[8 * n];
    sl[3] = dl * vzl;
    sl[4] = el + 0.5 * dl * (vxl * vxl + vyl * vyl + vzl * vzl);
    sr[0] = dr;
    sr[1] = br[7 * n];
    sr[2] = br[8 * n];
    sr[3] = dr * vzr;
    sr[4] = er + 0.5 * dr * (vxr * vxr + vyr * vyr + vzr * vzr);
    double vl, vr, pl1 = 0.0, pl2 = 0.0, pl3 = 0.0, pr1 = 0.0, pr2 = 0.0, pr3 = 0.0;
    if (dir == 0) {
        vl = vxl;
        vr = vxr;
        pl1 = pl * x / r;
        pr1 = pr * x / r;
        pl2 = -(pl * y);
        pr2 = -(pr * y);
    } else if (dir == 1) {
        vl = vyl;
        vr = vyr;
        pl1 = pl * y / r;
        pr1 = pr * y / r;
        pl2 = pl * x;
        pr2 = pr * x;
    } else {
        vl = vzl;
        vr = vzr;
        pl3 = pl;
        pr3 = pr;
    }
    fl[0] = sl[0] * vl;
    fl[1] = pl1 + sl[1] * vl;
    fl[2] = pl2 + sl[2] * vl;
    fl[3] = pl3 + sl[3] * vl;
    fl[4] = (sl[4] + pl) * vl;
    fr[0] = sr[0] * vr;
    fr[1] = pr1 + sr[1] * vr;
    fr[2] = pr2 + sr[2] * vr;
    fr[3] = pr3 + sr[3] * vr;
    fr[4] = (sr[4] + pr) * vr;
    // flux_hll, flux.f90:583-589
    const double cR = sqrt(gr * pr / dr), cL = sqrt(gl * pl / dl);
    const double ap = vmax(vmax(0.0, vr + cR), vl + cL);
    const double am = vmax(vmax(0.0, -(vr - cR)), -(vl - cL));
    for (int i = 0; i < 5; i++) F[i] = (ap * fl[i] + am * fr[i] - ap * am * (sr[i] - sl[i])) / (ap + am);
}

// flux_ang.f90:40-61 (and 245-268 inverted): cons(2:3), and cons_old(2:3)
// when avg>0, to r- and angular momentum about the z axis through y=-yoffset
__global__ __launch_bounds__(kThreads) void k_ang_to(int64_t n, int nx, int ny, double *__restrict__ cons,
                                                    double *__restrict__ cons_old, int avg, double dx, double dy,
                                                    double yoffset)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int ix = (int)(c % nx) + 1, iy = (int)((c / nx) % ny) + 1;
    const double x = centre(dx, ix, nx), y = centre(dy, iy, ny) + yoffset;
    const double r = sqrt(x * x + y * y);
    for (int k = 0; k < 1 + avg; k++) {
        double *q = k == 0 ? cons : cons_old;
        const double m2 = q[n + c], m3 = q[2 * n + c];
        const double angmom = x * m3 - y * m2;
        const double rmom = (m2 * x + m3 * y) / r;
        q[n + c] = rmom;
        q[2 * n + c] = angmom;
    }
}

__global__ __launch_bounds__(kThreads) void k_face_hll(int64_t n, int nx, int ny, int nz, bool flatz,
                                                      const double *__restrict__ cons, const double *__restrict__ u,
                                                      const double *__restrict__ p, const double *__restrict__ ad,
                                                      double *__restrict__ F, double dt, double dx, double dy,
                                                      double dz, double seps)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) face_hll_cell(n, c, nx, ny, nz, flatz, cons, u, p, ad, F, dt, dx, dy, dz, seps);
}

__global__ __launch_bounds__(kThreads) void k_face_update(int64_t n, int nx, int ny, int nz, bool flatz,
                                                         const double *__restrict__ F, double *__restrict__ cons,
                                                         const double *__restrict__ cons_old, int avg, double dt,
                                                         double ayz, double axz, double axy, double vol,
                                                         double srho, double seps)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) face_update_cell(n, c, nx, ny, nz, flatz, F, cons, cons_old, avg, dt, ayz, axz, axy, vol, srho, seps);
}

// ------------------------------------------------- flux: faces and update
// flux.f90:121-230: six HLL faces and fluxtmp on interior cells (fluxtmp=0 on
// ghosts, flux.f90:51), then the stage update and floors on every cell.
// ANG: flux_ang.f90:62-281 -- faces from face_flux_ang, the update in the
// transformed momenta, back to linear momenta (and cons_old when avg>0),
// then the floors.
template <bool ANG>
__global__ __launch_bounds__(kThreads) void k_flux_faces(int64_t n, int nx, int ny, int nz, bool flatz,
                                                        const double *__restrict__ st,
                                                        double *__restrict__ cons,
                                                        double *__restrict__ cons_old, int avg,
                                                        double dt, double ayz, double axz, double axy,
                                                        double vol, double srho, double seps, double dx,
                                                        double dy, double yoffset)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int ix, iy, iz;
    coords(c, nx, ny, ix, iy, iz);
    double ftmp[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (boundary_of(ix, iy, iz, nx, ny, nz) == 0) {
        double fx0[5], fx1[5], fy0[5], fy1[5], fz0[5], fz1[5];
        const int64_t sz = flatz ? 0 : (int64_t)nx * ny;
        if constexpr (ANG) {
            const double x = centre(dx, ix, nx), y = centre(dy, iy, ny);  // get_fluxes_ang: no yoffset
            const double r = sqrt(x * x + y * y);
            face_flux_ang<0>(st, n, c - 1, c, x, y, r, fx0);
            face_flux_ang<0>(st, n, c, c + 1, x, y, r, fx1);
            face_flux_ang<1>(st, n, c - nx, c, x, y, r, fy0);
            face_flux_ang<1>(st, n, c, c + nx, x, y, r, fy1);
            face_flux_ang<2>(st, n, c - sz, c, x, y, r, fz0);
            face_flux_ang<2>(st, n, c, c + sz, x, y, r, fz1);
        } else {
            dir_fluxes<0>(st, n, c, 1, fx0, fx1);
            dir_fluxes<1>(st, n, c, nx, fy0, fy1);
            dir_fluxes<2>(st, n, c, sz, fz0, fz1);
        }
        for (int i = 0; i < 5; i++)
            ftmp[i] = -((ayz * (fx1[i] - fx0[i]) + axz * (fy1[i] - fy0[i]) + axy * (fz1[i] - fz0[i])) * dt / vol);
    }
    double U[5];
    for (int i = 0; i < 5; i++) {
        const double ft = ftmp[i];
        const double cu = cons[(int64_t)i * n + c];
        U[i] = avg == 0 ? cu + ft : 0.5 * (cons_old[(int64_t)i * n + c] + cu + ft);
    }
    if constexpr (ANG) {  // flux_ang.f90:249-265
        const double x = centre(dx, ix, nx), y = centre(dy, iy, ny) + yoffset;
        const double r = sqrt(x * x + y * y);
        double angmom = U[2], rmom = U[1];
        U[1] = (rmom * x - angmom * y / r) / r;
        U[2] = (rmom * y + angmom * x / r) / r;
        if (avg > 0) {
            angmom = cons_old[2 * n + c];
            rmom = cons_old[n + c];
            cons_old[n + c] = (rmom * x - angmom * y / r) / r;
            cons_old[2 * n + c] = (rmom * y + angmom * x / r) / r;
        }
    }
    U[0] = vmax(U[0], srho);
    if (U[4] < 0.0) U[4] = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0] + seps;
    for (int i = 0; i < 5; i++) cons[(int64_t)i * n + c] = U[i];
}

// ------------------------------------------------------------------ avisc
// avisc.f90:13-24 with get_boundary (utils.f90:62-74).  Outer-layer cells
// (boundary=1) have only their donor in ineigh(1) in the reference and read
// ineigh(2:6)=0 (outside the array); here those neighbours are the cell itself.
__global__ __launch_bounds__(kThreads) void k_avisc(int64_t n, int nx, int ny, int nz, bool flatz,
                                                   const double *__restrict__ u, double *__restrict__ qq,
                                                   double ax, double ay, double az)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int ix, iy, iz;
    coords(c, nx, ny, ix, iy, iz);
    const int bnd = boundary_of(ix, iy, iz, nx, ny, nz);
    if (bnd == 0) {
        qq[c] = 0.0;
        qq[n + c] = 0.0;
        qq[2 * n + c] = 0.0;
        return;
    }
    const int64_t pl = (int64_t)nx * ny;
    int64_t nb[6];
    if (bnd == 2) {
        nb[0] = c + nx;
        nb[1] = c - nx;
        nb[2] = c + 1;
        nb[3] = c - 1;
        nb[4] = flatz ? c : c + pl;
        nb[5] = flatz ? c : c - pl;
    } else {
        for (int k = 0; k < 6; k++) nb[k] = c;
        if (ix == 1) nb[0] = c + 2;
        else if (ix == nx) nb[0] = c - 2;
        else if (iy == 1) nb[0] = c + 2 * (int64_t)nx;
        else if (iy == ny) nb[0] = c - 2 * (int64_t)nx;
        else if (iz == 1) nb[0] = flatz ? c : c + 2 * pl;
        else nb[0] = flatz ? c : c - 2 * pl;
    }
    for (int k = 0; k < 6; k++) {
        int jx, jy, jz;
        coords(nb[k], nx, ny, jx, jy, jz);
        if (boundary_of(jx, jy, jz, nx, ny, nz) > 0) nb[k] = c;
    }
    double t;
    t = vmin(0.0, u[nb[2]] - u[nb[3]]);
    qq[c] = ax * (t * t);
    t = vmin(0.0, u[n + nb[0]] - u[n + nb[1]]);
    qq[n + c] = ay * (t * t);
    t = vmin(0.0, u[2 * n + nb[4]] - u[2 * n + nb[5]]);
    qq[2 * n + c] = az * (t * t);
}

// ------------------------------------------------------------- launchers
int launch_state(bx_ctx *c)
{
    if (c->eos) return launch_state_tab(c);  // H2STAT >= 0 (bx_eos.hip)
    hipLaunchKernelGGL(k_state, dim3(blocks_for(c->n, kThreads)), dim3(kThreads), 0, c->stream, c->n, c->cons,
                       c->u, c->p, c->adindx, c->prm.gammafix, c->prm.small_eps, c->muc, c->rgas);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

int launch_velocity(bx_ctx *c, bool want_maxv)
{
    if (want_maxv) BX_HIP(hipMemsetAsync(&c->dsc->maxv_bits, 0, sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_velocity, dim3(blocks_for(c->n, kThreads)), dim3(kThreads), 0, c->stream, c->n,
                       c->cons, c->u, c->prm.vlimit, c->prm.small_eps,
                       want_maxv ? &c->dsc->maxv_bits : (unsigned long long *)nullptr);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

int launch_courant(bx_ctx *c)
{
    BX_HIP(hipMemsetAsync(&c->dsc->rate_bits[0], 0, sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_courant, dim3(blocks_for(c->n, kThreads)), dim3(kThreads), 0, c->stream, c->n, c->cons,
                       c->u, c->p, c->adindx, c->prm.dx, c->prm.dy, c->prm.dz, &c->dsc->rate_bits[0]);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

int launch_ghosts(bx_ctx *c, double *planes5)
{
    if (c->nghost == 0) return BX_OK;
    if (c->nscratch > 0) {
        hipLaunchKernelGGL(k_ghost_snapshot, dim3(blocks_for(c->nscratch, kThreads)), dim3(kThreads), 0, c->stream,
                           c->n, c->nscratch, c->ghost_scratch_src, planes5, c->ghost_scratch);
        BX_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_ghosts, dim3(blocks_for(c->nghost, kThreads)), dim3(kThreads), 0, c->stream, c->n,
                       c->nghost, c->ghost_dst, c->ghost_src, c->ghost_sign, planes5, c->ghost_scratch,
                       c->nscratch);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

int launch_source(bx_ctx *c, double dt)
{
    hipLaunchKernelGGL(k_source, dim3(blocks_for(c->n, kThreads)), dim3(kThreads), 0, c->stream, c->n, c->nx,
                       c->ny, c->nz, c->cons, c->u, dt, c->grav[0], c->grav[1], c->grav[2], c->prm.small_eps,
                       c->prm.fluxangmom, c->prm.dx, c->prm.dy, c->grav_flags ? c->gforce : (const double *)nullptr);
    BX_HIP(hipGetLastError());
    return launch_ghosts(c, c->cons);
}

int launch_flux(bx_ctx *c, int avg, double dt)
{
    const double dx = c->prm.dx, dy = c->prm.dy, dz = c->prm.dz;
    const bool ang = c->prm.fluxangmom != 0;
    const dim3 grid(blocks_for(c->n, kThreads));
    if (ang) {
        hipLaunchKernelGGL(k_ang_to, grid, dim3(kThreads), 0, c->stream, c->n, c->nx, c->ny, c->cons, c->cons_old,
                           avg, dx, dy, c->prm.yoffset);
        BX_HIP(hipGetLastError());
    }
    const double areaxy = dy * dx, areayz = dy * dz, areaxz = dx * dz, vol = dx * dy * dz;
    if (ang) {  // interface states of every cell, then the faces per cell (the pressure projection is per cell)
        hipLaunchKernelGGL(k_flux_states<true>, grid, dim3(kThreads), 0, c->stream, c->n, c->nx, c->ny, c->nz,
                           c->flatz, c->cons, c->u, c->p, c->adindx, c->states, dt, dx, dy, dz, c->prm.small_eps);
        BX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_flux_faces<true>, grid, dim3(kThreads), 0, c->stream, c->n, c->nx, c->ny, c->nz,
                           c->flatz, c->states, c->cons, c->cons_old, avg, dt, areayz, areaxz, areaxy, vol,
                           c->prm.small_rho, c->prm.small_eps, dx, dy, c->prm.yoffset);
    } else {  // each face's flux once, then the update
        hipLaunchKernelGGL(k_face_hll, grid, dim3(kThreads), 0, c->stream, c->n, c->nx, c->ny, c->nz, c->flatz,
                           c->cons, c->u, c->p, c->adindx, c->states, dt, dx, dy, dz, c->prm.small_eps);
        BX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_face_update, grid, dim3(kThreads), 0, c->stream, c->n, c->nx, c->ny, c->nz, c->flatz,
                           c->states, c->cons, c->cons_old, avg, dt, areayz, areaxz, areaxy, vol, c->prm.small_rho,
                           c->prm.small_eps);
    }
    BX_HIP(hipGetLastError());
    return launch_ghosts(c, c->cons);
}

int launch_avisc(bx_ctx *c)
{
    hipLaunchKernelGGL(k_avisc, dim3(blocks_for(c->n, kThreads)), dim3(kThreads), 0, c->stream, c->n, c->nx,
                       c->ny, c->nz, c->flatz, c->u, c->qq, c->prm.avmag[0], c->prm.avmag[1], c->prm.avmag[2]);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

}  // namespace bx
