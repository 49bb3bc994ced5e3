// bx_cells.h -- per-cell bodies of the exact call sequence (one reference
// subroutine's work on one cell), shared by the one-kernel-per-subroutine
// path (bx_exact.hip) and the batched boxes (bx_boxes.hip), so that both run
// the same arithmetic.  Arithmetic order follows the cited reference lines.
#pragma once

#include "bx_internal.h"
#include "bx_math.h"

namespace bx {

// 1-based (ix,iy,iz) of cell c (utils.f90:191-194 inverse)
__device__ __forceinline__ void coords(int64_t c, int nx, int ny, int &ix, int &iy, int &iz)
{
    int64_t plane = (int64_t)nx * ny;
    iz = (int)(c / plane) + 1;
    int64_t r = c - (int64_t)(iz - 1) * plane;
    iy = (int)(r / nx) + 1;
    ix = (int)(r - (int64_t)(iy - 1) * nx) + 1;
}

// init_grid.f90:81-107 ghost flag: 1 outer layer, 2 second layer, 0 interior
__device__ __forceinline__ int boundary_of(int ix, int iy, int iz, int nx, int ny, int nz)
{
    int f1 = (ix == 1 || ix == nx) || (iy == 1 || iy == ny) || (nz >= 6 && (iz == 1 || iz == nz));
    if (f1) return 1;
    int f2 = (ix == 2 || ix == nx - 1) || (iy == 2 || iy == ny - 1) || (nz >= 6 && (iz == 2 || iz == nz - 1));
    return f2 ? 2 : 0;
}

// centred coordinate of cell index i (1-based) on an axis of n cells (init_grid.f90:85-93)
__device__ __forceinline__ double centre(double d, int i, int n) { return d * ((double)i - ((double)n + 1.0) / 2.0); }

// state.f90:30-57, H2STAT=-1 branch
__device__ __forceinline__ void state_cell(int64_t n, int64_t c, double *__restrict__ cons,
                                           const double *__restrict__ u, double *__restrict__ p,
                                           double *__restrict__ ad, double gam, double seps, double muc, double rgas)
{
    const double rho = cons[c], E = cons[4 * n + c];
    const double u0 = u[c], u1 = u[n + c], u2 = u[2 * n + c];
    const double ekin = rho * 0.5 * (u0 * u0 + u1 * u1 + u2 * u2);
    double eps = vmax(E - ekin, seps);
    ad[c] = gam;
    if (rho < 1e-4) {
        const double tk = vmin(5e3, p[c] / (rho * rgas) * muc);
        const double pc = rho * tk * rgas / muc;
        p[c] = pc;
        eps = pc / (gam - 1.0);
        cons[4 * n + c] = eps + ekin;
    } else {
        p[c] = eps * (gam - 1.0);
    }
}

// velocity.f90:18-53: u = m/rho, vlimit clamp; returns max |u_d| before the clamp
__device__ __forceinline__ double velocity_cell(int64_t n, int64_t c, double *__restrict__ cons,
                                                double *__restrict__ u, double vlimit, double seps)
{
    double vm = 0.0;
    const double rho = cons[c];
    double m[3], v[3];
    for (int d = 0; d < 3; d++) {
        m[d] = cons[(1 + d) * n + c];
        v[d] = m[d] / rho;
    }
    int flag = 0;
    for (int d = 0; d < 3; d++) {
        vm = vmax(vm, fabs(v[d]));
        if (fabs(v[d]) > vlimit) flag = 1;
    }
    if (flag) {
        double E = cons[4 * n + c];
        double ekin = 0.5 * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / rho;
        double speed = 0.0;
        for (int d = 0; d < 3; d++) speed = speed + v[d] * v[d];
        speed = sqrt(speed);
        for (int d = 0; d < 3; d++) {
            v[d] = v[d] * vlimit / speed;
            m[d] = v[d] * rho;
            cons[(1 + d) * n + c] = m[d];
        }
        E = E - ekin;
        ekin = 0.5 * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / rho;
        cons[4 * n + c] = vmax(E, seps) + ekin;
    }
    for (int d = 0; d < 3; d++) u[d * n + c] = v[d];
    return vm;
}

// courant.f90:17-21: the cell's sum_d (c+|u_d|)/Delta_d
__device__ __forceinline__ double courant_cell(int64_t n, int64_t c, const double *__restrict__ cons,
                                               const double *__restrict__ u, const double *__restrict__ p,
                                               const double *__restrict__ ad, double dx, double dy, double dz)
{
    const double sound = sqrt(ad[c] * p[c] / cons[c]);
    return (sound + fabs(u[c])) / dx + (sound + fabs(u[n + c])) / dy + (sound + fabs(u[2 * n + c])) / dz;
}

// source.f90:148-182 on an interior cell with the gforce planes gf, or without
// them a uniform gravity vector g (the reference's g = -(phi(+)-phi(-))/(2d)
// = -0.0 for phi=0); with fluxangmom the terms of source.f90:73-90 come first
// (the pforce of source.f90:64-66 feeds nothing on the hydro path)
__device__ __forceinline__ void source_cell(int64_t n, int64_t c, int nx, int ny, int nz, double *__restrict__ cons,
                                            const double *__restrict__ u, double dt, double g0, double g1, double g2,
                                            double seps, int ang, double dx, double dy,
                                            const double *__restrict__ gf)
{
    int ix, iy, iz;
    coords(c, nx, ny, ix, iy, iz);
    if (boundary_of(ix, iy, iz, nx, ny, nz) > 0) return;
    const double rho = cons[c];
    double m0 = cons[n + c], m1 = cons[2 * n + c], m2 = cons[3 * n + c];
    const double ux = u[c], uy = u[n + c], uz = u[2 * n + c];
    if (ang) {
        const double x = centre(dx, ix, nx), y = centre(dy, iy, ny);
        const double r = sqrt(x * x + y * y);
        // r**4 as the reference's build evaluates it: a left-to-right product
        const double w = x * uy - y * ux, r4 = ((r * r) * r) * r;
        m1 = m1 + rho * (w * w) * y / r4 * dt;
        m0 = m0 + rho * (w * w) * x / r4 * dt;
    }
    if (gf) {  // gforce(1:3, igrid) of calc_gforce / external_grav
        g0 = gf[c];
        g1 = gf[n + c];
        g2 = gf[2 * n + c];
    }
    m1 = m1 + (rho * g1) * dt;
    m0 = m0 + (rho * g0) * dt;
    m2 = m2 + (g2 * rho) * dt;
    const double ekin = 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / rho;
    const double E = cons[4 * n + c];
    cons[n + c] = m0;
    cons[2 * n + c] = m1;
    cons[3 * n + c] = m2;
    cons[4 * n + c] = vmax(E - ekin + dt * rho * (ux * g0 + uy * g1 + uz * g2), seps) + ekin;
}

// utils.f90:314-403 as resolved on the host into (ghost, donor, momentum sign
// flips).  Donors that are ghosts are read from a snapshot taken first.
__device__ __forceinline__ void ghost_snapshot_one(int64_t n, int64_t ns, int64_t k, const int32_t *__restrict__ ssrc,
                                                   const double *__restrict__ planes, double *__restrict__ scratch)
{
    const int64_t s = ssrc[k];
    for (int f = 0; f < 5; f++) scratch[f * ns + k] = planes[f * n + s];
}

__device__ __forceinline__ void ghost_fill_one(int64_t n, int64_t k, const int32_t *__restrict__ dst,
                                               const int32_t *__restrict__ src, const int32_t *__restrict__ sign,
                                               double *__restrict__ planes, const double *__restrict__ scratch,
                                               int64_t ns)
{
    const int64_t d = dst[k];
    const int32_t s0 = src[k];
    const double *base;
    int64_t stride, s;
    if (s0 >= 0) {
        base = planes;
        stride = n;
        s = s0;
    } else {
        base = scratch;
        stride = ns;
        s = -(int64_t)s0 - 1;
    }
    const int sg = sign[k];
    planes[d] = base[s];
    for (int f = 1; f <= 3; f++) {
        double v = base[f * stride + s];
        planes[f * n + d] = ((sg >> (f - 1)) & 1) ? -v : v;
    }
    planes[4 * n + d] = base[4 * stride + s];
}

// ------------------------------------------ flux: one HLL flux per face
// The side of cell q in direction d (minus: pm=0, plus: pm=1) exactly as
// k_flux_states forms it (flux.f90:253-295 slopes, 312-482 states).
template <int d>
__device__ __forceinline__ Side<double> cell_side(int64_t q, int pm, int64_t n, int nx, int ny, int nz,
                                                 int64_t stride, const double *__restrict__ cons,
                                                 const double *__restrict__ u, const double *__restrict__ p,
                                                 const double *__restrict__ ad, double dt, double ds, double seps)
{
    int ix, iy, iz;
    coords(q, nx, ny, ix, iy, iz);
    const bool ghost = boundary_of(ix, iy, iz, nx, ny, nz) > 0;
    const double rho = cons[q], E = cons[4 * n + q];
    const double v[3] = {u[q], u[n + q], u[2 * n + q]};
    const double pc = p[q], gc = ad[q];
    const double eps = vmax(E - 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), seps);
    const double sound = sqrt(pc * gc / rho);
    double su[3] = {0.0, 0.0, 0.0}, sp = 0.0, sd = 0.0, se = 0.0, sgm = 0.0;
    if (!ghost) {
        const int64_t cm = q - stride, cp = q + stride;
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
    }
    const double lp = vmax(v[d] + sound, 0.0) * dt / ds;
    const double lm = vmin(v[d] - sound, 0.0) * dt / ds;
    const double zc = 0.0 * (dt * 0.5 / ds);  // transverse terms: flux.f90:391-479
    double mi[7], pl[7];
    muscl(rho, lm, lp, sd, mi[0], pl[0]);
    for (int j = 0; j < 3; j++) muscl(v[j], lm, lp, su[j], mi[1 + j], pl[1 + j]);
    muscl(pc, lm, lp, sp, mi[4], pl[4]);
    muscl(eps, lm, lp, se, mi[5], pl[5]);
    muscl(gc, lm, lp, sgm, mi[6], pl[6]);
    const double *o = pm ? pl : mi;
    Side<double> sd7;
    sd7.d = o[0] + zc;
    sd7.v[0] = o[1] + zc;
    sd7.v[1] = o[2] + zc;
    sd7.v[2] = o[3] + zc;
    sd7.p = o[4] + zc;
    sd7.e = o[5] + zc;
    sd7.g = o[6] + zc;
    return sd7;
}

// flux.f90:121-190 with each face computed once: F(d, c) is the HLL flux of
// the face between cell c and c + s_d (get_fluxes/flux_hll, flux.f90:486-590),
// for the faces some interior cell uses.  The reference forms every face
// twice, once from each side, from the same inputs and with the same result.
__device__ __forceinline__ void face_hll_cell(int64_t n, int64_t c, int nx, int ny, int nz, bool flatz,
                                              const double *__restrict__ cons, const double *__restrict__ u,
                                              const double *__restrict__ p, const double *__restrict__ ad,
                                              double *__restrict__ F, double dt, double dx, double dy, double dz,
                                              double seps)
{
    int ix, iy, iz;
    coords(c, nx, ny, ix, iy, iz);
    const bool in_c = boundary_of(ix, iy, iz, nx, ny, nz) == 0;
    const int64_t sz = flatz ? 0 : (int64_t)nx * ny;
    double f[5];
    // x face c+1/2: needed by c or c+1 when interior
    if (in_c || (ix < nx && boundary_of(ix + 1, iy, iz, nx, ny, nz) == 0)) {
        face_flux<double, 0>(cell_side<0>(c, 1, n, nx, ny, nz, 1, cons, u, p, ad, dt, dx, seps),
                             cell_side<0>(c + 1, 0, n, nx, ny, nz, 1, cons, u, p, ad, dt, dx, seps), f);
        for (int k = 0; k < 5; k++) F[(int64_t)k * n + c] = f[k];
    }
    if (in_c || (iy < ny && boundary_of(ix, iy + 1, iz, nx, ny, nz) == 0)) {
        face_flux<double, 1>(cell_side<1>(c, 1, n, nx, ny, nz, nx, cons, u, p, ad, dt, dy, seps),
                             cell_side<1>(c + nx, 0, n, nx, ny, nz, nx, cons, u, p, ad, dt, dy, seps), f);
        for (int k = 0; k < 5; k++) F[(int64_t)(5 + k) * n + c] = f[k];
    }
    if (in_c || (!flatz && iz < nz && boundary_of(ix, iy, iz + 1, nx, ny, nz) == 0)) {
        face_flux<double, 2>(cell_side<2>(c, 1, n, nx, ny, nz, sz, cons, u, p, ad, dt, dz, seps),
                             cell_side<2>(c + sz, 0, n, nx, ny, nz, sz, cons, u, p, ad, dt, dz, seps), f);
        for (int k = 0; k < 5; k++) F[(int64_t)(10 + k) * n + c] = f[k];
    }
}

// flux.f90:121-230 on the face fluxes: fluxtmp of an interior cell (zero on
// ghosts, flux.f90:51), the stage update and the floors
__device__ __forceinline__ void face_update_cell(int64_t n, int64_t c, int nx, int ny, int nz, bool flatz,
                                                 const double *__restrict__ F, double *__restrict__ cons,
                                                 const double *__restrict__ cons_old, int avg, double dt, double ayz,
                                                 double axz, double axy, double vol, double srho, double seps)
{
    int ix, iy, iz;
    coords(c, nx, ny, ix, iy, iz);
    double ftmp[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (boundary_of(ix, iy, iz, nx, ny, nz) == 0) {
        const int64_t sz = flatz ? 0 : (int64_t)nx * ny;
        for (int i = 0; i < 5; i++) {
            const double fx0 = F[(int64_t)i * n + c - 1], fx1 = F[(int64_t)i * n + c];
            const double fy0 = F[(int64_t)(5 + i) * n + c - nx], fy1 = F[(int64_t)(5 + i) * n + c];
            const double fz0 = F[(int64_t)(10 + i) * n + c - sz], fz1 = F[(int64_t)(10 + i) * n + c];
            ftmp[i] = -((ayz * (fx1 - fx0) + axz * (fy1 - fy0) + axy * (fz1 - fz0)) * dt / vol);
        }
    }
    double U[5];
    for (int i = 0; i < 5; i++) {
        const double ft = ftmp[i];
        const double cu = cons[(int64_t)i * n + c];
        U[i] = avg == 0 ? cu + ft : 0.5 * (cons_old[(int64_t)i * n + c] + cu + ft);
    }
    U[0] = vmax(U[0], srho);
    if (U[4] < 0.0) U[4] = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0] + seps;
    for (int i = 0; i < 5; i++) cons[(int64_t)i * n + c] = U[i];
}

}  // namespace bx
