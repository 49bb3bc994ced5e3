// bx_fused.hip -- the hydro step as two fused stencil kernels per step.
//
// The reference step (main.f90:211-555) runs ~25 passes over the grid per
// step.  On a "regular" state -- every ghost cell a (sign-flipped) copy of an
// interior donor, u == m/rho bitwise, no cell below the rho<1e-4 EOS branch,
// no speed above vlimit -- every pass except the two flux stages is pointwise,
// so they are folded into the stencil kernels:
//
//   stage 1:  U^n -> source(0) energy re-round (source.f90:176-182), state,
//             velocity -> MUSCL/HLL predictor flux(0) (flux.f90) -> floors ->
//             U*, and the stale-velocity pressure p* of main.f90:351-355
//   stage 2:  U*, p*, U^n -> velocity -> corrector flux(1) with
//             cons_old = source(0)(U^n) -> floors -> state, velocity,
//             source(1) re-round (main.f90:374-515) -> U^{n+1}, plus the
//             next step's Courant rate (courant.f90:17-23) as a max-reduction
//
// Every arithmetic operation is the one the exact sequence (bx_exact.hip)
// performs on the same values, so the fused step is bit-identical to it; a
// cell that would take an irregular branch raises a flag and the host redoes
// the step with the exact sequence.
//
// Work decomposition (gfx950, wave64):
//   * a wave holds 64 consecutive cells of one x-row; lanes 2..61 are output
//     cells, lanes 0,1,62,63 the x-halo.  x-neighbours come through wave
//     shuffles; each x-face is evaluated once (by the lane on its right).
//   * a block stacks TY output rows plus 2 halo rows above and below (TY+4
//     waves); y-neighbour data goes through LDS.
//   * the block marches in z over a chunk of planes, keeping the primitive
//     variables of planes k..k+2 and the z-face flux of plane k in registers.
#include <cmath>

#include "bx_internal.h"
#include "bx_math.h"

namespace bx {

constexpr int kSegOut = 60;  // output cells per wave row

struct StageArgs {
    int64_t n;
    int nx, ny, nz;
    int flatz;
    int zlo, zhi, lz;     // interior z range [zlo,zhi) and planes per block
    const double *uin;    // 5 planes: stage 1 U^n, stage 2 U*
    const double *uold;   // stage 2: U^n
    const double *pin;    // stage 2: p*
    double *uout;         // 5 planes
    double *pout;         // stage 1: p*
    double dt;
    double ds[3], inv_ds[3];
    int ds_pow2[3];
    double ayz, axz, axy, vol, inv_vol;
    int vol_pow2;
    double gam, gm1, seps, srho, vlimit;
    unsigned int *flags;
    unsigned long long *rate;  // stage 2: next step's Courant rate (max)
};

struct W6 {
    double d, u0, u1, u2, p, e;
};

__device__ __forceinline__ double div_by(double x, double d, double inv, int pow2)
{
    // x/d; for d a power of two x*(1/d) is the same correctly rounded value
    return pow2 ? x * inv : x / d;
}

__device__ __forceinline__ double shfl_d(double v, int src)
{
    return __shfl(v, src, 64);
}

__device__ __forceinline__ W6 shfl_w(const W6 &w, int src)
{
    W6 o;
    o.d = shfl_d(w.d, src);
    o.u0 = shfl_d(w.u0, src);
    o.u1 = shfl_d(w.u1, src);
    o.u2 = shfl_d(w.u2, src);
    o.p = shfl_d(w.p, src);
    o.e = shfl_d(w.e, src);
    return o;
}

__device__ __forceinline__ double comp(const W6 &w, int j)
{
    return j == 0 ? w.u0 : (j == 1 ? w.u1 : w.u2);
}

// interface states of one cell in direction d (flux.f90:253-351, 391-479):
// slopes from (wm, wc, wp) unless the cell is a ghost (zero slopes).
template <int d>
__device__ __forceinline__ void dir_states(const W6 &wm, const W6 &wc, const W6 &wp, double sound, bool ghost,
                                           const StageArgs &a, Side<double> &mi, Side<double> &pl)
{
    double su0 = 0.0, su1 = 0.0, su2 = 0.0, sp = 0.0, sd = 0.0, se = 0.0;
    if (!ghost) {
        su0 = slope(wm.u0, wc.u0, wp.u0);
        su1 = slope(wm.u1, wc.u1, wp.u1);
        su2 = slope(wm.u2, wc.u2, wp.u2);
        sp = slope(wm.p, wc.p, wp.p);
        sd = slope(wm.d, wc.d, wp.d);
        se = slope(wm.e, wc.e, wp.e);
    }
    const double ud = comp(wc, d);
    const double lp = div_by(vmax(ud + sound, 0.0) * a.dt, a.ds[d], a.inv_ds[d], a.ds_pow2[d]);
    const double lm = div_by(vmin(ud - sound, 0.0) * a.dt, a.ds[d], a.inv_ds[d], a.ds_pow2[d]);
    const double cm = 0.5 * (-1.0 - lm), cp = 0.5 * (1.0 - lp);
    const double z = 0.0;  // transverse terms: 0*cc with cc=dt*half/ds > 0
    mi.d = (wc.d + cm * sd) + z;
    pl.d = (wc.d + cp * sd) + z;
    mi.v[0] = (wc.u0 + cm * su0) + z;
    pl.v[0] = (wc.u0 + cp * su0) + z;
    mi.v[1] = (wc.u1 + cm * su1) + z;
    pl.v[1] = (wc.u1 + cp * su1) + z;
    mi.v[2] = (wc.u2 + cm * su2) + z;
    pl.v[2] = (wc.u2 + cp * su2) + z;
    mi.p = (wc.p + cm * sp) + z;
    pl.p = (wc.p + cp * sp) + z;
    mi.e = (wc.e + cm * se) + z;
    pl.e = (wc.e + cp * se) + z;
    mi.g = a.gam;  // slope(gamma,gamma,gamma)=0: gamma + (+-0) + 0 = gamma
    pl.g = a.gam;
}

__device__ __forceinline__ double sound_of(const W6 &w, double gam) { return sqrt(w.p * gam / w.d); }

// primitive variables of a loaded cell for each stage
template <int STAGE>
__device__ __forceinline__ W6 load_prims(const StageArgs &a, int64_t c, unsigned int &flg)
{
    const int64_t n = a.n;
    W6 w;
    const double r = a.uin[c], m0 = a.uin[n + c], m1 = a.uin[2 * n + c], m2 = a.uin[3 * n + c];
    const double E = a.uin[4 * n + c];
    w.d = r;
    w.u0 = m0 / r;
    w.u1 = m1 / r;
    w.u2 = m2 / r;
    if (STAGE == 1) {
        // source(0) energy re-round with g = -0 (source.f90:176-182), then
        // state at main.f90:314 with u = m/rho, velocity at :318
        const double ek = 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r;
        const double E1 = vmax(E - ek, a.seps) + ek;
        w.e = vmax(E1 - 0.5 * r * (w.u0 * w.u0 + w.u1 * w.u1 + w.u2 * w.u2), a.seps);
        w.p = w.e * a.gm1;
    } else {
        // velocity at main.f90:355; p* is the state of :353 (stale velocity)
        w.e = vmax(E - 0.5 * r * (w.u0 * w.u0 + w.u1 * w.u1 + w.u2 * w.u2), a.seps);
        w.p = a.pin[c];
    }
    if (fabs(w.u0) > a.vlimit || fabs(w.u1) > a.vlimit || fabs(w.u2) > a.vlimit) flg |= FLAG_CLAMP;
    return w;
}

template <int STAGE, int TY>
__global__ __launch_bounds__(64 * (TY + 4)) void k_stage(StageArgs a)
{
    constexpr int NR = TY + 4;
    __shared__ double bufW[6][NR][64];       // W(k) of all rows
    __shared__ double bufP[6][TY + 1][64];   // y-plus states of rows 1..TY+1
    __shared__ double bufF[5][TY + 1][64];   // y-face fluxes below rows 2..TY+2

    const int lane = threadIdx.x & 63;
    const int r = threadIdx.x >> 6;  // wave row, 0..TY+3
    const int nx = a.nx, ny = a.ny;
    const int64_t n = a.n, plane = (int64_t)nx * ny;

    const int x = blockIdx.x * kSegOut + lane;   // 0-based column
    const int y = blockIdx.y * TY + r;           // 0-based row
    const int xc = x < nx ? x : nx - 1;          // clamped for loads
    const int yc = y < ny ? y : ny - 1;
    const int64_t col = (int64_t)yc * nx + xc;
    const bool row_out = (r >= 2) && (r < TY + 2) && (y <= ny - 3);
    const bool is_out = row_out && lane >= 2 && lane < 62 && x >= 2 && x <= nx - 3;
    const bool xy_ghost = (x < 2) || (x > nx - 3) || (y < 2) || (y > ny - 3);

    const int kb = a.zlo + blockIdx.z * a.lz;
    const int ke = min(kb + a.lz, a.zhi);
    if (kb >= ke) return;  // uniform across the block

    unsigned int flg = 0;
    double rmax = 0.0;
    const int zs = a.flatz ? 0 : 1;  // z stride in planes

    auto zplane = [&](int k) -> int64_t { return (int64_t)k * plane; };
    auto zghost = [&](int k) -> bool { return !a.flatz && (k < 2 || k > a.nz - 3); };

    // z register queue: primitives of planes k, k+1, k+2 (flat z: all = k)
    W6 q0, q1, q2;
    q0 = load_prims<STAGE>(a, zplane(kb - 2 * zs) + col, flg);  // plane kb-2
    q1 = load_prims<STAGE>(a, zplane(kb - zs) + col, flg);      // kb-1
    q2 = load_prims<STAGE>(a, zplane(kb) + col, flg);           // kb
    // z-plus state of plane kb-1 and face kb-1/2 need planes kb-2..kb+1
    W6 q3 = load_prims<STAGE>(a, zplane(kb + zs) + col, flg);   // kb+1
    Side<double> zp_prev, zmi, zpl;
    double Fz_prev[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (!a.flatz) {
        const bool g1 = xy_ghost || zghost(kb - zs);
        dir_states<2>(q0, q1, q2, sound_of(q1, a.gam), g1, a, zmi, zpl);  // plane kb-1
        zp_prev = zpl;
        const bool g2 = xy_ghost || zghost(kb);
        dir_states<2>(q1, q2, q3, sound_of(q2, a.gam), g2, a, zmi, zpl);  // plane kb
        face_flux<double, 2>(zp_prev, zmi, Fz_prev);
        zp_prev = zpl;
    } else {
        zp_prev.d = zp_prev.p = zp_prev.e = zp_prev.g = 0.0;
        zp_prev.v[0] = zp_prev.v[1] = zp_prev.v[2] = 0.0;
    }
    // queue now: planes kb, kb+1 in (q2, q3); shift so q0=kb, q1=kb+1
    q0 = q2;
    q1 = q3;

    for (int k = kb; k < ke; k++) {
        // ---- leading plane k+2
        q2 = load_prims<STAGE>(a, zplane(k + 2 * zs) + col, flg);
        const bool ghost_k = xy_ghost || zghost(k);
        const double snd = sound_of(q0, a.gam);

        // ---- y direction through LDS
        bufW[0][r][lane] = q0.d;
        bufW[1][r][lane] = q0.u0;
        bufW[2][r][lane] = q0.u1;
        bufW[3][r][lane] = q0.u2;
        bufW[4][r][lane] = q0.p;
        bufW[5][r][lane] = q0.e;
        __syncthreads();
        Side<double> ymi, ypl;
        if (r >= 1 && r <= TY + 2) {
            W6 wm, wp;
            wm.d = bufW[0][r - 1][lane];
            wm.u0 = bufW[1][r - 1][lane];
            wm.u1 = bufW[2][r - 1][lane];
            wm.u2 = bufW[3][r - 1][lane];
            wm.p = bufW[4][r - 1][lane];
            wm.e = bufW[5][r - 1][lane];
            wp.d = bufW[0][r + 1][lane];
            wp.u0 = bufW[1][r + 1][lane];
            wp.u1 = bufW[2][r + 1][lane];
            wp.u2 = bufW[3][r + 1][lane];
            wp.p = bufW[4][r + 1][lane];
            wp.e = bufW[5][r + 1][lane];
            dir_states<1>(wm, q0, wp, snd, ghost_k, a, ymi, ypl);
            if (r <= TY + 1) {
                bufP[0][r - 1][lane] = ypl.d;
                bufP[1][r - 1][lane] = ypl.v[0];
                bufP[2][r - 1][lane] = ypl.v[1];
                bufP[3][r - 1][lane] = ypl.v[2];
                bufP[4][r - 1][lane] = ypl.p;
                bufP[5][r - 1][lane] = ypl.e;
            }
        }
        __syncthreads();
        if (r >= 2 && r <= TY + 2) {
            Side<double> L;
            L.d = bufP[0][r - 2][lane];
            L.v[0] = bufP[1][r - 2][lane];
            L.v[1] = bufP[2][r - 2][lane];
            L.v[2] = bufP[3][r - 2][lane];
            L.p = bufP[4][r - 2][lane];
            L.e = bufP[5][r - 2][lane];
            L.g = a.gam;
            double Fy[5];
            face_flux<double, 1>(L, ymi, Fy);
            for (int i = 0; i < 5; i++) bufF[i][r - 2][lane] = Fy[i];
        }
        __syncthreads();

        if (r >= 2 && r < TY + 2) {
            // ---- x direction through shuffles
            const int lm1 = (lane + 63) & 63, lp1 = (lane + 1) & 63;
            const W6 wxm = shfl_w(q0, lm1);
            const W6 wxp = shfl_w(q0, lp1);
            Side<double> xmi, xpl;
            dir_states<0>(wxm, q0, wxp, snd, ghost_k, a, xmi, xpl);
            Side<double> L;
            L.d = shfl_d(xpl.d, lm1);
            L.v[0] = shfl_d(xpl.v[0], lm1);
            L.v[1] = shfl_d(xpl.v[1], lm1);
            L.v[2] = shfl_d(xpl.v[2], lm1);
            L.p = shfl_d(xpl.p, lm1);
            L.e = shfl_d(xpl.e, lm1);
            L.g = a.gam;
            double Fx0[5];
            face_flux<double, 0>(L, xmi, Fx0);  // face x-1/2
            double Fx1[5];
            for (int i = 0; i < 5; i++) Fx1[i] = shfl_d(Fx0[i], lp1);  // face x+1/2

            // ---- z direction from the register queue: face k+1/2
            Side<double> zmi1, zpl1;
            double Fz1[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            if (!a.flatz) {
                const bool ghost_k1 = xy_ghost || zghost(k + zs);
                dir_states<2>(q0, q1, q2, sound_of(q1, a.gam), ghost_k1, a, zmi1, zpl1);
                face_flux<double, 2>(zp_prev, zmi1, Fz1);
            }

            if (is_out) {
                const int64_t c = zplane(k) + col;
                double Fy1[5];
                for (int i = 0; i < 5; i++) Fy1[i] = bufF[i][r - 1][lane];
                double Fy0[5];
                for (int i = 0; i < 5; i++) Fy0[i] = bufF[i][r - 2][lane];
                double ft[5];
                for (int i = 0; i < 5; i++) {
                    // flux.f90:188-190
                    const double dz = a.flatz ? 0.0 : (Fz1[i] - Fz_prev[i]);
                    const double s = (a.ayz * (Fx1[i] - Fx0[i]) + a.axz * (Fy1[i] - Fy0[i])) + a.axy * dz;
                    ft[i] = -div_by(s * a.dt, a.vol, a.inv_vol, a.vol_pow2);
                }
                double U[5];
                if (STAGE == 1) {
                    // cons = cons + fluxtmp (flux.f90:205) with cons = source(0)(U^n)
                    const double r0 = a.uin[c], m0 = a.uin[n + c], m1 = a.uin[2 * n + c], m2 = a.uin[3 * n + c];
                    const double ek = 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r0;
                    const double E1 = vmax(a.uin[4 * n + c] - ek, a.seps) + ek;
                    U[0] = r0 + ft[0];
                    U[1] = m0 + ft[1];
                    U[2] = m1 + ft[2];
                    U[3] = m2 + ft[3];
                    U[4] = E1 + ft[4];
                } else {
                    // cons = half*(cons_old+cons+fluxtmp) (flux.f90:213), cons_old = source(0)(U^n)
                    const double r0 = a.uold[c], m0 = a.uold[n + c], m1 = a.uold[2 * n + c], m2 = a.uold[3 * n + c];
                    const double ek = 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r0;
                    const double E1 = vmax(a.uold[4 * n + c] - ek, a.seps) + ek;
                    const double old[5] = {r0, m0, m1, m2, E1};
                    for (int i = 0; i < 5; i++) U[i] = 0.5 * (old[i] + a.uin[(int64_t)i * n + c] + ft[i]);
                }
                // floors, flux.f90:225-228
                U[0] = vmax(U[0], a.srho);
                if (U[4] < 0.0) U[4] = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0] + a.seps;
                if (STAGE == 1) {
                    // p* = state() at main.f90:353 on U* with the velocity of :318
                    if (U[0] < 1e-4) flg |= FLAG_LOWRHO;
                    const double ekin = U[0] * 0.5 * (q0.u0 * q0.u0 + q0.u1 * q0.u1 + q0.u2 * q0.u2);
                    a.pout[c] = vmax(U[4] - ekin, a.seps) * a.gm1;
                    for (int i = 0; i < 5; i++) a.uout[(int64_t)i * n + c] = U[i];
                } else {
                    // main.f90:374-515: state, velocity, source(1), state, velocity
                    if (U[0] < 1e-4) flg |= FLAG_LOWRHO;
                    const double v0 = U[1] / U[0], v1 = U[2] / U[0], v2 = U[3] / U[0];
                    if (fabs(v0) > a.vlimit || fabs(v1) > a.vlimit || fabs(v2) > a.vlimit) flg |= FLAG_CLAMP;
                    const double ek = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
                    U[4] = vmax(U[4] - ek, a.seps) + ek;
                    for (int i = 0; i < 5; i++) a.uout[(int64_t)i * n + c] = U[i];
                    // next step: state (main.f90:232) and courant rate (courant.f90:20-21)
                    const double ekin = U[0] * 0.5 * (v0 * v0 + v1 * v1 + v2 * v2);
                    const double p = vmax(U[4] - ekin, a.seps) * a.gm1;
                    const double snd1 = sqrt(a.gam * p / U[0]);
                    const double rt = (snd1 + fabs(v0)) / a.ds[0] + (snd1 + fabs(v1)) / a.ds[1] +
                                      (snd1 + fabs(v2)) / a.ds[2];
                    rmax = vmax(rmax, rt);
                }
            }
            if (!a.flatz) {
                for (int i = 0; i < 5; i++) Fz_prev[i] = Fz1[i];
                zp_prev = zpl1;
            }
        }
        q0 = q1;
        q1 = q2;
    }

    if (flg) atomicOr(a.flags, flg);
    if (STAGE == 2) {
        for (int off = 32; off > 0; off >>= 1) rmax = vmax(rmax, __shfl_xor(rmax, off, 64));
        if (lane == 0 && rmax > 0.0) atomicMax(a.rate, (unsigned long long)__double_as_longlong(rmax));
    }
}

// -------------------------------------------------------------- helpers
// u = m/rho, p = state() of a regular cell, adindx = gamma: the velocity and
// pressure planes the reference holds at the end of a step (main.f90:513-515)
__global__ __launch_bounds__(256) void k_materialize(int64_t n, const double *__restrict__ cons,
                                                    double *__restrict__ u, double *__restrict__ p,
                                                    double *__restrict__ ad, double gam, double seps)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double r = cons[c];
    const double v0 = cons[n + c] / r, v1 = cons[2 * n + c] / r, v2 = cons[3 * n + c] / r;
    u[c] = v0;
    u[n + c] = v1;
    u[2 * n + c] = v2;
    const double ekin = r * 0.5 * (v0 * v0 + v1 * v1 + v2 * v2);
    p[c] = vmax(cons[4 * n + c] - ekin, seps) * (gam - 1.0);
    ad[c] = gam;
}

// Regularity of the exact path's state: rho>=1e-4 and u == m/rho bitwise on
// every cell, plus the Courant rate of main.f90:232-241 for a regular state.
__global__ __launch_bounds__(256) void k_regular(int64_t n, const double *__restrict__ cons,
                                                const double *__restrict__ u, double gam, double seps,
                                                double dx, double dy, double dz, unsigned int *flags,
                                                unsigned long long *rate)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double rt = 0.0;
    unsigned int flg = 0;
    if (c < n) {
        const double r = cons[c];
        const double v0 = cons[n + c] / r, v1 = cons[2 * n + c] / r, v2 = cons[3 * n + c] / r;
        if (r < 1e-4) flg |= FLAG_LOWRHO;
        if (__double_as_longlong(v0) != __double_as_longlong(u[c]) ||
            __double_as_longlong(v1) != __double_as_longlong(u[n + c]) ||
            __double_as_longlong(v2) != __double_as_longlong(u[2 * n + c]))
            flg |= FLAG_CLAMP;
        const double ekin = r * 0.5 * (v0 * v0 + v1 * v1 + v2 * v2);
        const double p = vmax(cons[4 * n + c] - ekin, seps) * (gam - 1.0);
        const double snd = sqrt(gam * p / r);
        rt = (snd + fabs(v0)) / dx + (snd + fabs(v1)) / dy + (snd + fabs(v2)) / dz;
    }
    if (flg) atomicOr(flags, flg);
    for (int off = 32; off > 0; off >>= 1) rt = vmax(rt, __shfl_xor(rt, off, 64));
    if ((threadIdx.x & 63) == 0 && rt > 0.0) atomicMax(rate, (unsigned long long)__double_as_longlong(rt));
}

// ghost fill of the 5 conserved planes and (optionally) p*, shallow tables only
__global__ __launch_bounds__(256) void k_ghosts6(int64_t n, int64_t ng, const int32_t *__restrict__ dst,
                                                const int32_t *__restrict__ src,
                                                const int32_t *__restrict__ sign, double *__restrict__ planes,
                                                double *__restrict__ pstar)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ng) return;
    const int64_t d = dst[k], s = src[k];
    const int sg = sign[k];
    planes[d] = planes[s];
    for (int f = 1; f <= 3; f++) {
        const double v = planes[f * n + s];
        planes[f * n + d] = ((sg >> (f - 1)) & 1) ? -v : v;
    }
    planes[4 * n + d] = planes[4 * n + s];
    if (pstar) pstar[d] = pstar[s];
}

// ----------------------------------------------------------------- host
static bool is_pow2(double d)
{
    int e;
    double m = std::frexp(d, &e);
    return m == 0.5;
}

static StageArgs base_args(bx_ctx *c, double dt)
{
    StageArgs a{};
    a.n = c->n;
    a.nx = c->nx;
    a.ny = c->ny;
    a.nz = c->nz;
    a.flatz = c->flatz ? 1 : 0;
    a.zlo = c->flatz ? 0 : 2;
    a.zhi = c->flatz ? c->nz : c->nz - 2;
    a.lz = c->fused_lz;
    a.dt = dt;
    const double ds[3] = {c->prm.dx, c->prm.dy, c->prm.dz};
    for (int d = 0; d < 3; d++) {
        a.ds[d] = ds[d];
        a.inv_ds[d] = 1.0 / ds[d];
        a.ds_pow2[d] = is_pow2(ds[d]) ? 1 : 0;
    }
    const double dx = ds[0], dy = ds[1], dz = ds[2];
    a.axy = dy * dx;
    a.ayz = dy * dz;
    a.axz = dx * dz;
    a.vol = dx * dy * dz;
    a.inv_vol = 1.0 / a.vol;
    a.vol_pow2 = is_pow2(a.vol) ? 1 : 0;
    a.gam = c->prm.gammafix;
    a.gm1 = c->prm.gammafix - 1.0;
    a.seps = c->prm.small_eps;
    a.srho = c->prm.small_rho;
    a.vlimit = c->prm.vlimit;
    a.flags = &c->dsc->flags;
    return a;
}

template <int STAGE, int TY>
static int launch_stage_ty(bx_ctx *c, const StageArgs &a)
{
    const unsigned gx = (unsigned)((c->nx - 4 + kSegOut - 1) / kSegOut);
    const unsigned gy = (unsigned)((c->ny - 4 + TY - 1) / TY);
    const int nzi = a.zhi - a.zlo;
    const unsigned gz = (unsigned)((nzi + a.lz - 1) / a.lz);
    hipLaunchKernelGGL((k_stage<STAGE, TY>), dim3(gx, gy, gz), dim3(64 * (TY + 4)), 0, c->stream, a);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

template <int STAGE>
static int launch_stage(bx_ctx *c, const StageArgs &a)
{
    switch (c->fused_ty) {
    case 4: return launch_stage_ty<STAGE, 4>(c, a);
    case 8: return launch_stage_ty<STAGE, 8>(c, a);
    case 12: return launch_stage_ty<STAGE, 12>(c, a);
    default: return set_error(BX_ERR_ARG, "fused: unsupported row tile");
    }
}

int launch_materialize(bx_ctx *c)
{
    hipLaunchKernelGGL(k_materialize, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->cons, c->u, c->p,
                       c->adindx, c->prm.gammafix, c->prm.small_eps);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

// Ghost table shallow, allocation of the fused buffers.
int fused_prepare(bx_ctx *c)
{
    if (c->nscratch > 0)
        return set_error(BX_ERR_UNSUPPORTED, "fused step needs every ghost donor to be an interior cell");
    if (c->grav[0] != 0.0 || c->grav[1] != 0.0 || c->grav[2] != 0.0)
        return set_error(BX_ERR_UNSUPPORTED, "fused step assumes zero gravity");
    const size_t pb = sizeof(double) * (size_t)c->n;
    if (!c->ustar) BX_HIP(hipMalloc((void **)&c->ustar, 5 * pb));
    if (!c->unext) BX_HIP(hipMalloc((void **)&c->unext, 5 * pb));
    if (!c->pstar) BX_HIP(hipMalloc((void **)&c->pstar, pb));
    if (c->fused_ty == 0) c->fused_ty = 8;
    if (c->fused_lz == 0) c->fused_lz = 32;
    return BX_OK;
}

// Flags and Courant rate of the current (exact-path) state; ghosts must hold
// copies of their donors, which the exact sequence guarantees after a step.
int launch_regular_check(bx_ctx *c)
{
    BX_HIP(hipMemsetAsync(&c->dsc->flags, 0, sizeof(unsigned int), c->stream));
    BX_HIP(hipMemsetAsync(&c->dsc->rate_bits[0], 0, sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_regular, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->cons, c->u,
                       c->prm.gammafix, c->prm.small_eps, c->prm.dx, c->prm.dy, c->prm.dz, &c->dsc->flags,
                       &c->dsc->rate_bits[0]);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

static int ghosts6(bx_ctx *c, double *planes, double *pstar)
{
    if (c->nghost == 0) return BX_OK;
    hipLaunchKernelGGL(k_ghosts6, dim3(blocks_for(c->nghost, 256)), dim3(256), 0, c->stream, c->n, c->nghost,
                       c->ghost_dst, c->ghost_src, c->ghost_sign, planes, pstar);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

// One fused step from c->cons (U^n, regular) with full dt; result in c->unext.
// The Courant rate of U^{n+1} accumulates into rate_bits[0] (zeroed here) and
// irregular branches into flags (zeroed here).
int fused_step(bx_ctx *c, double dt)
{
    int rc;
    BX_HIP(hipMemsetAsync(&c->dsc->flags, 0, sizeof(unsigned int) + 0, c->stream));
    BX_HIP(hipMemsetAsync(&c->dsc->rate_bits[0], 0, sizeof(unsigned long long), c->stream));
    StageArgs a = base_args(c, dt);
    a.uin = c->cons;
    a.uout = c->ustar;
    a.pout = c->pstar;
    if (c->prof_on) BX_HIP(hipEventRecord(c->ev[0], c->stream));
    if ((rc = launch_stage<1>(c, a))) return rc;
    if (c->prof_on) BX_HIP(hipEventRecord(c->ev[1], c->stream));
    if ((rc = ghosts6(c, c->ustar, c->pstar))) return rc;
    StageArgs b = base_args(c, dt);
    b.uin = c->ustar;
    b.pin = c->pstar;
    b.uold = c->cons;
    b.uout = c->unext;
    b.rate = &c->dsc->rate_bits[0];
    if (c->prof_on) BX_HIP(hipEventRecord(c->ev[2], c->stream));
    if ((rc = launch_stage<2>(c, b))) return rc;
    if (c->prof_on) BX_HIP(hipEventRecord(c->ev[3], c->stream));
    if ((rc = ghosts6(c, c->unext, nullptr))) return rc;
    return BX_OK;
}

}  // namespace bx
