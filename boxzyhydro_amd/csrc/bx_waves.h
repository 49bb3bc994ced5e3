// bx_waves.h -- the wave bodies of the fused stage kernels (bx_fused.hip), as
// device functions of a wave's (virtual) block and wave coordinates, so that
// a kernel can run them for one launch grid (the stage kernels) or a
// workgroup can run them over a grid of its own (the batched grids).  The
// helpers of the fused step's host sequence (materialize, the regularity
// check, the shallow ghost fill) are here per cell for the same reason.
#pragma once

#include "bx_stage.h"

namespace bx {

// Per-wave reduction of a non-negative max into *dst; the atomic is skipped
// when the stored maximum is already larger (almost always, after warm-up).
__device__ __forceinline__ void wave_max_to(double v, unsigned long long *dst)
{
    for (int off = 32; off > 0; off >>= 1) v = vmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0 && v > 0.0) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(v);
        const unsigned long long cur = __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (b > cur) atomicMax(dst, b);
    }
}

// ---------------------------------------------------------------------------
// y-direction kernel: each wave holds 64 consecutive interior x-cells of one
// z-plane and marches in y, keeping the primitives of rows j..j+2, the y-plus
// state of row j and the face flux of j-1/2 in registers.  It writes the
// y-term of fluxtmp, E = areaxz*(fy(2)-fy(1)) (flux.f90:189), in the storage
// precision (FP64; FP32 in the mixed form, where the fluxes are FP32).
template <int STAGE, typename S>
__device__ __forceinline__ void ydir_wave(const StageArgs &a, S *__restrict__ Ey, int ly, unsigned bx, unsigned by,
                                          unsigned bz, int w)
{
    const int lane = threadIdx.x & 63;
    const int nx = a.nx, ny = a.ny;
    const int64_t n = a.n, plane = (int64_t)nx * ny;
    const int x = 2 + bx * 64 + lane;
    const bool xin = x <= nx - 3;
    const int xc = xin ? x : nx - 3;
    const int z = a.zlo + by * 4 + w;
    if (z >= a.zhi) return;  // whole wave; no block-level synchronisation in this kernel
    const int jb = 2 + bz * ly;
    const int je = min(jb + ly, ny - 2);
    if (jb >= je) return;
    const bool zg = !a.flatz && (z < 2 || z > a.nz - 3);
    const int64_t zbase = (int64_t)z * plane;
    const unsigned xoff = (unsigned)xc * (unsigned)sizeof(S);
    auto cell = [&](int j) -> int64_t { return zbase + (int64_t)j * nx; };  // uniform part of the cell offset
    auto yghost = [&](int j) -> bool { return zg || j < 2 || j > ny - 3; };
    unsigned int flg = 0;

    W6<S> q0 = load_prims<STAGE, S>(a, cell(jb - 2), xoff, flg);
    W6<S> q1 = load_prims<STAGE, S>(a, cell(jb - 1), xoff, flg);
    W6<S> q2 = load_prims<STAGE, S>(a, cell(jb), xoff, flg);
    W6<S> q3 = load_prims<STAGE, S>(a, cell(jb + 1), xoff, flg);
    Side<S> mi, pl, yp;
    dir_states<1>(q0, q1, q2, sound_of(q1, a.gam), yghost(jb - 1), a, mi, pl);
    yp = pl;
    dir_states<1>(q1, q2, q3, sound_of(q2, a.gam), yghost(jb), a, mi, pl);
    S F0[5];
    face_flux<S, 1>(yp, mi, F0);
    yp = pl;
    q0 = q2;
    q1 = q3;
    Raw<S> nxt = load_raw<STAGE, S>(a, cell(jb + 2), xoff);
    for (int j = jb; j < je; j++) {
        const Raw<S> cur = nxt;
        nxt = load_raw<STAGE, S>(a, cell(j + 3), xoff);  // row j+3 <= ny-1: prefetch for the next row
        q2 = prims<STAGE, S>(a, cur, flg);
        dir_states<1>(q0, q1, q2, sound_of(q1, a.gam), yghost(j + 1), a, mi, pl);
        S F1[5];
        face_flux<S, 1>(yp, mi, F1);
        if (xin) {
            const int64_t c = cell(j);
            for (int i = 0; i < 5; i++) elem<S>(Ey, (int64_t)i * n + c, xoff) = (S)(a.axz * ((double)F1[i] - (double)F0[i]));
        }
        for (int i = 0; i < 5; i++) F0[i] = F1[i];
        yp = pl;
        q0 = q1;
        q1 = q2;
    }
    if (flg) atomicOr(a.flags, flg);
}

// The stage update of one output cell c (uniform plane offset c + per-lane
// byte offset coff) from U (ui), U^n (uo, stage 2) and fluxtmp (ft), and
// everything pointwise up to the next stage: stage 1 writes U* and the
// stale-velocity pressure p* (q0: the cell's primitives of the stage input);
// stage 2 writes U^{n+1} and accumulates the next step's Courant rate.
//
// G (gravity, FP64): source() applies the cell's force from a.gf, whose
// planes the cell reads here again.  Stage 2 then also writes the pressure
// the exact sequence leaves (its state call after source(1), main.f90:511-513,
// takes the velocity from before the kick) to pout.
template <int STAGE, typename S, bool G = false>
__device__ __forceinline__ void stage_update(const StageArgs &a, const double *ui, const double *uo, const double *ft,
                                             const W6<S> &q0, int64_t c, unsigned coff, unsigned int &flg,
                                             double &rmax, bool park = false)
{
    const int64_t n = a.n;
    double U[5];
    double g[3] = {0.0, 0.0, 0.0};
    const double dth = a.dt * 0.5;
    if constexpr (G) {
        static_assert(sizeof(S) == 8, "gravity is FP64 only");
        for (int d = 0; d < 3; d++) g[d] = elem<double>(a.gf, (int64_t)d * n + c, coff);
    }
    if constexpr (G) {
        // cons_old = source(0)(U^n) with the velocity of the last velocity
        // call, u_old = m/rho (stage 1: cons = cons_old + fluxtmp, flux.f90:205;
        // stage 2: cons = half*(cons_old+cons+fluxtmp), flux.f90:213)
        const double *un = STAGE == 1 ? ui : uo;
        const double r0 = un[0], m[3] = {un[1], un[2], un[3]};
        double mk[3];
        grav_momenta(r0, m, g, dth, mk);
        const double num[4] = {m[0], m[1], m[2], 0.5 * (mk[0] * mk[0] + mk[1] * mk[1] + mk[2] * mk[2])};
        double quo[4];
        div_shared<4>(num, r0, quo);
        const double ek = quo[3];
        const double E1 = floor_at(grav_energy_arg(r0, un[4], ek, quo, g, dth), a.seps) + ek;
        const double old[5] = {r0, mk[0], mk[1], mk[2], E1};
        if (STAGE == 1)
            for (int i = 0; i < 5; i++) U[i] = old[i] + ft[i];
        else
            for (int i = 0; i < 5; i++) U[i] = 0.5 * (old[i] + ui[i] + ft[i]);
    } else if (STAGE == 1) {
        // cons = cons + fluxtmp (flux.f90:205) with cons = source(0)(U^n)
        const double r0 = ui[0], m0 = ui[1], m1 = ui[2], m2 = ui[3];
        const double ek = 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r0;
        const double E1 = floor_at(ui[4] - ek, a.seps) + ek;
        U[0] = r0 + ft[0];
        U[1] = m0 + ft[1];
        U[2] = m1 + ft[2];
        U[3] = m2 + ft[3];
        U[4] = E1 + ft[4];
    } else {
        // cons = half*(cons_old+cons+fluxtmp) (flux.f90:213), cons_old = source(0)(U^n)
        const double r0 = uo[0], m0 = uo[1], m1 = uo[2], m2 = uo[3];
        const double ek = 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r0;
        const double E1 = floor_at(uo[4] - ek, a.seps) + ek;
        const double old[5] = {r0, m0, m1, m2, E1};
        for (int i = 0; i < 5; i++) U[i] = 0.5 * (old[i] + ui[i] + ft[i]);
    }
    // floors, flux.f90:225-228
    if (!(U[0] > a.srho) || U[4] < 0.0) flg |= FLAG_FLOOR;
    U[0] = floor_at(U[0], a.srho);
    if (U[4] < 0.0) U[4] = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0] + a.seps;
    if constexpr (STAGE == 2 && sizeof(S) == 8) {
        // the donor of a two-level corner ghost: park the result before source(1)
        if (park) {
            const int64_t cell = c + (int64_t)(coff / sizeof(S));
            for (int j = 0; j < a.nd2; j++)
                if (a.d2cell[j] == cell)
                    for (int i = 0; i < 5; i++) a.d2ubar[6 * j + i] = U[i];
        }
    }
    if (STAGE == 1) {
        // p* = state() at main.f90:353 on U* with the velocity of :318
        if (U[0] < 1e-4) flg |= FLAG_LOWRHO;
        S Us[5];
        for (int i = 0; i < 5; i++) Us[i] = (S)U[i];
        const double r1 = Us[0], E1 = Us[4];
        const double w0 = q0.u0, w1 = q0.u1, w2 = q0.u2;
        const double ekin = r1 * 0.5 * (w0 * w0 + w1 * w1 + w2 * w2);
        if (!(E1 - ekin > a.seps)) flg |= FLAG_FLOOR;
        elem<S>(a.pout, c, coff) = (S)(floor_at(E1 - ekin, a.seps) * a.gm1);
        for (int i = 0; i < 5; i++) elem<S>(a.uout, (int64_t)i * n + c, coff) = Us[i];
    } else {
        // main.f90:374-515: state, velocity, source(1), state, velocity
        if (U[0] < 1e-4) flg |= FLAG_LOWRHO;
        double v0, v1, v2;
        if constexpr (G) {
            // velocity u = m/rho, the kick of source(1), and the velocity of
            // the kicked momenta: seven quotients by rho, one reciprocal
            const double m[3] = {U[1], U[2], U[3]};
            double mk[3];
            grav_momenta(U[0], m, g, dth, mk);
            const double num[7] = {m[0], m[1], m[2], mk[0], mk[1], mk[2],
                                   0.5 * (mk[0] * mk[0] + mk[1] * mk[1] + mk[2] * mk[2])};
            double quo[7];
            div_shared<7>(num, U[0], quo);
            if (fabs(quo[0]) > a.vlimit || fabs(quo[1]) > a.vlimit || fabs(quo[2]) > a.vlimit) flg |= FLAG_CLAMP;
            const double ek = quo[6];
            const double arg = grav_energy_arg(U[0], U[4], ek, quo, g, dth);
            if (!(arg > a.seps)) flg |= FLAG_FLOOR;
            U[4] = floor_at(arg, a.seps) + ek;
            for (int d = 0; d < 3; d++) U[1 + d] = mk[d];
            // the state after source(1) with the velocity from before it
            const double ekp = U[0] * 0.5 * (quo[0] * quo[0] + quo[1] * quo[1] + quo[2] * quo[2]);
            elem<S>(a.pout, c, coff) = floor_at(U[4] - ekp, a.seps) * a.gm1;
            v0 = quo[3];
            v1 = quo[4];
            v2 = quo[5];
            if (fabs(v0) > a.vlimit || fabs(v1) > a.vlimit || fabs(v2) > a.vlimit) flg |= FLAG_CLAMP;
        } else {
            // u = m/rho and the kinetic energy of the source re-round, quotients
            // by rho with one shared reciprocal
            const double num[4] = {U[1], U[2], U[3], 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3])};
            double quo[4];
            div_shared<4>(num, U[0], quo);
            v0 = quo[0];
            v1 = quo[1];
            v2 = quo[2];
            if (fabs(v0) > a.vlimit || fabs(v1) > a.vlimit || fabs(v2) > a.vlimit) flg |= FLAG_CLAMP;
            const double ek = quo[3];
            if (!(U[4] - ek > a.seps)) flg |= FLAG_FLOOR;
            U[4] = floor_at(U[4] - ek, a.seps) + ek;
        }
        S Us[5];
        for (int i = 0; i < 5; i++) {
            Us[i] = (S)U[i];
            elem<S>(a.uout, (int64_t)i * n + c, coff) = Us[i];
        }
        // next step: state (main.f90:232) and courant rate (courant.f90:20-21)
        double p, w0, w1, w2, rr;
        if constexpr (sizeof(S) == 8) {
            rr = U[0];
            w0 = v0;
            w1 = v1;
            w2 = v2;
            const double ekin = U[0] * 0.5 * (v0 * v0 + v1 * v1 + v2 * v2);
            p = floor_at(U[4] - ekin, a.seps) * a.gm1;
        } else {
            rr = Us[0];
            const double inv = 1.0 / rr;
            w0 = (double)Us[1] * inv;
            w1 = (double)Us[2] * inv;
            w2 = (double)Us[3] * inv;
            const double ekin = 0.5 * ((double)Us[1] * w0 + (double)Us[2] * w1 + (double)Us[3] * w2);
            p = floor_at((double)Us[4] - ekin, a.seps) * a.gm1;
        }
        const double sq = a.gam * p / rr;
        double snd1;
        sqrt_n<1>(&sq, &snd1);
        const double rt = div_by(snd1 + fabs(w0), a.ds[0], a.inv_ds[0], a.ds_pow2[0]) +
                          div_by(snd1 + fabs(w1), a.ds[1], a.inv_ds[1], a.ds_pow2[1]) +
                          div_by(snd1 + fabs(w2), a.ds[2], a.inv_ds[2], a.ds_pow2[2]);
        rmax = vmax(rmax, rt);
    }
}

// ---------------------------------------------------------------------------
// x/z kernel: each wave holds 64 consecutive cells of the row-major x-y plane
// (lanes 2..61 are outputs where interior, 0,1,62,63 the x-halo; a wave may
// run across the end of one row into the next, whose ghost columns separate
// them; x-neighbours through shuffles) and marches in z with the primitives of planes k..k+2, the z-plus
// state of plane k and the z-face flux of k-1/2 in registers.  It completes
// fluxtmp with the y-term from k_ydir in the reference's summation order and
// applies the stage update and everything pointwise up to the next stage.
template <int STAGE, typename S, bool PFE>
__device__ __forceinline__ void xz_wave(const StageArgs &a, const S *__restrict__ Ey, unsigned bx, unsigned bz, int w)
{
    const int lane = threadIdx.x & 63;
    const int nx = a.nx;
    const int64_t n = a.n, plane = (int64_t)nx * a.ny;

    const int64_t F0 = 2 * (int64_t)nx - 2 + (int64_t)(bx * 4 + w) * kSegOut;  // lane 0
    if (F0 + 2 >= (int64_t)(a.ny - 2) * nx) return;  // whole wave past the interior rows
    int64_t F = F0 + lane;
    if (F >= plane) F = plane - 1;  // halo lanes of the last wave: clamp for the loads
    const int y = (int)(F / nx), x = (int)(F - (int64_t)y * nx);
    const unsigned coff = (unsigned)F * (unsigned)sizeof(S);  // per-lane byte offset within a z-plane
    const bool x_ghost = (x < 2) || (x > nx - 3) || (y < 2) || (y > a.ny - 3);
    const bool is_out = lane >= 2 && lane < 62 && !x_ghost;

    const int kb = a.zlo + bz * a.lz;
    const int ke = min(kb + a.lz, a.zhi);
    if (kb >= ke) return;

    unsigned int flg = 0;
    double rmax = 0.0;
    const S gamS = S(a.gam);
    const int lm1 = (lane + 63) & 63, lp1 = (lane + 1) & 63;

    auto zplane = [&](int k) -> int64_t { return (int64_t)k * plane; };
    auto zghost = [&](int k) -> bool { return !a.flatz && (k < 2 || k > a.nz - 3); };

    W6<S> q0, q1, q2;
    S cs0 = S(0);  // sound speed of plane k (q0), carried from the z-states of the previous plane
    Side<S> zp_prev, zmi, zpl;
    S Fz_prev[5] = {S(0), S(0), S(0), S(0), S(0)};
    if (!a.flatz) {
        q0 = load_prims<STAGE, S>(a, zplane(kb - 2), coff, flg);
        q1 = load_prims<STAGE, S>(a, zplane(kb - 1), coff, flg);
        q2 = load_prims<STAGE, S>(a, zplane(kb), coff, flg);
        const W6<S> q3 = load_prims<STAGE, S>(a, zplane(kb + 1), coff, flg);
        dir_states<2>(q0, q1, q2, sound_of(q1, a.gam), x_ghost || zghost(kb - 1), a, zmi, zpl);
        zp_prev = zpl;
        cs0 = sound_of(q2, a.gam);
        dir_states<2>(q1, q2, q3, cs0, x_ghost || zghost(kb), a, zmi, zpl);
        face_flux<S, 2>(zp_prev, zmi, Fz_prev);
        zp_prev = zpl;
        q0 = q2;
        q1 = q3;
    } else {
        q0 = load_prims<STAGE, S>(a, zplane(kb), coff, flg);
        q1 = q0;
        zp_prev.d = zp_prev.p = zp_prev.e = zp_prev.g = S(0);
        zp_prev.v[0] = zp_prev.v[1] = zp_prev.v[2] = S(0);
    }

    auto lead = [&](int k) -> int { return a.flatz ? min(k + 1, ke - 1) : min(k + 2, a.nz - 1); };
    Raw<S> nxt = load_raw<STAGE, S>(a, zplane(lead(kb)), coff);
    for (int k = kb; k < ke; k++) {
        const Raw<S> cur = nxt;
        nxt = load_raw<STAGE, S>(a, zplane(lead(k + 1)), coff);  // prefetch the next leading plane
        q2 = prims<STAGE, S>(a, cur, flg);
        const bool ghost_k = x_ghost || zghost(k);
        // the y-term of this plane (written by k_ydir, not yet in cache): load early
        const int64_t cu = zplane(k);
        double ey[5];
        if (PFE)
            for (int i = 0; i < 5; i++) ey[i] = elem<S>(Ey, (int64_t)i * n + cu, coff);

        // ---- z face k+1/2 from the register queue
        S Fz1[5] = {S(0), S(0), S(0), S(0), S(0)};
        Side<S> zpl1;
        S cs1 = S(0);
        if (!a.flatz) {
            Side<S> zmi1;
            cs1 = sound_of(q1, a.gam);
            dir_states<2>(q0, q1, q2, cs1, x_ghost || zghost(k + 1), a, zmi1, zpl1);
            face_flux<S, 2>(zp_prev, zmi1, Fz1);
        } else {
            cs0 = sound_of(q0, a.gam);
        }

        // ---- x faces through shuffles
        S dFx[5];
        {
            const S snd = cs0;
            const W6<S> wxm = shfl_w(q0, lm1);
            const W6<S> wxp = shfl_w(q0, lp1);
            Side<S> xmi, xpl;
            dir_states<0>(wxm, q0, wxp, snd, ghost_k, a, xmi, xpl);
            Side<S> L;
            L.d = shfl_v(xpl.d, lm1);
            L.v[0] = shfl_v(xpl.v[0], lm1);
            L.v[1] = shfl_v(xpl.v[1], lm1);
            L.v[2] = shfl_v(xpl.v[2], lm1);
            L.p = shfl_v(xpl.p, lm1);
            L.e = shfl_v(xpl.e, lm1);
            L.g = gamS;
            S Fx0[5];
            face_flux<S, 0>(L, xmi, Fx0);  // face x-1/2
            for (int i = 0; i < 5; i++) dFx[i] = Fx0[i];
        }
        double ddx[5];
        for (int i = 0; i < 5; i++) {
            const S fx1 = shfl_v(dFx[i], lp1);  // face x+1/2
            ddx[i] = (double)fx1 - (double)dFx[i];
        }

        if (is_out) {
            const int64_t c = zplane(k);
            double ft[5];
            if (!PFE)
                for (int i = 0; i < 5; i++) ey[i] = elem<S>(Ey, (int64_t)i * n + c, coff);
            double ui[5], uo[5];
            for (int i = 0; i < 5; i++) {
                ui[i] = elem<S>(a.uin, (int64_t)i * n + c, coff);
                uo[i] = STAGE == 2 ? (double)elem<S>(a.uold, (int64_t)i * n + c, coff) : 0.0;
            }
            for (int i = 0; i < 5; i++) {
                // flux.f90:188-190 in the reference's order, accumulated in FP64
                const double dz = a.flatz ? 0.0 : ((double)Fz1[i] - (double)Fz_prev[i]);
                const double s = (a.ayz * ddx[i] + ey[i]) + a.axy * dz;
                ft[i] = -div_by(s * a.dt, a.vol, a.inv_vol, a.vol_pow2);
            }
            stage_update<STAGE, S>(a, ui, uo, ft, q0, c, coff, flg, rmax);
        }
        if (!a.flatz) {
            for (int i = 0; i < 5; i++) Fz_prev[i] = Fz1[i];
            zp_prev = zpl1;
            cs0 = cs1;
        }
        q0 = q1;
        q1 = q2;
        if (a.flatz) q0 = q2;
    }

    if (flg) atomicOr(a.flags, flg);
    if (STAGE == 2) wave_max_to(rmax, a.rate);
}

// u = m/rho, p = state() of a regular cell, adindx = gamma: the velocity and
// pressure planes the reference holds at the end of a step (main.f90:513-515)
__device__ __forceinline__ void materialize_cell(int64_t n, int64_t c, const double *__restrict__ cons,
                                                 double *__restrict__ u, double *__restrict__ p,
                                                 double *__restrict__ ad, double gam, double seps)
{
    const double r = cons[c];
    const double v0 = cons[n + c] / r, v1 = cons[2 * n + c] / r, v2 = cons[3 * n + c] / r;
    u[c] = v0;
    u[n + c] = v1;
    u[2 * n + c] = v2;
    const double ekin = r * 0.5 * (v0 * v0 + v1 * v1 + v2 * v2);
    if (p) p[c] = vmax(cons[4 * n + c] - ekin, seps) * (gam - 1.0);
    ad[c] = gam;
}

// Regularity of a cell of the exact path's state (rho >= 1e-4 and u == m/rho
// bitwise; flags returned) and its Courant rate (main.f90:232-241)
__device__ __forceinline__ double regular_cell(int64_t n, int64_t c, const double *__restrict__ cons,
                                               const double *__restrict__ u, double gam, double seps, double dx,
                                               double dy, double dz, unsigned int &flg)
{
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
    return (snd + fabs(v0)) / dx + (snd + fabs(v1)) / dy + (snd + fabs(v2)) / dz;
}

// ghost fill entry k of the 5 conserved planes and (optionally) p*, shallow tables only
template <typename S>
__device__ __forceinline__ void ghost6_one(int64_t n, int64_t k, const int32_t *__restrict__ dst,
                                           const int32_t *__restrict__ src, const int32_t *__restrict__ sign,
                                           S *__restrict__ planes, S *__restrict__ pstar)
{
    const int64_t d = dst[k], s = src[k];
    const int sg = sign[k];
    planes[d] = planes[s];
    for (int f = 1; f <= 3; f++) {
        const S v = planes[f * n + s];
        planes[f * n + d] = ((sg >> (f - 1)) & 1) ? -v : v;
    }
    planes[4 * n + d] = planes[4 * n + s];
    if (pstar) pstar[d] = pstar[s];
}

}  // namespace bx
