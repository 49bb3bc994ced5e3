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
// Work decomposition (gfx950, wave64), two kernels per stage, every wave
// independent (no LDS, no block barriers), each face evaluated once:
//   * k_ydir: a wave holds 64 interior x-cells of one z-plane and marches in y
//     with a register queue; it writes the y-term of fluxtmp (FP64 planes).
//   * k_xz: a wave holds 64 cells of one interior row (lanes 2..61 output,
//     0,1,62,63 the x-halo; x-neighbours through wave shuffles) and marches in
//     z with a register queue; it adds the x- and z-terms in the reference's
//     order and applies the stage update.
#include <cmath>

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
__global__ __launch_bounds__(256) void k_ydir(StageArgs a0, S *__restrict__ Ey, int ly)
{
    const StageArgs a = with_dt(a0);
    const int lane = threadIdx.x & 63;
    const int w = wave_id();
    const int nx = a.nx, ny = a.ny;
    const int64_t n = a.n, plane = (int64_t)nx * ny;
    const int x = 2 + blockIdx.x * 64 + lane;
    const bool xin = x <= nx - 3;
    const int xc = xin ? x : nx - 3;
    const int z = a.zlo + blockIdx.y * 4 + w;
    if (z >= a.zhi) return;  // whole wave; no block-level synchronisation in this kernel
    const int jb = 2 + blockIdx.z * ly;
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

// ---------------------------------------------------------------------------
// x/z kernel: each wave holds 64 consecutive cells of the row-major x-y plane
// (lanes 2..61 are outputs where interior, 0,1,62,63 the x-halo; a wave may
// run across the end of one row into the next, whose ghost columns separate
// them; x-neighbours through shuffles) and marches in z with the primitives of planes k..k+2, the z-plus
// state of plane k and the z-face flux of k-1/2 in registers.  It completes
// fluxtmp with the y-term from k_ydir in the reference's summation order and
// applies the stage update and everything pointwise up to the next stage.
template <int STAGE, typename S, int OCC, bool PFE>
__global__ __launch_bounds__(256, OCC) void k_xz(StageArgs a0, const S *__restrict__ Ey)
{
    const StageArgs a = with_dt(a0);
    const int lane = threadIdx.x & 63;
    const int w = wave_id();
    const int nx = a.nx;
    const int64_t n = a.n, plane = (int64_t)nx * a.ny;

    const int64_t F0 = 2 * (int64_t)nx - 2 + (int64_t)(blockIdx.x * 4 + w) * kSegOut;  // lane 0
    if (F0 + 2 >= (int64_t)(a.ny - 2) * nx) return;  // whole wave past the interior rows
    int64_t F = F0 + lane;
    if (F >= plane) F = plane - 1;  // halo lanes of the last wave: clamp for the loads
    const int y = (int)(F / nx), x = (int)(F - (int64_t)y * nx);
    const unsigned coff = (unsigned)F * (unsigned)sizeof(S);  // per-lane byte offset within a z-plane
    const bool x_ghost = (x < 2) || (x > nx - 3) || (y < 2) || (y > a.ny - 3);
    const bool is_out = lane >= 2 && lane < 62 && !x_ghost;

    const int kb = a.zlo + blockIdx.z * a.lz;
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
            double U[5];
            if (STAGE == 1) {
                // cons = cons + fluxtmp (flux.f90:205) with cons = source(0)(U^n)
                const double r0 = ui[0], m0 = ui[1], m1 = ui[2], m2 = ui[3];
                const double ek = 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r0;
                const double E1 = vmax(ui[4] - ek, a.seps) + ek;
                U[0] = r0 + ft[0];
                U[1] = m0 + ft[1];
                U[2] = m1 + ft[2];
                U[3] = m2 + ft[3];
                U[4] = E1 + ft[4];
            } else {
                // cons = half*(cons_old+cons+fluxtmp) (flux.f90:213), cons_old = source(0)(U^n)
                const double r0 = uo[0], m0 = uo[1], m1 = uo[2], m2 = uo[3];
                const double ek = 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r0;
                const double E1 = vmax(uo[4] - ek, a.seps) + ek;
                const double old[5] = {r0, m0, m1, m2, E1};
                for (int i = 0; i < 5; i++) U[i] = 0.5 * (old[i] + ui[i] + ft[i]);
            }
            // floors, flux.f90:225-228
            if (!(U[0] > a.srho) || U[4] < 0.0) flg |= FLAG_FLOOR;
            U[0] = vmax(U[0], a.srho);
            if (U[4] < 0.0) U[4] = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0] + a.seps;
            if (STAGE == 1) {
                // p* = state() at main.f90:353 on U* with the velocity of :318
                if (U[0] < 1e-4) flg |= FLAG_LOWRHO;
                S Us[5];
                for (int i = 0; i < 5; i++) Us[i] = (S)U[i];
                const double r1 = Us[0], E1 = Us[4];
                const double w0 = q0.u0, w1 = q0.u1, w2 = q0.u2;
                const double ekin = r1 * 0.5 * (w0 * w0 + w1 * w1 + w2 * w2);
                if (!(E1 - ekin > a.seps)) flg |= FLAG_FLOOR;
                elem<S>(a.pout, c, coff) = (S)(vmax(E1 - ekin, a.seps) * a.gm1);
                for (int i = 0; i < 5; i++) elem<S>(a.uout, (int64_t)i * n + c, coff) = Us[i];
            } else {
                // main.f90:374-515: state, velocity, source(1), state, velocity
                if (U[0] < 1e-4) flg |= FLAG_LOWRHO;
                const double v0 = U[1] / U[0], v1 = U[2] / U[0], v2 = U[3] / U[0];
                if (fabs(v0) > a.vlimit || fabs(v1) > a.vlimit || fabs(v2) > a.vlimit) flg |= FLAG_CLAMP;
                const double ek = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
                if (!(U[4] - ek > a.seps)) flg |= FLAG_FLOOR;
                U[4] = vmax(U[4] - ek, a.seps) + ek;
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
                    p = vmax(U[4] - ekin, a.seps) * a.gm1;
                } else {
                    rr = Us[0];
                    const double inv = 1.0 / rr;
                    w0 = (double)Us[1] * inv;
                    w1 = (double)Us[2] * inv;
                    w2 = (double)Us[3] * inv;
                    const double ekin = 0.5 * ((double)Us[1] * w0 + (double)Us[2] * w1 + (double)Us[3] * w2);
                    p = vmax((double)Us[4] - ekin, a.seps) * a.gm1;
                }
                const double snd1 = sqrt(a.gam * p / rr);
                const double rt =
                    (snd1 + fabs(w0)) / a.ds[0] + (snd1 + fabs(w1)) / a.ds[1] + (snd1 + fabs(w2)) / a.ds[2];
                rmax = vmax(rmax, rt);
            }
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
    wave_max_to(rt, rate);
}

// ghost fill of the 5 conserved planes and (optionally) p*, shallow tables only
template <typename S>
__global__ __launch_bounds__(256) void k_ghosts6(int64_t n, int64_t ng, const int32_t *__restrict__ dst,
                                                const int32_t *__restrict__ src,
                                                const int32_t *__restrict__ sign, S *__restrict__ planes,
                                                S *__restrict__ pstar)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ng) return;
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

// FP64 <-> FP32 copies of the 5 conserved planes (mixed form hand-over)
__global__ __launch_bounds__(256) void k_narrow(int64_t m, const double *__restrict__ a, float *__restrict__ b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) b[i] = (float)a[i];
}

__global__ __launch_bounds__(256) void k_widen(int64_t m, const float *__restrict__ a, double *__restrict__ b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) b[i] = (double)a[i];
}

// ----------------------------------------------------------------- host
static bool is_pow2(double d)
{
    int e;
    double m = std::frexp(d, &e);
    return m == 0.5;
}

StageArgs base_args(bx_ctx *c, double dt)
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

template <int STAGE, typename S>
static int launch_stage(bx_ctx *c, const StageArgs &a, hipEvent_t mid)
{
    const int nzi = a.zhi - a.zlo;
    const int ly = c->fused_ly;
    {
        const unsigned gx = (unsigned)((c->nx - 4 + 63) / 64);
        const unsigned gy = (unsigned)((nzi + 3) / 4);
        const unsigned gz = (unsigned)((c->ny - 4 + ly - 1) / ly);
        hipLaunchKernelGGL((k_ydir<STAGE, S>), dim3(gx, gy, gz), dim3(256), 0, c->stream, a, (S *)c->eflux, ly);
        BX_HIP(hipGetLastError());
    }
    if (mid) BX_HIP(hipEventRecord(mid, c->stream));
    {
        const int64_t nwaves = ((int64_t)(c->ny - 4) * c->nx + kSegOut - 1) / kSegOut;
        const unsigned gx = (unsigned)((nwaves + 3) / 4);
        const unsigned gy = 1;
        const unsigned gz = (unsigned)((nzi + a.lz - 1) / a.lz);
        const S *ey = (const S *)c->eflux;
        if (c->fused_occ == 4)
            hipLaunchKernelGGL((k_xz<STAGE, S, 4, false>), dim3(gx, gy, gz), dim3(256), 0, c->stream, a, ey);
        else if (c->fused_occ == 1)
            hipLaunchKernelGGL((k_xz<STAGE, S, 1, true>), dim3(gx, gy, gz), dim3(256), 0, c->stream, a, ey);
        else
            hipLaunchKernelGGL((k_xz<STAGE, S, 3, false>), dim3(gx, gy, gz), dim3(256), 0, c->stream, a, ey);
        BX_HIP(hipGetLastError());
    }
    return BX_OK;
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
    if (c->prm.fluxangmom) return set_error(BX_ERR_UNSUPPORTED, "fused step fluxes linear momentum (fluxangmom off)");
    if (c->grav_flags) return set_error(BX_ERR_UNSUPPORTED, "fused step assumes the zero-potential force");
    if (c->eos) return set_error(BX_ERR_UNSUPPORTED, "fused step assumes a fixed gamma (H2STAT=-1)");
    if (c->prm.precision == BX_PREC_MIXED) {
        const size_t pf = sizeof(float) * (size_t)c->n;
        if (!c->cons32) BX_HIP(hipMalloc((void **)&c->cons32, 5 * pf));
        if (!c->ustar32) BX_HIP(hipMalloc((void **)&c->ustar32, 5 * pf));
        if (!c->unext32) BX_HIP(hipMalloc((void **)&c->unext32, 5 * pf));
        if (!c->pstar32) BX_HIP(hipMalloc((void **)&c->pstar32, pf));
    }
    const size_t pb = sizeof(double) * (size_t)c->n;
    if (!c->eflux) BX_HIP(hipMalloc((void **)&c->eflux, 5 * pb));
    if (!c->ustar) BX_HIP(hipMalloc((void **)&c->ustar, 5 * pb));
    if (!c->unext) BX_HIP(hipMalloc((void **)&c->unext, 5 * pb));
    if (!c->pstar) BX_HIP(hipMalloc((void **)&c->pstar, pb));
    // defaults from the autotuner on the 256^3 box (FP64: 64 rows, 32 planes;
    // mixed: 64 rows, 16 planes, y-term loaded early)
    // default work split (autotuned on the 256^3 box: FP64 64 rows and 32 planes
    // per wave, mixed 16 planes with the y-term loaded early), shortened on
    // smaller grids until the y kernel has about 2 and the x/z kernel about 4
    // waves per SIMD to run
    const int64_t want = 2 * 1024;
    const int nzi = c->flatz ? c->nz : c->nz - 4;
    if (c->fused_ly == 0) {
        int ly = 64;
        while (ly > 8 && (int64_t)((c->nx - 4 + 63) / 64) * nzi * ((c->ny - 4 + ly - 1) / ly) < want) ly /= 2;
        c->fused_ly = ly;
    }
    if (c->fused_lz == 0) {
        int lz = c->prm.precision == BX_PREC_MIXED ? 16 : 32;
        const int64_t rowwaves = ((int64_t)(c->ny - 4) * c->nx + kSegOut - 1) / kSegOut;
        while (lz > 4 && rowwaves * ((nzi + lz - 1) / lz) < 2 * want) lz /= 2;
        c->fused_lz = lz;
        c->fused_occ = c->prm.precision == BX_PREC_MIXED ? 1 : 0;
    }
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

template <typename S>
static int ghosts6(bx_ctx *c, S *planes, S *pstar)
{
    if (c->nghost == 0) return BX_OK;
    hipLaunchKernelGGL(k_ghosts6<S>, dim3(blocks_for(c->nghost, 256)), dim3(256), 0, c->stream, c->n, c->nghost,
                       c->ghost_dst, c->ghost_src, c->ghost_sign, planes, pstar);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

int narrow_cons(bx_ctx *c)
{
    hipLaunchKernelGGL(k_narrow, dim3(blocks_for(5 * c->n, 256)), dim3(256), 0, c->stream, 5 * c->n, c->cons,
                       c->cons32);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

int widen_cons(bx_ctx *c)
{
    hipLaunchKernelGGL(k_widen, dim3(blocks_for(5 * c->n, 256)), dim3(256), 0, c->stream, 5 * c->n, c->cons32,
                       c->cons);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

// One fused step from U^n (c->cons, or c->cons32 in the mixed form; regular)
// with full dt; result in c->unext / c->unext32.  The Courant rate of U^{n+1}
// accumulates into rate_bits[0] and irregular branches into flags (both
// zeroed here).
template <typename S>
static int fused_step_t(bx_ctx *c, double dt, S *u0, S *us, S *ps, S *un, bool reset_flags = true,
                        const double *dtp = nullptr)
{
    int rc;
    if (reset_flags) BX_HIP(hipMemsetAsync(&c->dsc->flags, 0, sizeof(unsigned int), c->stream));
    BX_HIP(hipMemsetAsync(&c->dsc->rate_bits[0], 0, sizeof(unsigned long long), c->stream));
    StageArgs a = base_args(c, dt);
    a.dtp = dtp;
    a.uin = u0;
    a.uout = us;
    a.pout = ps;
    if (c->prof_on) BX_HIP(hipEventRecord(c->ev[0], c->stream));
    if ((rc = launch_stage<1, S>(c, a, c->prof_on ? c->ev[1] : nullptr))) return rc;
    if (c->prof_on) BX_HIP(hipEventRecord(c->ev[2], c->stream));
    if ((rc = ghosts6<S>(c, us, ps))) return rc;
    StageArgs b = base_args(c, dt);
    b.dtp = dtp;
    b.uin = us;
    b.pin = ps;
    b.uold = u0;
    b.uout = un;
    b.rate = &c->dsc->rate_bits[0];
    if (c->prof_on) BX_HIP(hipEventRecord(c->ev[3], c->stream));
    if ((rc = launch_stage<2, S>(c, b, c->prof_on ? c->ev[4] : nullptr))) return rc;
    if (c->prof_on) BX_HIP(hipEventRecord(c->ev[5], c->stream));
    if ((rc = ghosts6<S>(c, un, (S *)nullptr))) return rc;
    return BX_OK;
}

// ------------------------------------------------------ graph replays
// One step's scalars on the device (graph replays and the adjoint's Courant
// forward), in the operations of the host driver (step_dt in bx_capi.hip):
// first the previous step of the run advances time (main.f90:544-545), then
// the Courant rate of the current state gives dt (main.f90:266-281), which is
// logged; the rate is cleared for the corrector kernel to accumulate the next
// one.  The host adds the last step's dt to time after the run.
__global__ void k_step_dt(DevScalars *s, double cfl, double *log, int cap)
{
    double time = s->time;
    const int k = s->logk;
    if (k > 0) {
        time = time + s->dt;
        s->time = time;
    }
    const double rate = __longlong_as_double((long long)s->rate_bits[0]);
    const int step = s->step + 1;
    double dt = cfl / rate;  // courant.f90:29
    dt = dt * 1.0;
    const double dtold = s->dtold;
    if (step > 1 && dtold > 0.0) {
        const double cap_dt = dtold * 1.1;
        dt = cap_dt < dt ? cap_dt : dt;  // std::min(dt, cap_dt)
    }
    s->dtold = dt;
    if (dt + time > s->endtime) dt = s->endtime - time;
    dt = dt * 0.5;
    dt = dt * 2.0;
    s->dt = dt;
    s->step = step;
    s->rate_bits[0] = 0ull;
    if (k < cap) log[k] = dt;
    s->logk = k + 1;
}

// Two fused steps, u0 -> un -> u0, with dt from the device scalars.
template <typename S>
static int enqueue_pair(bx_ctx *c, S *u0, S *us, S *ps, S *un)
{
    int rc;
    for (int h = 0; h < 2; h++) {
        S *src = h == 0 ? u0 : un, *dst = h == 0 ? un : u0;
        hipLaunchKernelGGL(k_step_dt, dim3(1), dim3(1), 0, c->stream, c->dsc, c->prm.cfl, c->chunk_dts,
                           kChunkSteps);
        BX_HIP(hipGetLastError());
        StageArgs a = base_args(c, 0.0);
        a.dtp = &c->dsc->dt;
        a.uin = src;
        a.uout = us;
        a.pout = ps;
        if ((rc = launch_stage<1, S>(c, a, nullptr))) return rc;
        if ((rc = ghosts6<S>(c, us, ps))) return rc;
        StageArgs b = base_args(c, 0.0);
        b.dtp = &c->dsc->dt;
        b.uin = us;
        b.pin = ps;
        b.uold = src;
        b.uout = dst;
        b.rate = &c->dsc->rate_bits[0];
        if ((rc = launch_stage<2, S>(c, b, nullptr))) return rc;
        if ((rc = ghosts6<S>(c, dst, (S *)nullptr))) return rc;
    }
    return BX_OK;
}

void drop_step_graphs(bx_ctx *c)
{
    for (int i = 0; i < 2; i++) {
        if (c->sexec[i]) (void)hipGraphExecDestroy(c->sexec[i]);
        if (c->sgraph[i]) (void)hipGraphDestroy(c->sgraph[i]);
        c->sexec[i] = nullptr;
        c->sgraph[i] = nullptr;
        c->skey[i] = nullptr;
    }
}

// The captured pair that starts from the current state buffer (captured on
// first use; one per buffer parity and precision, two kept).
static int pair_exec(bx_ctx *c, hipGraphExec_t *out)
{
    const bool f32 = c->state_in_32;
    const void *key = f32 ? (const void *)c->cons32 : (const void *)c->cons;
    for (int i = 0; i < 2; i++)
        if (c->sexec[i] && c->skey[i] == key) {
            *out = c->sexec[i];
            return BX_OK;
        }
    const int slot = c->sexec[0] ? (c->sexec[1] ? 0 : 1) : 0;
    if (c->sexec[slot]) (void)hipGraphExecDestroy(c->sexec[slot]);
    if (c->sgraph[slot]) (void)hipGraphDestroy(c->sgraph[slot]);
    c->sexec[slot] = nullptr;
    c->sgraph[slot] = nullptr;
    c->skey[slot] = nullptr;
    if (!c->chunk_dts) BX_HIP(hipMalloc((void **)&c->chunk_dts, sizeof(double) * kChunkSteps));
    BX_HIP(hipStreamSynchronize(c->stream));
    BX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = f32 ? enqueue_pair<float>(c, c->cons32, c->ustar32, c->pstar32, c->unext32)
                       : enqueue_pair<double>(c, c->cons, c->ustar, c->pstar, c->unext);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    BX_HIP(e);
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_check(ei, "hipGraphInstantiate");
    }
    c->sgraph[slot] = g;
    c->sexec[slot] = x;
    c->skey[slot] = key;
    *out = x;
    return BX_OK;
}

int replay_step_pairs(bx_ctx *c, int pairs)
{
    hipGraphExec_t x = nullptr;
    if (int rc = pair_exec(c, &x)) return rc;
    for (int i = 0; i < pairs; i++) BX_HIP(hipGraphLaunch(x, c->stream));
    return BX_OK;
}

int fused_step_into(bx_ctx *c, double dt, double *u0, double *us, double *ps, double *un, bool reset_flags)
{
    return fused_step_t<double>(c, dt, u0, us, ps, un, reset_flags);
}

// A step whose dt is formed on the device from the Courant rate of its state
// (main.f90:266-281, endtime from the device scalars), logged into log[cap].
int fused_step_device_dt(bx_ctx *c, double *u0, double *us, double *ps, double *un, double *log, int cap)
{
    hipLaunchKernelGGL(k_step_dt, dim3(1), dim3(1), 0, c->stream, c->dsc, c->prm.cfl, log, cap);
    BX_HIP(hipGetLastError());
    return fused_step_t<double>(c, 0.0, u0, us, ps, un, false, &c->dsc->dt);
}

int fused_step(bx_ctx *c, double dt)
{
    if (c->prm.precision == BX_PREC_MIXED)
        return fused_step_t<float>(c, dt, c->cons32, c->ustar32, c->pstar32, c->unext32);
    return fused_step_t<double>(c, dt, c->cons, c->ustar, c->pstar, c->unext);
}

}  // namespace bx
