// bx_stage.h -- stencil building blocks shared by the fused step kernels and
// the batched tube kernel: launch arguments of one RK stage, the primitive
// variables of a cell, directional MUSCL interface states (flux.f90:253-351,
// 391-479) and wave-shuffle helpers.  Arithmetic follows the exact-path
// kernels operation for operation.
#pragma once

#include "bx_internal.h"
#include "bx_math.h"

namespace bx {

constexpr int kSegOut = 60;  // output cells per wave row

// Element of an array addressed as a wave-uniform element offset `u` from the
// base plus a per-lane 32-bit byte offset: the compiler then forms one uniform
// 64-bit base per plane (scalar registers) and keeps a single offset register
// per lane for every plane, instead of a 64-bit vector address per plane.
template <typename S>
__device__ __forceinline__ S &elem(const void *base, int64_t u, unsigned boff)
{
    return *reinterpret_cast<S *>(reinterpret_cast<char *>(const_cast<void *>(base)) + u * (int64_t)sizeof(S) + boff);
}

// the wave index of a thread, marked wave-uniform for the compiler
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }


struct StageArgs {
    int64_t n;
    int nx, ny, nz;
    int flatz;
    int zlo, zhi, lz;     // interior z range [zlo,zhi) and planes per block
    const void *uin;      // 5 planes: stage 1 U^n, stage 2 U*
    const void *uold;     // stage 2: U^n
    const void *pin;      // stage 2: p*
    void *uout;           // 5 planes
    void *pout;           // stage 1: p*
    double dt;
    double ds[3], inv_ds[3];
    int ds_pow2[3];
    double ayz, axz, axy, vol, inv_vol;
    int vol_pow2;
    double gam, gm1, seps, srho, vlimit;
    unsigned int *flags;
    unsigned long long *rate;  // stage 2: next step's Courant rate (max)
    const double *dtp;         // when set, the step's dt is read from here (graph replays), else dt
    // the donors of the two-level corner ghosts, whose corrector result
    // before source(1) is parked in d2ubar (6 per entry: that, then the
    // ghost's pressure)
    int nd2;
    const int32_t *d2cell;
    double *d2ubar;
    // gravity (the kernels' G variant): 3 planes of the force that source()
    // applies to each cell (source.f90:148-182), every ghost holding its
    // donor's force with the ghost's momentum sign flips, so that a ghost's
    // source result is its donor's up to those flips; null without gravity
    const double *gf;
};

// The arguments of a launch with dt resolved: a kernel reads it once at entry.
__device__ __forceinline__ StageArgs with_dt(const StageArgs &a0)
{
    StageArgs a = a0;
    if (a0.dtp) a.dt = *a0.dtp;
    return a;
}

// Stage arguments of a context for a step of full dt (bx_fused.hip)
StageArgs base_args(bx_ctx *c, double dt);

template <typename T>
struct W6 {
    T d, u0, u1, u2, p, e;
};

template <typename T>
__device__ __forceinline__ T div_by(T x, double d, double inv, int pow2)
{
    // x/d; for d a power of two x*(1/d) is the same correctly rounded value
    return pow2 ? x * T(inv) : Arith<T>::div(x, T(d));
}

// 32-bit whole-wave shift by one lane (DPP wave_shr:1 / wave_shl:1): the lane
// with no source lane (0 / 63) keeps `edge`, the others take their lower /
// higher neighbour's `v`.  No LDS round trip, unlike a lane permute.
template <int CTRL>
__device__ __forceinline__ int wave_shift(int v, int edge)
{
    return __builtin_amdgcn_update_dpp(edge, v, CTRL, 0xf, 0xf, false);
}

template <int CTRL, typename T>
__device__ __forceinline__ T wave_shift_t(T v, T edge)
{
    if constexpr (sizeof(T) == 8) {
        const long long vb = __double_as_longlong(v), eb = __double_as_longlong(edge);
        const int lo = wave_shift<CTRL>((int)vb, (int)eb);
        const int hi = wave_shift<CTRL>((int)(vb >> 32), (int)(eb >> 32));
        return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
    } else {
        return __int_as_float(wave_shift<CTRL>(__float_as_int(v), __float_as_int(edge)));
    }
}

// the value of the next lower / higher lane (a lane with no such lane keeps
// its own)
template <typename T>
__device__ __forceinline__ T lane_below(T v) { return wave_shift_t<0x138>(v, v); }
template <typename T>
__device__ __forceinline__ T lane_above(T v) { return wave_shift_t<0x130>(v, v); }

template <typename T>
__device__ __forceinline__ T shfl_v(T v, int src)
{
    return __shfl(v, src, 64);
}

template <typename T>
__device__ __forceinline__ W6<T> shfl_w(const W6<T> &w, int src)
{
    W6<T> o;
    o.d = shfl_v(w.d, src);
    o.u0 = shfl_v(w.u0, src);
    o.u1 = shfl_v(w.u1, src);
    o.u2 = shfl_v(w.u2, src);
    o.p = shfl_v(w.p, src);
    o.e = shfl_v(w.e, src);
    return o;
}

template <int CTRL, typename T>
__device__ __forceinline__ W6<T> wave_shift_w(const W6<T> &w)
{
    W6<T> o;
    o.d = wave_shift_t<CTRL>(w.d, w.d);
    o.u0 = wave_shift_t<CTRL>(w.u0, w.u0);
    o.u1 = wave_shift_t<CTRL>(w.u1, w.u1);
    o.u2 = wave_shift_t<CTRL>(w.u2, w.u2);
    o.p = wave_shift_t<CTRL>(w.p, w.p);
    o.e = wave_shift_t<CTRL>(w.e, w.e);
    return o;
}
template <typename T>
__device__ __forceinline__ W6<T> lane_below_w(const W6<T> &w) { return wave_shift_w<0x138>(w); }
template <typename T>
__device__ __forceinline__ W6<T> lane_above_w(const W6<T> &w) { return wave_shift_w<0x130>(w); }

template <typename T>
__device__ __forceinline__ T comp(const W6<T> &w, int j)
{
    return j == 0 ? w.u0 : (j == 1 ? w.u1 : w.u2);
}

// interface states of one cell in direction d (flux.f90:253-351, 391-479):
// slopes from (wm, wc, wp) unless the cell is a ghost (zero slopes).
// The reference forms each state as (q + c*s) + 0 (the transverse terms are
// exact zeros); that equals (q + 0) + c*s bit for bit, so the centre values
// are made canonical once (no -0) and a zero slope of either sign gives the
// same state (slope_fz).  For the same reason a ghost cell zeroes the two
// coefficients c instead of its six slopes (c*s is then a zero for the
// finite slopes of finite data).  lp and lm enter only as 1-lp and -1-lm,
// where a zero of either sign gives the same value, so the hardware maximum
// and minimum stand for vmax/vmin against 0.
template <int d, typename T>
__device__ __forceinline__ void dir_states(const W6<T> &wm, const W6<T> &wc, const W6<T> &wp, T sound, bool ghost,
                                           const StageArgs &a, Side<T> &mi, Side<T> &pl)
{
    const T su0 = slope_fz(wm.u0, wc.u0, wp.u0);
    const T su1 = slope_fz(wm.u1, wc.u1, wp.u1);
    const T su2 = slope_fz(wm.u2, wc.u2, wp.u2);
    const T sp = slope_fz(wm.p, wc.p, wp.p);
    const T sd = slope_fz(wm.d, wc.d, wp.d);
    const T se = slope_fz(wm.e, wc.e, wp.e);
    const T ud = comp(wc, d);
    const T dt = T(a.dt);
    const T lp = div_by(Arith<T>::fmax(ud + sound, T(0)) * dt, a.ds[d], a.inv_ds[d], a.ds_pow2[d]);
    const T lm = div_by(Arith<T>::fmin(ud - sound, T(0)) * dt, a.ds[d], a.inv_ds[d], a.ds_pow2[d]);
    const T cm = ghost ? T(0) : T(0.5) * (T(-1) - lm), cp = ghost ? T(0) : T(0.5) * (T(1) - lp);
    const T z = T(0);  // transverse terms: 0*cc with cc=dt*half/ds > 0
    const T d0 = wc.d + z, v0 = wc.u0 + z, v1 = wc.u1 + z, v2 = wc.u2 + z, p0 = wc.p + z, e0 = wc.e + z;
    mi.d = d0 + cm * sd;
    pl.d = d0 + cp * sd;
    mi.v[0] = v0 + cm * su0;
    pl.v[0] = v0 + cp * su0;
    mi.v[1] = v1 + cm * su1;
    pl.v[1] = v1 + cp * su1;
    mi.v[2] = v2 + cm * su2;
    pl.v[2] = v2 + cp * su2;
    mi.p = p0 + cm * sp;
    pl.p = p0 + cp * sp;
    mi.e = e0 + cm * se;
    pl.e = e0 + cp * se;
    mi.g = T(a.gam);  // slope(gamma,gamma,gamma)=0: gamma + (+-0) + 0 = gamma
    pl.g = T(a.gam);
}

template <typename T>
__device__ __forceinline__ T sound_of(const W6<T> &w, double gam)
{
    if constexpr (sizeof(T) == 8) {
        const double x = Arith<T>::div(w.p * T(gam), w.d);
        double r;
        sqrt_n<1>(&x, &r);
        return r;
    } else {
        return Arith<T>::root(Arith<T>::div(w.p * T(gam), w.d));
    }
}


// Raw values of one cell as stored (conserved variables, and p* in stage 2).
template <typename S>
struct Raw {
    S r, m0, m1, m2, E, p;
    S g0, g1, g2;  // stage 1 with gravity: the cell's force
};

// cell = uniform element offset u + per-lane byte offset boff
template <int STAGE, typename S, bool G = false>
__device__ __forceinline__ Raw<S> load_raw(const StageArgs &a, int64_t u, unsigned boff)
{
    const int64_t n = a.n;
    Raw<S> v;
    v.r = elem<S>(a.uin, u, boff);
    v.m0 = elem<S>(a.uin, n + u, boff);
    v.m1 = elem<S>(a.uin, 2 * n + u, boff);
    v.m2 = elem<S>(a.uin, 3 * n + u, boff);
    v.E = elem<S>(a.uin, 4 * n + u, boff);
    v.p = STAGE == 2 ? elem<S>(a.pin, u, boff) : S(0);
    if constexpr (G && STAGE == 1) {
        v.g0 = elem<S>(a.gf, u, boff);
        v.g1 = elem<S>(a.gf, n + u, boff);
        v.g2 = elem<S>(a.gf, 2 * n + u, boff);
    }
    return v;
}

// source.f90:148-182 at a cell (rho r, momenta m, energy E) with force g over
// the half step dth: the momenta m + (rho g) dth, then the energy re-round
// with the kinetic energy ek of the new momenta and the velocity u of the last
// velocity call (`vmax(E - ekin + dt*rho*(u.g), eps) + ekin`).  The caller
// forms ek (a quotient by rho it may share a reciprocal with).
__device__ __forceinline__ void grav_momenta(double r, const double m[3], const double g[3], double dth, double mk[3])
{
    mk[1] = m[1] + (r * g[1]) * dth;
    mk[0] = m[0] + (r * g[0]) * dth;
    mk[2] = m[2] + (g[2] * r) * dth;
}
__device__ __forceinline__ double grav_energy_arg(double r, double E, double ek, const double u[3], const double g[3],
                                                  double dth)
{
    return (E - ek) + (dth * r) * ((u[0] * g[0] + u[1] * g[1]) + u[2] * g[2]);
}

// primitive variables of a loaded cell for each stage.
// FP64 storage: the reference's own operations.  FP32 storage (mixed form):
// the kinetic energy, the energy/pressure subtraction and u are formed in FP64
// from the stored values, then rounded for the FP32 reconstruction.
template <int STAGE, typename S, bool G = false>
__device__ __forceinline__ W6<S> prims(const StageArgs &a, const Raw<S> &v, unsigned int &flg)
{
    W6<S> w;
    if constexpr (G && STAGE == 1) {
        static_assert(sizeof(S) == 8, "gravity is FP64 only");
        // source(0) (main.f90:289) with the cell's force: u_old = m/rho is the
        // velocity of the last velocity call (the previous step's), the state
        // at :314 takes it, the velocity at :318 the kicked momenta; seven
        // quotients by rho with one shared reciprocal
        const double r = v.r, E = v.E;
        const double m[3] = {v.m0, v.m1, v.m2}, g[3] = {v.g0, v.g1, v.g2};
        const double dth = a.dt * 0.5;
        double mk[3];
        grav_momenta(r, m, g, dth, mk);
        const double num[7] = {m[0], m[1], m[2], mk[0], mk[1], mk[2],
                               0.5 * (mk[0] * mk[0] + mk[1] * mk[1] + mk[2] * mk[2])};
        double quo[7];
        div_shared<7>(num, r, quo);
        const double uo[3] = {quo[0], quo[1], quo[2]};
        const double ek = quo[6];
        const double E1 = floor_at(grav_energy_arg(r, E, ek, uo, g, dth), a.seps) + ek;
        w.d = r;
        w.u0 = quo[3];
        w.u1 = quo[4];
        w.u2 = quo[5];
        // state (:314): pressure with u_old; the flux's eps with u (flux.f90)
        w.p = floor_at(E1 - 0.5 * r * (uo[0] * uo[0] + uo[1] * uo[1] + uo[2] * uo[2]), a.seps) * a.gm1;
        w.e = floor_at(E1 - 0.5 * r * (w.u0 * w.u0 + w.u1 * w.u1 + w.u2 * w.u2), a.seps);
        if (fabs(w.u0) > a.vlimit || fabs(w.u1) > a.vlimit || fabs(w.u2) > a.vlimit) flg |= FLAG_CLAMP;
    } else if constexpr (sizeof(S) == 8) {
        const double r = v.r, m0 = v.m0, m1 = v.m1, m2 = v.m2, E = v.E;
        w.d = r;
        // u = m/rho (velocity.f90:21) and, in stage 1, the kinetic energy of
        // the source re-round: quotients by rho with one shared reciprocal
        double num[4] = {m0, m1, m2, 0.0}, quo[4];
        if (STAGE == 1) {
            num[3] = 0.5 * (m0 * m0 + m1 * m1 + m2 * m2);
            div_shared<4>(num, r, quo);
        } else {
            div_shared<3>(num, r, quo);
        }
        w.u0 = quo[0];
        w.u1 = quo[1];
        w.u2 = quo[2];
        if (STAGE == 1) {
            // source(0) energy re-round with g = -0 (source.f90:176-182), then
            // state at main.f90:314 with u = m/rho, velocity at :318
            const double ek = quo[3];
            const double E1 = floor_at(E - ek, a.seps) + ek;
            w.e = floor_at(E1 - 0.5 * r * (w.u0 * w.u0 + w.u1 * w.u1 + w.u2 * w.u2), a.seps);
            w.p = w.e * a.gm1;
        } else {
            // velocity at main.f90:355; p* is the state of :353 (stale velocity)
            w.e = floor_at(E - 0.5 * r * (w.u0 * w.u0 + w.u1 * w.u1 + w.u2 * w.u2), a.seps);
            w.p = v.p;
        }
        if (fabs(w.u0) > a.vlimit || fabs(w.u1) > a.vlimit || fabs(w.u2) > a.vlimit) flg |= FLAG_CLAMP;
    } else {
        const double r = v.r, m0 = v.m0, m1 = v.m1, m2 = v.m2, E = v.E;
        const double inv = 1.0 / r;
        const double v0 = m0 * inv, v1 = m1 * inv, v2 = m2 * inv;
        const double ek = 0.5 * (m0 * v0 + m1 * v1 + m2 * v2);
        const double e = floor_at(E - ek, a.seps);
        w.d = (float)r;
        w.u0 = (float)v0;
        w.u1 = (float)v1;
        w.u2 = (float)v2;
        w.e = (float)e;
        w.p = STAGE == 1 ? (float)(e * a.gm1) : v.p;
        if (fabs(v0) > a.vlimit || fabs(v1) > a.vlimit || fabs(v2) > a.vlimit) flg |= FLAG_CLAMP;
    }
    return w;
}

template <int STAGE, typename S, bool G = false>
__device__ __forceinline__ W6<S> load_prims(const StageArgs &a, int64_t u, unsigned boff, unsigned int &flg)
{
    return prims<STAGE, S, G>(a, load_raw<STAGE, S, G>(a, u, boff), flg);
}


}  // namespace bx
