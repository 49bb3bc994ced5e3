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
};

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

template <typename T>
__device__ __forceinline__ T comp(const W6<T> &w, int j)
{
    return j == 0 ? w.u0 : (j == 1 ? w.u1 : w.u2);
}

// interface states of one cell in direction d (flux.f90:253-351, 391-479):
// slopes from (wm, wc, wp) unless the cell is a ghost (zero slopes).
template <int d, typename T>
__device__ __forceinline__ void dir_states(const W6<T> &wm, const W6<T> &wc, const W6<T> &wp, T sound, bool ghost,
                                           const StageArgs &a, Side<T> &mi, Side<T> &pl)
{
    T su0 = T(0), su1 = T(0), su2 = T(0), sp = T(0), sd = T(0), se = T(0);
    if (!ghost) {
        su0 = slope_finite(wm.u0, wc.u0, wp.u0);
        su1 = slope_finite(wm.u1, wc.u1, wp.u1);
        su2 = slope_finite(wm.u2, wc.u2, wp.u2);
        sp = slope_finite(wm.p, wc.p, wp.p);
        sd = slope_finite(wm.d, wc.d, wp.d);
        se = slope_finite(wm.e, wc.e, wp.e);
    }
    const T ud = comp(wc, d);
    const T dt = T(a.dt);
    const T lp = div_by(vmax(ud + sound, T(0)) * dt, a.ds[d], a.inv_ds[d], a.ds_pow2[d]);
    const T lm = div_by(vmin(ud - sound, T(0)) * dt, a.ds[d], a.inv_ds[d], a.ds_pow2[d]);
    const T cm = T(0.5) * (T(-1) - lm), cp = T(0.5) * (T(1) - lp);
    const T z = T(0);  // transverse terms: 0*cc with cc=dt*half/ds > 0
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
    mi.g = T(a.gam);  // slope(gamma,gamma,gamma)=0: gamma + (+-0) + 0 = gamma
    pl.g = T(a.gam);
}

template <typename T>
__device__ __forceinline__ T sound_of(const W6<T> &w, double gam)
{
    return Arith<T>::root(Arith<T>::div(w.p * T(gam), w.d));
}


}  // namespace bx
