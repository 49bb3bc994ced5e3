// bx_math.h -- device arithmetic of the hydro step, written to reproduce the
// reference's FP64 evaluation order exactly.  Build with -ffp-contract=off and
// without fast-math: every +, *, / and sqrt below is one IEEE-754 operation.
#pragma once

#include <hip/hip_runtime.h>

#ifndef BX_THETA
#define BX_THETA 1.25  // generalized-minmod theta: Makefile:20 -DSLOPE_THETA=1.25
#endif

namespace bx {

// Division and square root per arithmetic type: IEEE correctly rounded for
// double (the reference's arithmetic); for float (the mixed-precision
// reconstruction only) the hardware reciprocal and square root, ~1-2 ulp.
template <typename T>
struct Arith;
template <>
struct Arith<double> {
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double root(double x) { return sqrt(x); }
    static __device__ __forceinline__ double fmin(double a, double b) { return __builtin_fmin(a, b); }
    static __device__ __forceinline__ double fmax(double a, double b) { return __builtin_fmax(a, b); }
};
template <>
struct Arith<float> {
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float root(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float fmin(float a, float b) { return __builtin_fminf(a, b); }
    static __device__ __forceinline__ float fmax(float a, float b) { return __builtin_fmaxf(a, b); }
};

template <typename T>
__device__ __forceinline__ T vmax(T a, T b) { return a > b ? a : b; }
template <typename T>
__device__ __forceinline__ T vmin(T a, T b) { return a < b ? a : b; }

// vmax(x, c) for a floor c >= +0: x > c ? x : c and the hardware maximum
// give the same value for every x (a NaN x gives c either way; a zero x of
// either sign is not above c, and for c = +0 both give +0), so the single
// instruction is used.
template <typename T>
__device__ __forceinline__ T floor_at(T x, T c) { return Arith<T>::fmax(x, c); }

// FP64 square roots as the compiler lowers sqrt() (correctly rounded): for
// x below 2^-767 it scales x by 2^256 and the root by 2^-128, and it answers
// x itself for +-0 and +inf; otherwise the lowering is
//     y = rsq(x); s0 = x*y; h0 = 0.5*y; r = fma(-h0, s0, 0.5)
//     h1 = fma(h0, r, h0); s1 = fma(s0, r, s0)
//     s2 = fma(fma(-s1, s1, x), h1, s1); root = fma(fma(-s2, s2, x), h1, s2)
// which is taken here for x in [2^-767, inf); a lane with another x (zero,
// negative, infinite, NaN or tiny) takes sqrt() for all N.
template <int N>
__device__ __forceinline__ void sqrt_n(const double *x, double *r)
{
    bool ok = true;
    for (int i = 0; i < N; i++) {
        const double v = x[i];
        const double y = __builtin_amdgcn_rsq(v);
        const double s0 = v * y, h0 = y * 0.5;
        const double rr = __builtin_fma(-h0, s0, 0.5);
        const double h1 = __builtin_fma(h0, rr, h0), s1 = __builtin_fma(s0, rr, s0);
        const double s2 = __builtin_fma(__builtin_fma(-s1, s1, v), h1, s1);
        r[i] = __builtin_fma(__builtin_fma(-s2, s2, v), h1, s2);
        ok = ok && v >= 0x1p-767 && v < __builtin_huge_val();
    }
    if (__builtin_expect(!ok, 0))
        for (int i = 0; i < N; i++) r[i] = __builtin_sqrt(x[i]);
}

// utils.f90:696-707
template <typename T>
__device__ __forceinline__ T genminmod(T x1, T x2, T x3)
{
    if ((x1 > T(0)) && ((x2 > T(0)) && (x3 > T(0)))) {
        T tmp = vmin(x2, x3);
        return vmin(x1, tmp);
    } else if ((x1 < T(0)) && ((x2 < T(0)) && (x3 < T(0)))) {
        T tmp = vmax(x2, x3);
        return vmax(x1, tmp);
    }
    return T(0);
}

// utils.f90:738-741: slope(c0,c1,c2) = genminmod(tt*(c1-c0), half*(c2-c0), tt*(c2-c1))
template <typename T>
__device__ __forceinline__ T slope(T c0, T c1, T c2)
{
    const T tt = T(BX_THETA);
    return genminmod(tt * (c1 - c0), T(0.5) * (c2 - c0), tt * (c2 - c1));
}

// genminmod for finite (non-NaN) inputs with three min/max instructions:
// "all three > 0" is exactly "their minimum > 0" and "all three < 0" is
// "their maximum < 0"; min/max of same-signed nonzero values is exact.  For
// finite data the result equals genminmod() above bit for bit.
template <typename T>
__device__ __forceinline__ T genminmod_finite(T x1, T x2, T x3)
{
    const T mn = Arith<T>::fmin(x1, Arith<T>::fmin(x2, x3));
    const T mx = Arith<T>::fmax(x1, Arith<T>::fmax(x2, x3));
    return mn > T(0) ? mn : (mx < T(0) ? mx : T(0));
}

template <typename T>
__device__ __forceinline__ T slope_finite(T c0, T c1, T c2)
{
    const T tt = T(BX_THETA);
    return genminmod_finite(tt * (c1 - c0), T(0.5) * (c2 - c0), tt * (c2 - c1));
}

// slope_finite up to the sign of a zero result: max(mn,0) + min(mx,0) is mn
// when all three arguments are positive, mx when all are negative, and a zero
// otherwise.  For use where the slope only enters q + c*s with a canonical
// (never -0) q, which gives the same value for a zero slope of either sign.
template <typename T>
__device__ __forceinline__ T slope_fz(T c0, T c1, T c2)
{
    const T tt = T(BX_THETA);
    const T x1 = tt * (c1 - c0), x2 = T(0.5) * (c2 - c0), x3 = tt * (c2 - c1);
    const T mn = Arith<T>::fmin(x1, Arith<T>::fmin(x2, x3));
    const T mx = Arith<T>::fmax(x1, Arith<T>::fmax(x2, x3));
    return Arith<T>::fmax(mn, T(0)) + Arith<T>::fmin(mx, T(0));
}

// Primitive variables of one cell as the flux routine sees them.
template <typename T>
struct Prim {
    T d, u[3], p, e;  // rho, velocity, pressure, internal energy (eps)
};

// One directional interface state (flux.f90:336-351): rho, velocity, p, eps, gamma.
template <typename T>
struct Side {
    T d, v[3], p, e, g;
};

// flux.f90:299-308 interface_muscl
template <typename T>
__device__ __forceinline__ void muscl(T q, T lm, T lp, T s, T &minus, T &plus)
{
    minus = q + T(0.5) * (T(-1) - lm) * s;
    plus = q + T(0.5) * (T(1) - lp) * s;
}

// Correctly rounded quotients a[i]/b (i < N) that share one reciprocal, for
// FP64.  The hardware's IEEE division is v_div_scale (numerator and
// denominator), v_rcp, two Newton steps on the reciprocal, the quotient and
// its residual, v_div_fmas and v_div_fixup.  When b lies in [2^-300, 2^300]
// and each a is zero or lies in [2^-299, 2^302] in magnitude, v_div_scale
// leaves both operands as they are (exponent difference below 768, no
// denormal reciprocal or quotient, numerator exponent far above 53), VCC is
// clear and v_div_fixup returns its input for a nonzero quotient, so that
// sequence is
//     y0 = rcp(b); y1 = fma(y0, fma(-b, y0, 1), y0); y = fma(y1, fma(-b, y1, 1), y1)
//     q0 = a*y;    r = fma(-b, q0, a);              q = fma(r, y, q0)
// and y does not depend on a.  Here the residual is formed negated,
// r' = fma(b, q0, -a), and q = fma(-r', y, q0): for a nonzero the same bits
// (round-to-nearest is symmetric), and for a zero the signed zero that
// v_div_fixup gives (q0 = +-0, r' = +0, q = -0*y + q0 = q0).  A lane with an
// operand outside the range takes the plain division for all N.
template <int N>
__device__ __forceinline__ void div_shared(const double *a, double b, double *q)
{
    const double y0 = __builtin_amdgcn_rcp(b);
    const double y1 = __builtin_fma(y0, __builtin_fma(-b, y0, 1.0), y0);
    const double y = __builtin_fma(y1, __builtin_fma(-b, y1, 1.0), y1);
    double mag = 0.0;
    int emin = 0;
    for (int i = 0; i < N; i++) {
        const double q0 = a[i] * y;
        q[i] = __builtin_fma(-__builtin_fma(b, q0, -a[i]), y, q0);
        mag += __builtin_fabs(a[i]);
        const int e = a[i] == 0.0 ? 0 : __builtin_amdgcn_frexp_exp(a[i]);
        emin = i == 0 ? e : min(emin, e);
    }
    // |a[i]| <= sum <= 2^302 (false for NaN or inf), exponent(a[i]) >= -298
    // for a[i] != 0 (frexp: a = m 2^e, 0.5 <= |m| < 1), b in [2^-300, 2^300]
    const bool ok = mag <= 0x1p302 && emin >= -298 && b >= 0x1p-300 && b <= 0x1p300;
    if (__builtin_expect(!ok, 0))
        for (int i = 0; i < N; i++) q[i] = a[i] / b;
}

// a/b for double, where a zero numerator is answered without the division:
// for finite nonzero b, IEEE gives the signed zero of sign(a) xor sign(b),
// which is exactly a*b.  Used where most numerators are exactly +-0 (the
// transverse components of a one-dimensional flow), so whole waves skip the
// correctly rounded division sequence.
__device__ __forceinline__ double qdiv(double a, double b)
{
    if (a == 0.0 && b != 0.0 && __builtin_isfinite(b)) return a * b;
    return a / b;
}

// flux.f90:512-590 for one face: L = plus state of the cell below the face,
// R = minus state of the cell above it, direction dir (0,1,2).  ZSKIP: the
// final quotients go through qdiv (same values).
template <typename T, int dir, bool ZSKIP = false>
__device__ __forceinline__ void face_flux(const Side<T> &L, const Side<T> &R, T F[5])
{
    const T h = T(0.5);
    T sl[5], sr[5], fl[5], fr[5];
    sl[0] = L.d;
    sl[1] = L.d * L.v[0];
    sl[2] = L.d * L.v[1];
    sl[3] = L.d * L.v[2];
    sl[4] = L.e + h * L.d * (L.v[0] * L.v[0] + L.v[1] * L.v[1] + L.v[2] * L.v[2]);
    sr[0] = R.d;
    sr[1] = R.d * R.v[0];
    sr[2] = R.d * R.v[1];
    sr[3] = R.d * R.v[2];
    sr[4] = R.e + h * R.d * (R.v[0] * R.v[0] + R.v[1] * R.v[1] + R.v[2] * R.v[2]);
    const T vl = L.v[dir], vr = R.v[dir];
    const T zero = T(0);
    // fluxes_l = 0; fluxes_l(2+dir) = pl; then fluxes_l(k) = fluxes_l(k) + states_l(k)*vl
    fl[0] = sl[0] * vl;
    fl[1] = (dir == 0 ? L.p : zero) + sl[1] * vl;
    fl[2] = (dir == 1 ? L.p : zero) + sl[2] * vl;
    fl[3] = (dir == 2 ? L.p : zero) + sl[3] * vl;
    fl[4] = (sl[4] + L.p) * vl;
    fr[0] = sr[0] * vr;
    fr[1] = (dir == 0 ? R.p : zero) + sr[1] * vr;
    fr[2] = (dir == 1 ? R.p : zero) + sr[2] * vr;
    fr[3] = (dir == 2 ? R.p : zero) + sr[3] * vr;
    fr[4] = (sr[4] + R.p) * vr;
    // flux_hll, flux.f90:583-589
    T cr, cl;
    if constexpr (sizeof(T) == 8) {
        const double sq[2] = {Arith<T>::div(R.g * R.p, R.d), Arith<T>::div(L.g * L.p, L.d)};
        double rt[2];
        sqrt_n<2>(sq, rt);
        cr = rt[0];
        cl = rt[1];
    } else {
        cr = Arith<T>::root(Arith<T>::div(R.g * R.p, R.d));
        cl = Arith<T>::root(Arith<T>::div(L.g * L.p, L.d));
    }
    const T ap = vmax(vmax(zero, vr + cr), vl + cl);
    const T am = vmax(vmax(zero, -(vr - cr)), -(vl - cl));
    const T apam = ap * am;
    const T den = ap + am;
    if constexpr (sizeof(T) == 8 && !ZSKIP) {
        double num[5];
#pragma unroll
        for (int i = 0; i < 5; i++) num[i] = ap * fl[i] + am * fr[i] - apam * (sr[i] - sl[i]);
        div_shared<5>(num, den, F);
    } else {
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const T num = ap * fl[i] + am * fr[i] - apam * (sr[i] - sl[i]);
            if constexpr (ZSKIP && sizeof(T) == 8)
                F[i] = qdiv(num, den);
            else
                F[i] = Arith<T>::div(num, den);
        }
    }
}

}  // namespace bx
