// bx_adjoint.hip -- hand-written reverse mode of the FP64 fused step with a
// frozen dt sequence, and the checkpointed forward/backward drivers.
//
// The differentiated map is one step of main.f90:211-555 on a regular state
// (every ghost a copy of its donor, u = m/rho), as the fused kernels compute
// it, with dt taken as given:
//   U' = source(0)(U^n)                 energy re-round (source.f90:176-182)
//   W  = prims(U')                      rho, u = m/rho, e, p = e (gamma-1)
//   U* = U' + ft(W)                     predictor (flux.f90:205)
//   p* = (E* - rho* |u|^2/2)(gamma-1)   with the stage-1 velocity (main.f90:353)
//   W* = prims(U*, p*)
//   U^{n+1} = source(1)((U' + U* + ft(W*))/2)     corrector (flux.f90:213)
// where ft(W) = -((areayz dFx + areaxz dFy) + areaxy dFz) dt/vol over the
// MUSCL states (flux.f90:253-482) and HLL faces (flux.f90:487-590).
//
// Branches are differentiated as taken: the genminmod limiter passes the
// adjoint to the argument it selected, max/min of the HLL speed bounds and of
// the MUSCL lambda to the operand that won, and the internal-energy floors of
// the primitives and of the re-round by whether they acted.  The floors of the
// stage update itself (flux.f90:225-228) and of p* are not differentiated: the
// forward reports them (FLAG_FLOOR) and the adjoint refuses such a run.
//
// Work decomposition (gather form, no atomics): per stage, one kernel per
// direction turns the adjoint of fluxtmp into adjoints of every cell's
// primitives W -- y and z march along the direction with a register queue
// (faces, then interface states, then limiter slopes), x works through
// whole-wave DPP shifts and, running last, maps the complete primitive
// adjoint of every interior cell that donates to no ghost to the conserved
// one.  Ghost cells collect adjoints like any cell; two pointwise kernels over
// the donors map theirs and add each ghost's reverse to its donor (momentum
// signs flipped at walls).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "bx_stage.h"

namespace bx {

// ------------------------------------------------------------ building blocks
// 1/x for adjoint arithmetic (not the forward's correctly rounded quotients):
// the hardware reciprocal refined by one Newton step
__device__ __forceinline__ double arcp(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    const double e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ void wzero(W6<double> &w) { w.d = w.u0 = w.u1 = w.u2 = w.p = w.e = 0.0; }

__device__ __forceinline__ double &wref(W6<double> &w, int j) { return j == 0 ? w.u0 : (j == 1 ? w.u1 : w.u2); }

// coefficients (k0, k1, k2) of d slope / d (c0, c1, c2) for the argument the
// limiter selected (slope_finite, bx_math.h); all zero when it returned 0
__device__ __forceinline__ double slope_coef(double c0, double c1, double c2, double &k0, double &k1, double &k2)
{
    const double tt = BX_THETA;
    const double a1 = tt * (c1 - c0), a2 = 0.5 * (c2 - c0), a3 = tt * (c2 - c1);
    const double mn = __builtin_fmin(a1, __builtin_fmin(a2, a3));
    const double mx = __builtin_fmax(a1, __builtin_fmax(a2, a3));
    const bool pos = mn > 0.0, neg = !pos && mx < 0.0;
    const double r = pos ? mn : (neg ? mx : 0.0);
    const bool s1 = (pos || neg) && a1 == r;
    const bool s2 = (pos || neg) && !s1 && a2 == r;
    const bool s3 = (pos || neg) && !s1 && !s2;
    k0 = s1 ? -tt : (s2 ? -0.5 : 0.0);
    k1 = s1 ? tt : (s3 ? -tt : 0.0);
    k2 = s2 ? 0.5 : (s3 ? tt : 0.0);
    return r;
}

// reverse of dir_states<d> (bx_stage.h): adjoints of the minus/plus states of
// cell c accumulate into its neighbours' (gm, gp) and its own (gc) primitives
// and into the adjoint of its sound speed
template <int d>
__device__ __forceinline__ void dir_states_vjp(const W6<double> &wm, const W6<double> &wc, const W6<double> &wp,
                                               double snd, bool ghost, const StageArgs &a, const Side<double> &lmi,
                                               const Side<double> &lpl, W6<double> &gm, W6<double> &gc,
                                               W6<double> &gp, double &lsnd)
{
    const double ud = d == 0 ? wc.u0 : (d == 1 ? wc.u1 : wc.u2);
    const double k = a.dt * a.inv_ds[d];
    const double lp = vmax(ud + snd, 0.0) * k, lm = vmin(ud - snd, 0.0) * k;
    const double cm = 0.5 * (-1.0 - lm), cp = 0.5 * (1.0 - lp);
    double lcm = 0.0, lcp = 0.0;
    auto var = [&](double qm, double qc, double qp, double a_mi, double a_pl, double &g0, double &g1, double &g2) {
        g1 += a_mi + a_pl;
        if (!ghost) {
            double k0, k1, k2;
            const double s = slope_coef(qm, qc, qp, k0, k1, k2);
            lcm += s * a_mi;
            lcp += s * a_pl;
            const double ls = cm * a_mi + cp * a_pl;
            g0 += k0 * ls;
            g1 += k1 * ls;
            g2 += k2 * ls;
        }
    };
    var(wm.u0, wc.u0, wp.u0, lmi.v[0], lpl.v[0], gm.u0, gc.u0, gp.u0);
    var(wm.u1, wc.u1, wp.u1, lmi.v[1], lpl.v[1], gm.u1, gc.u1, gp.u1);
    var(wm.u2, wc.u2, wp.u2, lmi.v[2], lpl.v[2], gm.u2, gc.u2, gp.u2);
    var(wm.p, wc.p, wp.p, lmi.p, lpl.p, gm.p, gc.p, gp.p);
    var(wm.d, wc.d, wp.d, lmi.d, lpl.d, gm.d, gc.d, gp.d);
    var(wm.e, wc.e, wp.e, lmi.e, lpl.e, gm.e, gc.e, gp.e);
    // cm = (-1 - lm)/2, cp = (1 - lp)/2, lp = max(ud+c,0) dt/ds, lm = min(ud-c,0) dt/ds
    const double llp = -0.5 * lcp * k, llm = -0.5 * lcm * k;
    double lud = 0.0;
    if (ud + snd > 0.0) {
        lud += llp;
        lsnd += llp;
    }
    if (ud - snd < 0.0) {
        lud += llm;
        lsnd -= llm;
    }
    if (d == 0) gc.u0 += lud;
    else if (d == 1) gc.u1 += lud;
    else gc.u2 += lud;
}

// sound speed c = sqrt(p gamma / rho) (flux.f90:327): adjoint into (p, rho)
__device__ __forceinline__ void sound_vjp(const W6<double> &w, double snd, double gam, double lsnd, W6<double> &g)
{
    // dc/dp = gamma/(2 rho c) = c/(2p), dc/drho = -c/(2 rho)
    const double h = lsnd * 0.5 * snd;
    g.p += h * arcp(w.p);
    g.d -= h * arcp(w.d);
}

// conserved state s_i and physical flux f_i (flux.f90:512-557) of one side S
// for component i
template <int dir>
__device__ __forceinline__ void side_sf(const Side<double> &S, int i, double &s_i, double &f_i)
{
    const double vn = S.v[dir];
    if (i == 0) {
        s_i = S.d;
        f_i = s_i * vn;
    } else if (i < 4) {
        s_i = S.d * S.v[i - 1];
        f_i = (dir == i - 1 ? S.p : 0.0) + s_i * vn;
    } else {
        s_i = S.e + 0.5 * S.d * (S.v[0] * S.v[0] + S.v[1] * S.v[1] + S.v[2] * S.v[2]);
        f_i = (s_i + S.p) * vn;
    }
}

// adjoint of side S's component i: lf = adjoint of f_i, ls = adjoint of s_i
// (before the flux's own s_i vn term); accumulates into lS, lvn (normal velocity)
template <int dir>
__device__ __forceinline__ void side_acc(const Side<double> &S, int i, double s_i, double lf, double ls, Side<double> &lS,
                                         double &lvn)
{
    const double vn = S.v[dir];
    ls += lf * vn;
    if (i == 0) {
        lvn += lf * s_i;
        lS.d += ls;
    } else if (i < 4) {
        const int j = i - 1;
        lvn += lf * s_i;
        if (j == dir) lS.p += lf;
        lS.d += ls * S.v[j];
        lS.v[j] += ls * S.d;
    } else {
        lvn += lf * (s_i + S.p);
        lS.p += lf * vn;
        lS.e += ls;
        lS.d += ls * 0.5 * (S.v[0] * S.v[0] + S.v[1] * S.v[1] + S.v[2] * S.v[2]);
        for (int k = 0; k < 3; k++) lS.v[k] += ls * S.d * S.v[k];
    }
}

// reverse of face_flux<double, dir> (bx_math.h): adjoint of the 5 fluxes to
// the adjoints of the left (plus) and right (minus) interface states; the
// forward quantities are recomputed component by component
template <int dir>
__device__ __forceinline__ void face_vjp(const Side<double> &L, const Side<double> &R, const double lF[5],
                                         Side<double> &lL, Side<double> &lR)
{
    const double vl = L.v[dir], vr = R.v[dir];
    const double sq[2] = {(R.g * R.p) / R.d, (L.g * L.p) / L.d};
    double rt[2];
    sqrt_n<2>(sq, rt);
    const double cr = rt[0], cl = rt[1];
    const double a1 = vr + cr, a2 = vl + cl;
    const double t1 = vmax(0.0, a1);
    const double ap = vmax(t1, a2);
    const double b1 = -(vr - cr), b2 = -(vl - cl);
    const double t2 = vmax(0.0, b1);
    const double am = vmax(t2, b2);
    const double apam = ap * am, iden = arcp(ap + am);
    // pass 1: adjoints of ap, am (through the numerators and the denominator)
    double lN[5], lap = 0.0, lam = 0.0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        double sl, fl, sr, fr;
        side_sf<dir>(L, i, sl, fl);
        side_sf<dir>(R, i, sr, fr);
        const double dd = sr - sl;
        const double F = (ap * fl + am * fr - apam * dd) * iden;
        lN[i] = lF[i] * iden;
        lap += lN[i] * (fl - am * dd - F);
        lam += lN[i] * (fr - ap * dd - F);
    }
    lL.d = lL.v[0] = lL.v[1] = lL.v[2] = lL.p = lL.e = lL.g = 0.0;
    lR = lL;
    double lvl = 0.0, lvr = 0.0, lcl = 0.0, lcr = 0.0;
    // ap = vmax(vmax(0, a1), a2), vmax(x, y) = x > y ? x : y
    if (!(t1 > a2)) {
        lvl += lap;
        lcl += lap;
    } else if (!(0.0 > a1)) {
        lvr += lap;
        lcr += lap;
    }
    // am = vmax(vmax(0, b1), b2), b = c - v
    if (!(t2 > b2)) {
        lvl -= lam;
        lcl += lam;
    } else if (!(0.0 > b1)) {
        lvr -= lam;
        lcr += lam;
    }
    // pass 2: the two sides' states and fluxes
#pragma unroll
    for (int i = 0; i < 5; i++) {
        double sl, fl, sr, fr;
        side_sf<dir>(L, i, sl, fl);
        side_sf<dir>(R, i, sr, fr);
        side_acc<dir>(L, i, sl, ap * lN[i], apam * lN[i], lL, lvl);
        side_acc<dir>(R, i, sr, am * lN[i], -apam * lN[i], lR, lvr);
    }
    lL.v[dir] += lvl;
    lR.v[dir] += lvr;
    // c = sqrt(g p / d): dc/dp = c/(2p), dc/dd = -c/(2d)
    const double hl = lcl * 0.5 * cl, hr = lcr * 0.5 * cr;
    lL.p += hl * arcp(L.p);
    lL.d -= hl * arcp(L.d);
    lR.p += hr * arcp(R.p);
    lR.d -= hr * arcp(R.d);
}

// primitive adjoints of a cell: the first direction that reaches a cell
// assigns them, later directions add (cells no direction reaches stay zero)
__device__ __forceinline__ void add_to(double *__restrict__ LW, int64_t n, int64_t u, unsigned boff,
                                       const W6<double> &g, bool assign)
{
    const double v[6] = {g.d, g.u0, g.u1, g.u2, g.p, g.e};
    for (int k = 0; k < 6; k++) {
        double &t = elem<double>(LW, (int64_t)k * n + u, boff);
        t = assign ? v[k] : t + v[k];
    }
}

// ------------------------------------------------- direction kernels (y, z)
// A wave holds 64 interior x-cells of one interior line along direction d
// (d=1: y lines of a z-plane, d=2: z lines of a y-row) and marches along it
// over the cells of chunk [jb, je), emitting the primitive adjoints of those
// cells: faces j+1/2 for j in [jb-2, je], state reverses for [jb-1, je].
// g = gscale * lam is the adjoint of fluxtmp (zero on ghost cells).  The
// y kernel runs first and assigns its cells' primitive adjoints; the z kernel
// assigns on z-ghost cells (no y line passes through them) and adds elsewhere.
template <int STAGE, int d>
__global__ __launch_bounds__(256) void k_adj_march(StageArgs a, const double *__restrict__ lam, double gscale,
                                                   double kappa, double *__restrict__ LW, int chunk)
{
    const int lane = threadIdx.x & 63;
    const int w = wave_id();
    const int nx = a.nx, ny = a.ny;
    const int64_t n = a.n, plane = (int64_t)nx * ny;
    const int x = 2 + blockIdx.x * 64 + lane;
    const bool xin = x <= nx - 3;
    const int xc = xin ? x : nx - 3;
    const unsigned xoff = (unsigned)xc * 8u;
    int64_t base, stride;
    int nd;
    if (d == 1) {
        const int z = a.zlo + blockIdx.y * 4 + w;
        if (z >= a.zhi) return;
        base = (int64_t)z * plane;
        stride = nx;
        nd = ny;
    } else {
        const int y = 2 + blockIdx.y * 4 + w;
        if (y > ny - 3) return;
        base = (int64_t)y * nx;
        stride = plane;
        nd = a.nz;
    }
    const int jb = max(1, (int)blockIdx.z * chunk), je = min(nd - 1, (int)blockIdx.z * chunk + chunk);
    if (jb >= je) return;
    auto U = [&](int j) -> int64_t { return base + (int64_t)j * stride; };
    auto ghost = [&](int j) -> bool { return j < 2 || j > nd - 3; };
    unsigned int flg = 0;
    auto W = [&](int j) -> W6<double> { return load_prims<STAGE, double>(a, U(min(max(j, 0), nd - 1)), xoff, flg); };
    auto G = [&](int j, double g[5]) {
        for (int k = 0; k < 5; k++) g[k] = gscale * elem<double>(lam, (int64_t)k * n + U(j), xoff);
    };
    const int s0 = max(1, jb - 1), s1 = min(nd - 2, je);
    W6<double> w0 = W(s0 - 1), w1 = W(s0), w2 = W(s0 + 1);
    double c1 = sound_of(w1, a.gam);
    Side<double> mi_s, pl_s, lmi_s;
    dir_states<d>(w0, w1, w2, c1, ghost(s0), a, mi_s, pl_s);
    lmi_s.d = lmi_s.v[0] = lmi_s.v[1] = lmi_s.v[2] = lmi_s.p = lmi_s.e = lmi_s.g = 0.0;
    double gs[5], gn[5];
    G(s0, gs);
    if (s0 - 1 >= 1) {
        // face (s0-1)+1/2 gives the adjoint of cell s0's minus state
        const W6<double> wmm = W(s0 - 2);
        Side<double> mi0, pl0, dummy;
        dir_states<d>(wmm, w0, w1, sound_of(w0, a.gam), ghost(s0 - 1), a, mi0, pl0);
        double gp[5], lF[5];
        G(s0 - 1, gp);
        for (int k = 0; k < 5; k++) lF[k] = kappa * (gs[k] - gp[k]);
        face_vjp<d>(pl0, mi_s, lF, dummy, lmi_s);
    }
    W6<double> acc0, acc1, acc2;
    wzero(acc0);
    wzero(acc1);
    wzero(acc2);
    // the raw values of cell s+2 are loaded one step ahead
    auto R = [&](int j) -> Raw<double> { return load_raw<STAGE, double>(a, U(min(max(j, 0), nd - 1)), xoff); };
    Raw<double> nraw = R(s0 + 2);
    for (int s = s0; s <= s1; s++) {
        // queue: w0..w2 = W(s-1..s+1), c1 = sound(s); states of s, adjoint of its minus state
        const W6<double> w3 = prims<STAGE, double>(a, nraw, flg);
        nraw = R(s + 3);
        const double c2 = sound_of(w2, a.gam);
        Side<double> lpl_s, lmi_n;
        lpl_s.d = lpl_s.v[0] = lpl_s.v[1] = lpl_s.v[2] = lpl_s.p = lpl_s.e = lpl_s.g = 0.0;
        lmi_n = lpl_s;
        Side<double> mi_n, pl_n;
        if (s <= nd - 3) {
            dir_states<d>(w1, w2, w3, c2, ghost(s + 1), a, mi_n, pl_n);
            G(s + 1, gn);
            double lF[5];
            for (int k = 0; k < 5; k++) lF[k] = kappa * (gn[k] - gs[k]);
            face_vjp<d>(pl_s, mi_n, lF, lpl_s, lmi_n);
        }
        // the z march adds cell s-1's adjoint to the y march's: its primitive
        // adjoints are loaded before the state reverse, so that the load runs
        // under it
        const bool emit = xin && s - 1 >= jb && s - 1 < je;
        const bool assign = d == 1 || ghost(s - 1);
        double prev[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (emit && !assign)
            for (int k = 0; k < 6; k++) prev[k] = elem<double>(LW, (int64_t)k * n + U(s - 1), xoff);
        double lsnd = 0.0;
        dir_states_vjp<d>(w0, w1, w2, c1, ghost(s), a, lmi_s, lpl_s, acc0, acc1, acc2, lsnd);
        sound_vjp(w1, c1, a.gam, lsnd, acc1);
        if (emit) {
            const double v[6] = {acc0.d, acc0.u0, acc0.u1, acc0.u2, acc0.p, acc0.e};
            for (int k = 0; k < 6; k++) elem<double>(LW, (int64_t)k * n + U(s - 1), xoff) = assign ? v[k] : prev[k] + v[k];
        }
        // shift the queue
        w0 = w1;
        w1 = w2;
        w2 = w3;
        c1 = c2;
        mi_s = mi_n;
        pl_s = pl_n;
        lmi_s = lmi_n;
        acc0 = acc1;
        acc1 = acc2;
        wzero(acc2);
        for (int k = 0; k < 5; k++) gs[k] = gn[k];
    }
    // cell s1 is complete when s1+1 took no state reverse (s1 = nd-2)
    if (xin && s1 >= jb && s1 < je && s1 == nd - 2) add_to(LW, n, U(s1), xoff, acc0, d == 1 || ghost(s1));
}

// ------------------------------------------------------- pointwise maps
// Ghost gather: the ghosts whose donor is cell c (inverse of the ghost table),
// slot[c] = -1 when none.  Each ghost's contribution is evaluated at the
// ghost's own values and added with its momentum signs.
struct GhostInv {
    const int32_t *slot, *start, *gidx, *gsign;
};

__device__ __forceinline__ void add_signed(double acc[5], const double v[5], int sg)
{
    acc[0] += v[0];
    for (int f = 1; f <= 3; f++) acc[f] += ((sg >> (f - 1)) & 1) ? -v[f] : v[f];
    acc[4] += v[4];
}

// stage 2 at one cell: primitive adjoints lw of W* = (rho*, m*/rho*, p*, e*)
// at U* = (r, m, E) -> adjoint of U* (out[0..4]) and of p* (out[5])
__device__ __forceinline__ void fin2_core(double r, const double m[3], double E, const double lw[6], double seps,
                                          double out[6])
{
    const double u[3] = {m[0] / r, m[1] / r, m[2] / r};
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const bool eact = E - 0.5 * r * uu > seps;
    const double le = lw[5];
    double lr = lw[0] + (eact ? -0.5 * uu * le : 0.0);
    const double ir = arcp(r);
    for (int j = 0; j < 3; j++) {
        const double lu = lw[1 + j] + (eact ? -r * u[j] * le : 0.0);
        out[1 + j] = lu * ir;
        lr -= lu * u[j] * ir;
    }
    out[0] = lr;
    out[4] = eact ? le : 0.0;
    out[5] = lw[4];
}

__device__ __forceinline__ void fin2_cell(int64_t n, int64_t c, double seps, const double *__restrict__ us,
                                          const double *__restrict__ LW, double out[6])
{
    const double m[3] = {us[n + c], us[2 * n + c], us[3 * n + c]};
    const double lw[6] = {LW[c], LW[n + c], LW[2 * n + c], LW[3 * n + c], LW[4 * n + c], LW[5 * n + c]};
    fin2_core(us[c], m, us[4 * n + c], lw, seps, out);
}

// the rest of stage 2's reverse at an interior cell, from a[6] (the W*
// reverse of the cell and of the ghosts it donates to) and hl = lam/2: the
// direct halves of the corrector, and the reverse of
// p* = max(E* - rho* |u'|^2/2, eps)(gamma-1) with the stage-1 velocity
// u' = m'/rho' (main.f90:353); U^n = (r0, m0) and U* = (r1, ., E1).  Gives the
// adjoint of U* (lus) and the sum of that and the adjoint of
// U' = source(0)(U^n) (lsum), the two terms stage 1's reverse adds.
__device__ __forceinline__ void mid2_tail(double a[6], const double hl[5], double r0, const double m0[3], double r1,
                                          double E1, double gm1, double seps, double lus[5], double lsum[5])
{
    for (int f = 0; f < 5; f++) a[f] += hl[f];
    const double w[3] = {m0[0] / r0, m0[1] / r0, m0[2] / r0};
    const double ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double lp = a[5];
    double lw[3] = {0.0, 0.0, 0.0};
    if (E1 - r1 * 0.5 * ww > seps) {
        a[4] += gm1 * lp;
        a[0] -= gm1 * lp * 0.5 * ww;
        for (int j = 0; j < 3; j++) lw[j] = -gm1 * lp * r1 * w[j];
    }
    double lup[5];
    const double ir0 = arcp(r0);
    double lr0 = 0.0;
    for (int j = 0; j < 3; j++) {
        lup[1 + j] = hl[1 + j] + lw[j] * ir0;
        lr0 -= lw[j] * w[j] * ir0;
    }
    lup[0] = hl[0] + lr0;
    lup[4] = hl[4];
    for (int f = 0; f < 5; f++) {
        lus[f] = a[f];
        lsum[f] = a[f] + lup[f];
    }
}

// stage 1 at one cell: primitive adjoints lw of W = prims(source(0)(U^n))
// at U^n = (r, m, E) plus the direct terms ls (U* = U' + fluxtmp; the
// corrector's U'; zero on ghosts), through the re-round to the adjoint of U^n
__device__ __forceinline__ void fin1_core(double r, const double m[3], double E, const double lw[6],
                                          const double ls[5], double gm1, double seps, double out[5])
{
    const double ek = 0.5 * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / r;
    const bool ract = E - ek > seps;
    const double E1 = vmax(E - ek, seps) + ek;
    const double u[3] = {m[0] / r, m[1] / r, m[2] / r};
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const bool eact = E1 - 0.5 * r * uu > seps;
    const double le = lw[5] + gm1 * lw[4];  // p = e (gamma-1)
    double lr = lw[0] + (eact ? -0.5 * uu * le : 0.0);
    double lE1 = eact ? le : 0.0;
    double lm[3];
    const double ir = arcp(r);
    for (int j = 0; j < 3; j++) {
        const double lu = lw[1 + j] + (eact ? -r * u[j] * le : 0.0);
        lm[j] = lu * ir;
        lr -= lu * u[j] * ir;
    }
    lr += ls[0];
    for (int j = 0; j < 3; j++) lm[j] += ls[1 + j];
    lE1 += ls[4];
    // E1 = max(E - ek, eps) + ek
    double lE = lE1;
    if (!ract) {
        lE = 0.0;
        for (int j = 0; j < 3; j++) lm[j] += lE1 * m[j] * ir;
        lr -= lE1 * ek * ir;
    }
    out[0] = lr;
    for (int j = 0; j < 3; j++) out[1 + j] = lm[j];
    out[4] = lE;
}

__device__ __forceinline__ void fin1_cell(int64_t n, int64_t c, double gm1, double seps, const double *__restrict__ u0,
                                          const double *__restrict__ lsum, const double *__restrict__ LW, double out[5])
{
    const double m[3] = {u0[n + c], u0[2 * n + c], u0[3 * n + c]};
    const double lw[6] = {LW[c], LW[n + c], LW[2 * n + c], LW[3 * n + c], LW[4 * n + c], LW[5 * n + c]};
    const double ls[5] = {lsum[c], lsum[n + c], lsum[2 * n + c], lsum[3 * n + c], lsum[4 * n + c]};
    fin1_core(u0[c], m, u0[4 * n + c], lw, ls, gm1, seps, out);
}

// Where the x kernel ends a stage's reverse for the interior cells that
// donate to no ghost (the donors gather their ghosts' adjoints in the
// pointwise kernels below, after every direction): stage 2 writes lus and
// lsum, stage 1 the adjoint of U^n.
struct AdjTail {
    const int32_t *slot;  // per cell: its donor entry, or -1
    double *lus, *lsum;   // stage 2: outputs; stage 1: lsum is an input
    double *lout;         // stage 1: output
};

// ----------------------------------------------------- direction kernel (x)
// A wave holds 64 consecutive cells of the row-major x-y plane of one z-plane
// (as the forward x/z kernel; ghost columns separate rows and carry zero
// slopes and zero fluxtmp adjoint); lanes 3..60 emit, assigning on x-ghost
// cells (no y or z line passes through them) and adding elsewhere.
constexpr int kAdjOut = 58;

template <int STAGE>
__global__ __launch_bounds__(256) void k_adj_x(StageArgs a, const double *__restrict__ lam, double gscale,
                                               double kappa, double *__restrict__ LW, AdjTail tail)
{
    const int lane = threadIdx.x & 63;
    const int w = wave_id();
    const int nx = a.nx;
    const int64_t n = a.n, plane = (int64_t)nx * a.ny;
    const int64_t F0 = 2 * (int64_t)nx - 3 + (int64_t)(blockIdx.x * 4 + w) * kAdjOut;
    if (F0 + 3 >= (int64_t)(a.ny - 2) * nx) return;
    int64_t F = F0 + lane;
    if (F >= plane) F = plane - 1;
    const int y = (int)(F / nx), x = (int)(F - (int64_t)y * nx);
    const int k = a.zlo + blockIdx.y;
    const bool zg = !a.flatz && (k < 2 || k > a.nz - 3);
    const bool ghost = zg || x < 2 || x > nx - 3 || y < 2 || y > a.ny - 3;
    const unsigned coff = (unsigned)F * 8u;
    const int64_t zb = (int64_t)k * plane;
    unsigned int flg = 0;
    const W6<double> q = load_prims<STAGE, double>(a, zb, coff, flg);
    const double snd = sound_of(q, a.gam);
    double g[5];
    for (int i = 0; i < 5; i++) g[i] = gscale * elem<double>(lam, (int64_t)i * n + zb, coff);
    const W6<double> wm = lane_below_w(q), wp = lane_above_w(q);
    Side<double> mi, pl;
    dir_states<0>(wm, q, wp, snd, ghost, a, mi, pl);
    // face x-1/2 of this lane: L = plus state of lane-1, R = own minus state
    Side<double> L;
    L.d = lane_below(pl.d);
    L.v[0] = lane_below(pl.v[0]);
    L.v[1] = lane_below(pl.v[1]);
    L.v[2] = lane_below(pl.v[2]);
    L.p = lane_below(pl.p);
    L.e = lane_below(pl.e);
    L.g = a.gam;
    double lF[5];
    for (int i = 0; i < 5; i++) lF[i] = kappa * (g[i] - lane_below(g[i]));
    Side<double> lL, lmi;
    face_vjp<0>(L, mi, lF, lL, lmi);
    Side<double> lpl;  // the face x+1/2 belongs to lane+1
    lpl.d = lane_above(lL.d);
    lpl.v[0] = lane_above(lL.v[0]);
    lpl.v[1] = lane_above(lL.v[1]);
    lpl.v[2] = lane_above(lL.v[2]);
    lpl.p = lane_above(lL.p);
    lpl.e = lane_above(lL.e);
    lpl.g = 0.0;
    // the y and z marches' primitive adjoints of this cell, loaded before the
    // state reverse so that the load runs under it (x-ghost cells assign)
    const bool emit = lane >= 3 && lane <= 60 && F0 + lane < (int64_t)(a.ny - 2) * nx;
    const bool assign = x < 2 || x > nx - 3;
    double prev[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (emit && !assign)
        for (int k = 0; k < 6; k++) prev[k] = elem<double>(LW, (int64_t)k * n + zb, coff);
    W6<double> gm, gc, gp;
    wzero(gm);
    wzero(gc);
    wzero(gp);
    double lsnd = 0.0;
    dir_states_vjp<0>(wm, q, wp, snd, ghost, a, lmi, lpl, gm, gc, gp, lsnd);
    sound_vjp(q, snd, a.gam, lsnd, gc);
    // own adjoint + the wp part of lane-1 + the wm part of lane+1
    const W6<double> fromL = lane_below_w(gp), fromR = lane_above_w(gm);
    gc.d += fromL.d + fromR.d;
    gc.u0 += fromL.u0 + fromR.u0;
    gc.u1 += fromL.u1 + fromR.u1;
    gc.u2 += fromL.u2 + fromR.u2;
    gc.p += fromL.p + fromR.p;
    gc.e += fromL.e + fromR.e;
    if (emit) {
        const double v[6] = {gc.d, gc.u0, gc.u1, gc.u2, gc.p, gc.e};
        double lw[6];
        for (int k = 0; k < 6; k++) lw[k] = assign ? v[k] : prev[k] + v[k];
        if (ghost || tail.slot[zb + F] >= 0) {
            for (int k = 0; k < 6; k++) elem<double>(LW, (int64_t)k * n + zb, coff) = lw[k];
        } else {
            // an interior cell that donates to no ghost: the rest of the stage
            // (the bases laundered so that these loads are made here, not kept
            // in registers from the kernel's first loads of the same cells)
            const double *U = (const double *)a.uin;
            const double *lamr = lam;
            __asm__ volatile("" : "+s"(U), "+s"(lamr));
            const double r = elem<double>(U, zb, coff), E = elem<double>(U, 4 * n + zb, coff);
            const double m[3] = {elem<double>(U, n + zb, coff), elem<double>(U, 2 * n + zb, coff),
                                 elem<double>(U, 3 * n + zb, coff)};
            if constexpr (STAGE == 2) {
                double out[6], ls[5], lsum[5];
                fin2_core(r, m, E, lw, a.seps, out);
                const double *U0 = (const double *)a.uold;
                const double m0[3] = {elem<double>(U0, n + zb, coff), elem<double>(U0, 2 * n + zb, coff),
                                      elem<double>(U0, 3 * n + zb, coff)};
                double hl[5];  // lam/2, loaded again rather than held through the reverse
                for (int f = 0; f < 5; f++) hl[f] = gscale * elem<double>(lamr, (int64_t)f * n + zb, coff);
                mid2_tail(out, hl, elem<double>(U0, zb, coff), m0, r, E, a.gm1, a.seps, ls, lsum);
                for (int f = 0; f < 5; f++) {
                    elem<double>(tail.lus, (int64_t)f * n + zb, coff) = ls[f];
                    elem<double>(tail.lsum, (int64_t)f * n + zb, coff) = lsum[f];
                }
            } else {
                double ls[5], out[5];
                for (int f = 0; f < 5; f++) ls[f] = elem<double>(tail.lsum, (int64_t)f * n + zb, coff);
                fin1_core(r, m, E, lw, ls, a.gm1, a.seps, out);
                for (int f = 0; f < 5; f++) elem<double>(tail.lout, (int64_t)f * n + zb, coff) = out[f];
            }
        }
    }
}

// ------------------------------------------------------- pointwise kernels
// The donors (interior cells some ghost copies), after the stage-2
// directions: their own W* reverse plus their ghosts', then mid2_tail.
// Ghost entries of lus and lsum stay zero.
__global__ __launch_bounds__(256) void k_adj_mid2(int64_t n, int ndonor, const int32_t *__restrict__ donor,
                                                 double gm1, double seps, const double *__restrict__ u0,
                                                 const double *__restrict__ us, const double *__restrict__ lam,
                                                 const double *__restrict__ LW, GhostInv gi,
                                                 double *__restrict__ lus, double *__restrict__ lsum)
{
    const int kd = blockIdx.x * blockDim.x + threadIdx.x;
    if (kd >= ndonor) return;
    const int64_t c = donor[kd];
    double a[6];
    fin2_cell(n, c, seps, us, LW, a);
    for (int t = gi.start[kd]; t < gi.start[kd + 1]; t++) {
        double b[6];
        fin2_cell(n, gi.gidx[t], seps, us, LW, b);
        add_signed(a, b, gi.gsign[t]);
        a[5] += b[5];
    }
    double hl[5], ls[5], lsm[5];
    for (int f = 0; f < 5; f++) hl[f] = 0.5 * lam[f * n + c];
    const double m0[3] = {u0[n + c], u0[2 * n + c], u0[3 * n + c]};
    mid2_tail(a, hl, u0[c], m0, us[c], us[4 * n + c], gm1, seps, ls, lsm);
    for (int f = 0; f < 5; f++) {
        lus[f * n + c] = ls[f];
        lsum[f * n + c] = lsm[f];
    }
}

// The donors after the stage-1 directions: the adjoint of U^n at the cell
// plus that of the ghosts it donates to (ghost entries of lout stay zero).
__global__ __launch_bounds__(256) void k_adj_fin1(int64_t n, int ndonor, const int32_t *__restrict__ donor,
                                                 double gm1, double seps, const double *__restrict__ u0,
                                                 const double *__restrict__ lsum, const double *__restrict__ LW,
                                                 GhostInv gi, double *__restrict__ lout)
{
    const int kd = blockIdx.x * blockDim.x + threadIdx.x;
    if (kd >= ndonor) return;
    const int64_t c = donor[kd];
    double a[5];
    fin1_cell(n, c, gm1, seps, u0, lsum, LW, a);
    for (int t = gi.start[kd]; t < gi.start[kd + 1]; t++) {
        double b[5];
        fin1_cell(n, gi.gidx[t], gm1, seps, u0, lsum, LW, b);
        add_signed(a, b, gi.gsign[t]);
    }
    for (int f = 0; f < 5; f++) lout[f * n + c] = a[f];
}

// ghost fold: every donor gathers its ghosts' adjoints (momentum k negated
// where the ghost's sign bit k is set) and clears them; planes 5.. (p*) unsigned
__global__ __launch_bounds__(256) void k_adj_fold(int64_t n, int ndonor, const int32_t *__restrict__ donor,
                                                 const int32_t *__restrict__ start, const int32_t *__restrict__ gidx,
                                                 const int32_t *__restrict__ gsign, double *__restrict__ L5,
                                                 double *__restrict__ L1)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ndonor) return;
    const int64_t d = donor[k];
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int t = start[k]; t < start[k + 1]; t++) {
        const int64_t g = gidx[t];
        const int sg = gsign[t];
        for (int f = 0; f < 5; f++) {
            const double v = L5[f * n + g];
            acc[f] += (f >= 1 && f <= 3 && ((sg >> (f - 1)) & 1)) ? -v : v;
            L5[f * n + g] = 0.0;
        }
        if (L1) {
            acc[5] += L1[g];
            L1[g] = 0.0;
        }
    }
    for (int f = 0; f < 5; f++) L5[f * n + d] += acc[f];
    if (L1) L1[d] += acc[5];
}

}  // namespace bx

// ======================================================= host orchestration
struct bx_adj {
    int nsteps = 0, every = 0;
    std::vector<double> dts;
    double *ckpt = nullptr;  // checkpoints: U at steps 0, every, 2 every, ... (5 planes each)
    // segment stores, each U_0..U_every (5 planes), U*_s (5), p*_s (1) of one
    // segment, and the segment each holds (-1: none)
    std::vector<double *> stores;
    std::vector<int> store_seg;
    size_t store_bytes = 0;     // bytes of each store
    int ckpt_count = 0;         // checkpoints the checkpoint buffer holds
    int64_t store_budget = -1;  // bytes for segment stores; -1: half the free device memory
    double *lam[2] = {nullptr, nullptr};
    double *lus = nullptr;  // 5 planes: adjoint of U* (ghost entries zero)
    double *lsum = nullptr;  // 5 planes: adjoint of U* plus that of U' (ghost entries zero)
    double *lw = nullptr;   // 6 planes: primitive adjoints
    int32_t *fold = nullptr;  // donor, start, ghost, sign arrays (one allocation)
    int32_t *slot = nullptr;  // per cell: index of its donor entry, or -1
    int ndonor = 0, nfold = 0;
    int chunk_y = 64, chunk_z = 32;
    int64_t recomputed = 0;  // segments recomputed by backward sweeps
    double *dlog = nullptr;  // device log of the Courant dts of a forward
    int dlog_cap = 0;
};

namespace bx {

void adjoint_free(bx_ctx *c)
{
    if (!c->adj) return;
    bx_adj *j = c->adj;
    void *p[] = {j->ckpt, j->lam[0], j->lam[1], j->lus, j->lsum, j->lw, j->fold, j->slot, j->dlog};
    for (void *q : p)
        if (q) (void)hipFree(q);
    for (double *q : j->stores)
        if (q) (void)hipFree(q);
    delete j;
    c->adj = nullptr;
}

static int adj_init(bx_ctx *c)
{
    if (c->adj) return BX_OK;
    bx_adj *j = new (std::nothrow) bx_adj();
    if (!j) return set_error(BX_ERR_ARG, "adjoint: out of host memory");
    c->adj = j;
    // chunk lengths: shortened on small grids until the march kernels have
    // about 2 waves per SIMD to run
    {
        const int64_t want = 2 * 1024, bx = (c->nx - 4 + 63) / 64;
        const int nzi = c->flatz ? c->nz : c->nz - 4;
        while (j->chunk_y > 8 && bx * ((nzi + 3) / 4) * 4 * ((c->ny + j->chunk_y - 1) / j->chunk_y) < want) j->chunk_y /= 2;
        while (j->chunk_z > 8 && bx * (c->ny - 4) * ((c->nz + j->chunk_z - 1) / j->chunk_z) < want) j->chunk_z /= 2;
    }
    const size_t pb = sizeof(double) * (size_t)c->n;
    BX_HIP(hipMalloc((void **)&j->lam[0], 5 * pb));
    BX_HIP(hipMalloc((void **)&j->lam[1], 5 * pb));
    BX_HIP(hipMalloc((void **)&j->lus, 5 * pb));
    BX_HIP(hipMalloc((void **)&j->lsum, 5 * pb));
    BX_HIP(hipMalloc((void **)&j->lw, 6 * pb));
    // ghost entries of the adjoint planes are never written and stay zero, and
    // so do the primitive adjoints of cells no direction reaches
    for (double *q : {j->lam[0], j->lam[1], j->lus, j->lsum}) BX_HIP(hipMemsetAsync(q, 0, 5 * pb, c->stream));
    BX_HIP(hipMemsetAsync(j->lw, 0, 6 * pb, c->stream));
    // inverse ghost table: donors with the ghosts that copy them
    const int64_t ng = c->nghost;
    std::vector<int32_t> dst(ng), src(ng), sgn(ng);
    if (ng) {
        BX_HIP(hipMemcpy(dst.data(), c->ghost_dst, sizeof(int32_t) * ng, hipMemcpyDeviceToHost));
        // the table of the fused step (a pair of set_ghost_cells calls composed)
        const int32_t *ts = c->nscratch ? c->fghost_src : c->ghost_src;
        const int32_t *tg = c->nscratch ? c->fghost_sign : c->ghost_sign;
        if (!ts || !tg) return set_error(BX_ERR_UNSUPPORTED, "adjoint: ghost donors that are ghosts are not supported");
        BX_HIP(hipMemcpy(src.data(), ts, sizeof(int32_t) * ng, hipMemcpyDeviceToHost));
        BX_HIP(hipMemcpy(sgn.data(), tg, sizeof(int32_t) * ng, hipMemcpyDeviceToHost));
    }
    std::vector<int64_t> order(ng);
    for (int64_t k = 0; k < ng; k++) {
        if (src[k] < 0) return set_error(BX_ERR_UNSUPPORTED, "adjoint: ghost donors that are ghosts are not supported");
        order[k] = k;
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return src[a] < src[b]; });
    std::vector<int32_t> donor, start, gidx, gs;
    for (int64_t t = 0; t < ng; t++) {
        const int64_t k = order[t];
        if (donor.empty() || donor.back() != src[k]) {
            donor.push_back(src[k]);
            start.push_back((int32_t)t);
        }
        gidx.push_back(dst[k]);
        gs.push_back(sgn[k]);
    }
    start.push_back((int32_t)ng);
    j->ndonor = (int)donor.size();
    j->nfold = (int)ng;
    {
        std::vector<int32_t> slot((size_t)c->n, -1);
        for (size_t k = 0; k < donor.size(); k++) slot[donor[k]] = (int32_t)k;
        BX_HIP(hipMalloc((void **)&j->slot, sizeof(int32_t) * (size_t)c->n));
        BX_HIP(hipMemcpy(j->slot, slot.data(), sizeof(int32_t) * (size_t)c->n, hipMemcpyHostToDevice));
    }
    const size_t tot = donor.size() + start.size() + 2 * (size_t)ng;
    BX_HIP(hipMalloc((void **)&j->fold, sizeof(int32_t) * std::max<size_t>(tot, 1)));
    int32_t *f = j->fold;
    BX_HIP(hipMemcpy(f, donor.data(), sizeof(int32_t) * donor.size(), hipMemcpyHostToDevice));
    BX_HIP(hipMemcpy(f + donor.size(), start.data(), sizeof(int32_t) * start.size(), hipMemcpyHostToDevice));
    if (ng) {
        BX_HIP(hipMemcpy(f + donor.size() + start.size(), gidx.data(), sizeof(int32_t) * ng, hipMemcpyHostToDevice));
        BX_HIP(hipMemcpy(f + donor.size() + start.size() + ng, gs.data(), sizeof(int32_t) * ng,
                         hipMemcpyHostToDevice));
    }
    return BX_OK;
}

static int fold(bx_ctx *c, double *L5, double *L1)
{
    bx_adj *j = c->adj;
    if (j->ndonor == 0) return BX_OK;
    const int32_t *f = j->fold;
    const int32_t *donor = f, *start = f + j->ndonor, *gidx = start + j->ndonor + 1, *gs = gidx + j->nfold;
    hipLaunchKernelGGL(k_adj_fold, dim3(blocks_for(j->ndonor, 256)), dim3(256), 0, c->stream, c->n, j->ndonor, donor,
                       start, gidx, gs, L5, L1);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

// adjoint of fluxtmp (gscale*lam) -> primitive adjoints in lw, every direction
template <int STAGE>
static int directions(bx_ctx *c, const StageArgs &a, const double *lam, double gscale, const AdjTail &tail)
{
    bx_adj *j = c->adj;
    const double kx = a.ayz * a.dt / a.vol, ky = a.axz * a.dt / a.vol, kz = a.axy * a.dt / a.vol;
    const int nzi = a.zhi - a.zlo;
    const unsigned bx = (unsigned)((c->nx - 4 + 63) / 64);
    {
        dim3 grid(bx, (unsigned)((nzi + 3) / 4), (unsigned)((c->ny + j->chunk_y - 1) / j->chunk_y));
        hipLaunchKernelGGL((k_adj_march<STAGE, 1>), grid, dim3(256), 0, c->stream, a, lam, gscale, ky, j->lw,
                           j->chunk_y);
        BX_HIP(hipGetLastError());
    }
    if (!c->flatz) {  // flat z: the z terms are exactly zero
        dim3 grid(bx, (unsigned)((c->ny - 4 + 3) / 4), (unsigned)((c->nz + j->chunk_z - 1) / j->chunk_z));
        hipLaunchKernelGGL((k_adj_march<STAGE, 2>), grid, dim3(256), 0, c->stream, a, lam, gscale, kz, j->lw,
                           j->chunk_z);
        BX_HIP(hipGetLastError());
    }
    {
        const int64_t span = (int64_t)(c->ny - 4) * c->nx + 6;
        const unsigned waves = (unsigned)((span + kAdjOut - 1) / kAdjOut);
        dim3 grid((waves + 3) / 4, (unsigned)nzi, 1);
        hipLaunchKernelGGL((k_adj_x<STAGE>), grid, dim3(256), 0, c->stream, a, lam, gscale, kx, j->lw, tail);
        BX_HIP(hipGetLastError());
    }
    return BX_OK;
}

static GhostInv ghost_inv(bx_ctx *c)
{
    bx_adj *j = c->adj;
    GhostInv g;
    g.slot = j->slot;
    g.start = j->fold + j->ndonor;
    g.gidx = g.start + j->ndonor + 1;
    g.gsign = g.gidx + j->nfold;
    return g;
}

// one reverse step: lam (adjoint of U^{n+1}, ghosts clear) -> lout (adjoint of U^n, ghosts clear)
static int adj_step(bx_ctx *c, double dt, const double *un, const double *us, const double *ps, const double *lam,
                    double *lout)
{
    bx_adj *j = c->adj;
    int rc;
    const int64_t n = c->n;
    const GhostInv gi = ghost_inv(c);
    const int32_t *donor = j->fold;
    const unsigned nbd = blocks_for(j->ndonor, 256);
    StageArgs b = base_args(c, dt);
    b.uin = us;
    b.pin = ps;
    b.uold = un;
    // the x kernel ends the reverse of every interior cell but the donors
    if ((rc = directions<2>(c, b, lam, 0.5, AdjTail{j->slot, j->lus, j->lsum, nullptr}))) return rc;
    if (j->ndonor) {
        hipLaunchKernelGGL(k_adj_mid2, dim3(nbd), dim3(256), 0, c->stream, n, j->ndonor, donor,
                           c->prm.gammafix - 1.0, c->prm.small_eps, un, us, lam, j->lw, gi, j->lus, j->lsum);
        BX_HIP(hipGetLastError());
    }
    StageArgs a = base_args(c, dt);
    a.uin = un;
    if ((rc = directions<1>(c, a, j->lus, 1.0, AdjTail{j->slot, nullptr, j->lsum, lout}))) return rc;
    if (j->ndonor) {
        hipLaunchKernelGGL(k_adj_fin1, dim3(nbd), dim3(256), 0, c->stream, n, j->ndonor, donor,
                           c->prm.gammafix - 1.0, c->prm.small_eps, un, j->lsum, j->lw, gi, lout);
        BX_HIP(hipGetLastError());
    }
    return BX_OK;
}

static int read_flags_rate(bx_ctx *c, unsigned *flags, double *rate)
{
    BX_HIP(hipMemcpyAsync(c->hsc, c->dsc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
    BX_HIP(hipStreamSynchronize(c->stream));
    *flags = c->hsc->flags;
    unsigned long long b = c->hsc->rate_bits[0];
    std::memcpy(rate, &b, sizeof(double));
    return BX_OK;
}

}  // namespace bx

using namespace bx;

extern "C" int bx_adjoint_forward(bx_ctx *c, int32_t nsteps, int32_t every, const double *dts_in, double *dts_out)
{
    if (!c || nsteps < 0 || every < 0) return set_error(BX_ERR_ARG, "bx_adjoint_forward: bad arguments");
    if (c->prm.precision != BX_PREC_FP64) return set_error(BX_ERR_UNSUPPORTED, "bx_adjoint_forward: FP64 contexts only");
    if (c->eos) return set_error(BX_ERR_UNSUPPORTED, "bx_adjoint_forward: fixed-gamma EOS (H2STAT=-1) only");
    if (c->prm.fluxangmom) return set_error(BX_ERR_UNSUPPORTED, "bx_adjoint_forward: fluxangmom off only");
    if (c->grav_flags) return set_error(BX_ERR_UNSUPPORTED, "bx_adjoint_forward: no gravity configuration");
    if (c->nd2 > 0)
        return set_error(BX_ERR_UNSUPPORTED, "bx_adjoint_forward: two-level corner ghosts (walls in y and z) are "
                                             "not differentiated");
    BX_HIP(hipSetDevice(c->device));
    int rc;
    if (!c->fused_ready) {
        if ((rc = fused_prepare(c))) return rc;
        c->fused_ready = true;
    }
    if ((rc = adj_init(c))) return rc;
    bx_adj *j = c->adj;
    j->nsteps = 0;  // set again only when this forward completes: a refused run leaves nothing to sweep
    // a regular state: ghosts filled from their donors, u = m/rho, p = state()
    if ((rc = launch_ghosts(c, c->cons))) return rc;
    if ((rc = launch_materialize(c))) return rc;
    if ((rc = launch_regular_check(c))) return rc;
    unsigned flags;
    double rate;
    if ((rc = read_flags_rate(c, &flags, &rate))) return rc;
    if (flags) return set_error(BX_ERR_UNSUPPORTED, "bx_adjoint_forward: the state is not regular (rho<1e-4 or u != m/rho)");
    c->ghosts_filled = true;
    c->fused_valid = true;
    c->aux_stale = false;
    const size_t pb = sizeof(double) * (size_t)c->n;
    const int64_t n = c->n;
    j->every = every;
    // Checkpoints every `every` steps; the last `nkeep` segments also keep
    // their per-step states (U, U*, p*) from this forward, as many as the
    // store budget holds (at least one store, which the backward needs to
    // recompute the others).  Writing them costs no traffic: the steps write
    // their stage results there instead of into the scratch planes.
    const int nseg = every > 0 && nsteps > 0 ? (nsteps + every - 1) / every : 0;
    const size_t seg_bytes = pb * (size_t)(5 * (every + 1) + 6 * every);
    int nkeep = 0;
    if (nseg > 0) {
        // buffers are kept between forwards while their sizes still fit
        if (j->ckpt && j->ckpt_count < nseg) {
            (void)hipFree(j->ckpt);
            j->ckpt = nullptr;
        }
        if (!j->ckpt) {
            BX_HIP(hipMalloc((void **)&j->ckpt, 5 * pb * (size_t)nseg));
            j->ckpt_count = nseg;
        }
        if (j->store_bytes != seg_bytes) {
            for (double *q : j->stores)
                if (q) (void)hipFree(q);
            j->stores.clear();
            j->store_bytes = seg_bytes;
        }
        int64_t budget = j->store_budget;
        if (budget < 0) {
            size_t fr = 0, tot = 0;
            BX_HIP(hipMemGetInfo(&fr, &tot));
            budget = (int64_t)((fr + j->stores.size() * seg_bytes) / 2);
        }
        nkeep = (int)std::max<int64_t>(1, std::min<int64_t>(nseg, budget / (int64_t)seg_bytes));
        while ((int)j->stores.size() > nkeep) {
            (void)hipFree(j->stores.back());
            j->stores.pop_back();
        }
        while ((int)j->stores.size() < nkeep) {
            double *q = nullptr;
            const hipError_t e = hipMalloc((void **)&q, seg_bytes);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                if (j->stores.empty()) return hip_check(e, "hipMalloc adjoint segment store");
                nkeep = (int)j->stores.size();  // fewer stores than the budget allows: recompute the rest
                break;
            }
            j->stores.push_back(q);
        }
    }
    j->store_seg.assign(j->stores.size(), -1);
    const int first_kept = nseg - nkeep;  // segments [first_kept, nseg) are kept
    j->dts.assign(nsteps, 0.0);
    // Nothing is read back per step: the flags accumulate over the run and are
    // checked once at its end.  Without a given dt sequence each step's
    // Courant dt is formed on the device (main.f90:266-281 with no endtime)
    // and the sequence is read back once.
    const bool frozen = dts_in != nullptr;
    if (frozen) {
        BX_HIP(hipMemsetAsync(&c->dsc->flags, 0, sizeof(unsigned int), c->stream));
    } else if (nsteps > 0) {
        if (j->dlog_cap < nsteps) {
            if (j->dlog) (void)hipFree(j->dlog);
            j->dlog = nullptr;
            BX_HIP(hipMalloc((void **)&j->dlog, sizeof(double) * (size_t)nsteps));
            j->dlog_cap = nsteps;
        }
        DevScalars *h = c->hsc;  // rate_bits[0]: the Courant rate of the regular check above
        h->time = 0.0;
        h->dtold = 0.0;
        h->step = 0;
        h->endtime = std::numeric_limits<double>::infinity();
        h->flags = 0;
        h->logk = 0;
        BX_HIP(hipMemcpyAsync(c->dsc, h, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream));
    }
    // Once the steps have started, the context's state, its u/p/cons_old planes
    // and the Courant rate no longer go together: an error or a refusal marks
    // them stale (the next bx_step checks the state again) and, when a
    // checkpoint of the start was kept, puts the starting state back.
    auto invalidate = [&](bool restore) {
        c->fused_valid = false;
        c->aux_stale = true;
        c->old_stale = true;
        if (restore && nseg > 0)
            (void)hipMemcpyAsync(c->cons, j->ckpt, 5 * pb, hipMemcpyDeviceToDevice, c->stream);
        (void)hipStreamSynchronize(c->stream);
    };
    auto refuse = [&](unsigned f) -> int {
        invalidate(true);
        if (f & FLAG_IRREGULAR) return set_error(BX_ERR_UNSUPPORTED, "bx_adjoint_forward: the run met rho<1e-4 or |u|>vlimit");
        return set_error(BX_ERR_UNSUPPORTED, "bx_adjoint_forward: the run met a floor of the update (not differentiated)");
    };
    auto U = [&](int i, int t) { return j->stores[i] + (size_t)5 * n * t; };
    auto Sx = [&](int i, int t) { return j->stores[i] + (size_t)5 * n * (every + 1) + (size_t)5 * n * t; };
    auto Px = [&](int i, int t) { return j->stores[i] + (size_t)5 * n * (every + 1) + (size_t)5 * n * every + (size_t)n * t; };
    for (int k = 0; k < nsteps; k++) {
        const int sg = every > 0 ? k / every : 0, t = every > 0 ? k % every : 0;
        if (every > 0 && t == 0)
            BX_HIP(hipMemcpyAsync(j->ckpt + (size_t)sg * 5 * n, sg > first_kept ? U(sg - 1 - first_kept, every) : c->cons,
                                  5 * pb, hipMemcpyDeviceToDevice, c->stream));
        const bool kept = nseg > 0 && sg >= first_kept;
        if (kept && t == 0)  // the segment's first state: the scratch state, or the previous store's last
            BX_HIP(hipMemcpyAsync(U(sg - first_kept, 0), sg > first_kept ? U(sg - 1 - first_kept, every) : c->cons,
                                  5 * pb, hipMemcpyDeviceToDevice, c->stream));
        const int i = kept ? sg - first_kept : 0;
        double *u0 = kept ? U(i, t) : c->cons, *us = kept ? Sx(i, t) : c->ustar;
        double *ps = kept ? Px(i, t) : c->pstar, *un = kept ? U(i, t + 1) : c->unext;
        if (frozen) {
            if ((rc = fused_step_into(c, dts_in[k], u0, us, ps, un, false))) {
                invalidate(false);
                return rc;
            }
            j->dts[k] = dts_in[k];
        } else if ((rc = fused_step_device_dt(c, u0, us, ps, un, j->dlog, nsteps))) {
            invalidate(false);
            return rc;
        }
        if (!kept) std::swap(c->cons, c->unext);
    }
    if (!frozen && nsteps > 0)
        BX_HIP(hipMemcpyAsync(j->dts.data(), j->dlog, sizeof(double) * (size_t)nsteps, hipMemcpyDeviceToHost,
                              c->stream));
    if (nkeep > 0)  // the final state, from the end of the last kept segment
        BX_HIP(hipMemcpyAsync(c->cons, U(nkeep - 1, nsteps - (nseg - 1) * every), 5 * pb, hipMemcpyDeviceToDevice,
                              c->stream));
    if ((rc = read_flags_rate(c, &flags, &rate))) {  // also completes the dt read-back
        invalidate(false);
        return rc;
    }
    if (flags & (FLAG_IRREGULAR | FLAG_FLOOR)) return refuse(flags);
    if (dts_out)
        for (int k = 0; k < nsteps; k++) dts_out[k] = j->dts[k];
    for (int i = 0; i < nkeep; i++) j->store_seg[i] = first_kept + i;
    j->nsteps = nsteps;
    c->aux_stale = true;
    c->old_stale = true;
    BX_HIP(hipStreamSynchronize(c->stream));
    return BX_OK;
}

extern "C" int bx_adjoint_backward(bx_ctx *c, const double *lam_T, double *lam_0)
{
    if (!c || !lam_T || !lam_0) return set_error(BX_ERR_ARG, "bx_adjoint_backward: null argument");
    if (!c->adj || c->adj->every <= 0 || !c->adj->ckpt || c->adj->stores.empty() || c->adj->nsteps <= 0)
        return set_error(BX_ERR_STATE,
                         "bx_adjoint_backward: no completed checkpointed forward (bx_adjoint_forward with nsteps > 0, "
                         "every > 0)");
    BX_HIP(hipSetDevice(c->device));
    bx_adj *j = c->adj;
    const int64_t n = c->n;
    const size_t pb = sizeof(double) * (size_t)n;
    int rc;
    BX_HIP(hipMemcpyAsync(j->lam[0], lam_T, 5 * pb, hipMemcpyDeviceToDevice, c->stream));
    if ((rc = fold(c, j->lam[0], nullptr))) return rc;
    int cur = 0;
    const int E = j->every;
    const int ns = (int)j->stores.size();
    for (int s0 = ((j->nsteps - 1) / E) * E; s0 >= 0; s0 -= E) {
        const int sg = s0 / E, len = std::min(E, j->nsteps - s0);
        int i = -1;
        for (int q = 0; q < ns; q++)
            if (j->store_seg[q] == sg) i = q;
        if (i < 0) {
            // recompute the segment from its checkpoint with the frozen dts, into a
            // store whose segment this sweep has passed, else the one needed last
            for (int q = 0; q < ns && i < 0; q++)
                if (j->store_seg[q] < 0 || j->store_seg[q] > sg) i = q;
            if (i < 0) {
                i = 0;
                for (int q = 1; q < ns; q++)
                    if (j->store_seg[q] < j->store_seg[i]) i = q;
            }
            double *Useg = j->stores[i], *Sseg = Useg + (size_t)5 * n * (E + 1), *Pseg = Sseg + (size_t)5 * n * E;
            BX_HIP(hipMemcpyAsync(Useg, j->ckpt + (size_t)sg * 5 * n, 5 * pb, hipMemcpyDeviceToDevice, c->stream));
            for (int t = 0; t < len; t++)
                if ((rc = fused_step_into(c, j->dts[s0 + t], Useg + (size_t)5 * n * t, Sseg + (size_t)5 * n * t,
                                          Pseg + (size_t)n * t, Useg + (size_t)5 * n * (t + 1), true)))
                    return rc;
            j->store_seg[i] = sg;
            j->recomputed++;
        }
        double *Useg = j->stores[i], *Sseg = Useg + (size_t)5 * n * (E + 1), *Pseg = Sseg + (size_t)5 * n * E;
        for (int t = len - 1; t >= 0; t--) {
            if ((rc = adj_step(c, j->dts[s0 + t], Useg + (size_t)5 * n * t, Sseg + (size_t)5 * n * t,
                               Pseg + (size_t)n * t, j->lam[cur], j->lam[1 - cur])))
                return rc;
            cur = 1 - cur;
        }
    }
    BX_HIP(hipMemcpyAsync(lam_0, j->lam[cur], 5 * pb, hipMemcpyDeviceToDevice, c->stream));
    BX_HIP(hipStreamSynchronize(c->stream));
    return BX_OK;
}

extern "C" int bx_adjoint_tune(bx_ctx *c, int32_t y_chunk, int32_t z_chunk)
{
    if (!c || y_chunk < 1 || z_chunk < 1) return set_error(BX_ERR_ARG, "bx_adjoint_tune: bad arguments");
    BX_HIP(hipSetDevice(c->device));
    int rc;
    if ((rc = adj_init(c))) return rc;
    c->adj->chunk_y = y_chunk;
    c->adj->chunk_z = z_chunk;
    return BX_OK;
}

extern "C" int bx_adjoint_store(bx_ctx *c, int64_t bytes, int64_t *recomputed)
{
    if (!c || bytes < -1) return set_error(BX_ERR_ARG, "bx_adjoint_store: bad arguments");
    BX_HIP(hipSetDevice(c->device));
    int rc;
    if ((rc = adj_init(c))) return rc;
    c->adj->store_budget = bytes;
    if (recomputed) *recomputed = c->adj->recomputed;
    return BX_OK;
}

extern "C" int bx_cons_device(bx_ctx *c, double *planes, int32_t to_ctx)
{
    if (!c || !planes) return set_error(BX_ERR_ARG, "bx_cons_device: null argument");
    if (c->prm.precision != BX_PREC_FP64) return set_error(BX_ERR_UNSUPPORTED, "bx_cons_device: FP64 contexts only");
    BX_HIP(hipSetDevice(c->device));
    const size_t bytes = sizeof(double) * 5 * (size_t)c->n;
    if (to_ctx) {
        BX_HIP(hipMemcpyAsync(c->cons, planes, bytes, hipMemcpyDeviceToDevice, c->stream));
        c->fused_valid = false;
        c->ghosts_filled = false;
        c->aux_stale = true;
        c->old_stale = true;
    } else {
        BX_HIP(hipMemcpyAsync(planes, c->cons, bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    BX_HIP(hipStreamSynchronize(c->stream));
    return BX_OK;
}
