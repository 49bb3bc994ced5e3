// bx_batch.hip -- many independent shock tubes in one persistent launch.
//
// A tube is an nx x ny x nz grid with nz < 6 (flat z, init_grid.f90:241-247),
// outflow on every side, and a set-up that is the same in every (y,z) row:
// each cell holds its row's value, except the outermost ghost layer, the
// "frame" (boundary=1, init_grid.f90:81-107), which holds rho=small_rho, m=0,
// E=small_eps with the row's u and p (init_conditions.f90:60-68).
//
// On such a grid the reference keeps every row bit-identical to every other
// row for the whole run: identical inputs go through identical operations;
// the two y faces (and the two flat-z faces) of a cell are evaluated on the
// same canonicalised states, so fy(2)-fy(1) and fz(2)-fz(1) are exactly +0
// and fluxtmp (flux.f90:187-191) is -(((areayz*dfx + 0) + 0)*dt)/vol.  A
// tube is therefore advanced as one x-row of nx cells.  Ghost cells are read
// through their donors (outflow: c<2 -> 2, c>nx-3 -> nx-3): after the first
// ghost fill a ghost's cons, u and p equal its donor's at every state() and
// velocity() call, because both calls are pointwise functions of equal inputs.
//
// The first step is the one exception, as in the reference: its Courant
// maximum (main.f90:232-241) runs over the set-up, frame cells included, and
// those take the low-density branch of state.f90:52-56.  The frame tuples
// carry the row's u and p, so the maximum is over {row cell c, 0<c<nx-1} and
// {frame tuple at c} for every c.  bx_batch_check verifies what the row form
// needs of the set-up (density >= 1e-4 outside the frame, ghost velocities
// equal to their donors').
//
// Every state() (state.f90:109-139, the low-density branch and its energy
// reset included) and velocity() call (velocity.f90:18-53, the vlimit clamp
// included) is evaluated as the reference makes it, with the pressure and
// velocity the reference holds at that point of main.f90:211-555.
//
// Work decomposition: one workgroup advances one tube at a time, taking tubes
// from a queue ordered by estimated cost (largest first), so the long tubes
// start first and the short ones fill in behind them.  A step is two
// barrier-separated phases over the row (predictor and corrector) and one
// workgroup max-reduction of the Courant rate.  In a phase each wave takes
// segments of 64 consecutive cells (60 outputs, x-neighbours through wave
// shuffles) and carries each cell from its stored values through every
// pointwise call up to the stage's fluxes, then through the stage update and
// every pointwise call up to the next stage.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "bx_stage.h"

namespace bx {

enum : unsigned int {
    FLAG_BADDT = 4u,  // a non-finite or non-positive dt: the tube was stopped there
};

struct Tube {
    int64_t off;  // first cell of the tube's row in the concatenated rows
    int nx;
    int dx_pow2, vol_pow2, maxsteps;
    double dx, dy, dz, inv_dx;
    double ayz, vol, inv_vol;
    double gam, gm1, cfl, endtime, srho, seps, vlimit, muc, rgas;
};

// Rows are planes of `total` doubles, cell-contiguous per tube:
//   ic   9 planes: set-up rho, mx, my, mz, E, u0, u1, u2, p
//   st   9 planes: end-of-step state (same order), the input of the next step
//   ss   9 planes: predictor result U*, u* (velocity at main.f90:355), p* (:353)
//   old  5 planes: cons_old (main.f90:327-333)
struct BatchArgs {
    int ntube;
    int maxlog;
    int64_t total;
    const Tube *tubes;
    const int *order;
    int *queue;
    const double *ic;
    double *st;
    double *ss;
    double *old;
    double *t_out;
    int *steps_out;
    unsigned *status;
    double *dtlog;
    int nqueue;                     // tubes in the queue (order[0..nqueue))
    int ngang, gang_p;              // gangs and members per gang (gang_p > 1: gangs on)
    const int *gang_tube;           // the tube of each gang
    unsigned *gang_ctr;             // ngang arrival counters
    unsigned long long *gang_rate;  // 3 per gang
    unsigned *gang_fail;
};

// Element c of a plane, addressed as the plane's (workgroup-uniform) base plus
// a 32-bit byte offset: the compiler then keeps one offset register per cell
// for all planes instead of a 64-bit address per plane.
#define PL(base, c) \
    (*reinterpret_cast<double *>(reinterpret_cast<char *>(const_cast<double *>(base)) + ((unsigned)(c) << 3)))

// a value every lane holds, marked uniform for the compiler
__device__ __forceinline__ double uniform_d(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ int donor_of(int c, int nx) { return c < 2 ? 2 : (c > nx - 3 ? nx - 3 : c); }

// state.f90:109-139 (H2STAT=-1): pressure from (rho, E, u) and the pressure
// pold held before the call; the low-density branch resets E.
__device__ __forceinline__ double eos_state(const Tube &t, double rho, double &E, double u0, double u1, double u2,
                                            double pold, unsigned &flg)
{
    const double ekin = rho * 0.5 * (u0 * u0 + u1 * u1 + u2 * u2);
    double eps = floor_at(E - ekin, t.seps);
    if (rho < 1e-4) {
        flg |= FLAG_LOWRHO;
        const double tk = vmin(5e3, pold / (rho * t.rgas) * t.muc);
        const double pc = rho * tk * t.rgas / t.muc;
        eps = pc / t.gm1;
        E = eps + ekin;
        return pc;
    }
    return eps * t.gm1;
}

// velocity.f90:18-53 on one cell: u = m/rho, then the vlimit clamp
__device__ __forceinline__ void velocity_of(const Tube &t, double U[5], double v[3], unsigned &flg)
{
    const double rho = U[0];
    v[0] = qdiv(U[1], rho);
    v[1] = qdiv(U[2], rho);
    v[2] = qdiv(U[3], rho);
    if (fabs(v[0]) > t.vlimit || fabs(v[1]) > t.vlimit || fabs(v[2]) > t.vlimit) {
        flg |= FLAG_CLAMP;
        double ekin = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / rho;
        double speed = 0.0;
        for (int d = 0; d < 3; d++) speed = speed + v[d] * v[d];
        speed = sqrt(speed);
        for (int d = 0; d < 3; d++) {
            v[d] = v[d] * t.vlimit / speed;
            U[1 + d] = v[d] * rho;
        }
        const double E = U[4] - ekin;
        ekin = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / rho;
        U[4] = floor_at(E, t.seps) + ekin;
    }
}

// source(): the energy re-round with the gravity vector -0 (source.f90:161-182)
__device__ __forceinline__ double source_E(const Tube &t, const double U[5])
{
    const double ek = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
    return floor_at(U[4] - ek, t.seps) + ek;
}

// courant.f90:20-21 for one cell
__device__ __forceinline__ double rate_of(const Tube &t, double rho, double p, double u0, double u1, double u2)
{
    const double sq = t.gam * p / rho;
    double sound;
    sqrt_n<1>(&sq, &sound);
    const double r0 = (sound + fabs(u0)) / t.dx;
    const double r1 = (sound + fabs(u1)) / t.dy;
    // the same operands give the same quotient: the z term of a cube cell with |u1|==|u2|
    const double r2 = (t.dz == t.dy && fabs(u2) == fabs(u1)) ? r1 : (sound + fabs(u2)) / t.dz;
    return r0 + r1 + r2;
}

// fluxtmp of one component of a row cell (flux.f90:187-191): y and z terms are +0
__device__ __forceinline__ double flux_tmp(const Tube &t, double dFx, double dt)
{
    const double s = (t.ayz * dFx + 0.0) + 0.0;
    const double q = s * dt;
    return -(t.vol_pow2 ? q * t.inv_vol : qdiv(q, t.vol));
}

__device__ __forceinline__ double block_max(double v, double *red, int nw)
{
    for (int off = 32; off > 0; off >>= 1) v = vmax(v, __shfl_xor(v, off, 64));
    __syncthreads();  // earlier readers of red are done; this phase's stores are visible
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = red[0];
    for (int k = 1; k < nw; k++) m = vmax(m, red[k]);
    return m;
}

// Only lanes 1..62 of a segment use their neighbours' values (lanes 0, 1, 62
// and 63 are its halo), so whole-wave shifts serve for the x-neighbours.
// x interface states of every lane's cell and the flux of its face x-1/2
// (the face x+1/2 is the next lane's)
__device__ __forceinline__ void x_faces(const StageArgs &sa, const W6<double> &q, double snd, bool ghost,
                                        double Fx[5])
{
    const W6<double> wm = lane_below_w(q), wp = lane_above_w(q);
    Side<double> mi, pl;
    dir_states<0>(wm, q, wp, snd, ghost, sa, mi, pl);
    Side<double> L;
    L.d = lane_below(pl.d);
    L.v[0] = lane_below(pl.v[0]);
    L.v[1] = lane_below(pl.v[1]);
    L.v[2] = lane_below(pl.v[2]);
    L.p = lane_below(pl.p);
    L.e = lane_below(pl.e);
    L.g = sa.gam;
    face_flux<double, 0, true>(L, mi, Fx);
}

// ------------------------------------------------------------------ gangs
// A tube too long for one CU can be split over a gang of P workgroups, each
// taking a contiguous range of its segments.  The predictor of a member also
// covers the two cells past each shared edge (the neighbour computes them too,
// with the same operations on the same values), so only the end-of-step state
// of the four cells on either side of an edge crosses between members: those
// are written and read with device-coherent (agent-scope) accesses, and the
// members meet at one barrier on a device counter per step, after the
// corrector, with the Courant maximum.  Every other cell stays with its
// member.  The arithmetic per cell is the single-workgroup form's, and the
// maximum is exact, so a gang gives the same bits.  A barrier that waits
// longer than kGangSpins sleeps (members not resident together) marks the
// batch failed and the member leaves the tube; the host then runs the batch
// again without gangs.
constexpr unsigned kGangSpins = 1u << 24;

struct GangCtx {
    unsigned *ctr;             // arrivals, monotonic within a run
    unsigned long long *rate;  // three rotating slots for the Courant maximum (bits of a double >= 0)
    unsigned *fail;
    int member, p;
    unsigned epoch;            // barriers passed (the same in every thread)
    int exch;                  // maxima exchanged
};

__device__ __forceinline__ double load_c(const double *p)
{
    return __longlong_as_double(__hip_atomic_load((const long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ void store_c(double *p, double v)
{
    __hip_atomic_store((long long *)p, __double_as_longlong(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// false: the wait gave up
__device__ __forceinline__ bool gang_barrier(GangCtx &g, int *s_ok)
{
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores and atomics are complete
    __syncthreads();
    g.epoch++;
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(g.ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned target = g.epoch * (unsigned)g.p;
        unsigned spins = 0;
        int ok = 1;
        while (__hip_atomic_load(g.ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(4);
            if (++spins > kGangSpins) {
                ok = 0;
                __hip_atomic_fetch_or(g.fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

// the gang's maximum of each member's r (every thread holds its member's r)
__device__ __forceinline__ bool gang_max(GangCtx &g, double &r, int *s_ok, double *s_r)
{
    const int slot = g.exch % 3;
    if (threadIdx.x == 0)
        __hip_atomic_fetch_max(&g.rate[slot], (unsigned long long)__double_as_longlong(r), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    if (!gang_barrier(g, s_ok)) return false;
    if (threadIdx.x == 0) {
        *s_r = __longlong_as_double(
            (long long)__hip_atomic_load(&g.rate[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        // the slot two exchanges on: every member read it before two barriers ago
        if (g.member == 0)
            __hip_atomic_store(&g.rate[(slot + 1) % 3], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    r = *s_r;
    g.exch++;
    return true;
}

// One tube ti; GANG: this workgroup is member g.member of g.p, else the whole
// tube.  Every thread of the workgroup calls it.
template <int W, bool GANG>
__device__ __forceinline__ void run_tube(const BatchArgs &a, const Tube &t, int ti, GangCtx &g, double *red, int *s_ok,
                         double *s_r, unsigned *sflag, double (*ext)[9][2])
{
    const int lane = threadIdx.x & 63, wv = wave_id();  // wave-uniform for the compiler
    const int64_t T = a.total;
    const int nx = t.nx;
    const int nseg = (nx - 4 + kSegOut - 1) / kSegOut;
    // segments [s0, s1) and output cells [c_lo, c_hi) of this workgroup
    int s0 = 0, s1 = nseg, c_lo = 0, c_hi = nx;
    if (GANG) {
        const int per = (nseg + g.p - 1) / g.p;
        s0 = min(nseg, g.member * per);
        s1 = min(nseg, s0 + per);
        c_lo = g.member == 0 ? 0 : s0 * kSegOut + 2;
        c_hi = g.member == g.p - 1 ? nx : s1 * kSegOut + 2;
    }
    // The predictor also takes the two cells beyond each shared edge of the
    // range (the segment before and after, results kept to [w_lo, w_hi)), so
    // that the corrector needs no other member's predictor results: the
    // members meet once per step, and exchange the four end-of-step cells on
    // either side of each shared edge that those extra cells read.  The extra
    // cells' predictor results stay in LDS (ext[side][plane][cell]): the
    // neighbour writes the same cells' results to the rows, and two workgroups
    // writing one line through different caches could leave either's older
    // bytes in memory.
    const bool left = GANG && g.member > 0, right = GANG && g.member < g.p - 1;
    const int sp0 = s0 - (left ? 1 : 0), sp1 = s1 + (right ? 1 : 0);
    const int w_lo = c_lo - (left ? 2 : 0), w_hi = c_hi + (right ? 2 : 0);
    // a cell another member writes, or one of ours another member reads
    auto foreign = [&](int c) { return GANG && (c < c_lo || c >= c_hi); };
    auto shared_out = [&](int c) { return (left && c < c_lo + 4) || (right && c >= c_hi - 4); };
    auto ld = [&](const double *base, int c) -> double {
        if (GANG && (foreign(c) || shared_out(c))) return load_c(&PL(base, c));
        return PL(base, c);
    };
    auto st = [&](double *base, int c, double v) {
        if (GANG && shared_out(c)) store_c(&PL(base, c), v);
        else PL(base, c) = v;
    };
    const double *icr = a.ic + t.off;
    double *str = a.st + t.off;
    double *ssr = a.ss + t.off;
    double *oldr = a.old + t.off;

    StageArgs sa{};
    sa.ds[0] = t.dx;
    sa.inv_ds[0] = t.inv_dx;
    sa.ds_pow2[0] = t.dx_pow2;
    sa.gam = t.gam;
    unsigned flg = 0;

    // Courant maximum of the set-up (main.f90:232-241): every row cell
    // with its own values, and the frame tuple at every c
    double r0 = 0.0;
    for (int c = c_lo + (int)threadIdx.x; c < c_hi; c += 64 * W) {
        const double u0 = PL(icr + 5 * T, c), u1 = PL(icr + 6 * T, c), u2 = PL(icr + 7 * T, c), pi = PL(icr + 8 * T, c);
        unsigned fset = 0;
        if (c > 0 && c < nx - 1) {
            const double rho = PL(icr, c);
            double E = PL(icr + 4 * T, c);
            const double p = eos_state(t, rho, E, u0, u1, u2, pi, fset);
            r0 = vmax(r0, rate_of(t, rho, p, u0, u1, u2));
        }
        double Ef = t.seps;
        const double pf = eos_state(t, t.srho, Ef, u0, u1, u2, pi, fset);
        r0 = vmax(r0, rate_of(t, t.srho, pf, u0, u1, u2));
    }
    double rate = uniform_d(block_max(r0, red, W));
    if (GANG && !gang_max(g, rate, s_ok, s_r)) return;

    double time = 0.0, dtold = 0.0;
    int step = 0;
    unsigned status = 0;
    for (;;) {
        if (time >= t.endtime || step >= t.maxsteps) break;  // main.f90:212
        step++;
        // main.f90:266-281; every thread forms the same dt
        double dt = t.cfl / rate;
        dt = dt * 1.0;
        if (step > 1 && dtold > 0.0) dt = vmin(dt, dtold * 1.1);
        dtold = dt;
        if (dt + time > t.endtime) dt = t.endtime - time;
        dt = dt * 0.5;
        dt = dt * 2.0;
        if (!(dt > 0.0) || !__builtin_isfinite(dt)) {
            status |= FLAG_BADDT;
            step--;
            break;
        }
        if (a.dtlog && threadIdx.x == 0 && (!GANG || g.member == 0) && step <= a.maxlog)
            a.dtlog[(int64_t)ti * a.maxlog + step - 1] = dt;
        sa.dt = dt;
        const double *src = step == 1 ? icr : str;

        // ---------------- predictor: main.f90:232-355
        for (int s = sp0 + wv; s < sp1; s += W) {
            const int c = s * kSegOut + lane;  // lane 2 is the segment's first output
            const int d = donor_of(min(c, nx - 1), nx);
            const bool ghost = c < 2 || c > nx - 3;
            double U[5], v[3];
            for (int k = 0; k < 5; k++) U[k] = ld(src + k * T, d);
            const double w0 = ld(src + 5 * T, d), w1 = ld(src + 6 * T, d), w2 = ld(src + 7 * T, d);
            const double p0 = ld(src + 8 * T, d);
            // in the segments past an edge only the cells the kept outputs read
            // count (the rest may hold an older state)
            const bool count = !GANG || (s >= s0 && s < s1) || (c >= c_lo - 4 && c < c_hi + 4);
            unsigned pf = 0;
            const double p232 = eos_state(t, U[0], U[4], w0, w1, w2, p0, pf);    // :232
            U[4] = source_E(t, U);                                               // :289
            const double p314 = eos_state(t, U[0], U[4], w0, w1, w2, p232, pf);  // :314
            velocity_of(t, U, v, pf);                                            // :318
            if (count) flg |= pf;
            W6<double> q;
            q.d = U[0];
            q.u0 = v[0];
            q.u1 = v[1];
            q.u2 = v[2];
            q.p = p314;
            q.e = floor_at(U[4] - 0.5 * U[0] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), t.seps);
            double Fx[5];
            x_faces(sa, q, sound_of(q, t.gam), ghost, Fx);
            double dF[5];
            for (int k = 0; k < 5; k++) dF[k] = lane_above(Fx[k]) - Fx[k];
            __asm__ volatile("" ::: "memory");  // keep the update's loads after the faces (register pressure)
            if (lane >= 2 && lane < 62 && !ghost && c >= w_lo && c < w_hi) {
                double Us[5], us[3];
                const bool mine = !GANG || (c >= c_lo && c < c_hi);
                for (int k = 0; k < 5; k++) {
                    if (mine) PL(oldr + k * T, c) = U[k];         // cons_old, main.f90:327-333
                    Us[k] = U[k] + flux_tmp(t, dF[k], dt);      // flux.f90:205
                }
                Us[0] = floor_at(Us[0], t.srho);                // flux.f90:225-228
                if (Us[4] < 0.0) Us[4] = 0.5 * (Us[1] * Us[1] + Us[2] * Us[2] + Us[3] * Us[3]) / Us[0] + t.seps;
                const double ps = eos_state(t, Us[0], Us[4], v[0], v[1], v[2], p314, flg);  // :353
                velocity_of(t, Us, us, flg);                                             // :355
                if (mine) {
                    for (int k = 0; k < 5; k++) PL(ssr + k * T, c) = Us[k];
                    for (int k = 0; k < 3; k++) PL(ssr + (5 + k) * T, c) = us[k];
                    PL(ssr + 8 * T, c) = ps;
                } else {
                    const int side = c < c_lo ? 0 : 1, j = c < c_lo ? c - w_lo : c - c_hi;
                    for (int k = 0; k < 5; k++) ext[side][k][j] = Us[k];
                    for (int k = 0; k < 3; k++) ext[side][5 + k][j] = us[k];
                    ext[side][8][j] = ps;
                }
            }
        }
        __syncthreads();

        // ---------------- corrector: main.f90:368-515, then :232-241 of the next step
        double rn = 0.0;
        for (int s = s0 + wv; s < s1; s += W) {
            const int c = s * kSegOut + lane;
            const int d = donor_of(min(c, nx - 1), nx);
            const bool ghost = c < 2 || c > nx - 3;
            W6<double> q;
            // the predictor results: the rows, or LDS past a shared edge
            const bool xe = GANG && (d < c_lo || d >= c_hi);
            const int xs = d < c_lo ? 0 : 1, xj = min(max(d < c_lo ? d - w_lo : d - c_hi, 0), 1);
            auto sv = [&](int k) { return xe ? ext[xs][k][xj] : PL(ssr + k * T, d); };
            q.d = sv(0);
            const double E = sv(4);
            q.u0 = sv(5);
            q.u1 = sv(6);
            q.u2 = sv(7);
            q.p = sv(8);
            q.e = floor_at(E - 0.5 * q.d * (q.u0 * q.u0 + q.u1 * q.u1 + q.u2 * q.u2), t.seps);
            double Fx[5];
            x_faces(sa, q, sound_of(q, t.gam), ghost, Fx);
            double dF[5];
            for (int k = 0; k < 5; k++) dF[k] = lane_above(Fx[k]) - Fx[k];
            __asm__ volatile("" ::: "memory");  // keep the update's loads after the faces (register pressure)
            if (lane >= 2 && lane < 62 && !ghost) {
                double U[5], v[3], v2[3];
                for (int k = 0; k < 5; k++)
                    U[k] = 0.5 * (PL(oldr + k * T, c) + PL(ssr + k * T, c) + flux_tmp(t, dF[k], dt));  // flux.f90:213
                U[0] = floor_at(U[0], t.srho);
                if (U[4] < 0.0) U[4] = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0] + t.seps;
                const double p376 = eos_state(t, U[0], U[4], q.u0, q.u1, q.u2, q.p, flg);    // :376
                velocity_of(t, U, v, flg);                                                  // :378
                U[4] = source_E(t, U);                                                      // :490
                const double p513 = eos_state(t, U[0], U[4], v[0], v[1], v[2], p376, flg);  // :513
                velocity_of(t, U, v2, flg);                                                 // :515
                for (int k = 0; k < 5; k++) st(str + k * T, c, U[k]);
                for (int k = 0; k < 3; k++) st(str + (5 + k) * T, c, v2[k]);
                st(str + 8 * T, c, p513);
                // the next step's state (:232) and Courant rate (:241)
                double E2 = U[4];
                unsigned fn = 0;
                const double p2 = eos_state(t, U[0], E2, v2[0], v2[1], v2[2], p513, fn);
                rn = vmax(rn, rate_of(t, U[0], p2, v2[0], v2[1], v2[2]));
            }
        }
        rate = uniform_d(block_max(rn, red, W));  // its barriers also order this phase's stores before the next
        if (GANG && !gang_max(g, rate, s_ok, s_r)) return;
        time = time + dt;              // main.f90:545
    }
    if (step == 0) {
        // no step taken: the state is the set-up
        for (int c = c_lo + (int)threadIdx.x; c < c_hi; c += 64 * W)
            for (int k = 0; k < 9; k++) st(str + k * T, c, PL(icr + k * T, c));
    }
    if (flg) atomicOr(sflag, flg);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (!GANG || g.member == 0) {
            a.t_out[ti] = time;
            a.steps_out[ti] = step;
        }
        atomicOr(&a.status[ti], *sflag | status);
    }
}

template <int W>
__global__ __launch_bounds__(64 * W) void k_tubes(BatchArgs a, const Tube *__restrict__ tubes, const int *__restrict__ order)
{
    __shared__ double red[W];
    __shared__ int cur;
    __shared__ unsigned sflag;
    __shared__ int s_ok;
    __shared__ double s_r;
    __shared__ double ext[2][9][2];
    // the gangs first: workgroups [0, ngang * gang_p), gang blockIdx / gang_p
    if (a.gang_p > 1 && (int)blockIdx.x < a.ngang * a.gang_p) {
        if (threadIdx.x == 0) sflag = 0;
        __syncthreads();
        GangCtx g;
        const int gi = (int)blockIdx.x / a.gang_p;
        g.ctr = a.gang_ctr + gi;
        g.rate = a.gang_rate + 3 * gi;
        g.fail = a.gang_fail;
        g.member = (int)blockIdx.x - gi * a.gang_p;
        g.p = a.gang_p;
        g.epoch = 0;
        g.exch = 0;
        const int ti = __builtin_amdgcn_readfirstlane(a.gang_tube[gi]);
        const Tube t = tubes[ti];
        run_tube<W, true>(a, t, ti, g, red, &s_ok, &s_r, &sflag, ext);
    }
    GangCtx none{};
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            cur = atomicAdd(a.queue, 1);
            sflag = 0;
        }
        __syncthreads();
        const int qi = __builtin_amdgcn_readfirstlane(cur);
        if (qi >= a.nqueue) break;  // the whole workgroup leaves together
        const int ti = __builtin_amdgcn_readfirstlane(order[qi]);
        const Tube t = tubes[ti];
        run_tube<W, false>(a, t, ti, none, red, &s_ok, &s_r, &sflag, ext);
    }
}

// Clears the queue counter and the gangs' counters, maxima and failure word
// (a kernel, not memsets, so that the launch can be captured).
__global__ void k_tubes_reset(int *queue, int ngang, unsigned *ctr, unsigned long long *rate, unsigned *fail,
                              unsigned *status, int ntube)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *queue = 0;
        if (fail) *fail = 0u;
    }
    if (i < ngang) ctr[i] = 0u;
    if (i < 3 * ngang) rate[i] = 0ull;
    if (i < ntube) status[i] = 0u;
}


}  // namespace bx

// ===================================================================== C-ABI
struct bx_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int ntube = 0;
    int64_t total = 0, interior = 0;
    std::vector<bx::Tube> htubes;
    std::vector<int> horder;
    bx::Tube *tubes = nullptr;
    int *order = nullptr;
    int *queue = nullptr;
    double *ic = nullptr, *st = nullptr, *ss = nullptr, *old = nullptr;
    double *t_out = nullptr;
    int *steps_out = nullptr;
    unsigned *status = nullptr;
    double *dtlog = nullptr;
    int maxlog = 0;
    int waves = 16, blocks_per_cu = 1, ncu = 0;
    // gangs: the ngang costliest tubes, gang_p workgroups each (gang_p < 2: off)
    int gang_p = 0, ngang = 0;
    std::vector<int> hsteps;  // steps of each tube in the last finished run (empty: none yet)
    int *gang_tube = nullptr;
    unsigned *gang_ctr = nullptr;
    unsigned long long *gang_rate = nullptr;
    unsigned *gang_fail = nullptr;
    int *qorder = nullptr;  // the queue order without the gang tubes
    int nqueue = 0;
    int gang_fallbacks = 0;
    int table_maxsteps = -1;  // maxsteps held by the device tube table
    bool graph_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool timed = false;
};

namespace {

constexpr int kTubeRowMin = 4;  // interior cells per row

bool is_pow2(double d)
{
    int e;
    return d > 0 && std::frexp(d, &e) == 0.5;
}

// eos.f90:144 with parameters.f90:55 (the H2STAT=-1 mean molecular weight)
double muc_of(const bx_params &p) { return 1.0 / (p.xabun * 0.5 + p.yabun * 0.25 + p.zabun / p.mu_z); }

// units.f90:10-17
double rgas_of()
{
    const double vel = 1.496e13 / ((3600.0 * 365.25 * 24.0) / (2.0 * 3.14159265358979323846));
    return 8.254e7 / (vel * vel);
}

bool check_tube(const bx_params &p, int64_t off, const double *cons, const double *u, std::string &why)
{
    const int nx = p.nx;
    if (p.nz >= 6) {
        why = "nz >= 6 (the row form needs a flat z direction)";
        return false;
    }
    if (p.ny < 5 || p.nz < 1 || nx < kTubeRowMin + 4) {
        why = "grid too small (ny >= 5, nz >= 1 and nx >= 8 are needed)";
        return false;
    }
    if (!(p.dx > 0 && p.dy > 0 && p.dz > 0) || !(p.gammafix > 1.0)) {
        why = "spacings must be > 0 and gammafix > 1";
        return false;
    }
    for (int k = 0; k < 6; k++)
        if (p.no_outflow[k]) {
            why = "a reflecting side (the row form is for outflow tubes)";
            return false;
        }
    if (p.precision != BX_PREC_FP64) {
        why = "precision other than FP64";
        return false;
    }
    for (int c = 1; c < nx - 1; c++)
        if (!(cons[(off + c) * 5] >= 1e-4)) {
            why = "a set-up density below 1e-4 outside the frame";
            return false;
        }
    const int gh[4] = {0, 1, nx - 2, nx - 1};
    for (int g : gh) {
        const int d = g < 2 ? 2 : nx - 3;
        for (int k = 0; k < 3; k++)
            if (std::memcmp(&u[(off + g) * 3 + k], &u[(off + d) * 3 + k], sizeof(double)) != 0) {
                why = "a ghost cell's set-up velocity differs from its donor's";
                return false;
            }
    }
    return true;
}

int ensure_table(bx_batch *b, int maxsteps)
{
    if (b->table_maxsteps == maxsteps) return 0;
    for (auto &t : b->htubes) t.maxsteps = maxsteps;
    BX_HIP(hipStreamSynchronize(b->stream));
    BX_HIP(hipMemcpy(b->tubes, b->htubes.data(), sizeof(bx::Tube) * b->ntube, hipMemcpyHostToDevice));
    b->table_maxsteps = maxsteps;
    return 0;
}

int enqueue(bx_batch *b, bool gangs = true)
{
    const bool g = gangs && b->gang_p > 1 && b->ngang > 0;
    bx::BatchArgs a;
    a.ntube = b->ntube;
    a.maxlog = b->maxlog;
    a.total = b->total;
    a.tubes = b->tubes;
    a.order = b->order;
    a.queue = b->queue;
    a.ic = b->ic;
    a.st = b->st;
    a.ss = b->ss;
    a.old = b->old;
    a.t_out = b->t_out;
    a.steps_out = b->steps_out;
    a.status = b->status;
    a.dtlog = b->dtlog;
    a.nqueue = g ? b->nqueue : b->ntube;
    a.ngang = g ? b->ngang : 0;
    a.gang_p = g ? b->gang_p : 0;
    a.gang_tube = b->gang_tube;
    a.gang_ctr = b->gang_ctr;
    a.gang_rate = b->gang_rate;
    a.gang_fail = b->gang_fail;
    const int *order = g ? b->qorder : b->order;
    const int gwg = a.ngang * a.gang_p;
    const unsigned grid = (unsigned)std::max(std::max(1, gwg), std::min(a.nqueue + gwg, b->ncu * b->blocks_per_cu));
    // the counters are reset by a kernel (not memsets), so the pair can be captured
    const int nreset = std::max(std::max(b->ntube, 3 * a.ngang), 1);
    hipLaunchKernelGGL(bx::k_tubes_reset, dim3((nreset + 255) / 256), dim3(256), 0, b->stream, b->queue, a.ngang,
                       b->gang_ctr, b->gang_rate, b->gang_fail, b->status, b->ntube);
    switch (b->waves) {
    case 4:
        hipLaunchKernelGGL(bx::k_tubes<4>, dim3(grid), dim3(256), 0, b->stream, a, b->tubes, order);
        break;
    case 8:
        hipLaunchKernelGGL(bx::k_tubes<8>, dim3(grid), dim3(512), 0, b->stream, a, b->tubes, order);
        break;
    default:
        hipLaunchKernelGGL(bx::k_tubes<16>, dim3(grid), dim3(1024), 0, b->stream, a, b->tubes, order);
        break;
    }
    BX_HIP(hipGetLastError());
    return 0;
}

void drop_graph(bx_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

int capture(bx_batch *b)
{
    drop_graph(b);
    BX_HIP(hipStreamSynchronize(b->stream));
    BX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    BX_HIP(e);
    b->graph = g;
    BX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_valid = true;
    return 0;
}

void batch_free(bx_batch *b)
{
    drop_graph(b);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    void *ptrs[] = {b->tubes, b->order, b->queue, b->ic, b->st, b->ss, b->old, b->t_out, b->steps_out, b->status,
                    b->dtlog, b->gang_tube, b->gang_ctr, b->gang_rate, b->gang_fail, b->qorder};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // namespace

extern "C" {

int bx_batch_check(int32_t ntube, const bx_params *prm, const double *cons_rows, const double *u_rows, int32_t *bad)
{
    if (bad) *bad = -1;
    if (ntube <= 0 || !prm || !cons_rows || !u_rows)
        return bx::set_error(BX_ERR_ARG, "bx_batch_check: bad arguments");
    int64_t off = 0;
    for (int i = 0; i < ntube; i++) {
        std::string why;
        if (!check_tube(prm[i], off, cons_rows, u_rows, why)) {
            if (bad) *bad = i;
            return bx::set_error(BX_ERR_UNSUPPORTED, "bx_batch: tube " + std::to_string(i) + ": " + why);
        }
        off += prm[i].nx;
    }
    return 0;
}

int bx_batch_create(int32_t ntube, const bx_params *prm, const double *endtime, const double *cons_rows,
                    const double *u_rows, const double *p_rows, int32_t device, bx_batch **out)
{
    if (!out || !endtime || !p_rows) return bx::set_error(BX_ERR_ARG, "bx_batch_create: null argument");
    *out = nullptr;
    if (int rc = bx_batch_check(ntube, prm, cons_rows, u_rows, nullptr)) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return bx::set_error(BX_ERR_ARG, "bx_batch_create: no such device");
    bx_batch *b = new (std::nothrow) bx_batch();
    if (!b) return bx::set_error(BX_ERR_ARG, "bx_batch_create: out of host memory");
    b->device = device;
    b->ntube = ntube;
    const double rgas = rgas_of();
    std::vector<double> cost(ntube);
    int64_t off = 0;
    for (int i = 0; i < ntube; i++) {
        const bx_params &p = prm[i];
        bx::Tube t{};
        t.off = off;
        t.nx = p.nx;
        t.dx = p.dx;
        t.dy = p.dy;
        t.dz = p.dz;
        t.inv_dx = 1.0 / p.dx;
        t.dx_pow2 = is_pow2(p.dx);
        t.ayz = p.dy * p.dz;         // flux.f90:28
        t.vol = p.dx * p.dy * p.dz;  // flux.f90:31
        t.inv_vol = 1.0 / t.vol;
        t.vol_pow2 = is_pow2(t.vol);
        t.gam = p.gammafix;
        t.gm1 = p.gammafix - 1.0;
        t.cfl = p.cfl;
        t.endtime = endtime[i];
        t.srho = p.small_rho;
        t.seps = p.small_eps;
        t.vlimit = p.vlimit;
        t.muc = muc_of(p);
        t.rgas = rgas;
        t.maxsteps = 1000000;
        b->htubes.push_back(t);
        // cost estimate for the queue order: cells x steps, steps from the set-up's signal speed
        double smax = 0.0;
        for (int c = 2; c < p.nx - 2; c++) {
            const double *U = &cons_rows[(off + c) * 5];
            const double cs = std::sqrt(p.gammafix * std::max(p_rows[off + c], 0.0) / U[0]);
            smax = std::max(smax, std::fabs(U[1] / U[0]) + cs);
        }
        cost[i] = (double)(p.nx - 4) * std::max(endtime[i], 0.0) * smax / (p.cfl * p.dx);
        b->interior += (int64_t)(p.nx - 4) * (p.ny - 4) * p.nz;
        off += p.nx;
    }
    b->total = off;
    b->table_maxsteps = 1000000;
    b->horder.resize(ntube);
    std::iota(b->horder.begin(), b->horder.end(), 0);
    std::stable_sort(b->horder.begin(), b->horder.end(), [&](int x, int y) { return cost[x] > cost[y]; });

    auto fail = [&](int rc) {
        batch_free(b);
        return rc;
    };
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return fail(bx::hip_check(e, "hipSetDevice"));
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return fail(bx::hip_check(e, "hipGetDeviceProperties"));
    b->ncu = prop.multiProcessorCount;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(bx::hip_check(e, "hipStreamCreate"));
    const int64_t T = b->total;
    struct {
        void **p;
        size_t bytes;
    } allocs[] = {
        {(void **)&b->tubes, sizeof(bx::Tube) * ntube}, {(void **)&b->order, sizeof(int) * ntube},
        {(void **)&b->queue, 256},                       {(void **)&b->ic, sizeof(double) * 9 * T},
        {(void **)&b->st, sizeof(double) * 9 * T},       {(void **)&b->ss, sizeof(double) * 9 * T},
        {(void **)&b->old, sizeof(double) * 5 * T},      {(void **)&b->t_out, sizeof(double) * ntube},
        {(void **)&b->steps_out, sizeof(int) * ntube},   {(void **)&b->status, sizeof(unsigned) * ntube},
    };
    for (auto &al : allocs)
        if ((e = hipMalloc(al.p, al.bytes)) != hipSuccess) return fail(bx::hip_check(e, "hipMalloc (batch)"));
    for (auto &ev : b->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return fail(bx::hip_check(e, "hipEventCreate"));
    // set-up rows as planes
    std::vector<double> h((size_t)9 * T);
    for (int64_t c = 0; c < T; c++) {
        for (int k = 0; k < 5; k++) h[k * T + c] = cons_rows[c * 5 + k];
        for (int k = 0; k < 3; k++) h[(5 + k) * T + c] = u_rows[c * 3 + k];
        h[8 * T + c] = p_rows[c];
    }
    if ((e = hipMemcpy(b->ic, h.data(), sizeof(double) * 9 * T, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->st, h.data(), sizeof(double) * 9 * T, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->order, b->horder.data(), sizeof(int) * ntube, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->tubes, b->htubes.data(), sizeof(bx::Tube) * ntube, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(bx::hip_check(e, "hipMemcpy (batch set-up)"));
    *out = b;
    return 0;
}

int bx_batch_destroy(bx_batch *b)
{
    if (!b) return 0;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    return 0;
}

int bx_batch_cells(const bx_batch *b, int64_t *row_cells, int64_t *interior_cells)
{
    if (!b) return bx::set_error(BX_ERR_ARG, "bx_batch_cells: null batch");
    if (row_cells) *row_cells = b->total;
    if (interior_cells) *interior_cells = b->interior;
    return 0;
}

int bx_batch_order(const bx_batch *b, int32_t *order)
{
    if (!b || !order) return bx::set_error(BX_ERR_ARG, "bx_batch_order: null argument");
    std::copy(b->horder.begin(), b->horder.end(), order);
    return 0;
}

int bx_batch_tune(bx_batch *b, int32_t waves_per_block, int32_t blocks_per_cu)
{
    if (!b) return bx::set_error(BX_ERR_ARG, "bx_batch_tune: null batch");
    if (waves_per_block != 4 && waves_per_block != 8 && waves_per_block != 16)
        return bx::set_error(BX_ERR_ARG, "bx_batch_tune: waves_per_block must be 4, 8 or 16");
    if (blocks_per_cu < 1 || blocks_per_cu > 8)
        return bx::set_error(BX_ERR_ARG, "bx_batch_tune: blocks_per_cu must be in [1,8]");
    if (b->waves != waves_per_block || b->blocks_per_cu != blocks_per_cu) b->graph_valid = false;
    b->waves = waves_per_block;
    b->blocks_per_cu = blocks_per_cu;
    return 0;
}

int bx_batch_log_dt(bx_batch *b, int32_t maxlog)
{
    if (!b || maxlog < 0) return bx::set_error(BX_ERR_ARG, "bx_batch_log_dt: bad arguments");
    BX_HIP(hipSetDevice(b->device));
    BX_HIP(hipStreamSynchronize(b->stream));
    if (b->dtlog) (void)hipFree(b->dtlog);
    b->dtlog = nullptr;
    b->maxlog = 0;
    b->graph_valid = false;
    if (maxlog > 0) {
        BX_HIP(hipMalloc(&b->dtlog, sizeof(double) * (size_t)maxlog * b->ntube));
        b->maxlog = maxlog;
    }
    return 0;
}

int bx_batch_run(bx_batch *b, int32_t maxsteps, int32_t use_graph)
{
    if (!b || maxsteps < 0) return bx::set_error(BX_ERR_ARG, "bx_batch_run: bad arguments");
    BX_HIP(hipSetDevice(b->device));
    if (int rc = ensure_table(b, maxsteps)) return rc;
    if (use_graph && !b->graph_valid)
        if (int rc = capture(b)) return rc;
    BX_HIP(hipEventRecord(b->ev[0], b->stream));
    if (use_graph) {
        BX_HIP(hipGraphLaunch(b->exec, b->stream));
    } else if (int rc = enqueue(b)) {
        return rc;
    }
    BX_HIP(hipEventRecord(b->ev[1], b->stream));
    b->timed = true;
    return 0;
}

int bx_batch_sync(bx_batch *b, double *ms)
{
    if (!b) return bx::set_error(BX_ERR_ARG, "bx_batch_sync: null batch");
    BX_HIP(hipSetDevice(b->device));
    BX_HIP(hipStreamSynchronize(b->stream));
    if (b->gang_p > 1 && b->ngang > 0 && b->timed) {
        unsigned fail = 0;
        BX_HIP(hipMemcpy(&fail, b->gang_fail, sizeof(unsigned), hipMemcpyDeviceToHost));
        if (fail) {
            // a gang's members were not resident together: the batch again, without gangs
            b->gang_fallbacks++;
            BX_HIP(hipEventRecord(b->ev[0], b->stream));
            if (int rc = enqueue(b, false)) return rc;
            BX_HIP(hipEventRecord(b->ev[1], b->stream));
            BX_HIP(hipStreamSynchronize(b->stream));
        }
    }
    if (b->timed) {  // the steps of each tube, the measured costs for bx_batch_gang
        b->hsteps.resize(b->ntube);
        BX_HIP(hipMemcpy(b->hsteps.data(), b->steps_out, sizeof(int) * b->ntube, hipMemcpyDeviceToHost));
    }
    if (ms) {
        float f = 0.0f;
        if (b->timed) BX_HIP(hipEventElapsedTime(&f, b->ev[0], b->ev[1]));
        *ms = f;
    }
    return 0;
}

int bx_batch_gang(bx_batch *b, int32_t members, int32_t ntubes)
{
    if (!b) return bx::set_error(BX_ERR_ARG, "bx_batch_gang: null batch");
    if (members < 0 || ntubes < 0 || members > 64)
        return bx::set_error(BX_ERR_ARG, "bx_batch_gang: members in [0,64] and ntubes >= 0");
    BX_HIP(hipSetDevice(b->device));
    BX_HIP(hipStreamSynchronize(b->stream));
    b->graph_valid = false;
    if (members < 2 || ntubes == 0) {
        b->gang_p = 0;
        b->ngang = 0;
        return 0;
    }
    if (ntubes > b->ntube) return bx::set_error(BX_ERR_ARG, "bx_batch_gang: more gangs than tubes");
    if ((int64_t)members * ntubes > b->ncu)
        return bx::set_error(BX_ERR_ARG, "bx_batch_gang: members x ntubes must not exceed the CUs "
                                         "(the members of a gang run together)");
    // the costliest tubes: measured (steps x interior cells of the last run) when a
    // run has finished, else the set-up estimate's order
    std::vector<int> ord(b->horder);
    if ((int)b->hsteps.size() == b->ntube) {
        std::vector<double> cost(b->ntube);
        for (int i = 0; i < b->ntube; i++) cost[i] = (double)b->hsteps[i] * (double)(b->htubes[i].nx - 4);
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    }
    std::vector<int> gt(ord.begin(), ord.begin() + ntubes), q(ord.begin() + ntubes, ord.end());
    for (void *p : {(void *)b->gang_tube, (void *)b->gang_ctr, (void *)b->gang_rate, (void *)b->gang_fail,
                    (void *)b->qorder})
        if (p) (void)hipFree(p);
    b->gang_tube = b->qorder = nullptr;
    b->gang_ctr = b->gang_fail = nullptr;
    b->gang_rate = nullptr;
    b->gang_p = b->ngang = 0;
    BX_HIP(hipMalloc((void **)&b->gang_tube, sizeof(int) * ntubes));
    BX_HIP(hipMalloc((void **)&b->gang_ctr, sizeof(unsigned) * ntubes));
    BX_HIP(hipMalloc((void **)&b->gang_rate, sizeof(unsigned long long) * 3 * ntubes));
    BX_HIP(hipMalloc((void **)&b->gang_fail, sizeof(unsigned)));
    BX_HIP(hipMalloc((void **)&b->qorder, sizeof(int) * std::max<size_t>(q.size(), 1)));
    BX_HIP(hipMemcpy(b->gang_tube, gt.data(), sizeof(int) * ntubes, hipMemcpyHostToDevice));
    if (!q.empty()) BX_HIP(hipMemcpy(b->qorder, q.data(), sizeof(int) * q.size(), hipMemcpyHostToDevice));
    b->gang_p = members;
    b->ngang = ntubes;
    b->nqueue = (int)q.size();
    return 0;
}

int bx_batch_gang_fallbacks(const bx_batch *b, int32_t *n)
{
    if (!b || !n) return bx::set_error(BX_ERR_ARG, "bx_batch_gang_fallbacks: null argument");
    *n = b->gang_fallbacks;
    return 0;
}

int bx_batch_result(bx_batch *b, double *cons_rows, double *u_rows, double *p_rows, double *time, int32_t *steps,
                    uint32_t *status)
{
    if (!b) return bx::set_error(BX_ERR_ARG, "bx_batch_result: null batch");
    BX_HIP(hipSetDevice(b->device));
    BX_HIP(hipStreamSynchronize(b->stream));
    const int64_t T = b->total;
    std::vector<int> hsteps(b->ntube);
    BX_HIP(hipMemcpy(hsteps.data(), b->steps_out, sizeof(int) * b->ntube, hipMemcpyDeviceToHost));
    if (cons_rows || u_rows || p_rows) {
        std::vector<double> h((size_t)9 * T);
        BX_HIP(hipMemcpy(h.data(), b->st, sizeof(double) * 9 * T, hipMemcpyDeviceToHost));
        // after a step, ghost cells hold their donors' values (the reference's
        // last ghost fill); a tube that took no step returns its set-up row
        for (int i = 0; i < b->ntube; i++) {
            const bx::Tube &t = b->htubes[i];
            const bool stepped = hsteps[i] > 0;
            for (int c = 0; c < t.nx; c++) {
                const int dc = stepped ? (c < 2 ? 2 : (c > t.nx - 3 ? t.nx - 3 : c)) : c;
                const int64_t g = t.off + c, d = t.off + dc;
                if (cons_rows)
                    for (int k = 0; k < 5; k++) cons_rows[g * 5 + k] = h[k * T + d];
                if (u_rows)
                    for (int k = 0; k < 3; k++) u_rows[g * 3 + k] = h[(5 + k) * T + d];
                if (p_rows) p_rows[g] = h[8 * T + d];
            }
        }
    }
    if (time) BX_HIP(hipMemcpy(time, b->t_out, sizeof(double) * b->ntube, hipMemcpyDeviceToHost));
    if (steps) std::copy(hsteps.begin(), hsteps.end(), steps);
    if (status) BX_HIP(hipMemcpy(status, b->status, sizeof(unsigned) * b->ntube, hipMemcpyDeviceToHost));
    return 0;
}

int bx_batch_dts(bx_batch *b, double *dts)
{
    if (!b || !dts) return bx::set_error(BX_ERR_ARG, "bx_batch_dts: null argument");
    if (!b->dtlog) return bx::set_error(BX_ERR_STATE, "bx_batch_dts: dt logging is off (bx_batch_log_dt)");
    BX_HIP(hipSetDevice(b->device));
    BX_HIP(hipStreamSynchronize(b->stream));
    BX_HIP(hipMemcpy(dts, b->dtlog, sizeof(double) * (size_t)b->maxlog * b->ntube, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
