// bx_boxes.hip -- many independent grids of uneven size (3-D boxes or flat-z
// slabs) in one persistent launch.
//
// Each grid is what the reference runs in one process: main.f90:211-555 on
// an nx x ny x nz grid with its own walls, spacings, gamma, CFL number and
// endtime, with the fixed-gamma EOS (H2STAT=-1), the plain flux (no
// flux_ang) and the uniform force of a zero potential.  One workgroup steps
// one grid as the single-grid step (bx_step, fused mode) steps it:
//   - the exact call sequence, subroutine by subroutine, with the per-cell
//     bodies of the one-kernel-per-subroutine path (bx_cells.h), for the first
//     step and for any step from a state that is not regular;
//   - on a regular state whose ghost donors are all interior cells, fused
//     steps: the wave bodies of the fused stage kernels (bx_waves.h) over the
//     grid, the workgroup's waves taking the virtual waves of those kernels'
//     grids; a step that meets an irregular branch is redone with the exact
//     sequence from U^n.
// Both give the single-grid step's bits, and so the reference's.  A pass
// that reads neighbours (ghost fills, faces, stage updates) is separated from
// its producer by a workgroup barrier; pointwise passes run back to back on
// the same cell-to-lane mapping.  The Courant maximum is a workgroup
// reduction and the dt logic of main.f90:266-281 runs in every lane on that
// uniform value.
//
// Work decomposition: grids are taken from a queue ordered by estimated cost
// (cells x steps, largest first), so the long runs start first and the short
// ones fill in behind them.  A grid's planes (240 B per cell) are its own
// contiguous region; a grid of up to ~16k cells stays in its XCD's L2.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "bx_cells.h"
#include "bx_waves.h"

namespace bx {

enum : unsigned int {
    BOX_BADDT = 4u,  // a non-finite or non-positive dt: the grid was stopped before that step
};

// Per-grid region, planes of n doubles: cons 5 | u 3 | p | adindx | cons_old 5
// | face fluxes 15, then the ghost-donor snapshot (5 planes of nscratch).  The
// first 9 planes are the result (cons, u, p).
constexpr int kRegionPlanes = 30;
constexpr int kResultPlanes = 9;
// Set-up, planes of n doubles: cons 5 | u 3 | p
constexpr int kSetupPlanes = 9;

struct Box {
    int64_t n, ic_off, reg_off, g_off, x_off;
    int64_t ng, ns;
    int nx, ny, nz, flatz;
    int maxsteps;
    double dx, dy, dz, gam, cfl, endtime, srho, seps, vlimit, muc, rgas;
    double g0, g1, g2;
    double ayz, axz, axy, vol;
    // fused steps (every ghost donor an interior cell): the stage constants
    // as the single-grid step forms them (base_args in bx_fused.hip)
    int fusable, zlo, zhi, ly, lz, vol_pow2;
    int ds_pow2[3];
    double inv_ds[3], inv_vol, gm1;
};

struct BoxArgs {
    int nbox;
    int maxlog;
    const Box *boxes;
    const int *order;
    int *queue;
    const double *ic;
    double *reg;
    const int32_t *gdst, *gsrc, *gsign, *gxsrc;
    double *t_out;
    int *steps_out;
    unsigned *status;
    double *dtlog;
};

// The cell index of a grid-stride loop, hidden from loop strength reduction:
// without this the compiler keeps one 64-bit induction pointer per plane
// access in flight (some sixty planes per step) and spills.
__device__ __forceinline__ int64_t opaque(int64_t v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// A uniform value made opaque at the start of each pass, so that the plane
// addresses derived from it are formed per pass instead of being hoisted out
// of the step loop (where they would outnumber the scalar registers).
__device__ __forceinline__ int64_t opaque_uniform(int64_t v)
{
    asm volatile("" : "+v"(v));
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// A value every lane of the wave holds (read from LDS after a barrier), marked
// uniform for the compiler so that the step loop's control flow and the stage
// arguments stay in scalar registers.
__device__ __forceinline__ double uniform_double(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// one set_ghost_cells call (utils.f90:305-480) on the grid's cons planes,
// entered and left with every lane's writes visible to the workgroup
__device__ __forceinline__ void box_ghosts(const Box &b, double *cons, double *scratch, const int32_t *dst,
                                           const int32_t *src, const int32_t *sign, const int32_t *xsrc, int tid,
                                           int nt)
{
    if (b.ns > 0) {
#pragma unroll 1
        for (int64_t k0 = tid; k0 < b.ns; k0 += nt) ghost_snapshot_one(b.n, b.ns, opaque(k0), xsrc, cons, scratch);
        __syncthreads();
    }
#pragma unroll 1
    for (int64_t k0 = tid; k0 < b.ng; k0 += nt) ghost_fill_one(b.n, opaque(k0), dst, src, sign, cons, scratch, b.ns);
    __syncthreads();
}

// every cell c of the grid, lane tid of nt
// (the pass's own n shadows the grid's)
#define BOX_CELLS(c)                                                                        \
    _Pragma("unroll 1") for (int64_t n = opaque_uniform(b.n), c##0 = tid; c##0 < n; c##0 += nt) \
        for (int64_t c = opaque(c##0), c##1 = 0; c##1 < 1; c##1++)

template <int W>
__global__ __launch_bounds__(64 * W) void k_boxes(BoxArgs a)
{
    __shared__ double red[W];
    __shared__ int cur;
    __shared__ unsigned int sflags;       // irregular branches met by a fused step or the regularity check
    __shared__ unsigned long long srate;  // Courant rate (IEEE bits) from the corrector or the check
    const int tid = threadIdx.x, nt = 64 * W;
    const int lane = tid & 63, wv = wave_id();  // wave-uniform for the compiler
    for (;;) {
        __syncthreads();
        if (tid == 0) cur = atomicAdd(a.queue, 1);
        __syncthreads();
        const int qi = __builtin_amdgcn_readfirstlane(cur);
        if (qi >= a.nbox) break;  // the whole workgroup leaves together
        const int bi = __builtin_amdgcn_readfirstlane(a.order[qi]);
        const Box b = a.boxes[bi];
        const int64_t n = b.n;
        double *cons = a.reg + b.reg_off;
        double *u = cons + 5 * n;
        double *p = cons + 8 * n;
        double *ad = cons + 9 * n;
        double *old = cons + 10 * n;
        double *F = cons + 15 * n;
        double *scratch = cons + (int64_t)kRegionPlanes * n;
        const double *ic = a.ic + b.ic_off;
        const int32_t *gd = a.gdst + b.g_off, *gs = a.gsrc + b.g_off, *gg = a.gsign + b.g_off;
        const int32_t *gx = a.gxsrc + b.x_off;
        const bool flatz = b.flatz != 0;

        // the set-up: cons, u, p as uploaded (adindx is set by the first state call)
        BOX_CELLS(c)
        {
            for (int f = 0; f < 5; f++) cons[f * n + c] = ic[f * n + c];
            for (int f = 0; f < 3; f++) u[f * n + c] = ic[(5 + f) * n + c];
            p[c] = ic[8 * n + c];
        }
        __syncthreads();

        double time = 0.0, dtold = 0.0;
        int step = 0;
        unsigned st = 0;
        // Fused steps as bx_step(BX_STEP_FUSED) takes them: on a regular state
        // (after an exact step, ghosts filled, rho >= 1e-4, u = m/rho), two
        // stages of the fused kernels' wave bodies over this grid; a step that
        // meets the low-density branch or the clamp is redone with the exact
        // sequence from U^n.  Buffers of the fused step reuse cons_old and the
        // face-flux planes: U* = old, p* = F[0], U^{n+1} = F[1..5], y-term = F[6..10].
        bool regular = false, aux_stale = false;
        double rate_f = 0.0;
        double *us = old, *ps = F, *un = F + n, *ey = F + 6 * n;
        while (step < b.maxsteps && time < b.endtime) {  // main.f90:212
            double dt_step = 0.0;
            bool fused_done = false;
            if (b.fusable && regular) {
                // main.f90:266-281 on the rate of the current state
                double dt = b.cfl / rate_f;
                dt = dt * 1.0;
                if (step + 1 > 1 && dtold > 0.0) {
                    const double cap = dtold * 1.1;
                    dt = cap < dt ? cap : dt;
                }
                const double dtcap = dt;
                if (dt + time > b.endtime) dt = b.endtime - time;
                if (!(dt > 0.0) || !(dt <= 1.7976931348623157e308)) {
                    st |= BOX_BADDT;
                    break;
                }
                dt = dt * 0.5;
                dt = dt * 2.0;
                StageArgs sa = {};
                sa.n = n;
                sa.nx = b.nx;
                sa.ny = b.ny;
                sa.nz = b.nz;
                sa.flatz = b.flatz;
                sa.zlo = b.zlo;
                sa.zhi = b.zhi;
                sa.lz = b.lz;
                sa.dt = dt;
                sa.ds[0] = b.dx;
                sa.ds[1] = b.dy;
                sa.ds[2] = b.dz;
                for (int d = 0; d < 3; d++) {
                    sa.inv_ds[d] = b.inv_ds[d];
                    sa.ds_pow2[d] = b.ds_pow2[d];
                }
                sa.ayz = b.ayz;
                sa.axz = b.axz;
                sa.axy = b.axy;
                sa.vol = b.vol;
                sa.inv_vol = b.inv_vol;
                sa.vol_pow2 = b.vol_pow2;
                sa.gam = b.gam;
                sa.gm1 = b.gm1;
                sa.seps = b.seps;
                sa.srho = b.srho;
                sa.vlimit = b.vlimit;
                sa.flags = &sflags;
                if (tid == 0) {
                    sflags = 0u;
                    srate = 0ull;
                }
                __syncthreads();
                const int nzi = b.zhi - b.zlo;
                const unsigned ygx = (unsigned)((b.nx - 4 + 63) / 64), ygy = (unsigned)((nzi + 3) / 4);
                const unsigned ygz = (unsigned)((b.ny - 4 + b.ly - 1) / b.ly);
                const unsigned xgx = (unsigned)((((int64_t)(b.ny - 4) * b.nx + kSegOut - 1) / kSegOut + 3) / 4);
                const unsigned xgz = (unsigned)((nzi + b.lz - 1) / b.lz);
                for (int stage = 1; stage <= 2; stage++) {
                    if (stage == 1) {
                        sa.uin = cons;
                        sa.uout = us;
                        sa.pout = ps;
                    } else {
                        sa.uin = us;
                        sa.pin = ps;
                        sa.uold = cons;
                        sa.uout = un;
                        sa.pout = nullptr;
                        sa.rate = &srate;
                    }
                    // the y-direction waves, then the x/z waves (each wave of the
                    // workgroup takes virtual waves of the stage kernels' grids)
                    for (unsigned v = (unsigned)wv; v < ygx * ygy * ygz * 4u; v += (unsigned)W) {
                        const unsigned blk = v >> 2;
                        if (stage == 1) ydir_wave<1, double>(sa, ey, b.ly, blk % ygx, (blk / ygx) % ygy, blk / (ygx * ygy), (int)(v & 3u));
                        else ydir_wave<2, double>(sa, ey, b.ly, blk % ygx, (blk / ygx) % ygy, blk / (ygx * ygy), (int)(v & 3u));
                    }
                    __syncthreads();
                    for (unsigned v = (unsigned)wv; v < xgx * xgz * 4u; v += (unsigned)W) {
                        const unsigned blk = v >> 2;
                        if (stage == 1) xz_wave<1, double, false>(sa, ey, blk % xgx, blk / xgx, (int)(v & 3u));
                        else xz_wave<2, double, false>(sa, ey, blk % xgx, blk / xgx, (int)(v & 3u));
                    }
                    __syncthreads();
                    for (int64_t k0 = tid; k0 < b.ng; k0 += nt)
                        ghost6_one<double>(n, opaque(k0), gd, gs, gg, stage == 1 ? us : un, stage == 1 ? ps : nullptr);
                    __syncthreads();
                }
                if ((__builtin_amdgcn_readfirstlane(sflags) & FLAG_IRREGULAR) == 0u) {
                    BOX_CELLS(c)
                    for (int f = 0; f < 5; f++) cons[f * n + c] = un[f * n + c];
                    __syncthreads();
                    step += 1;
                    dtold = dtcap;
                    rate_f = uniform_double(__longlong_as_double((long long)srate));
                    aux_stale = true;
                    fused_done = true;
                    dt_step = dt;
                } else {
                    regular = false;  // redo this step with the exact sequence from U^n
                }
            }
            if (!fused_done) {
            if (aux_stale) {  // u, p, adindx of the last fused step's state (main.f90:513-515)
                BOX_CELLS(c) materialize_cell(n, c, cons, u, p, ad, b.gam, b.seps);
                __syncthreads();
                aux_stale = false;
            }
            // state (main.f90:232) and the Courant maximum (main.f90:241)
            double r = 0.0;
            BOX_CELLS(c)
            {
                state_cell(n, c, cons, u, p, ad, b.gam, b.seps, b.muc, b.rgas);
                r = vmax(r, courant_cell(n, c, cons, u, p, ad, b.dx, b.dy, b.dz));
            }
            for (int off = 32; off > 0; off >>= 1) r = vmax(r, __shfl_xor(r, off, 64));
            if (lane == 0) red[wv] = r;
            __syncthreads();
            double rate = red[0];
            for (int k = 1; k < W; k++) rate = vmax(rate, red[k]);
            rate = uniform_double(rate);
            // main.f90:266-281, the operations of the host driver
            double dt = b.cfl / rate;
            dt = dt * 1.0;
            if (step + 1 > 1 && dtold > 0.0) {
                const double cap = dtold * 1.1;
                dt = cap < dt ? cap : dt;
            }
            const double dtcap = dt;
            if (dt + time > b.endtime) dt = b.endtime - time;
            if (!(dt > 0.0) || !(dt <= 1.7976931348623157e308)) {  // NaN, zero or infinite: stop here
                st |= BOX_BADDT;
                break;
            }
            step += 1;
            dtold = dtcap;
            dt = dt * 0.5;

            // source(0) (main.f90:289) and its ghost fill (source.f90:199), then main.f90:312
            BOX_CELLS(c) source_cell(n, c, b.nx, b.ny, b.nz, cons, u, dt, b.g0, b.g1, b.g2, b.seps, 0, b.dx, b.dy, nullptr);
            __syncthreads();
            box_ghosts(b, cons, scratch, gd, gs, gg, gx, tid, nt);
            if (b.ns > 0) box_ghosts(b, cons, scratch, gd, gs, gg, gx, tid, nt);
            // state, velocity and copy_old (main.f90:314-333)
            BOX_CELLS(c)
            {
                state_cell(n, c, cons, u, p, ad, b.gam, b.seps, b.muc, b.rgas);
                velocity_cell(n, c, cons, u, b.vlimit, b.seps);
                for (int f = 0; f < 5; f++) old[f * n + c] = cons[f * n + c];
            }
            dt = dt * 2.0;
            // two flux stages (main.f90:346-374), each followed by ghosts, state and velocity
#pragma unroll 1
            for (int avg = 0; avg < 2; avg++) {
                __syncthreads();
                BOX_CELLS(c) face_hll_cell(n, c, b.nx, b.ny, b.nz, flatz, cons, u, p, ad, F, dt, b.dx, b.dy, b.dz, b.seps);
                __syncthreads();
                BOX_CELLS(c)
                face_update_cell(n, c, b.nx, b.ny, b.nz, flatz, F, cons, old, avg, dt, b.ayz, b.axz, b.axy, b.vol,
                                 b.srho, b.seps);
                __syncthreads();
                box_ghosts(b, cons, scratch, gd, gs, gg, gx, tid, nt);
                if (b.ns > 0) box_ghosts(b, cons, scratch, gd, gs, gg, gx, tid, nt);
                BOX_CELLS(c)
                {
                    state_cell(n, c, cons, u, p, ad, b.gam, b.seps, b.muc, b.rgas);
                    velocity_cell(n, c, cons, u, b.vlimit, b.seps);
                }
            }
            dt = 0.5 * dt;
            // source(1) (main.f90:490), ghosts (main.f90:511), state and velocity
            BOX_CELLS(c) source_cell(n, c, b.nx, b.ny, b.nz, cons, u, dt, b.g0, b.g1, b.g2, b.seps, 0, b.dx, b.dy, nullptr);
            __syncthreads();
            box_ghosts(b, cons, scratch, gd, gs, gg, gx, tid, nt);
            if (b.ns > 0) box_ghosts(b, cons, scratch, gd, gs, gg, gx, tid, nt);
            BOX_CELLS(c)
            {
                state_cell(n, c, cons, u, p, ad, b.gam, b.seps, b.muc, b.rgas);
                velocity_cell(n, c, cons, u, b.vlimit, b.seps);
            }
            dt = dt * 2.0;
            dt_step = dt;
            regular = false;
            if (b.fusable) {
                // regularity and Courant rate of the new state (k_regular of the single-grid step)
                if (tid == 0) {
                    sflags = 0u;
                    srate = 0ull;
                }
                __syncthreads();
                double rr = 0.0;
                unsigned flg = 0;
                BOX_CELLS(c) rr = vmax(rr, regular_cell(n, c, cons, u, b.gam, b.seps, b.dx, b.dy, b.dz, flg));
                if (flg) atomicOr(&sflags, flg);
                wave_max_to(rr, &srate);
                __syncthreads();
                regular = __builtin_amdgcn_readfirstlane(sflags) == 0u;
                rate_f = uniform_double(__longlong_as_double((long long)srate));
            }
            }
            time = time + dt_step;  // main.f90:545
            if (tid == 0 && step - 1 < a.maxlog) a.dtlog[(int64_t)bi * a.maxlog + step - 1] = dt_step;
        }
        if (aux_stale) {
            BOX_CELLS(c) materialize_cell(n, c, cons, u, p, ad, b.gam, b.seps);
            __syncthreads();
        }
        if (tid == 0) {
            a.t_out[bi] = time;
            a.steps_out[bi] = step;
            a.status[bi] = st;
        }
    }
}

#undef BOX_CELLS

__global__ void k_boxes_reset(int *queue) { *queue = 0; }

}  // namespace bx

// ===================================================================== C-ABI
struct bx_boxes {
    int device = 0;
    hipStream_t stream = nullptr;
    int nbox = 0;
    int64_t cells = 0, interior = 0, reg_total = 0, ic_total = 0;
    std::vector<bx::Box> hbox;
    std::vector<int> horder;
    bx::Box *boxes = nullptr;
    int *order = nullptr;
    int *queue = nullptr;
    double *ic = nullptr, *reg = nullptr;
    int32_t *gdst = nullptr, *gsrc = nullptr, *gsign = nullptr, *gxsrc = nullptr;
    double *t_out = nullptr;
    int *steps_out = nullptr;
    unsigned *status = nullptr;
    double *dtlog = nullptr;
    int maxlog = 0;
    int waves = 8, blocks_per_cu = 1, ncu = 0;
    int table_maxsteps = -1;
    bool graph_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool timed = false;
};

namespace {

using bx::set_error;

// units.f90:10-17
double rgas_units()
{
    const double pi = 3.1415926535897931;
    const double time = (3600.0 * 365.25 * 24.0) / (2.0 * pi);
    const double vel = 1.496e13 / time;
    return 8.254e7 / ((vel * vel) / 1.0);
}

// d a power of two (then x/d == x*(1/d) exactly)
bool is_pow2(double d)
{
    int e;
    return d > 0 && std::frexp(d, &e) == 0.5;
}

bool check_box(const bx_params &p, std::string &why)
{
    if (p.nx < 5 || p.ny < 5 || p.nz < 1) {
        why = "need nx,ny >= 5 and nz >= 1 (two ghost layers per side)";
        return false;
    }
    if ((int64_t)p.nx * p.ny * p.nz > (int64_t)INT32_MAX) {
        why = "grid exceeds 2^31 cells";
        return false;
    }
    if (!(p.dx > 0 && p.dy > 0 && p.dz > 0) || !(p.gammafix > 1.0)) {
        why = "spacings must be > 0 and gammafix > 1";
        return false;
    }
    if (p.precision != BX_PREC_FP64) {
        why = "precision other than FP64";
        return false;
    }
    if (p.fluxangmom) {
        why = "fluxangmom (flux_ang runs on single grids)";
        return false;
    }
    return true;
}

int ensure_table(bx_boxes *b, int maxsteps)
{
    if (b->table_maxsteps == maxsteps) return 0;
    for (auto &x : b->hbox) x.maxsteps = maxsteps;
    BX_HIP(hipStreamSynchronize(b->stream));
    BX_HIP(hipMemcpy(b->boxes, b->hbox.data(), sizeof(bx::Box) * b->nbox, hipMemcpyHostToDevice));
    b->table_maxsteps = maxsteps;
    return 0;
}

int enqueue(bx_boxes *b)
{
    bx::BoxArgs a;
    a.nbox = b->nbox;
    a.maxlog = b->maxlog;
    a.boxes = b->boxes;
    a.order = b->order;
    a.queue = b->queue;
    a.ic = b->ic;
    a.reg = b->reg;
    a.gdst = b->gdst;
    a.gsrc = b->gsrc;
    a.gsign = b->gsign;
    a.gxsrc = b->gxsrc;
    a.t_out = b->t_out;
    a.steps_out = b->steps_out;
    a.status = b->status;
    a.dtlog = b->dtlog;
    const unsigned grid = (unsigned)std::max(1, std::min(b->nbox, b->ncu * b->blocks_per_cu));
    // the queue counter is reset by a kernel (not a memset), so the pair can be captured
    hipLaunchKernelGGL(bx::k_boxes_reset, dim3(1), dim3(1), 0, b->stream, b->queue);
    if (b->waves == 4) hipLaunchKernelGGL(bx::k_boxes<4>, dim3(grid), dim3(256), 0, b->stream, a);
    else if (b->waves == 16) hipLaunchKernelGGL(bx::k_boxes<16>, dim3(grid), dim3(1024), 0, b->stream, a);
    else hipLaunchKernelGGL(bx::k_boxes<8>, dim3(grid), dim3(512), 0, b->stream, a);
    BX_HIP(hipGetLastError());
    return 0;
}

void drop_graph(bx_boxes *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

int capture(bx_boxes *b)
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

void boxes_free(bx_boxes *b)
{
    drop_graph(b);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    void *ptrs[] = {b->boxes, b->order, b->queue, b->ic, b->reg, b->gdst, b->gsrc, b->gsign, b->gxsrc,
                    b->t_out, b->steps_out, b->status, b->dtlog};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // namespace

extern "C" {

int bx_boxes_create(int32_t nbox, const bx_params *prm, const double *endtime, const double *cons, const double *u,
                    const double *p, int32_t device, bx_boxes **out)
{
    if (!out) return set_error(BX_ERR_ARG, "bx_boxes_create: null argument");
    *out = nullptr;
    if (nbox < 1 || !prm || !endtime || !cons || !u || !p)
        return set_error(BX_ERR_ARG, "bx_boxes_create: null argument or no grids");
    std::vector<bx::Box> hb((size_t)nbox);
    std::vector<int32_t> gd, gs, gg, gx;
    std::vector<double> cost((size_t)nbox);
    int64_t cells = 0, interior = 0, reg = 0, icn = 0;
    const double rgas = rgas_units();
    for (int i = 0; i < nbox; i++) {
        bx_params q = prm[i];
        std::string why;
        if (!check_box(q, why)) return set_error(BX_ERR_ARG, "bx_boxes_create: grid " + std::to_string(i) + ": " + why);
        if (q.nz < 6) q.no_outflow[4] = q.no_outflow[5] = 0;  // init_grid.f90:76-79
        int64_t ng = 0, ns = 0;
        if (int rc = bx_ghost_table(&q, &ng, &ns, nullptr, nullptr, nullptr, nullptr)) return rc;
        const size_t g0 = gd.size(), x0 = gx.size();
        gd.resize(g0 + ng);
        gs.resize(g0 + ng);
        gg.resize(g0 + ng);
        gx.resize(x0 + ns);
        if (int rc = bx_ghost_table(&q, &ng, &ns, gd.data() + g0, gs.data() + g0, gg.data() + g0, gx.data() + x0))
            return rc;
        bx::Box &b = hb[i];
        b.n = (int64_t)q.nx * q.ny * q.nz;
        b.ic_off = icn;
        b.reg_off = reg;
        b.g_off = (int64_t)g0;
        b.x_off = (int64_t)x0;
        b.ng = ng;
        b.ns = ns;
        b.nx = q.nx;
        b.ny = q.ny;
        b.nz = q.nz;
        b.flatz = q.nz < 6;
        b.maxsteps = 1000000;
        b.dx = q.dx;
        b.dy = q.dy;
        b.dz = q.dz;
        b.gam = q.gammafix;
        b.cfl = q.cfl;
        b.endtime = endtime[i];
        b.srho = q.small_rho;
        b.seps = q.small_eps;
        b.vlimit = q.vlimit;
        b.muc = 1.0 / (q.xabun * 0.5 + q.yabun * 0.25 + q.zabun / q.mu_z);  // eos.f90:144
        b.rgas = rgas;
        // calc_gforce with phi=0 (utils.f90:566-568): -(0-0)/(2*d) = -0.0
        b.g0 = -((0.0 - 0.0) / (2.0 * q.dx));
        b.g1 = -((0.0 - 0.0) / (2.0 * q.dy));
        b.g2 = -((0.0 - 0.0) / (2.0 * q.dz));
        // flux.f90:41-46
        b.axy = q.dy * q.dx;
        b.ayz = q.dy * q.dz;
        b.axz = q.dx * q.dz;
        b.vol = q.dx * q.dy * q.dz;
        // fused steps need every ghost donor to be an interior cell (one fill pass)
        b.fusable = ns == 0 ? 1 : 0;
        b.zlo = b.flatz ? 0 : 2;
        b.zhi = b.flatz ? q.nz : q.nz - 2;
        b.ly = 32;  // rows and planes per virtual wave (32x32 measured fastest of 8..32)
        b.lz = 32;
        const double ds[3] = {q.dx, q.dy, q.dz};
        for (int d = 0; d < 3; d++) {
            b.inv_ds[d] = 1.0 / ds[d];
            b.ds_pow2[d] = is_pow2(ds[d]) ? 1 : 0;
        }
        b.inv_vol = 1.0 / b.vol;
        b.vol_pow2 = is_pow2(b.vol) ? 1 : 0;
        b.gm1 = q.gammafix - 1.0;
        const int64_t in_x = q.nx - 4, in_y = q.ny - 4, in_z = q.nz < 6 ? q.nz : q.nz - 4;
        const int64_t off = cells;  // first cell of this grid in the concatenated host arrays
        cells += b.n;
        interior += std::max<int64_t>(0, in_x) * std::max<int64_t>(0, in_y) * std::max<int64_t>(0, in_z);
        icn += (int64_t)bx::kSetupPlanes * b.n;
        reg += (int64_t)bx::kRegionPlanes * b.n + 5 * ns;
        // cost estimate for the queue order: cells x steps, steps from the set-up's signal speeds
        double rate = 0.0;
        for (int64_t c = 0; c < b.n; c++) {
            const double rho = cons[(off + c) * 5], pc = p[off + c];
            if (!(rho > 0) || !(pc > 0)) continue;
            const double cs = std::sqrt(q.gammafix * pc / rho);
            const double r = (cs + std::fabs(u[(off + c) * 3])) / q.dx + (cs + std::fabs(u[(off + c) * 3 + 1])) / q.dy +
                             (cs + std::fabs(u[(off + c) * 3 + 2])) / q.dz;
            if (r > rate && std::isfinite(r)) rate = r;
        }
        const double steps = q.cfl > 0 ? std::max(1.0, endtime[i] * rate / q.cfl) : 1.0;
        cost[i] = (double)b.n * std::min(steps, 1e6);
    }

    int ndev = 0;
    BX_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return set_error(BX_ERR_ARG, "bx_boxes_create: no such device");
    BX_HIP(hipSetDevice(device));
    bx_boxes *b = new (std::nothrow) bx_boxes();
    if (!b) return set_error(BX_ERR_ARG, "bx_boxes_create: out of host memory");
    b->device = device;
    b->nbox = nbox;
    b->cells = cells;
    b->interior = interior;
    b->reg_total = reg;
    b->ic_total = icn;
    b->hbox = hb;
    b->horder.resize(nbox);
    std::iota(b->horder.begin(), b->horder.end(), 0);
    std::stable_sort(b->horder.begin(), b->horder.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    auto fail = [&](hipError_t e, const char *what) {
        const int r = bx::hip_check(e, what);
        boxes_free(b);
        return r;
    };
    hipError_t e;
    if ((e = hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
        return fail(e, "hipDeviceGetAttribute");
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(e, "hipStreamCreate");
    struct {
        void **ptr;
        size_t bytes;
    } allocs[] = {{(void **)&b->boxes, sizeof(bx::Box) * nbox},
                  {(void **)&b->order, sizeof(int) * nbox},
                  {(void **)&b->queue, 256},
                  {(void **)&b->ic, sizeof(double) * (size_t)icn},
                  {(void **)&b->reg, sizeof(double) * (size_t)reg},
                  {(void **)&b->gdst, sizeof(int32_t) * std::max<size_t>(1, gd.size())},
                  {(void **)&b->gsrc, sizeof(int32_t) * std::max<size_t>(1, gs.size())},
                  {(void **)&b->gsign, sizeof(int32_t) * std::max<size_t>(1, gg.size())},
                  {(void **)&b->gxsrc, sizeof(int32_t) * std::max<size_t>(1, gx.size())},
                  {(void **)&b->t_out, sizeof(double) * nbox},
                  {(void **)&b->steps_out, sizeof(int) * nbox},
                  {(void **)&b->status, sizeof(unsigned) * nbox},
                  {(void **)&b->dtlog, sizeof(double)}};
    for (auto &x : allocs)
        if ((e = hipMalloc(x.ptr, x.bytes)) != hipSuccess) return fail(e, "hipMalloc boxes");
    if ((e = hipMemset(b->reg, 0, sizeof(double) * (size_t)reg)) != hipSuccess) return fail(e, "hipMemset");
    // set-up: AoS (k, n) host arrays -> k planes of n per grid
    {
        std::vector<double> tmp((size_t)icn);
        int64_t cell0 = 0;
        for (int i = 0; i < nbox; i++) {
            const int64_t n = hb[i].n;
            double *d = tmp.data() + hb[i].ic_off;
            for (int64_t c = 0; c < n; c++) {
                for (int f = 0; f < 5; f++) d[f * n + c] = cons[(cell0 + c) * 5 + f];
                for (int f = 0; f < 3; f++) d[(5 + f) * n + c] = u[(cell0 + c) * 3 + f];
                d[8 * n + c] = p[cell0 + c];
            }
            cell0 += n;
        }
        if ((e = hipMemcpy(b->ic, tmp.data(), sizeof(double) * tmp.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy set-up");
    }
    auto put = [&](void *dst, const void *src, size_t bytes) {
        return bytes ? hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    if ((e = put(b->gdst, gd.data(), sizeof(int32_t) * gd.size())) != hipSuccess ||
        (e = put(b->gsrc, gs.data(), sizeof(int32_t) * gs.size())) != hipSuccess ||
        (e = put(b->gsign, gg.data(), sizeof(int32_t) * gg.size())) != hipSuccess ||
        (e = put(b->gxsrc, gx.data(), sizeof(int32_t) * gx.size())) != hipSuccess ||
        (e = put(b->order, b->horder.data(), sizeof(int) * nbox)) != hipSuccess)
        return fail(e, "hipMemcpy tables");
    if (int rc = ensure_table(b, 1000000)) {
        boxes_free(b);
        return rc;
    }
    *out = b;
    return BX_OK;
}

int bx_boxes_destroy(bx_boxes *b)
{
    if (!b) return BX_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    boxes_free(b);
    return BX_OK;
}

int bx_boxes_cells(const bx_boxes *b, int64_t *cells, int64_t *interior_cells)
{
    if (!b) return set_error(BX_ERR_ARG, "bx_boxes_cells: null handle");
    if (cells) *cells = b->cells;
    if (interior_cells) *interior_cells = b->interior;
    return BX_OK;
}

int bx_boxes_order(const bx_boxes *b, int32_t *order)
{
    if (!b || !order) return set_error(BX_ERR_ARG, "bx_boxes_order: null argument");
    std::copy(b->horder.begin(), b->horder.end(), order);
    return BX_OK;
}

int bx_boxes_tune(bx_boxes *b, int32_t waves_per_block, int32_t blocks_per_cu)
{
    if (!b) return set_error(BX_ERR_ARG, "bx_boxes_tune: null handle");
    // the 4- and 8-wave builds hold about 210 registers per lane (two waves per
    // SIMD); the 16-wave build is held to 128 and spills to scratch
    if (waves_per_block != 4 && waves_per_block != 8 && waves_per_block != 16)
        return set_error(BX_ERR_ARG, "bx_boxes_tune: waves per workgroup must be 4, 8 or 16");
    if (blocks_per_cu < 1 || waves_per_block * blocks_per_cu > 16)
        return set_error(BX_ERR_ARG, "bx_boxes_tune: at most 16 waves per CU");
    b->waves = waves_per_block;
    b->blocks_per_cu = blocks_per_cu;
    drop_graph(b);
    return BX_OK;
}

int bx_boxes_log_dt(bx_boxes *b, int32_t maxlog)
{
    if (!b || maxlog < 0) return set_error(BX_ERR_ARG, "bx_boxes_log_dt: bad argument");
    BX_HIP(hipSetDevice(b->device));
    BX_HIP(hipStreamSynchronize(b->stream));
    // logging off and the captured launch (which holds the old log's address)
    // dropped first, so a failed allocation below leaves a handle that runs
    // without a log
    drop_graph(b);
    b->maxlog = 0;
    if (b->dtlog) BX_HIP(hipFree(b->dtlog));
    b->dtlog = nullptr;
    if (maxlog == 0) return BX_OK;
    const size_t cnt = (size_t)maxlog * (size_t)b->nbox;
    BX_HIP(hipMalloc((void **)&b->dtlog, sizeof(double) * cnt));
    BX_HIP(hipMemset(b->dtlog, 0, sizeof(double) * cnt));
    b->maxlog = maxlog;
    return BX_OK;
}

int bx_boxes_run(bx_boxes *b, int32_t maxsteps, int32_t use_graph)
{
    if (!b || maxsteps < 0) return set_error(BX_ERR_ARG, "bx_boxes_run: bad argument");
    BX_HIP(hipSetDevice(b->device));
    if (int rc = ensure_table(b, maxsteps)) return rc;
    if (!b->ev[0]) {
        BX_HIP(hipEventCreate(&b->ev[0]));
        BX_HIP(hipEventCreate(&b->ev[1]));
    }
    BX_HIP(hipEventRecord(b->ev[0], b->stream));
    if (use_graph) {
        if (!b->graph_valid)
            if (int rc = capture(b)) return rc;
        BX_HIP(hipGraphLaunch(b->exec, b->stream));
    } else if (int rc = enqueue(b)) {
        return rc;
    }
    BX_HIP(hipEventRecord(b->ev[1], b->stream));
    b->timed = true;
    return BX_OK;
}

int bx_boxes_sync(bx_boxes *b, double *ms)
{
    if (!b) return set_error(BX_ERR_ARG, "bx_boxes_sync: null handle");
    BX_HIP(hipSetDevice(b->device));
    BX_HIP(hipStreamSynchronize(b->stream));
    if (ms) {
        float t = 0.f;
        if (b->timed) BX_HIP(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
        *ms = t;
    }
    return BX_OK;
}

int bx_boxes_result(bx_boxes *b, double *cons, double *u, double *p, double *time, int32_t *steps, uint32_t *status)
{
    if (!b) return set_error(BX_ERR_ARG, "bx_boxes_result: null handle");
    BX_HIP(hipSetDevice(b->device));
    BX_HIP(hipStreamSynchronize(b->stream));
    const int nb = b->nbox;
    if (cons || u || p) {
        // the result planes of each grid (cons, u, p), gathered into one host buffer
        std::vector<double> res((size_t)(b->cells * bx::kResultPlanes));
        int64_t cell0 = 0;
        for (int i = 0; i < nb; i++) {
            const int64_t n = b->hbox[i].n;
            BX_HIP(hipMemcpyAsync(res.data() + cell0 * bx::kResultPlanes, b->reg + b->hbox[i].reg_off,
                                  sizeof(double) * bx::kResultPlanes * n, hipMemcpyDeviceToHost, b->stream));
            cell0 += n;
        }
        BX_HIP(hipStreamSynchronize(b->stream));
        cell0 = 0;
        for (int i = 0; i < nb; i++) {
            const int64_t n = b->hbox[i].n;
            const double *r = res.data() + cell0 * bx::kResultPlanes;
            for (int64_t c = 0; c < n; c++) {
                if (cons)
                    for (int f = 0; f < 5; f++) cons[(cell0 + c) * 5 + f] = r[f * n + c];
                if (u)
                    for (int f = 0; f < 3; f++) u[(cell0 + c) * 3 + f] = r[(5 + f) * n + c];
                if (p) p[cell0 + c] = r[8 * n + c];
            }
            cell0 += n;
        }
    }
    if (time) BX_HIP(hipMemcpy(time, b->t_out, sizeof(double) * nb, hipMemcpyDeviceToHost));
    if (steps) BX_HIP(hipMemcpy(steps, b->steps_out, sizeof(int) * nb, hipMemcpyDeviceToHost));
    if (status) BX_HIP(hipMemcpy(status, b->status, sizeof(unsigned) * nb, hipMemcpyDeviceToHost));
    return BX_OK;
}

int bx_boxes_dts(bx_boxes *b, double *dts)
{
    if (!b || !dts) return set_error(BX_ERR_ARG, "bx_boxes_dts: null argument");
    if (b->maxlog == 0) return set_error(BX_ERR_STATE, "bx_boxes_dts: dt logging is off (bx_boxes_log_dt)");
    BX_HIP(hipSetDevice(b->device));
    BX_HIP(hipStreamSynchronize(b->stream));
    BX_HIP(hipMemcpy(dts, b->dtlog, sizeof(double) * (size_t)b->maxlog * b->nbox, hipMemcpyDeviceToHost));
    return BX_OK;
}

}  // extern "C"
