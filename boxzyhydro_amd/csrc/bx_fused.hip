// bx_fused.hip -- the hydro step as one fused stencil kernel per RK stage.
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
// Work decomposition (gfx950, wave64): by default each stage is one kernel,
// k_stage, a workgroup holding a 64 x TY tile of outputs with its +-2 halo in
// LDS and marching in z (bx_tile.h).  The tuning variants 0, 1 and 4
// (bx_tune) run the earlier form, two kernels per stage with every wave
// independent (no LDS, no block barriers):
//   * k_ydir: a wave holds 64 interior x-cells of one z-plane and marches in y
//     with a register queue; it writes the y-term of fluxtmp (FP64 planes).
//   * k_xz: a wave holds 64 cells of one interior row (lanes 2..61 output,
//     0,1,62,63 the x-halo; x-neighbours through wave shuffles) and marches in
//     z with a register queue; it adds the x- and z-terms in the reference's
//     order and applies the stage update.
// Both give the same bits.
#include <cmath>

#include "bx_tile.h"

namespace bx {

// One stage as one kernel: x-y tile of 64 x TY outputs with its +-2 halo in
// LDS, marching in z (bx_tile.h: tile_stage).  P2: every spacing and the
// cell volume are powers of two (their divisions are multiplications), known
// at compile time, which keeps the spacings themselves out of the registers.
template <int STAGE, typename S, int TY, bool P2, bool G = false>
__global__ __launch_bounds__(64 * (TY + 4)) void k_stage(StageArgs a0)
{
    __shared__ TileLds<S, TY> lds;
    StageArgs a = with_dt(a0);
    if constexpr (P2) {
        a.ds_pow2[0] = a.ds_pow2[1] = a.ds_pow2[2] = 1;
        a.vol_pow2 = 1;
    }
    // Workgroups go to the 8 XCDs in turn (workgroup id mod 8).  Give each
    // XCD a contiguous range of tiles, x fastest, so that the tiles in flight
    // on one XCD are neighbours and share their halo rows and boundary cache
    // lines in that XCD's L2.
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned T = gx * gy * gridDim.z;
    const unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned xcd = L & 7u, slot = L >> 3, q = T >> 3, rem = T & 7u;
    const unsigned t = xcd * q + (xcd < rem ? xcd : rem) + slot;
    tile_stage<STAGE, S, TY, G>(a, lds, t % gx, (t / gx) % gy, t / (gx * gy));
}


// y-direction kernel: one launch over the interior (bx_waves.h: ydir_wave)
template <int STAGE, typename S>
__global__ __launch_bounds__(256) void k_ydir(StageArgs a0, S *__restrict__ Ey, int ly)
{
    const StageArgs a = with_dt(a0);
    ydir_wave<STAGE, S>(a, Ey, ly, blockIdx.x, blockIdx.y, blockIdx.z, wave_id());
}

// x/z kernel: one launch over the interior (bx_waves.h: xz_wave)
template <int STAGE, typename S, int OCC, bool PFE>
__global__ __launch_bounds__(256, OCC) void k_xz(StageArgs a0, const S *__restrict__ Ey)
{
    const StageArgs a = with_dt(a0);
    xz_wave<STAGE, S, PFE>(a, Ey, blockIdx.x, blockIdx.z, wave_id());
}

// -------------------------------------------------------------- helpers
// p: null keeps the pressure plane (written by the fused corrector with gravity)
__global__ __launch_bounds__(256) void k_materialize(int64_t n, const double *__restrict__ cons,
                                                    double *__restrict__ u, double *__restrict__ p,
                                                    double *__restrict__ ad, double gam, double seps)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) materialize_cell(n, c, cons, u, p, ad, gam, seps);
}

__global__ __launch_bounds__(256) void k_regular(int64_t n, const double *__restrict__ cons,
                                                const double *__restrict__ u, double gam, double seps,
                                                double dx, double dy, double dz, unsigned int *flags,
                                                unsigned long long *rate)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned int flg = 0;
    const double rt = c < n ? regular_cell(n, c, cons, u, gam, seps, dx, dy, dz, flg) : 0.0;
    if (flg) atomicOr(flags, flg);
    wave_max_to(rt, rate);
}

template <typename S>
__global__ __launch_bounds__(256) void k_ghosts6(int64_t n, int64_t ng, const int32_t *__restrict__ dst,
                                                const int32_t *__restrict__ src,
                                                const int32_t *__restrict__ sign, S *__restrict__ planes,
                                                S *__restrict__ pstar)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ng) ghost6_one<S>(n, k, dst, src, sign, planes, pstar);
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
    a.nd2 = c->nd2;
    a.d2cell = c->d2_src;
    a.d2ubar = c->d2_ubar;
    a.gf = c->grav_flags ? c->gfused : nullptr;
    return a;
}

template <int STAGE, typename S, int TY>
static int launch_tile(bx_ctx *c, const StageArgs &a)
{
    const bool p2 = a.ds_pow2[0] && a.ds_pow2[1] && a.ds_pow2[2] && a.vol_pow2;
    const int nzi = a.zhi - a.zlo;
    const unsigned gx = (unsigned)((c->nx - 4 + 63) / 64);
    const unsigned gy = (unsigned)((c->ny - 4 + TY - 1) / TY);
    const unsigned gz = (unsigned)((nzi + a.lz - 1) / a.lz);
    const dim3 grid(gx, gy, gz), block(64 * (TY + 4));
    if constexpr (sizeof(S) == 8 && TY == 8) {
        if (a.gf) {
            // the corrector's power-of-two form spills 5 VGPRs at 168; its
            // general form (161, no spill) gives the same bits and is ~1 % faster
            if (p2 && STAGE == 1) hipLaunchKernelGGL((k_stage<STAGE, S, TY, true, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((k_stage<STAGE, S, TY, false, true>), grid, block, 0, c->stream, a);
            BX_HIP(hipGetLastError());
            return BX_OK;
        }
    }
    if (a.gf) return set_error(BX_ERR_UNSUPPORTED, "fused step with gravity: FP64 64 x 8 tiles only");
    if (p2)
        hipLaunchKernelGGL((k_stage<STAGE, S, TY, true>), grid, block, 0, c->stream, a);
    else
        hipLaunchKernelGGL((k_stage<STAGE, S, TY, false>), grid, block, 0, c->stream, a);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

template <int STAGE, typename S>
static int launch_stage(bx_ctx *c, const StageArgs &a, hipEvent_t mid, S *eflux)
{
    if (c->grav_flags && !a.gf) return set_error(BX_ERR_STATE, "fused step with gravity before its force planes");
    if (c->fused_occ >= 10) {
        if (mid) BX_HIP(hipEventRecord(mid, c->stream));
        switch (c->fused_occ) {
        case 10: return launch_tile<STAGE, S, 4>(c, a);
        default: return launch_tile<STAGE, S, 8>(c, a);
        }
    }
    if (a.gf) return set_error(BX_ERR_UNSUPPORTED, "fused step with gravity: the tiled kernel only (variant 11)");
    const int nzi = a.zhi - a.zlo;
    const int ly = c->fused_ly;
    {
        const unsigned gx = (unsigned)((c->nx - 4 + 63) / 64);
        const unsigned gy = (unsigned)((nzi + 3) / 4);
        const unsigned gz = (unsigned)((c->ny - 4 + ly - 1) / ly);
        hipLaunchKernelGGL((k_ydir<STAGE, S>), dim3(gx, gy, gz), dim3(256), 0, c->stream, a, eflux, ly);
        BX_HIP(hipGetLastError());
    }
    if (mid) BX_HIP(hipEventRecord(mid, c->stream));
    {
        const int64_t nwaves = ((int64_t)(c->ny - 4) * c->nx + kSegOut - 1) / kSegOut;
        const unsigned gx = (unsigned)((nwaves + 3) / 4);
        const unsigned gy = 1;
        const unsigned gz = (unsigned)((nzi + a.lz - 1) / a.lz);
        const S *ey = eflux;
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

// the pressure plane of the materialised state at the corner ghosts (the
// value the last fused step's fix-up recorded)
__global__ __launch_bounds__(64) void k_corner_pressure(int nd2, const int32_t *__restrict__ dst,
                                                        const double *__restrict__ ubar, double *__restrict__ p)
{
    const int j = threadIdx.x;
    if (j < nd2) p[dst[j]] = ubar[6 * j + 5];
}

int launch_materialize(bx_ctx *c)
{
    hipLaunchKernelGGL(k_materialize, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->cons, c->u,
                       c->p_fused ? (double *)nullptr : c->p, c->adindx, c->prm.gammafix, c->prm.small_eps);
    BX_HIP(hipGetLastError());
    if (c->nd2 > 0 && c->prm.precision == BX_PREC_FP64) {  // the mixed form copies the donors
        hipLaunchKernelGGL(k_corner_pressure, dim3(1), dim3(64), 0, c->stream, c->nd2, c->d2_dst, c->d2_ubar, c->p);
        BX_HIP(hipGetLastError());
    }
    return BX_OK;
}

// Whether the fused kernels compute this context's step (no allocation).
int fused_supported(const bx_ctx *c)
{
    if (c->nscratch > 0 && !c->fghost_ok)
        return set_error(BX_ERR_UNSUPPORTED, "fused step needs the ghost donors of a pair of set_ghost_cells calls "
                                             "to resolve to interior cells within two levels");
    if (c->grav[0] != 0.0 || c->grav[1] != 0.0 || c->grav[2] != 0.0)
        return set_error(BX_ERR_UNSUPPORTED, "fused step assumes zero gravity");
    if (c->prm.fluxangmom) return set_error(BX_ERR_UNSUPPORTED, "fused step fluxes linear momentum (fluxangmom off)");
    if (c->grav_flags & BX_GRAV_EXTERNAL)
        return set_error(BX_ERR_UNSUPPORTED, "fused step with gravity: a given potential only (no point mass, whose "
                                             "force needs the corrector's mass moments before source(1))");
    if (c->eos) return set_error(BX_ERR_UNSUPPORTED, "fused step assumes a fixed gamma (H2STAT=-1)");
    return BX_OK;
}

// Default work split of the stage kernels: one kernel per stage, FP64 on
// 64 x 8 tiles (variant 11) with z-chunks of 64 planes, the mixed form on
// 64 x 4 tiles (variant 10, its FP32 tiles being small enough for several
// workgroups per CU) with chunks of 32; the chunks are shortened until the
// grid has two workgroups per CU.  The y chunk (fused_ly) only serves the
// two-kernel variants.  Allocates nothing.
void fused_defaults(bx_ctx *c)
{
    const int64_t want = 2 * 1024;
    const int nzi = c->flatz ? c->nz : c->nz - 4;
    if (c->fused_ly == 0) {
        int ly = 64;
        while (ly > 8 && (int64_t)((c->nx - 4 + 63) / 64) * nzi * ((c->ny - 4 + ly - 1) / ly) < want) ly /= 2;
        c->fused_ly = ly;
    }
    if (c->fused_lz == 0) {
        if (c->prm.precision == BX_PREC_MIXED) {
            int lz = 32;
            const int64_t tiles = (int64_t)((c->nx - 4 + 63) / 64) * ((c->ny - 4 + 3) / 4);
            while (lz > 4 && tiles * ((nzi + lz - 1) / lz) < 512) lz /= 2;
            c->fused_lz = lz;
            c->fused_occ = 10;
        } else {
            // One workgroup per CU at a time (LDS): the chunk that minimises
            // the rounds of workgroups times the planes each one marches
            // (its chunk and the 4 halo planes); ties go to the smaller chunk
            int ncu = 0;
            if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
                ncu < 1)
                ncu = 256;
            const int64_t tiles = (int64_t)((c->nx - 4 + 63) / 64) * ((c->ny - 4 + 7) / 8);
            int lz = 4;
            int64_t best = INT64_MAX;
            for (int cand = 4; cand <= 256; cand *= 2) {
                const int64_t chunks = (nzi + cand - 1) / cand;
                const int64_t cost = (tiles * chunks + ncu - 1) / ncu * ((cand < nzi ? cand : nzi) + 4);
                if (cost < best) {
                    best = cost;
                    lz = cand;
                }
            }
            c->fused_lz = lz;
            c->fused_occ = 11;
        }
    }
}

// Ghost table shallow, allocation of the fused buffers.
int fused_prepare(bx_ctx *c)
{
    if (int rc = fused_supported(c)) return rc;
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
    fused_defaults(c);
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
                       c->ghost_dst, c->nscratch ? c->fghost_src : c->ghost_src,
                       c->nscratch ? c->fghost_sign : c->ghost_sign, planes, pstar);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

// The two-level corner ghosts after the corrector's ghost fill.  The exact
// sequence leaves such a ghost holding +-(its interior cell v's corrector
// result before source(1)); that value reaches the results only through the
// next step's Courant rate (courant.f90:11-15 runs over every cell) and the
// stored state.  One lane per ghost evaluates v's corrector update again
// from the corrector's inputs with the tiled kernel's operations (its six
// faces, flux.f90:121-190, and the update of flux.f90:213-228), parks the
// result before source(1), writes the ghost and adds its Courant rate (the
// state and velocity of main.f90:232 and of the step's last velocity call,
// state.f90:30-57, velocity.f90:18-21).
__global__ __launch_bounds__(64) void k_corner_ghosts(StageArgs a0, const int32_t *__restrict__ dst,
                                                      const int32_t *__restrict__ sign)
{
    const StageArgs a = with_dt(a0);
    const int j = threadIdx.x;
    double rt = 0.0;
    unsigned int flg = 0;
    if (j < a.nd2) {
        const int64_t v = a.d2cell[j];
        const int64_t plane = (int64_t)a.nx * a.ny;
        const int ix = (int)(v % a.nx), iy = (int)((v / a.nx) % a.ny), iz = (int)(v / plane);
        auto ghost = [&](int x, int y, int z) {
            return x < 2 || x > a.nx - 3 || y < 2 || y > a.ny - 3 || (!a.flatz && (z < 2 || z > a.nz - 3));
        };
        double F[3][2][5];
        for (int d = 0; d < 3; d++)
            for (int h = 0; h < 2; h++)
                for (int i = 0; i < 5; i++) F[d][h][i] = 0.0;
        W6<double> q0{};
        auto faces = [&](auto dirc) {
            constexpr int d = decltype(dirc)::value;
            const int64_t e = d == 0 ? 1 : (d == 1 ? a.nx : plane);
            W6<double> w[5];
            for (int i = 0; i < 5; i++) w[i] = load_prims<2, double>(a, v + (i - 2) * e, 0u, flg);
            if (d == 0) q0 = w[2];
            Side<double> mi[3], pl[3];
            for (int i = 1; i <= 3; i++) {
                const int o = i - 2;
                const bool gh = ghost(ix + (d == 0 ? o : 0), iy + (d == 1 ? o : 0), iz + (d == 2 ? o : 0));
                dir_states<d>(w[i - 1], w[i], w[i + 1], sound_of(w[i], a.gam), gh, a, mi[i - 1], pl[i - 1]);
            }
            face_flux<double, d>(pl[0], mi[1], F[d][0]);
            face_flux<double, d>(pl[1], mi[2], F[d][1]);
        };
        faces(std::integral_constant<int, 0>{});
        faces(std::integral_constant<int, 1>{});
        if (!a.flatz) faces(std::integral_constant<int, 2>{});
        double ui[5], uo[5], ft[5];
        for (int i = 0; i < 5; i++) {
            ui[i] = elem<double>(a.uin, (int64_t)i * a.n + v, 0u);
            uo[i] = elem<double>(a.uold, (int64_t)i * a.n + v, 0u);
            const double dx = F[0][1][i] - F[0][0][i];
            const double ey = a.axz * (F[1][1][i] - F[1][0][i]);
            const double dz = a.flatz ? 0.0 : F[2][1][i] - F[2][0][i];
            const double sm = (a.ayz * dx + ey) + a.axy * dz;
            ft[i] = -div_by(sm * a.dt, a.vol, a.inv_vol, a.vol_pow2);
        }
        double rv = 0.0;
        // rewrites U^{n+1} of v with the bits the tiled kernel wrote and parks
        // the result before source(1) in d2ubar (with gravity also p of v)
        if (a.gf) stage_update<2, double, true>(a, ui, uo, ft, q0, v, 0u, flg, rv, true);
        else stage_update<2, double>(a, ui, uo, ft, q0, v, 0u, flg, rv, true);
        double U[5];
        for (int i = 0; i < 5; i++) U[i] = a.d2ubar[6 * j + i];
        // the pressure the exact sequence leaves in the ghost: the state call
        // after the last ghost pair (main.f90:511-513) takes the velocity of
        // the previous velocity call, when the ghost held h2's value from
        // before the third pair, 0.5*(cons_old + U* + 0) of v up to signs
        // (flux.f90:213, cons_old = U^n after source(0), which leaves rho as
        // it is and, without gravity, m too; with gravity m is kicked by v's
        // force); the signs drop out of |u|^2
        {
            double rc = 0.5 * ((uo[0] + ui[0]) + 0.0);
            rc = vmax(rc, a.srho);
            double mo[3] = {uo[1], uo[2], uo[3]};
            if (a.gf) {
                const double g[3] = {a.gf[v], a.gf[a.n + v], a.gf[2 * a.n + v]};
                const double m[3] = {uo[1], uo[2], uo[3]};
                grav_momenta(uo[0], m, g, a.dt * 0.5, mo);
            }
            double uc[3];
            for (int d = 0; d < 3; d++) uc[d] = (0.5 * ((mo[d] + ui[1 + d]) + 0.0)) / rc;
            const double ek = U[0] * 0.5 * (uc[0] * uc[0] + uc[1] * uc[1] + uc[2] * uc[2]);
            a.d2ubar[6 * j + 5] = vmax(U[4] - ek, a.seps) * a.gm1;
        }
        const int sg = sign[j];
        for (int f = 1; f <= 3; f++)
            if ((sg >> (f - 1)) & 1) U[f] = -U[f];
        const int64_t gcell = dst[j];
        double *un = (double *)a.uout;
        for (int i = 0; i < 5; i++) un[(int64_t)i * a.n + gcell] = U[i];
        const double v0 = U[1] / U[0], v1 = U[2] / U[0], v2 = U[3] / U[0];
        const double ekin = U[0] * 0.5 * (v0 * v0 + v1 * v1 + v2 * v2);
        const double p = vmax(U[4] - ekin, a.seps) * a.gm1;
        const double snd = sqrt(a.gam * p / U[0]);
        rt = div_by(snd + fabs(v0), a.ds[0], a.inv_ds[0], a.ds_pow2[0]) +
             div_by(snd + fabs(v1), a.ds[1], a.inv_ds[1], a.ds_pow2[1]) +
             div_by(snd + fabs(v2), a.ds[2], a.inv_ds[2], a.ds_pow2[2]);
    }
    if (flg) atomicOr(a.flags, flg);
    wave_max_to(rt, a.rate);
}

// after the corrector's ghost fill of un (b: the corrector's arguments)
static int corner_ghosts(bx_ctx *c, const StageArgs &b)
{
    if (c->nd2 == 0) return BX_OK;
    hipLaunchKernelGGL(k_corner_ghosts, dim3(1), dim3(64), 0, c->stream, b, c->d2_dst, c->d2_sign);
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

// With gravity the corrector also writes the pressure plane (the state after
// source(1), which takes the velocity from before the kick); returns that
// plane for the ghost fill, null without gravity.
template <typename S>
static S *grav_pressure(bx_ctx *c, StageArgs &b)
{
    if constexpr (sizeof(S) == 8) {
        if (b.gf) {
            b.pout = c->p;
            return c->p;
        }
    }
    return nullptr;
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
    if ((rc = launch_stage<1, S>(c, a, c->prof_on ? c->ev[1] : nullptr, (S *)c->eflux))) return rc;
    if (c->prof_on) BX_HIP(hipEventRecord(c->ev[2], c->stream));
    if ((rc = ghosts6<S>(c, us, ps))) return rc;
    StageArgs b = base_args(c, dt);
    b.dtp = dtp;
    b.uin = us;
    b.pin = ps;
    b.uold = u0;
    b.uout = un;
    b.rate = &c->dsc->rate_bits[0];
    S *pfin = grav_pressure<S>(c, b);
    if (c->prof_on) BX_HIP(hipEventRecord(c->ev[3], c->stream));
    if ((rc = launch_stage<2, S>(c, b, c->prof_on ? c->ev[4] : nullptr, (S *)c->eflux))) return rc;
    if (c->prof_on) BX_HIP(hipEventRecord(c->ev[5], c->stream));
    if ((rc = ghosts6<S>(c, un, pfin))) return rc;
    if constexpr (sizeof(S) == 8) return corner_ghosts(c, b);
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
        if ((rc = launch_stage<1, S>(c, a, nullptr, (S *)c->eflux))) return rc;
        if ((rc = ghosts6<S>(c, us, ps))) return rc;
        StageArgs b = base_args(c, 0.0);
        b.dtp = &c->dsc->dt;
        b.uin = us;
        b.pin = ps;
        b.uold = src;
        b.uout = dst;
        b.rate = &c->dsc->rate_bits[0];
        S *pfin = grav_pressure<S>(c, b);
        if ((rc = launch_stage<2, S>(c, b, nullptr, (S *)c->eflux))) return rc;
        if ((rc = ghosts6<S>(c, dst, pfin))) return rc;
        if constexpr (sizeof(S) == 8)
            if ((rc = corner_ghosts(c, b))) return rc;
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

__global__ void k_clear_scalars(unsigned int *flags, unsigned long long *rate)
{
    *flags = 0u;
    *rate = 0ull;
}

// One FP64 fused step u0 -> us, ps -> un (ey: the y-term planes) on the
// context's stream with the flags and the Courant rate accumulated into the
// given device words, which
// a kernel clears first (no memset nodes, so the sequence can be captured).
// Allocates nothing; the tile defaults must be set (fused_defaults).
int fused_step_words(bx_ctx *c, double dt, const double *u0, double *us, double *ps, double *un, double *ey,
                     unsigned int *flags, unsigned long long *rate)
{
    if (c->grav_flags) return set_error(BX_ERR_UNSUPPORTED, "this fused step entry takes no gravity configuration");
    int rc;
    hipLaunchKernelGGL(k_clear_scalars, dim3(1), dim3(1), 0, c->stream, flags, rate);
    BX_HIP(hipGetLastError());
    StageArgs a = base_args(c, dt);
    a.flags = flags;
    a.uin = u0;
    a.uout = us;
    a.pout = ps;
    if ((rc = launch_stage<1, double>(c, a, nullptr, ey))) return rc;
    if ((rc = ghosts6<double>(c, us, ps))) return rc;
    StageArgs b = base_args(c, dt);
    b.flags = flags;
    b.uin = us;
    b.pin = ps;
    b.uold = u0;
    b.uout = un;
    b.rate = rate;
    if ((rc = launch_stage<2, double>(c, b, nullptr, ey))) return rc;
    if ((rc = ghosts6<double>(c, un, (double *)nullptr))) return rc;
    return corner_ghosts(c, b);
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
