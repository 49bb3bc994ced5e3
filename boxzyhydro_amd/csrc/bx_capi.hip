// bx_capi.hip -- the extern "C" entry points of libboxzy_hip.so (include/boxzy.h):
// context set-up (the device counterpart of init_grid.f90), host<->device
// transfers in the reference's Fortran layout, one entry point per reference
// subroutine, and the step driver of main.f90:211-555.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "bx_internal.h"

namespace {
thread_local std::string g_last_error;
}

namespace bx {

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int hip_check(hipError_t e, const char *what)
{
    return set_error(BX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Host-side restatement of one set_ghost_cells call (utils.f90:200-403) over
// symbolic values: each cell holds (source cell, momentum sign flips), and the
// ghost list is walked in the reference's serial order (the OMP DO of
// utils.f90:314 is order-dependent only where a donor is itself a ghost, and
// every call site repeats the call, which settles that order).  The result is
// a gather table: ghost <- +-(value of a cell before the call).
static int build_ghost_table(const bx_params *prm, std::vector<int32_t> &dst, std::vector<int32_t> &src,
                             std::vector<int32_t> &sgn, std::vector<int32_t> &scratch_src)
{
    const int nx = prm->nx, ny = prm->ny, nz = prm->nz;
    const int64_t n = (int64_t)nx * ny * nz;
    auto id = [&](int ix, int iy, int iz) -> int32_t {
        return (int32_t)((ix - 1) + (int64_t)(iy - 1) * nx + (int64_t)(iz - 1) * nx * ny);
    };
    auto bnd = [&](int ix, int iy, int iz) -> int {
        bool f1 = ix == 1 || ix == nx || iy == 1 || iy == ny || (nz >= 6 && (iz == 1 || iz == nz));
        if (f1) return 1;
        bool f2 = ix == 2 || ix == nx - 1 || iy == 2 || iy == ny - 1 || (nz >= 6 && (iz == 2 || iz == nz - 1));
        return f2 ? 2 : 0;
    };
    // utils.f90:200-281 (iz>=6 test at line 215 kept)
    auto check_corner = [&](int ix, int iy, int iz, int32_t ineigh) -> int32_t {
        if (ix < 3) {
            if (iy < 3) {
                if (nz >= 6) {
                    if (iz < 3) ineigh = id(3, 3, 3);
                    else if (iz > nz - 2) ineigh = id(3, 3, nz - 2);
                    else ineigh = id(3, 3, iz);
                } else ineigh = id(3, 3, iz);
            } else if (iy > ny - 2) {
                if (iz >= 6) {
                    if (iz < 3) ineigh = id(3, ny - 2, 3);
                    else if (iz > nz - 2) ineigh = id(3, ny - 2, nz - 2);
                    else ineigh = id(3, ny - 2, iz);
                } else ineigh = id(3, ny - 2, iz);
            } else {
                if (nz >= 6) {
                    if (iz < 3) ineigh = id(3, iy, 3);
                    else if (iz > nz - 2) ineigh = id(3, iy, nz - 2);
                } else ineigh = id(3, iy, iz);
            }
        } else if (ix > nx - 2) {
            if (iy < 3) {
                if (nz >= 6) {
                    if (iz < 3) ineigh = id(nx - 2, 3, 3);
                    else if (iz > nz - 2) ineigh = id(nx - 2, 3, nz - 2);
                    else ineigh = id(nx - 2, 3, iz);
                } else ineigh = id(nx - 2, 3, iz);
            } else if (iy > ny - 2) {
                if (nz >= 6) {
                    if (iz < 3) ineigh = id(nx - 2, ny - 2, 3);
                    else if (iz > nz - 2) ineigh = id(nx - 2, ny - 2, nz - 2);
                    else ineigh = id(nx - 2, ny - 2, iz);
                } else ineigh = id(nx - 2, ny - 2, iz);
            } else {
                if (nz >= 6) {
                    if (iz < 3) ineigh = id(nx - 2, iy, 3);
                    else if (iz > nz - 2) ineigh = id(nx - 2, iy, nz - 2);
                } else ineigh = id(nx - 2, iy, iz);
            }
        } else {
            if (iy < 3) {
                if (nz >= 6) {
                    if (iz < 3) ineigh = id(ix, 3, 3);
                    else if (iz > nz - 2) ineigh = id(ix, 3, nz - 2);
                } else ineigh = id(ix, 3, iz);
            } else if (iy > ny - 2) {
                if (nz >= 6) {
                    if (iz < 3) ineigh = id(ix, ny - 2, 3);
                    else if (iz > nz - 2) ineigh = id(ix, ny - 2, nz - 2);
                } else ineigh = id(ix, ny - 2, iz);
            }
        }
        return ineigh;
    };

    // ghost list in cell order (init_grid.f90:166-206)
    std::vector<int32_t> ghosts;
    for (int iz = 1; iz <= nz; iz++)
        for (int iy = 1; iy <= ny; iy++)
            for (int ix = 1; ix <= nx; ix++)
                if (bnd(ix, iy, iz) > 0) ghosts.push_back(id(ix, iy, iz));

    std::vector<int32_t> vsrc(n);
    std::vector<uint8_t> vsg(n, 0);
    for (int64_t k = 0; k < n; k++) vsrc[k] = (int32_t)k;
    const int32_t *nof = prm->no_outflow;
    const bool zwall = nz >= 6;  // init_grid.f90:76-79
    auto pass = [&]() {
        int32_t ineigh = 0;
        for (int32_t g : ghosts) {
            int ix = g % nx + 1, iy = (g / nx) % ny + 1, iz = g / (nx * ny) + 1;
            const int b = bnd(ix, iy, iz);
            const int32_t pl = nx * ny;
            if (b == 1) {
                if (ix == 1) ineigh = g + 2;
                else if (ix == nx) ineigh = g - 2;
                else if (iy == 1) ineigh = g + 2 * nx;
                else if (iy == ny) ineigh = g - 2 * nx;
                else if (iz == 1) ineigh = nz < 6 ? g : g + 2 * pl;
                else ineigh = nz < 6 ? g : g - 2 * pl;
            } else {
                if (ix == 2) ineigh = g + 1;
                else if (ix == nx - 1) ineigh = g - 1;
                else if (iy == 2) ineigh = g + nx;
                else if (iy == ny - 1) ineigh = g - nx;
                else if (iz == 2) ineigh = nz < 6 ? g : g + pl;
                else if (iz == nz - 1) ineigh = nz < 6 ? g : g - pl;
            }
            ineigh = check_corner(ix, iy, iz, ineigh);
            vsrc[g] = vsrc[ineigh];
            vsg[g] = vsg[ineigh];
            auto mirror = [&](int32_t ip, int bit) {
                vsrc[g] = vsrc[ip];
                vsg[g] = (uint8_t)(vsg[ip] ^ (1u << bit));
            };
            if (nof[0]) {
                if (ix == 1) mirror(id(ix + 3, iy, iz), 0);
                else if (ix == 2) mirror(id(ix + 1, iy, iz), 0);
            }
            if (nof[1]) {
                if (ix == nx) mirror(id(ix - 3, iy, iz), 0);
                else if (ix == nx - 1) mirror(id(ix - 1, iy, iz), 0);
            }
            if (nof[2]) {
                if (iy == 1) mirror(id(ix, iy + 3, iz), 1);
                else if (iy == 2) mirror(id(ix, iy + 1, iz), 1);
            }
            if (nof[3]) {
                if (iy == ny) mirror(id(ix, iy - 3, iz), 1);
                else if (iy == ny - 1) mirror(id(ix, iy - 1, iz), 1);
            }
            if (nof[4] && zwall) {
                if (iz == 1) mirror(id(ix, iy, iz + 3), 2);
                else if (iz == 2) mirror(id(ix, iy, iz + 1), 2);
            }
            if (nof[5] && zwall) {
                if (iz == nz) mirror(id(ix, iy, iz - 3), 2);
                else if (iz == nz - 1) mirror(id(ix, iy, iz - 1), 2);
            }
        }
    };
    pass();
    // Donors that are themselves ghosts (read before the pass reached them)
    // refer to their value from before the call; those values are snapshotted
    // into a scratch list first.
    dst.clear();
    src.clear();
    sgn.clear();
    scratch_src.clear();
    std::vector<int32_t> slot(n, -1);
    for (int32_t g : ghosts) {
        int32_t s = vsrc[g];
        int sx = s % nx + 1, sy = (s / nx) % ny + 1, sz = s / (nx * ny) + 1;
        if (bnd(sx, sy, sz) != 0) {
            if (slot[s] < 0) {
                slot[s] = (int32_t)scratch_src.size();
                scratch_src.push_back(s);
            }
            s = -(slot[s] + 1);
        }
        dst.push_back(g);
        src.push_back(s);
        sgn.push_back(vsg[g]);
    }
    return BX_OK;
}

}  // namespace bx

using namespace bx;

extern "C" {

const char *bx_last_error(void) { return g_last_error.c_str(); }

int bx_create(const bx_params *prm, int32_t device, bx_ctx **out)
{
    if (!prm || !out) return set_error(BX_ERR_ARG, "bx_create: null argument");
    *out = nullptr;
    if (prm->nx < 5 || prm->ny < 5 || prm->nz < 1)
        return set_error(BX_ERR_ARG, "bx_create: need nx,ny >= 5 and nz >= 1 (two ghost layers per side)");
    if (!(prm->dx > 0 && prm->dy > 0 && prm->dz > 0)) return set_error(BX_ERR_ARG, "bx_create: spacings must be > 0");
    if (!(prm->gammafix > 1.0)) return set_error(BX_ERR_ARG, "bx_create: gammafix must be > 1");
    if (prm->precision != BX_PREC_FP64 && prm->precision != BX_PREC_MIXED)
        return set_error(BX_ERR_ARG, "bx_create: unknown precision");
    const int64_t n = (int64_t)prm->nx * prm->ny * prm->nz;
    if (n > (int64_t)INT32_MAX) return set_error(BX_ERR_ARG, "bx_create: grid exceeds 2^31 cells");

    int ndev = 0;
    BX_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return set_error(BX_ERR_ARG, "bx_create: no such device");
    BX_HIP(hipSetDevice(device));

    bx_ctx *c = new (std::nothrow) bx_ctx();
    if (!c) return set_error(BX_ERR_ARG, "bx_create: out of host memory");
    c->prm = *prm;
    c->device = device;
    c->nx = prm->nx;
    c->ny = prm->ny;
    c->nz = prm->nz;
    c->n = n;
    c->flatz = prm->nz < 6;
    if (c->flatz) {
        c->prm.no_outflow[4] = 0;
        c->prm.no_outflow[5] = 0;
    }
    c->muc = 1.0 / (prm->xabun * 0.5 + prm->yabun * 0.25 + prm->zabun / prm->mu_z);  // eos.f90:144
    {
        const double pi = 3.1415926535897931;  // units.f90:10-17
        const double time = (3600.0 * 365.25 * 24.0) / (2.0 * pi);
        const double vel = 1.496e13 / time;
        c->rgas = 8.254e7 / ((vel * vel) / 1.0);
    }
    // calc_gforce with phi=0 (utils.f90:566-568): -(0-0)/(2*d) = -0.0
    c->grav[0] = -((0.0 - 0.0) / (2.0 * prm->dx));
    c->grav[1] = -((0.0 - 0.0) / (2.0 * prm->dy));
    c->grav[2] = -((0.0 - 0.0) / (2.0 * prm->dz));

    std::vector<int32_t> gd, gs, gg, gx;
    int rc = build_ghost_table(&c->prm, gd, gs, gg, gx);
    if (rc) {
        delete c;
        return rc;
    }
    c->nghost = (int64_t)gd.size();
    c->nscratch = (int64_t)gx.size();
    // The pair of calls composed: a ghost whose donor is the pre-call value of
    // ghost h gets h's value after the first call, which is h's donor when
    // that is an interior cell (sign flips compose).  When h's donor is the
    // pre-call value of ghost h2 the pair leaves h2's value from before it;
    // the fused step supports that when h2's own pair gives an interior cell
    // v (the corner ghosts of a box walled in y and z): such a ghost ends a
    // step as +-(v's corrector result before source(1)) and is fixed up after
    // the ghost fill.  Deeper chains take the exact sequence.
    std::vector<int32_t> fs(gs), fg(gg), d2dst, d2src, d2sign;
    c->fghost_ok = true;
    if (!gx.empty()) {
        std::vector<int64_t> entry((size_t)n, -1);
        for (size_t k = 0; k < gd.size(); k++) entry[(size_t)gd[k]] = (int64_t)k;
        // the pair's value of entry k: -1 with the interior cell and sign, or
        // the entry of the ghost whose pre-pair value it is (sign: its sign)
        auto pair_of = [&](size_t k, int32_t &cell, int32_t &sgn) -> int64_t {
            if (gs[k] >= 0) {
                cell = gs[k];
                sgn = gg[k];
                return -1;
            }
            const int64_t h = entry[(size_t)gx[(size_t)(-gs[k] - 1)]];
            if (h < 0) return -2;
            sgn = gg[k] ^ gg[(size_t)h];
            if (gs[(size_t)h] >= 0) {
                cell = gs[(size_t)h];
                return -1;
            }
            const int64_t h2 = entry[(size_t)gx[(size_t)(-gs[(size_t)h] - 1)]];
            return h2 < 0 ? -2 : h2;
        };
        for (size_t k = 0; k < gd.size() && c->fghost_ok; k++) {
            int32_t cell = 0, sgn = 0;
            const int64_t r = pair_of(k, cell, sgn);
            if (r == -1) {
                fs[k] = cell;
                fg[k] = sgn;
                continue;
            }
            int32_t cell2 = 0, sgn2 = 0;
            if (r < 0 || pair_of((size_t)r, cell2, sgn2) != -1) {
                c->fghost_ok = false;
                break;
            }
            fs[k] = cell2;
            fg[k] = sgn ^ sgn2;
            d2dst.push_back(gd[k]);
            d2src.push_back(cell2);
            d2sign.push_back(sgn ^ sgn2);
        }
        if ((int)d2dst.size() > kMaxD2) c->fghost_ok = false;
        if (!c->fghost_ok) d2dst.clear();
    }
    c->nd2 = (int)d2dst.size();

    auto fail = [&](hipError_t e, const char *what) {
        int r = hip_check(e, what);
        bx_destroy(c);
        return r;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(e, "hipStreamCreate");
    const size_t pb = sizeof(double) * (size_t)n;
    struct {
        double **ptr;
        size_t planes;
    } allocs[] = {{&c->cons, 5}, {&c->cons_old, 5}, {&c->u, 3}, {&c->p, 1},
                  {&c->adindx, 1}, {&c->qq, 3}};
    for (auto &a : allocs) {
        if ((e = hipMalloc((void **)a.ptr, pb * a.planes)) != hipSuccess) return fail(e, "hipMalloc state");
        if ((e = hipMemsetAsync(*a.ptr, 0, pb * a.planes, c->stream)) != hipSuccess) return fail(e, "hipMemset");
    }
    const size_t ng = std::max<size_t>(1, gd.size());
    if ((e = hipMalloc((void **)&c->ghost_dst, sizeof(int32_t) * ng)) != hipSuccess) return fail(e, "hipMalloc ghosts");
    if ((e = hipMalloc((void **)&c->ghost_src, sizeof(int32_t) * ng)) != hipSuccess) return fail(e, "hipMalloc ghosts");
    if ((e = hipMalloc((void **)&c->ghost_sign, sizeof(int32_t) * ng)) != hipSuccess) return fail(e, "hipMalloc ghosts");
    if (!gd.empty()) {
        if ((e = hipMemcpy(c->ghost_dst, gd.data(), sizeof(int32_t) * gd.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy ghosts");
        if ((e = hipMemcpy(c->ghost_src, gs.data(), sizeof(int32_t) * gs.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy ghosts");
        if ((e = hipMemcpy(c->ghost_sign, gg.data(), sizeof(int32_t) * gg.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy ghosts");
    }
    if (!gx.empty() && c->fghost_ok) {
        if ((e = hipMalloc((void **)&c->fghost_src, sizeof(int32_t) * ng)) != hipSuccess)
            return fail(e, "hipMalloc ghosts");
        if ((e = hipMalloc((void **)&c->fghost_sign, sizeof(int32_t) * ng)) != hipSuccess)
            return fail(e, "hipMalloc ghosts");
        if ((e = hipMemcpy(c->fghost_src, fs.data(), sizeof(int32_t) * fs.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy ghosts");
        if ((e = hipMemcpy(c->fghost_sign, fg.data(), sizeof(int32_t) * fg.size(), hipMemcpyHostToDevice)) !=
            hipSuccess)
            return fail(e, "hipMemcpy ghosts");
    }
    if (c->nd2 > 0) {
        const size_t b = sizeof(int32_t) * (size_t)c->nd2;
        if ((e = hipMalloc((void **)&c->d2_dst, b)) != hipSuccess) return fail(e, "hipMalloc corner ghosts");
        if ((e = hipMalloc((void **)&c->d2_src, b)) != hipSuccess) return fail(e, "hipMalloc corner ghosts");
        if ((e = hipMalloc((void **)&c->d2_sign, b)) != hipSuccess) return fail(e, "hipMalloc corner ghosts");
        if ((e = hipMalloc((void **)&c->d2_ubar, sizeof(double) * 6 * (size_t)c->nd2)) != hipSuccess)
            return fail(e, "hipMalloc corner ghosts");
        if ((e = hipMemcpy(c->d2_dst, d2dst.data(), b, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy corner ghosts");
        if ((e = hipMemcpy(c->d2_src, d2src.data(), b, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy corner ghosts");
        if ((e = hipMemcpy(c->d2_sign, d2sign.data(), b, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy corner ghosts");
    }
    if (!gx.empty()) {
        if ((e = hipMalloc((void **)&c->ghost_scratch_src, sizeof(int32_t) * gx.size())) != hipSuccess)
            return fail(e, "hipMalloc ghost scratch");
        if ((e = hipMalloc((void **)&c->ghost_scratch, sizeof(double) * 5 * gx.size())) != hipSuccess)
            return fail(e, "hipMalloc ghost scratch");
        if ((e = hipMemcpy(c->ghost_scratch_src, gx.data(), sizeof(int32_t) * gx.size(), hipMemcpyHostToDevice)) !=
            hipSuccess)
            return fail(e, "hipMemcpy ghost scratch");
    }
    if ((e = hipMalloc((void **)&c->dsc, sizeof(DevScalars))) != hipSuccess) return fail(e, "hipMalloc scalars");
    if ((e = hipMemsetAsync(c->dsc, 0, sizeof(DevScalars), c->stream)) != hipSuccess) return fail(e, "hipMemset");
    if ((e = hipHostMalloc((void **)&c->hsc, sizeof(DevScalars), hipHostMallocDefault)) != hipSuccess)
        return fail(e, "hipHostMalloc");
    std::memset(c->hsc, 0, sizeof(DevScalars));
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return fail(e, "hipStreamSynchronize");
    *out = c;
    return BX_OK;
}

int bx_destroy(bx_ctx *c)
{
    if (!c) return BX_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    adjoint_free(c);
    eos_free(c);
    gravity_free(c);
    drop_step_graphs(c);
    void *dev[] = {c->cons, c->cons_old, c->u, c->p, c->adindx, c->qq, c->states, c->ghost_dst, c->ghost_src,
                   c->ghost_sign, c->ghost_scratch_src, c->ghost_scratch, c->fghost_src, c->fghost_sign,
                   c->d2_dst, c->d2_src, c->d2_sign, c->d2_ubar, c->ustar, c->pstar, c->unext, c->dsc,
                   c->cons32, c->ustar32, c->unext32, c->pstar32, c->eflux, c->chunk0, c->chunk_dts, c->stage};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (c->hsc) (void)hipHostFree(c->hsc);
    for (int i = 0; i < 6; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BX_OK;
}

int bx_ghost_table(const bx_params *prm, int64_t *nghost, int64_t *nscratch, int32_t *dst, int32_t *src,
                   int32_t *sign, int32_t *scratch_src)
{
    if (!prm || !nghost || !nscratch) return set_error(BX_ERR_ARG, "bx_ghost_table: null argument");
    if (prm->nx < 5 || prm->ny < 5 || prm->nz < 1) return set_error(BX_ERR_ARG, "bx_ghost_table: bad grid");
    bx_params p = *prm;
    if (p.nz < 6) p.no_outflow[4] = p.no_outflow[5] = 0;
    std::vector<int32_t> gd, gs, gg, gx;
    int rc = build_ghost_table(&p, gd, gs, gg, gx);
    if (rc) return rc;
    *nghost = (int64_t)gd.size();
    *nscratch = (int64_t)gx.size();
    if (dst) std::copy(gd.begin(), gd.end(), dst);
    if (src) std::copy(gs.begin(), gs.end(), src);
    if (sign) std::copy(gg.begin(), gg.end(), sign);
    if (scratch_src) std::copy(gx.begin(), gx.end(), scratch_src);
    return BX_OK;
}

int bx_ngrid(const bx_ctx *c, int64_t *ngrid, int64_t *nghost)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_ngrid: null context");
    if (ngrid) *ngrid = c->n;
    if (nghost) *nghost = c->nghost;
    return BX_OK;
}

int bx_sync(bx_ctx *c)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_sync: null context");
    BX_HIP(hipSetDevice(c->device));
    BX_HIP(hipStreamSynchronize(c->stream));
    return BX_OK;
}

// Host arrays in the reference's layout, cons(5,n) etc. (field fastest), and
// the context's planes (one field after another): the transfer moves the
// caller's bytes as they are through a device staging buffer, and a kernel
// reorders them on the device.  Element e of a chunk of cells [n0, n0+cnt/k)
// is field e%k of cell n0+e/k.
__global__ __launch_bounds__(256) void k_rows_to_planes(int64_t n0, int64_t cnt, int k, int64_t n,
                                                        const double *__restrict__ rows, double *__restrict__ planes)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= cnt) return;
    const int64_t cell = e / k;
    planes[(int64_t)(e - cell * k) * n + n0 + cell] = rows[e];
}

__global__ __launch_bounds__(256) void k_planes_to_rows(int64_t n0, int64_t cnt, int k, int64_t n,
                                                        const double *__restrict__ planes, double *__restrict__ rows)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= cnt) return;
    const int64_t cell = e / k;
    rows[e] = planes[(int64_t)(e - cell * k) * n + n0 + cell];
}

constexpr int64_t kStageDoubles = int64_t(1) << 23;  // 64 MiB

static int ensure_stage(bx_ctx *c)
{
    if (c->stage) return BX_OK;
    const int64_t len = std::min<int64_t>(kStageDoubles, 5 * c->n);  // at least one cell of any k <= 5
    BX_HIP(hipMalloc((void **)&c->stage, sizeof(double) * (size_t)len));
    c->stage_len = len;
    return BX_OK;
}

// host array (k, n) -> k device planes
static int upload_aos(bx_ctx *c, const double *h, int k, double *d)
{
    const int64_t n = c->n;
    if (k == 1) {
        BX_HIP(hipMemcpyAsync(d, h, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
        BX_HIP(hipStreamSynchronize(c->stream));
        return BX_OK;
    }
    if (int rc = ensure_stage(c)) return rc;
    for (int64_t n0 = 0; n0 < n;) {
        const int64_t m = std::min<int64_t>(n - n0, c->stage_len / k), cnt = m * k;
        BX_HIP(hipMemcpyAsync(c->stage, h + n0 * k, sizeof(double) * (size_t)cnt, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_rows_to_planes, dim3(blocks_for(cnt, 256)), dim3(256), 0, c->stream, n0, cnt, k, n,
                           c->stage, d);
        BX_HIP(hipGetLastError());
        n0 += m;
    }
    BX_HIP(hipStreamSynchronize(c->stream));
    return BX_OK;
}

// k device planes -> host array (k, n)
static int download_aos(bx_ctx *c, const double *d, int k, double *h)
{
    const int64_t n = c->n;
    if (k == 1) {
        BX_HIP(hipMemcpyAsync(h, d, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        BX_HIP(hipStreamSynchronize(c->stream));
        return BX_OK;
    }
    if (int rc = ensure_stage(c)) return rc;
    for (int64_t n0 = 0; n0 < n;) {
        const int64_t m = std::min<int64_t>(n - n0, c->stage_len / k), cnt = m * k;
        hipLaunchKernelGGL(k_planes_to_rows, dim3(blocks_for(cnt, 256)), dim3(256), 0, c->stream, n0, cnt, k, n, d,
                           c->stage);
        BX_HIP(hipGetLastError());
        BX_HIP(hipMemcpyAsync(h + n0 * k, c->stage, sizeof(double) * (size_t)cnt, hipMemcpyDeviceToHost, c->stream));
        n0 += m;
    }
    BX_HIP(hipStreamSynchronize(c->stream));
    return BX_OK;
}

int bx_upload(bx_ctx *c, const double *cons, const double *u, const double *p, const double *adindx)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_upload: null context");
    BX_HIP(hipSetDevice(c->device));
    int rc;
    if (cons) {
        if ((rc = upload_aos(c, cons, 5, c->cons))) return rc;
        BX_HIP(hipMemcpyAsync(c->cons_old, c->cons, sizeof(double) * 5 * c->n, hipMemcpyDeviceToDevice, c->stream));
    }
    if (u && (rc = upload_aos(c, u, 3, c->u))) return rc;
    if (p && (rc = upload_aos(c, p, 1, c->p))) return rc;
    if (adindx && (rc = upload_aos(c, adindx, 1, c->adindx))) return rc;
    BX_HIP(hipStreamSynchronize(c->stream));
    c->fused_valid = false;
    c->aux_stale = false;
    c->old_stale = false;
    c->p_fused = false;
    if (cons) c->ghosts_filled = false;
    return BX_OK;
}

int bx_download(bx_ctx *c, double *cons, double *u, double *p, double *adindx)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_download: null context");
    BX_HIP(hipSetDevice(c->device));
    int rc;
    if (c->state_in_32) {
        if ((rc = widen_cons(c))) return rc;
        c->state_in_32 = false;
        c->aux_stale = true;
        c->fused_valid = false;
    }
    if (c->aux_stale && (u || p || adindx)) {
        if ((rc = launch_materialize(c))) return rc;
        c->aux_stale = false;
    }
    if (cons && (rc = download_aos(c, c->cons, 5, cons))) return rc;
    if (u && (rc = download_aos(c, c->u, 3, u))) return rc;
    if (p && (rc = download_aos(c, c->p, 1, p))) return rc;
    if (adindx && (rc = download_aos(c, c->adindx, 1, adindx))) return rc;
    return BX_OK;
}

int bx_download_cons_old(bx_ctx *c, double *cons_old)
{
    if (!c || !cons_old) return set_error(BX_ERR_ARG, "bx_download_cons_old: null argument");
    if (c->old_stale)
        return set_error(BX_ERR_STATE, "bx_download_cons_old: cons_old is not kept by the fused step");
    BX_HIP(hipSetDevice(c->device));
    return download_aos(c, c->cons_old, 5, cons_old);
}

// The per-subroutine entry points operate on the full reference state; after
// fused steps the velocity and pressure planes are first rebuilt from cons.
static int ensure_aux(bx_ctx *c)
{
    BX_HIP(hipSetDevice(c->device));
    if (c->state_in_32) {  // mixed form: bring the FP32 state back to the FP64 planes
        int rc = widen_cons(c);
        if (rc) return rc;
        c->state_in_32 = false;
        c->aux_stale = true;
    }
    if (c->aux_stale) {
        int rc = launch_materialize(c);
        if (rc) return rc;
        c->aux_stale = false;
    }
    c->fused_valid = false;
    return BX_OK;
}

int bx_state(bx_ctx *c)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_state: null context");
    int rc = ensure_aux(c);
    return rc ? rc : launch_state(c);
}

int bx_velocity(bx_ctx *c, double *maxv)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_velocity: null context");
    int rc = ensure_aux(c);
    if (rc || (rc = launch_velocity(c, maxv != nullptr))) return rc;
    if (maxv) {
        BX_HIP(hipMemcpyAsync(&c->hsc->maxv_bits, &c->dsc->maxv_bits, sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, c->stream));
        BX_HIP(hipStreamSynchronize(c->stream));
        unsigned long long b = c->hsc->maxv_bits;
        std::memcpy(maxv, &b, sizeof(double));
    }
    return BX_OK;
}

int bx_courant(bx_ctx *c, double *dt)
{
    if (!c || !dt) return set_error(BX_ERR_ARG, "bx_courant: null argument");
    int rc = ensure_aux(c);
    if (rc || (rc = launch_courant(c))) return rc;
    BX_HIP(hipMemcpyAsync(&c->hsc->rate_bits[0], &c->dsc->rate_bits[0], sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, c->stream));
    BX_HIP(hipStreamSynchronize(c->stream));
    double rate;
    unsigned long long b = c->hsc->rate_bits[0];
    std::memcpy(&rate, &b, sizeof(double));
    *dt = c->prm.cfl / rate;  // courant.f90:29
    return BX_OK;
}

int bx_source(bx_ctx *c, double tphase, int32_t siter, double dt)
{
    (void)tphase;  // used only by EXTERNALGRAV/ROTATE builds of the reference
    (void)siter;
    if (!c) return set_error(BX_ERR_ARG, "bx_source: null context");
    int rc = ensure_aux(c);
    return rc ? rc : launch_source(c, dt);
}

int bx_set_ghost_cells(bx_ctx *c)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_set_ghost_cells: null context");
    int rc = ensure_aux(c);
    return rc ? rc : launch_ghosts(c, c->cons);
}

int bx_flux(bx_ctx *c, int32_t avg, double dt)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_flux: null context");
    if (avg != 0 && avg != 1) return set_error(BX_ERR_ARG, "bx_flux: avg must be 0 or 1");
    int rc = ensure_aux(c);
    if (rc) return rc;
    if (!c->states) BX_HIP(hipMalloc((void **)&c->states, sizeof(double) * state_planes(c) * (size_t)c->n));
    return launch_flux(c, avg, dt);
}

int bx_copy_old(bx_ctx *c)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_copy_old: null context");
    int rc = ensure_aux(c);
    if (rc) return rc;
    BX_HIP(hipMemcpyAsync(c->cons_old, c->cons, sizeof(double) * 5 * c->n, hipMemcpyDeviceToDevice, c->stream));
    c->old_stale = false;
    return BX_OK;
}

int bx_avisc(bx_ctx *c, double *qq)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_avisc: null context");
    int rc = ensure_aux(c);
    if (rc || (rc = launch_avisc(c))) return rc;
    if (qq) return download_aos(c, c->qq, 3, qq);
    return BX_OK;
}

int bx_fallback_count(const bx_ctx *c, int64_t *count)
{
    if (!c || !count) return set_error(BX_ERR_ARG, "bx_fallback_count: null argument");
    *count = c->fallbacks;
    return BX_OK;
}

int bx_graph_step_count(const bx_ctx *c, int64_t *count)
{
    if (!c || !count) return set_error(BX_ERR_ARG, "bx_graph_step_count: null argument");
    *count = c->graph_steps;
    return BX_OK;
}

int bx_step_counts(const bx_ctx *c, int64_t *fused, int64_t *exact)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_step_counts: null context");
    if (fused) *fused = c->nfused;
    if (exact) *exact = c->nexact;
    return BX_OK;
}

}  // extern "C"

namespace bx {

// The second of each pair of set_ghost_cells calls; a no-op when every donor
// is an interior cell (one call then already gives the fixed point).
static int ghosts_again(bx_ctx *c) { return c->nscratch > 0 ? launch_ghosts(c, c->cons) : BX_OK; }

// One step of main.f90:211-555 with the reference's kernel sequence.
// dt logic (main.f90:266-281) runs on the host in the same IEEE operations.
static int exact_step(bx_ctx *c, double endtime, double *time, double *dtold, int32_t *step, double *dt_out)
{
    int rc;
    if (!c->states) BX_HIP(hipMalloc((void **)&c->states, sizeof(double) * state_planes(c) * (size_t)c->n));
    *step += 1;
    if ((rc = launch_state(c))) return rc;  // :232
    double dt;
    {
        if ((rc = launch_courant(c))) return rc;  // :241
        BX_HIP(hipMemcpyAsync(&c->hsc->rate_bits[0], &c->dsc->rate_bits[0], sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, c->stream));
        BX_HIP(hipStreamSynchronize(c->stream));
        double rate;
        unsigned long long b = c->hsc->rate_bits[0];
        std::memcpy(&rate, &b, sizeof(double));
        dt = c->prm.cfl / rate;
    }
    const double f1 = 1.0;
    dt = dt * f1;
    if (*step > 1 && *dtold > 0.0) dt = std::min(dt, *dtold * 1.1);
    *dtold = dt;
    if (dt + *time > endtime) dt = endtime - *time;
    dt = dt * 0.5;
    if ((rc = launch_source(c, dt))) return rc;  // :289 (+ ghosts, source.f90:199)
    if ((rc = ghosts_again(c))) return rc;       // :312
    if ((rc = launch_state(c))) return rc;       // :314
    if ((rc = launch_velocity(c, false))) return rc;  // :318
    BX_HIP(hipMemcpyAsync(c->cons_old, c->cons, sizeof(double) * 5 * c->n, hipMemcpyDeviceToDevice,
                          c->stream));  // :327-333
    dt = dt * 2.0;
    if ((rc = launch_flux(c, 0, dt))) return rc;  // :346 (+ ghosts, flux.f90:231)
    if ((rc = ghosts_again(c))) return rc;        // :351
    if ((rc = launch_state(c))) return rc;
    if ((rc = launch_velocity(c, false))) return rc;
    if ((rc = launch_flux(c, 1, dt))) return rc;  // :368 (+ ghosts, flux.f90:231)
    if ((rc = ghosts_again(c))) return rc;        // :374
    if ((rc = launch_state(c))) return rc;
    if ((rc = launch_velocity(c, false))) return rc;
    dt = 0.5 * dt;
    if ((c->grav_flags & BX_GRAV_POTENTIAL) && (rc = launch_calc_gforce(c))) return rc;   // :481
    if ((c->grav_flags & BX_GRAV_EXTERNAL) && (rc = launch_external_grav(c))) return rc;  // :487
    if ((rc = launch_source(c, dt))) return rc;  // :490 (+ ghosts, source.f90:199)
    if ((rc = ghosts_again(c))) return rc;       // :511
    if ((rc = launch_state(c))) return rc;
    if ((rc = launch_velocity(c, false))) return rc;
    dt = dt * 2.0;
    *time = *time + dt;
    *dt_out = dt;
    c->old_stale = false;
    c->p_fused = false;
    c->nexact++;
    return BX_OK;
}

}  // namespace bx

namespace bx {

static double rate_from_bits(unsigned long long b)
{
    double r;
    std::memcpy(&r, &b, sizeof(double));
    return r;
}

// main.f90:266-281 on the host (same IEEE operations as the reference)
static double step_dt(double rate, double cfl, double endtime, double time, double *dtold, int32_t step)
{
    double dt = cfl / rate;  // courant.f90:29
    dt = dt * 1.0;
    if (step > 1 && *dtold > 0.0) dt = std::min(dt, *dtold * 1.1);
    *dtold = dt;
    if (dt + time > endtime) dt = endtime - time;
    dt = dt * 0.5;
    return dt * 2.0;  // full dt of the flux stages (main.f90:342)
}

static int read_scalars(bx_ctx *c)
{
    BX_HIP(hipMemcpyAsync(c->hsc, c->dsc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
    BX_HIP(hipStreamSynchronize(c->stream));
    return BX_OK;
}

// Steps (an even number) that can be replayed from the current state with
// none of them reaching endtime: with dtold > 0 each dt is at most 1.1x the
// previous one (main.f90:271-273), so the times stay below these bounds (taken
// with a margin far above the rounding of the bound itself).  Then every step
// of the chunk is taken, none is clipped, and the loop condition holds before
// each.
static int chunk_steps(double time, double dtold, double endtime, int32_t step, int32_t left)
{
    if (!(dtold > 0.0) || !std::isfinite(dtold) || !std::isfinite(time) || step < 1) return 0;
    const double m = 1.0 + 1e-9;
    double t = time, d = dtold;
    int k = 0;
    while (k < kChunkSteps && k < left && step + k < 1000000) {
        d = d * 1.1 * m;
        t = (t + d) * m;
        if (!(t < endtime)) break;
        k++;
    }
    return k & ~1;
}

// K steps replayed from the captured pair (BX_STEP_GRAPH).  The state at the
// chunk's start is kept; if any step met an irregular branch the chunk is
// undone (*taken = 0) and the caller takes those steps one at a time.
static int graph_chunk(bx_ctx *c, int K, double endtime, double *time, double *dtold, int32_t *step, double *dts,
                       int32_t *taken)
{
    *taken = 0;
    const size_t bytes = (c->state_in_32 ? sizeof(float) : sizeof(double)) * 5 * (size_t)c->n;
    void *cur = c->state_in_32 ? (void *)c->cons32 : (void *)c->cons;
    if (!c->chunk0) BX_HIP(hipMalloc(&c->chunk0, sizeof(double) * 5 * (size_t)c->n));
    BX_HIP(hipMemcpyAsync(c->chunk0, cur, bytes, hipMemcpyDeviceToDevice, c->stream));
    const unsigned long long rate0 = c->hsc->rate_bits[0];  // Courant rate of the current state
    DevScalars *h = c->hsc;
    h->time = *time;
    h->dtold = *dtold;
    h->step = *step;
    h->endtime = endtime;
    h->flags = 0;
    h->logk = 0;
    BX_HIP(hipMemcpyAsync(c->dsc, h, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream));
    if (int rc = replay_step_pairs(c, K / 2)) return rc;
    double log[kChunkSteps];
    BX_HIP(hipMemcpyAsync(log, c->chunk_dts, sizeof(double) * K, hipMemcpyDeviceToHost, c->stream));
    if (int rc = read_scalars(c)) return rc;
    if ((h->flags & FLAG_IRREGULAR) || h->step != *step + K) {
        BX_HIP(hipMemcpyAsync(cur, c->chunk0, bytes, hipMemcpyDeviceToDevice, c->stream));
        h->rate_bits[0] = rate0;
        return BX_OK;
    }
    *time = h->time + h->dt;  // the last step's main.f90:545 (the device advanced time for the others)
    *dtold = h->dtold;
    *step = h->step;
    if (dts)
        for (int i = 0; i < K; i++) dts[i] = log[i];
    c->aux_stale = true;
    c->old_stale = true;
    c->p_fused = c->grav_flags != 0;
    c->graph_steps += K;
    c->nfused += K;
    *taken = K;
    return BX_OK;
}

// Fused stepping (see include/boxzy.h): returns after `nsteps` steps or
// endtime.  With `graph`, runs of steps that cannot reach endtime are replayed
// from a captured pair of steps, with dt formed on the device.
static int fused_loop(bx_ctx *c, int32_t nsteps, double endtime, double *time, double *dtold, int32_t *step,
                      double *dts, int32_t *taken, bool graph)
{
    int rc;
    int32_t k = 0;
    int32_t single = 0;  // steps to take one at a time (after an undone chunk or an irregular step)
    // While fused_valid holds, hsc->rate_bits[0] is the Courant rate of the
    // current state (from the regularity check or the last fused step).
    while (k < nsteps && *time < endtime && *step < 1000000) {
        if (!c->fused_valid) {
            // the state comes from the exact path or an upload: check it
            if (c->aux_stale) {
                if ((rc = launch_materialize(c))) return rc;
                c->aux_stale = false;
            }
            // with a potential, gforce must be the force of the current phi
            // (an exact step computes it), so that source(0) and source(1)
            // of every fused step see the same force
            bool ok = c->ghosts_filled && grav_settled(c);
            if (ok) {
                if ((rc = launch_regular_check(c))) return rc;
                if ((rc = read_scalars(c))) return rc;
                ok = c->hsc->flags == 0;
            }
            if (ok && (rc = grav_fused_planes(c))) return rc;
            if (!ok) {
                double dt;
                if ((rc = exact_step(c, endtime, time, dtold, step, &dt))) return rc;
                c->ghosts_filled = true;
                if (dts) dts[k] = dt;
                k++;
                continue;
            }
            c->fused_valid = true;
            if (c->prm.precision == BX_PREC_MIXED && !c->state_in_32) {
                if ((rc = narrow_cons(c))) return rc;
                c->state_in_32 = true;
            }
        }
        if (graph && single == 0 && !c->prof_on) {
            const int K = chunk_steps(*time, *dtold, endtime, *step, nsteps - k);
            if (K >= 2) {
                int32_t got = 0;
                if ((rc = graph_chunk(c, K, endtime, time, dtold, step, dts ? dts + k : nullptr, &got))) return rc;
                if (got) {
                    k += got;
                    continue;
                }
                single = K;
            }
        }
        if (single > 0) single--;
        const double rate = rate_from_bits(c->hsc->rate_bits[0]);
        const double time0 = *time, dtold0 = *dtold;
        const int32_t step0 = *step;
        *step += 1;
        const double dt = step_dt(rate, c->prm.cfl, endtime, *time, dtold, *step);
        if ((rc = fused_step(c, dt))) return rc;
        if ((rc = read_scalars(c))) return rc;
        if (c->hsc->flags & FLAG_IRREGULAR) {
            // an irregular branch: redo this step with the exact sequence from U^n
            // (replays resume after a full chunk's worth of regular steps)
            c->fallbacks++;
            if (graph) single = kChunkSteps;
            *time = time0;
            *dtold = dtold0;
            *step = step0;
            if (c->state_in_32) {
                if ((rc = widen_cons(c))) return rc;
                c->state_in_32 = false;
            }
            if ((rc = launch_materialize(c))) return rc;
            c->aux_stale = false;
            c->fused_valid = false;
            double dtx;
            if ((rc = exact_step(c, endtime, time, dtold, step, &dtx))) return rc;
            if (dts) dts[k] = dtx;
            k++;
            continue;
        }
        if (c->prof_on) {
            const int pairs[4][2] = {{0, 1}, {1, 2}, {3, 4}, {4, 5}};
            for (int q = 0; q < 4; q++) {
                float t = 0.f;
                BX_HIP(hipEventElapsedTime(&t, c->ev[pairs[q][0]], c->ev[pairs[q][1]]));
                c->prof_ms[q] += t;
            }
            c->prof_steps++;
        }
        if (c->state_in_32) std::swap(c->cons32, c->unext32);
        else std::swap(c->cons, c->unext);
        c->nfused++;
        c->aux_stale = true;
        c->p_fused = c->grav_flags != 0;
        c->old_stale = true;
        *time = *time + dt;  // main.f90:544-545 (dt*half*two == dt)
        if (dts) dts[k] = dt;
        k++;
    }
    *taken = k;
    return BX_OK;
}

}  // namespace bx

extern "C" int bx_step(bx_ctx *c, int32_t mode, int32_t nsteps, double endtime, double *time, double *dtold,
                       int32_t *step, double *dts, int32_t *taken)
{
    if (!c || !time || !dtold || !step) return set_error(BX_ERR_ARG, "bx_step: null argument");
    if (mode != BX_STEP_EXACT && mode != BX_STEP_FUSED && mode != BX_STEP_GRAPH)
        return set_error(BX_ERR_ARG, "bx_step: unknown mode");
    BX_HIP(hipSetDevice(c->device));
    if (c->prm.precision != BX_PREC_FP64 && mode == BX_STEP_EXACT)
        return set_error(BX_ERR_UNSUPPORTED, "bx_step: the mixed form runs in fused mode only");
    const bool graph = mode == BX_STEP_GRAPH;
    if (graph) mode = BX_STEP_FUSED;
    int32_t k = 0;
    int rc = BX_OK;
    if (mode == BX_STEP_FUSED && !c->fused_ready && !c->fused_unsupported) {
        rc = fused_prepare(c);
        if (rc == BX_ERR_UNSUPPORTED && c->prm.precision != BX_PREC_FP64) return rc;
        if (rc == BX_ERR_UNSUPPORTED) c->fused_unsupported = true;  // deep ghost chains: exact steps
        else if (rc) return rc;
        else c->fused_ready = true;
        rc = BX_OK;
    }
    if (mode == BX_STEP_FUSED && c->fused_ready) {
        rc = fused_loop(c, nsteps, endtime, time, dtold, step, dts, &k, graph);
    } else {
        if ((rc = ensure_aux(c))) return rc;
        while (k < nsteps && *time < endtime && *step < 1000000) {
            double dt;
            if ((rc = exact_step(c, endtime, time, dtold, step, &dt))) break;
            c->ghosts_filled = true;
            if (dts) dts[k] = dt;
            k++;
        }
    }
    if (taken) *taken = k;
    if (rc) return rc;
    BX_HIP(hipStreamSynchronize(c->stream));
    return BX_OK;
}

extern "C" int bx_tune(bx_ctx *c, int32_t y_chunk, int32_t z_chunk, int32_t occupancy)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_tune: null context");
    if (y_chunk < 1 || z_chunk < 1) return set_error(BX_ERR_ARG, "bx_tune: chunks must be >= 1");
    if (occupancy != 0 && occupancy != 1 && occupancy != 4 && !(occupancy >= 10 && occupancy <= 11))
        return set_error(BX_ERR_ARG, "bx_tune: variant 0, 1, 4, 10 or 11");
    if (c->grav_flags && occupancy != 11)
        return set_error(BX_ERR_UNSUPPORTED, "bx_tune: with gravity the fused step has the 64 x 8 tiled kernel only "
                                             "(variant 11)");
    c->fused_ly = y_chunk;
    c->fused_lz = z_chunk;
    c->fused_occ = occupancy;
    drop_step_graphs(c);
    return BX_OK;
}

extern "C" int bx_profile(bx_ctx *c, int32_t enable, double *kernel_ms, int64_t *steps)
{
    if (!c) return set_error(BX_ERR_ARG, "bx_profile: null context");
    BX_HIP(hipSetDevice(c->device));
    if (enable && !c->ev[0])
        for (int i = 0; i < 6; i++) BX_HIP(hipEventCreate(&c->ev[i]));
    if (kernel_ms)
        for (int q = 0; q < 4; q++) kernel_ms[q] = c->prof_ms[q];
    if (steps) *steps = c->prof_steps;
    if (enable >= 0) {
        c->prof_on = enable != 0;
        for (int q = 0; q < 4; q++) c->prof_ms[q] = 0.0;
        c->prof_steps = 0;
    }
    return BX_OK;
}
