/*
 * boxzy_oracle.c -- CPU restatement of the BOXZYHYDRO hydro step (oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py, to check against.  The product path never calls it.
 *
 * It restates, in C with OpenMP, the reference's Fortran subroutines with the
 * same arithmetic order, the same Fortran-ordered arrays (cons(5,ngrid) ...),
 * the same neighbour tables and the same pass structure, so that
 *   (a) it is bit-identical to the reference built from its own sources
 *       (checked in tests against golden vectors from oracle/_ref), and
 *   (b) its wall time is representative of the reference's OpenMP step.
 * Build flags must keep IEEE semantics: -O2 -ffp-contract=off, no fast-math,
 * baseline x86-64 (no FMA), exactly like the reference's gfortran -O3 build.
 *
 * Reference files (relative to the reference root): see each function.
 */
#include "boxzy_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HALF 0.5
#define SLOPE_THETA 1.25 /* Makefile:20 -DSLOPE_THETA=1.25 */

struct bxo_grid {
    bxo_params prm;
    int nx, ny, nz, ngrid;
    double muc, rgas;
    /* geometry (init_grid.f90:76-264); ids are 0-based, -1 = unset (ref: 0) */
    int *ix, *iy, *iz, *boundary;
    int *ineigh; /* 6 per cell: +y, -y, +x, -x, +z, -z */
    int nghost, nbound;
    int *indx_ghost, *indx_bound;
    /* caller-owned state */
    double *cons, *cons_old, *u, *p, *adindx, *qq;
    /* owned work arrays */
    double *fluxtmp, *gforce, *pforce;
    double *slope_u, *slope_p, *slope_d, *slope_e, *slope_g;
    double *state_u_p, *state_u_m;
    double *state_p_p, *state_p_m, *state_d_p, *state_d_m;
    double *state_e_p, *state_e_m, *state_g_p, *state_g_m;
};

static inline double fmax2(double a, double b) { return a > b ? a : b; }
static inline double fmin2(double a, double b) { return a < b ? a : b; }

/* utils.f90:191-194 get_id (1-based ix,iy,iz -> 0-based id) */
static inline int get_id(const bxo_grid *g, int ix, int iy, int iz)
{
    return (ix - 1) + (iy - 1) * g->nx + (iz - 1) * g->nx * g->ny;
}

/* units.f90:10-17: scale%rgas = rgasCGS/(vel**2/kelvin) */
static double units_rgas(void)
{
    const double pi = 3.1415926535897931;
    double length = 1.496e13;
    double time = (3600.0 * 365.25 * 24.0) / (2.0 * pi);
    double vel = length / time;
    double kelvin = 1.0;
    return 8.254e7 / ((vel * vel) / kelvin);
}

bxo_grid *bxo_create(const bxo_params *prm)
{
    bxo_grid *g = (bxo_grid *)calloc(1, sizeof(bxo_grid));
    g->prm = *prm;
    g->nx = prm->nx;
    g->ny = prm->ny;
    g->nz = prm->nz;
    int nx = g->nx, ny = g->ny, nz = g->nz;
    int n = nx * ny * nz;
    g->ngrid = n;
    /* init_grid.f90:76-79: no z walls on a degenerate z axis */
    if (nz < 6) {
        g->prm.no_outflow[4] = 0;
        g->prm.no_outflow[5] = 0;
    }
    /* eos.f90:144 */
    g->muc = 1.0 / (prm->xabun * HALF + prm->yabun * 0.25 + prm->zabun / prm->mu_z);
    g->rgas = units_rgas();

    g->ix = (int *)malloc(sizeof(int) * n);
    g->iy = (int *)malloc(sizeof(int) * n);
    g->iz = (int *)malloc(sizeof(int) * n);
    g->boundary = (int *)malloc(sizeof(int) * n);
    g->ineigh = (int *)malloc(sizeof(int) * 6 * (size_t)n);

    /* init_grid.f90:81-118 boundary flags */
    int nghost = 0, nbound = 0;
    for (int iz = 1; iz <= nz; iz++) {
        int f1z = 0, f2z = 0;
        if (nz >= 6) {
            if (iz == 1 || iz == nz) f1z = 1;
            if (iz == 2 || iz == nz - 1) f2z = 1;
        }
        for (int iy = 1; iy <= ny; iy++) {
            int f1y = (iy == 1 || iy == ny), f2y = (iy == 2 || iy == ny - 1);
            for (int ix = 1; ix <= nx; ix++) {
                int f1x = (ix == 1 || ix == nx), f2x = (ix == 2 || ix == nx - 1);
                int c = get_id(g, ix, iy, iz);
                g->ix[c] = ix;
                g->iy[c] = iy;
                g->iz[c] = iz;
                int b = 0;
                if (f1x + f1y + f1z > 0) {
                    b = 1;
                    nbound++;
                    nghost++;
                } else if (f2x + f2y + f2z > 0) {
                    b = 2;
                    nghost++;
                }
                g->boundary[c] = b;
            }
        }
    }
    g->nghost = nghost;
    g->nbound = nbound;
    g->indx_ghost = (int *)malloc(sizeof(int) * (nghost > 0 ? nghost : 1));
    g->indx_bound = (int *)malloc(sizeof(int) * (nbound > 0 ? nbound : 1));

    /* init_grid.f90:166-264 neighbour and donor tables */
    int ib = 0, ig = 0;
    for (int c = 0; c < n; c++) {
        int *nb = g->ineigh + 6 * (size_t)c;
        int ix = g->ix[c], iy = g->iy[c], iz = g->iz[c];
        if (g->boundary[c] == 1) {
            g->indx_bound[ib++] = c;
            g->indx_ghost[ig++] = c;
            for (int k = 0; k < 6; k++) nb[k] = -1;
            if (ix == 1) nb[0] = c + 2;
            else if (ix == nx) nb[0] = c - 2;
            else if (iy == 1) nb[0] = c + 2 * nx;
            else if (iy == ny) nb[0] = c - 2 * nx;
            else if (iz == 1) nb[0] = (nz < 6) ? c : c + 2 * nx * ny;
            else if (iz == nz) nb[0] = (nz < 6) ? c : c - 2 * nx * ny;
            else abort(); /* init_grid.f90:200-202 */
        } else {
            if (g->boundary[c] == 2) g->indx_ghost[ig++] = c;
            nb[0] = c + nx;
            nb[1] = c - nx;
            nb[2] = c + 1;
            nb[3] = c - 1;
            if (nz < 6) {
                nb[4] = c;
                nb[5] = c;
            } else {
                nb[4] = c + nx * ny;
                nb[5] = c - nx * ny;
            }
        }
    }

    size_t N = (size_t)n;
    g->fluxtmp = (double *)calloc(5 * N, sizeof(double));
    g->gforce = (double *)calloc(3 * N, sizeof(double));
    g->pforce = (double *)calloc(3 * N, sizeof(double));
    g->slope_u = (double *)calloc(9 * N, sizeof(double));
    g->state_u_p = (double *)calloc(9 * N, sizeof(double));
    g->state_u_m = (double *)calloc(9 * N, sizeof(double));
    double **three[] = {&g->slope_p, &g->slope_d, &g->slope_e, &g->slope_g,
                        &g->state_p_p, &g->state_p_m, &g->state_d_p, &g->state_d_m,
                        &g->state_e_p, &g->state_e_m, &g->state_g_p, &g->state_g_m};
    for (unsigned k = 0; k < sizeof(three) / sizeof(three[0]); k++)
        *three[k] = (double *)calloc(3 * N, sizeof(double));

    /* utils.f90:549-571 calc_gforce with phi=0 (init_conditions.f90:29,58):
     * gforce = -(0-0)/(2*d) = -0.0 on interior cells, 0 elsewhere. */
    for (int c = 0; c < n; c++) {
        if (g->boundary[c] > 0) continue;
        double phi = 0.0;
        g->gforce[3 * c + 0] = -((phi - phi) / (2.0 * prm->dx));
        g->gforce[3 * c + 1] = -((phi - phi) / (2.0 * prm->dy));
        g->gforce[3 * c + 2] = -((phi - phi) / (2.0 * prm->dz));
    }
    return g;
}

void bxo_destroy(bxo_grid *g)
{
    if (!g) return;
    void *ptrs[] = {g->ix, g->iy, g->iz, g->boundary, g->ineigh, g->indx_ghost, g->indx_bound,
                    g->fluxtmp, g->gforce, g->pforce, g->slope_u, g->slope_p, g->slope_d,
                    g->slope_e, g->slope_g, g->state_u_p, g->state_u_m, g->state_p_p,
                    g->state_p_m, g->state_d_p, g->state_d_m, g->state_e_p, g->state_e_m,
                    g->state_g_p, g->state_g_m};
    for (unsigned k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
    free(g);
}

void bxo_bind(bxo_grid *g, double *cons, double *cons_old, double *u, double *p, double *adindx,
              double *qq)
{
    g->cons = cons;
    g->cons_old = cons_old;
    g->u = u;
    g->p = p;
    g->adindx = adindx;
    g->qq = qq;
}

int bxo_ngrid(const bxo_grid *g) { return g->ngrid; }
int bxo_nghost(const bxo_grid *g) { return g->nghost; }
double bxo_muc(const bxo_grid *g) { return g->muc; }
double bxo_rgas(const bxo_grid *g) { return g->rgas; }

void bxo_geometry(const bxo_grid *g, int *boundary, int *ineigh)
{
    for (int c = 0; c < g->ngrid; c++) {
        boundary[c] = g->boundary[c];
        for (int k = 0; k < 6; k++) ineigh[6 * c + k] = g->ineigh[6 * c + k] + 1;
    }
}

void bxo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int bxo_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------- state */
/* state.f90:29-87, H2STAT=-1 branch (ideal gas, low-density temperature cap) */
void bxo_state(bxo_grid *g)
{
    const double gam = g->prm.gammafix, seps = g->prm.small_eps;
    const double muc = g->muc, rgas = g->rgas;
    double *cons = g->cons, *u = g->u, *p = g->p, *ad = g->adindx;
    const int n = g->ngrid;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
        double *U = cons + 5 * (size_t)c;
        const double *v = u + 3 * (size_t)c;
        double ekin = U[0] * HALF * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); /* :127 */
        double eps = fmax2(U[4] - ekin, seps);                                  /* :128 */
        ad[c] = gam;                                                            /* :130 */
        if (U[0] < 1e-4) {                                                      /* :132 */
            double tk = fmin2(5e3, p[c] / (U[0] * rgas) * muc);                /* :133 */
            p[c] = U[0] * tk * rgas / muc;                                      /* :134 */
            eps = p[c] / (gam - 1.0);                                           /* :135 */
            U[4] = eps + ekin;                                                  /* :136 */
        } else {
            p[c] = eps * (gam - 1.0);                                           /* :138 */
        }
    }
}

/* ------------------------------------------------------------- velocity */
/* velocity.f90:18-53 (u=m/rho, then the vlimit speed clamp) */
void bxo_velocity(bxo_grid *g)
{
    double *cons = g->cons, *u = g->u;
    const double vlimit = g->prm.vlimit, seps = g->prm.small_eps;
    const int n = g->ngrid;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
        const double *U = cons + 5 * (size_t)c;
        double *v = u + 3 * (size_t)c;
        for (int d = 0; d < 3; d++) v[d] = U[d + 1] / U[0]; /* :21 */
    }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
        double *U = cons + 5 * (size_t)c;
        double *v = u + 3 * (size_t)c;
        int flag = 0;
        for (int d = 0; d < 3; d++)
            if (fabs(v[d]) > vlimit) flag = 1; /* :33 */
        if (flag) {
            double ekin = HALF * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]; /* :37 */
            double speed = 0.0;
            for (int d = 0; d < 3; d++) speed = speed + v[d] * v[d];
            speed = sqrt(speed);
            for (int d = 0; d < 3; d++) {
                v[d] = v[d] * vlimit / speed;
                U[d + 1] = v[d] * U[0];
            }
            U[4] = U[4] - ekin;
            ekin = HALF * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
            U[4] = fmax2(U[4], seps) + ekin; /* :49 */
        }
    }
}

/* -------------------------------------------------------------- courant */
/* courant.f90:17-29: dtrate = max over ALL cells; dt = cfl/dtrate */
double bxo_courant(bxo_grid *g)
{
    const double *cons = g->cons, *u = g->u, *p = g->p, *ad = g->adindx;
    const double dx = g->prm.dx, dy = g->prm.dy, dz = g->prm.dz;
    const int n = g->ngrid;
    double dtrate = 0.0;
#pragma omp parallel for schedule(static) reduction(max : dtrate)
    for (int c = 0; c < n; c++) {
        const double *v = u + 3 * (size_t)c;
        double sound = sqrt(ad[c] * p[c] / cons[5 * (size_t)c]);
        double r0 = (sound + fabs(v[0])) / dx + (sound + fabs(v[1])) / dy + (sound + fabs(v[2])) / dz;
        dtrate = fmax2(dtrate, r0);
    }
    return g->prm.cfl / dtrate;
}

/* ------------------------------------------------------- set_ghost_cells */
/* utils.f90:200-281 check_corner, including its iz>=6 test at line 215 */
static int check_corner(const bxo_grid *g, int ix, int iy, int iz, int ineigh)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    if (ix < 3) {
        if (iy < 3) {
            if (nz >= 6) {
                if (iz < 3) ineigh = get_id(g, 3, 3, 3);
                else if (iz > nz - 2) ineigh = get_id(g, 3, 3, nz - 2);
                else ineigh = get_id(g, 3, 3, iz);
            } else ineigh = get_id(g, 3, 3, iz);
        } else if (iy > ny - 2) {
            if (iz >= 6) {
                if (iz < 3) ineigh = get_id(g, 3, ny - 2, 3);
                else if (iz > nz - 2) ineigh = get_id(g, 3, ny - 2, nz - 2);
                else ineigh = get_id(g, 3, ny - 2, iz);
            } else ineigh = get_id(g, 3, ny - 2, iz);
        } else {
            if (nz >= 6) {
                if (iz < 3) ineigh = get_id(g, 3, iy, 3);
                else if (iz > nz - 2) ineigh = get_id(g, 3, iy, nz - 2);
            } else ineigh = get_id(g, 3, iy, iz);
        }
    } else if (ix > nx - 2) {
        if (iy < 3) {
            if (nz >= 6) {
                if (iz < 3) ineigh = get_id(g, nx - 2, 3, 3);
                else if (iz > nz - 2) ineigh = get_id(g, nx - 2, 3, nz - 2);
                else ineigh = get_id(g, nx - 2, 3, iz);
            } else ineigh = get_id(g, nx - 2, 3, iz);
        } else if (iy > ny - 2) {
            if (nz >= 6) {
                if (iz < 3) ineigh = get_id(g, nx - 2, ny - 2, 3);
                else if (iz > nz - 2) ineigh = get_id(g, nx - 2, ny - 2, nz - 2);
                else ineigh = get_id(g, nx - 2, ny - 2, iz);
            } else ineigh = get_id(g, nx - 2, ny - 2, iz);
        } else {
            if (nz >= 6) {
                if (iz < 3) ineigh = get_id(g, nx - 2, iy, 3);
                else if (iz > nz - 2) ineigh = get_id(g, nx - 2, iy, nz - 2);
            } else ineigh = get_id(g, nx - 2, iy, iz);
        }
    } else {
        if (iy < 3) {
            if (nz >= 6) {
                if (iz < 3) ineigh = get_id(g, ix, 3, 3);
                else if (iz > nz - 2) ineigh = get_id(g, ix, 3, nz - 2);
            } else ineigh = get_id(g, ix, 3, iz);
        } else if (iy > ny - 2) {
            if (nz >= 6) {
                if (iz < 3) ineigh = get_id(g, ix, ny - 2, 3);
                else if (iz > nz - 2) ineigh = get_id(g, ix, ny - 2, nz - 2);
            } else ineigh = get_id(g, ix, ny - 2, iz);
        }
    }
    return ineigh;
}

/* utils.f90:287-296 */
static inline void ghost_helper(double *cons, int idx, int ip, const double fac[3])
{
    double *d = cons + 5 * (size_t)idx;
    const double *s = cons + 5 * (size_t)ip;
    d[0] = s[0];
    d[1] = s[1] * fac[0];
    d[2] = s[2] * fac[1];
    d[3] = s[3] * fac[2];
    d[4] = s[4];
}

/* utils.f90:314-403.  The reference loops with OMP DO; every call site calls
 * it twice in a row, which makes the result equal to that of this serial
 * loop in ghost-list order, so the serial order is used here.              */
void bxo_set_ghost_cells(bxo_grid *g)
{
    double *cons = g->cons;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int *nof = g->prm.no_outflow;
    int ineigh = -1;
    for (int k = 0; k < g->nghost; k++) {
        int idx = g->indx_ghost[k];
        int ix = g->ix[idx], iy = g->iy[idx], iz = g->iz[idx];
        const int *nb = g->ineigh + 6 * (size_t)idx;
        if (g->boundary[idx] == 1) ineigh = nb[0];
        else {
            if (ix == 2) ineigh = nb[2];
            else if (ix == nx - 1) ineigh = nb[3];
            else if (iy == 2) ineigh = nb[0];
            else if (iy == ny - 1) ineigh = nb[1];
            else if (iz == 2) ineigh = nb[4];
            else if (iz == nz - 1) ineigh = nb[5];
        }
        ineigh = check_corner(g, ix, iy, iz, ineigh);
        for (int d = 0; d < 5; d++) cons[5 * (size_t)idx + d] = cons[5 * (size_t)ineigh + d];

        const int plane = nx * ny;
        double fac[3];
        if (nof[0]) { /* :338-348 */
            fac[0] = -1.0; fac[1] = 1.0; fac[2] = 1.0;
            if (ix == 1) ghost_helper(cons, idx, plane * (iz - 1) + nx * (iy - 1) + ix + 3 - 1, fac);
            else if (ix == 2) ghost_helper(cons, idx, plane * (iz - 1) + nx * (iy - 1) + ix + 1 - 1, fac);
        }
        if (nof[1]) { /* :349-359 */
            fac[0] = -1.0; fac[1] = 1.0; fac[2] = 1.0;
            if (ix == nx) ghost_helper(cons, idx, plane * (iz - 1) + nx * (iy - 1) + ix - 3 - 1, fac);
            else if (ix == nx - 1) ghost_helper(cons, idx, plane * (iz - 1) + nx * (iy - 1) + ix - 1 - 1, fac);
        }
        if (nof[2]) { /* :360-370 */
            fac[0] = 1.0; fac[1] = -1.0; fac[2] = 1.0;
            if (iy == 1) ghost_helper(cons, idx, plane * (iz - 1) + nx * ((iy + 3) - 1) + ix - 1, fac);
            else if (iy == 2) ghost_helper(cons, idx, plane * (iz - 1) + nx * ((iy + 1) - 1) + ix - 1, fac);
        }
        if (nof[3]) { /* :371-381 */
            fac[0] = 1.0; fac[1] = -1.0; fac[2] = 1.0;
            if (iy == ny) ghost_helper(cons, idx, plane * (iz - 1) + nx * ((iy - 3) - 1) + ix - 1, fac);
            else if (iy == ny - 1) ghost_helper(cons, idx, plane * (iz - 1) + nx * ((iy - 1) - 1) + ix - 1, fac);
        }
        if (nof[4]) { /* :382-392 */
            fac[0] = 1.0; fac[1] = 1.0; fac[2] = -1.0;
            if (iz == 1) ghost_helper(cons, idx, plane * (iz + 3 - 1) + nx * (iy - 1) + ix - 1, fac);
            else if (iz == 2) ghost_helper(cons, idx, plane * (iz + 1 - 1) + nx * (iy - 1) + ix - 1, fac);
        }
        if (nof[5]) { /* :393-403 */
            fac[0] = 1.0; fac[1] = 1.0; fac[2] = -1.0;
            if (iz == nz) ghost_helper(cons, idx, plane * (iz - 3 - 1) + nx * (iy - 1) + ix - 1, fac);
            else if (iz == nz - 1) ghost_helper(cons, idx, plane * (iz - 1 - 1) + nx * (iy - 1) + ix - 1, fac);
        }
    }
}

/* --------------------------------------------------------------- source */
/* source.f90:53-199 with every optional-physics flag off */
void bxo_source(bxo_grid *g, double dt)
{
    double *cons = g->cons, *u = g->u, *p = g->p, *gf = g->gforce, *pf = g->pforce;
    const double dx = g->prm.dx, dy = g->prm.dy, dz = g->prm.dz, seps = g->prm.small_eps;
    const int n = g->ngrid;
    /* :58-105 pressure force (only consumed by particles; kept for pass fidelity) */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
        if (g->boundary[c] > 0) continue;
        const int *b = g->ineigh + 6 * (size_t)c; /* get_boundary_noanchor: no anchors */
        double r = cons[5 * (size_t)c];
        pf[3 * c + 0] = -((p[b[2]] - p[b[3]]) / (r * 2.0 * dx));
        pf[3 * c + 1] = -((p[b[0]] - p[b[1]]) / (r * 2.0 * dy));
        pf[3 * c + 2] = -((p[b[4]] - p[b[5]]) / (r * 2.0 * dz));
    }
    /* :148-191 gravity momentum and energy update */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
        if (g->boundary[c] > 0) continue;
        double *U = cons + 5 * (size_t)c;
        const double *G = gf + 3 * (size_t)c;
        double ux = u[3 * c], uy = u[3 * c + 1], uz = u[3 * c + 2];
        U[2] = U[2] + (U[0] * G[1]) * dt;
        U[1] = U[1] + (U[0] * G[0]) * dt;
        U[3] = U[3] + (G[2] * U[0]) * dt;
        double ekin = HALF * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
        U[4] = fmax2(U[4] - ekin + dt * U[0] * (ux * G[0] + uy * G[1] + uz * G[2]), seps) + ekin;
    }
    bxo_set_ghost_cells(g); /* :199 */
}

/* ------------------------------------------------------------------ flux */
/* utils.f90:696-707 */
static inline double genminmod(double x1, double x2, double x3)
{
    if ((x1 > 0.0) && ((x2 > 0.0) && (x3 > 0.0))) {
        double tmp = fmin2(x2, x3);
        return fmin2(x1, tmp);
    } else if ((x1 < 0.0) && ((x2 < 0.0) && (x3 < 0.0))) {
        double tmp = fmax2(x2, x3);
        return fmax2(x1, tmp);
    }
    return 0.0;
}

/* utils.f90:738-741 */
static inline double slope(double c0, double c1, double c2)
{
    const double tt = SLOPE_THETA;
    return genminmod(tt * (c1 - c0), HALF * (c2 - c0), tt * (c2 - c1));
}

/* flux.f90:253-295, idir 0..2 */
static void calculate_slopes(bxo_grid *g, const int isten[3], int idir)
{
    const double *cons = g->cons, *u = g->u;
    const int i1 = isten[0], i2 = isten[1], i3 = isten[2];
    for (int idim = 0; idim < 3; idim++)
        g->slope_u[idir + 3 * idim + 9 * (size_t)i2] =
            slope(u[3 * (size_t)i1 + idim], u[3 * (size_t)i2 + idim], u[3 * (size_t)i3 + idim]);
    double eps[3];
    for (int k = 0; k < 3; k++) {
        const double *U = cons + 5 * (size_t)isten[k];
        const double *v = u + 3 * (size_t)isten[k];
        eps[k] = fmax2(U[4] - HALF * U[0] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), g->prm.small_eps);
    }
    size_t o = idir + 3 * (size_t)i2;
    g->slope_p[o] = slope(g->p[i1], g->p[i2], g->p[i3]);
    g->slope_d[o] = slope(cons[5 * (size_t)i1], cons[5 * (size_t)i2], cons[5 * (size_t)i3]);
    g->slope_e[o] = slope(eps[0], eps[1], eps[2]);
    g->slope_g[o] = slope(g->adindx[i1], g->adindx[i2], g->adindx[i3]);
}

/* flux.f90:299-308 */
static inline void interface_muscl(double q, double lm, double lp, double s, double *minus, double *plus)
{
    *minus = q + HALF * (-1.0 - lm) * s;
    *plus = q + HALF * (1.0 - lp) * s;
}

/* flux.f90:312-482 (loops over all three directions for cell isten(2)) */
static void calculate_states(bxo_grid *g, int idx, double dt)
{
    const double *cons = g->cons, *u = g->u, *p = g->p, *ad = g->adindx;
    const double *U = cons + 5 * (size_t)idx;
    const double *v = u + 3 * (size_t)idx;
    double sound = sqrt(p[idx] * ad[idx] / U[0]);
    double ds[3] = {g->prm.dx, g->prm.dy, g->prm.dz};
    for (int d = 0; d < 3; d++) {
        double lambda_p = fmax2(v[d] + sound, 0.0) * dt / ds[d];
        double lambda_m = fmin2(v[d] - sound, 0.0) * dt / ds[d];
        for (int j = 0; j < 3; j++) {
            size_t o = d + 3 * j + 9 * (size_t)idx;
            interface_muscl(v[j], lambda_m, lambda_p, g->slope_u[o], &g->state_u_m[o], &g->state_u_p[o]);
        }
        double eps = fmax2(U[4] - HALF * U[0] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), g->prm.small_eps);
        size_t o = d + 3 * (size_t)idx;
        interface_muscl(p[idx], lambda_m, lambda_p, g->slope_p[o], &g->state_p_m[o], &g->state_p_p[o]);
        interface_muscl(U[0], lambda_m, lambda_p, g->slope_d[o], &g->state_d_m[o], &g->state_d_p[o]);
        interface_muscl(eps, lambda_m, lambda_p, g->slope_e[o], &g->state_e_m[o], &g->state_e_p[o]);
        interface_muscl(ad[idx], lambda_m, lambda_p, g->slope_g[o], &g->state_g_m[o], &g->state_g_p[o]);
    }
    /* :391-479 transverse terms: identically zero, added as 0*cc */
    const double zero = 0.0;
    for (int d = 0; d < 3; d++) {
        double cc = dt * HALF / ds[d];
        size_t o = d + 3 * (size_t)idx;
        g->state_d_p[o] = g->state_d_p[o] + zero * cc;
        for (int j = 0; j < 3; j++) g->state_u_p[d + 3 * j + 9 * (size_t)idx] += zero * cc;
        g->state_p_p[o] = g->state_p_p[o] + zero * cc;
        g->state_g_p[o] = g->state_g_p[o] + zero * cc;
        g->state_e_p[o] = g->state_e_p[o] + zero * cc;
        g->state_d_m[o] = g->state_d_m[o] + zero * cc;
        for (int j = 0; j < 3; j++) g->state_u_m[d + 3 * j + 9 * (size_t)idx] += zero * cc;
        g->state_p_m[o] = g->state_p_m[o] + zero * cc;
        g->state_e_m[o] = g->state_e_m[o] + zero * cc;
        g->state_g_m[o] = g->state_g_m[o] + zero * cc;
    }
}

/* flux.f90:575-590 */
static inline void flux_hll(const double fl[5], const double fr[5], const double ul[5], const double ur[5],
                            double gl, double gr, double dl, double dr, double pl, double pr, double vl,
                            double vr, double fhll[5])
{
    double cr = sqrt(gr * pr / dr);
    double cl = sqrt(gl * pl / dl);
    double ap = fmax2(fmax2(0.0, vr + cr), vl + cl);
    double am = fmax2(fmax2(0.0, -(vr - cr)), -(vl - cl));
    for (int i = 0; i < 5; i++) fhll[i] = (ap * fl[i] + am * fr[i] - ap * am * (ur[i] - ul[i])) / (ap + am);
}

/* flux.f90:487-573, idir 0..2; fhll[2][5] */
static void get_fluxes(const bxo_grid *g, const int isten[3], int idir, double fhll[2][5])
{
    for (int mp = 0; mp <= 1; mp++) {
        int L = isten[mp], R = isten[mp + 1];
        size_t oL = idir + 3 * (size_t)L, oR = idir + 3 * (size_t)R;
        double dl = g->state_d_p[oL], pl = g->state_p_p[oL], el = g->state_e_p[oL], gl = g->state_g_p[oL];
        double vxl = g->state_u_p[idir + 0 + 9 * (size_t)L];
        double vyl = g->state_u_p[idir + 3 + 9 * (size_t)L];
        double vzl = g->state_u_p[idir + 6 + 9 * (size_t)L];
        double sl[5] = {dl, dl * vxl, dl * vyl, dl * vzl, el + HALF * dl * (vxl * vxl + vyl * vyl + vzl * vzl)};
        double dr = g->state_d_m[oR], pr = g->state_p_m[oR], er = g->state_e_m[oR], gr = g->state_g_m[oR];
        double vxr = g->state_u_m[idir + 0 + 9 * (size_t)R];
        double vyr = g->state_u_m[idir + 3 + 9 * (size_t)R];
        double vzr = g->state_u_m[idir + 6 + 9 * (size_t)R];
        double sr[5] = {dr, dr * vxr, dr * vyr, dr * vzr, er + HALF * dr * (vxr * vxr + vyr * vyr + vzr * vzr)};
        double fl[5] = {0, 0, 0, 0, 0}, fr[5] = {0, 0, 0, 0, 0};
        double vl, vr;
        if (idir == 0) { vr = vxr; vl = vxl; }
        else if (idir == 1) { vr = vyr; vl = vyl; }
        else { vr = vzr; vl = vzl; }
        fl[1 + idir] = pl;
        fr[1 + idir] = pr;
        fl[0] = sl[0] * vl;
        fl[1] = fl[1] + sl[1] * vl;
        fl[2] = fl[2] + sl[2] * vl;
        fl[3] = fl[3] + sl[3] * vl;
        fl[4] = (sl[4] + pl) * vl;
        fr[0] = sr[0] * vr;
        fr[1] = fr[1] + sr[1] * vr;
        fr[2] = fr[2] + sr[2] * vr;
        fr[3] = fr[3] + sr[3] * vr;
        fr[4] = (sr[4] + pr) * vr;
        flux_hll(fl, fr, sl, sr, gl, gr, dl, dr, pl, pr, vl, vr, fhll[mp]);
    }
}

/* flux.f90:6-235 */
void bxo_flux(bxo_grid *g, int avg, double dt)
{
    const int n = g->ngrid;
    const double dx = g->prm.dx, dy = g->prm.dy, dz = g->prm.dz;
    const double areaxy = dy * dx, areayz = dy * dz, areaxz = dx * dz, vol = dx * dy * dz;
    double *cons = g->cons;
    /* :47-81 slopes */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
        for (int k = 0; k < 9; k++) g->slope_u[9 * (size_t)c + k] = 0.0;
        for (int k = 0; k < 3; k++) {
            g->slope_p[3 * (size_t)c + k] = 0.0;
            g->slope_d[3 * (size_t)c + k] = 0.0;
            g->slope_e[3 * (size_t)c + k] = 0.0;
            g->slope_g[3 * (size_t)c + k] = 0.0;
        }
        for (int k = 0; k < 5; k++) g->fluxtmp[5 * (size_t)c + k] = 0.0;
        int b[6];
        if (g->boundary[c] > 0) for (int k = 0; k < 6; k++) b[k] = c;
        else for (int k = 0; k < 6; k++) b[k] = g->ineigh[6 * (size_t)c + k];
        int sx[3] = {b[3], c, b[2]}, sy[3] = {b[1], c, b[0]}, sz[3] = {b[5], c, b[4]};
        calculate_slopes(g, sx, 0);
        calculate_slopes(g, sy, 1);
        calculate_slopes(g, sz, 2);
    }
    /* :87-116 states (the reference calls calculate_states three times per
     * cell with the same centre cell; once is equivalent: each call assigns
     * the same values and then adds exact zeros) */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
        calculate_states(g, c, dt);
        calculate_states(g, c, dt);
        calculate_states(g, c, dt);
    }
    /* :121-193 fluxes on interior cells */
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
        if (g->boundary[c] > 0) continue;
        const int *b = g->ineigh + 6 * (size_t)c;
        double fx[2][5], fy[2][5], fz[2][5];
        int sx[3] = {b[3], c, b[2]}, sy[3] = {b[1], c, b[0]}, sz[3] = {b[5], c, b[4]};
        get_fluxes(g, sx, 0, fx);
        get_fluxes(g, sy, 1, fy);
        get_fluxes(g, sz, 2, fz);
        for (int i = 0; i < 5; i++)
            g->fluxtmp[5 * (size_t)c + i] =
                -((areayz * (fx[1][i] - fx[0][i]) + areaxz * (fy[1][i] - fy[0][i]) + areaxy * (fz[1][i] - fz[0][i])) *
                  dt / vol);
    }
    /* :201-217 stage update */
    if (avg == 0) {
#pragma omp parallel for schedule(static)
        for (int c = 0; c < n; c++)
            for (int i = 0; i < 5; i++) cons[5 * (size_t)c + i] = cons[5 * (size_t)c + i] + g->fluxtmp[5 * (size_t)c + i];
    } else {
#pragma omp parallel for schedule(static)
        for (int c = 0; c < n; c++)
            for (int i = 0; i < 5; i++)
                cons[5 * (size_t)c + i] =
                    HALF * (g->cons_old[5 * (size_t)c + i] + cons[5 * (size_t)c + i] + g->fluxtmp[5 * (size_t)c + i]);
    }
    /* :223-230 floors */
    const double srho = g->prm.small_rho, seps = g->prm.small_eps;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
        double *U = cons + 5 * (size_t)c;
        U[0] = fmax2(U[0], srho);
        if (U[4] < 0.0) U[4] = HALF * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0] + seps;
    }
    bxo_set_ghost_cells(g); /* :231 */
}

/* main.f90:327-333 */
void bxo_copy_old(bxo_grid *g)
{
    const size_t n5 = 5 * (size_t)g->ngrid;
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < n5; k++) g->cons_old[k] = g->cons[k];
}

/* -------------------------------------------------------------- avisc */
/* avisc.f90:13-24.  For boundary=1 cells the reference's get_boundary reads
 * ineigh(2:6)=0, i.e. grid(0), outside the array; that case is undefined in
 * the reference and is defined here as "neighbour = self" (qq = 0).        */
void bxo_avisc(bxo_grid *g)
{
    const int n = g->ngrid;
    const double *u = g->u;
    double *qq = g->qq;
    const double ax = g->prm.avmag[0], ay = g->prm.avmag[1], az = g->prm.avmag[2];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
        if (g->boundary[c] > 0) {
            int b[6];
            for (int k = 0; k < 6; k++) {
                int nb = g->ineigh[6 * (size_t)c + k];
                b[k] = (nb < 0 || g->boundary[nb] > 0) ? c : nb; /* utils.f90:62-74 */
            }
            double t;
            t = fmin2(0.0, u[3 * (size_t)b[2] + 0] - u[3 * (size_t)b[3] + 0]);
            qq[3 * (size_t)c + 0] = ax * (t * t);
            t = fmin2(0.0, u[3 * (size_t)b[0] + 1] - u[3 * (size_t)b[1] + 1]);
            qq[3 * (size_t)c + 1] = ay * (t * t);
            t = fmin2(0.0, u[3 * (size_t)b[4] + 2] - u[3 * (size_t)b[5] + 2]);
            qq[3 * (size_t)c + 2] = az * (t * t);
        } else {
            qq[3 * (size_t)c + 0] = 0.0;
            qq[3 * (size_t)c + 1] = 0.0;
            qq[3 * (size_t)c + 2] = 0.0;
        }
    }
}

/* ---------------------------------------------- scalar hooks for tests */
double bxo_slope(double c0, double c1, double c2) { return slope(c0, c1, c2); }

/* one HLL face (flux.f90:512-590): L, R = {rho, vx, vy, vz, p, eps, gamma} */
void bxo_face(int idir, const double L[7], const double R[7], double F[5])
{
    /* reuse get_fluxes through a 2-cell scratch grid of interface states */
    double fl[5] = {0, 0, 0, 0, 0}, fr[5] = {0, 0, 0, 0, 0};
    double dl = L[0], vxl = L[1], vyl = L[2], vzl = L[3], pl = L[4], el = L[5], gl = L[6];
    double dr = R[0], vxr = R[1], vyr = R[2], vzr = R[3], pr = R[4], er = R[5], gr = R[6];
    double sl[5] = {dl, dl * vxl, dl * vyl, dl * vzl, el + HALF * dl * (vxl * vxl + vyl * vyl + vzl * vzl)};
    double sr[5] = {dr, dr * vxr, dr * vyr, dr * vzr, er + HALF * dr * (vxr * vxr + vyr * vyr + vzr * vzr)};
    double vl = idir == 0 ? vxl : (idir == 1 ? vyl : vzl);
    double vr = idir == 0 ? vxr : (idir == 1 ? vyr : vzr);
    fl[1 + idir] = pl;
    fr[1 + idir] = pr;
    fl[0] = sl[0] * vl;
    fl[1] = fl[1] + sl[1] * vl;
    fl[2] = fl[2] + sl[2] * vl;
    fl[3] = fl[3] + sl[3] * vl;
    fl[4] = (sl[4] + pl) * vl;
    fr[0] = sr[0] * vr;
    fr[1] = fr[1] + sr[1] * vr;
    fr[2] = fr[2] + sr[2] * vr;
    fr[3] = fr[3] + sr[3] * vr;
    fr[4] = (sr[4] + pr) * vr;
    flux_hll(fl, fr, sl, sr, gl, gr, dl, dr, pl, pr, vl, vr, F);
}

/* ------------------------------------------------------- diagnostics */
/* utils.f90:1021-1098 conservation_diagnostic + get_etot (output only) */
void bxo_diagnostics(bxo_grid *g, double out[12])
{
    const int n = g->ngrid;
    const double dx = g->prm.dx, dy = g->prm.dy, dz = g->prm.dz;
    const double *cons = g->cons, *u = g->u;
    double mass = 0, ekin = 0, etot = 0, amom = 0, momx = 0, momy = 0, momz = 0;
    double xcom = 0, ycom = 0, zcom = 0, em = 0, ef = 0;
    const double x0 = (g->nx + 1.0) / 2.0, y0 = (g->ny + 1.0) / 2.0, z0 = (g->nz + 1.0) / 2.0;
#pragma omp parallel for schedule(static) reduction(+ : mass, ekin, etot, amom, momx, momy, momz, xcom, ycom, zcom, em, ef)
    for (int c = 0; c < n; c++) {
        if (g->boundary[c] > 0) continue;
        const double *U = cons + 5 * (size_t)c;
        double x = dx * (g->ix[c] - x0), y = dy * (g->iy[c] - y0), z = dz * (g->iz[c] - z0);
        double angle = atan2(y, x);
        amom += (U[2] * cos(angle) - U[1] * sin(angle)) * sqrt(x * x + y * y) * dx * dy * dz;
        momx += U[1] * dz * dx * dy;
        momy += U[2] * dz * dx * dy;
        momz += U[3] * dz * dx * dy;
        mass += U[0] * dz * dx * dy;
        xcom += x * U[0] * dz * dx * dy;
        ycom += y * U[0] * dz * dx * dy;
        zcom += z * U[0] * dz * dx * dy;
        etot += U[4] * dz * dx * dy;
        ekin += (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / (2.0 * U[0]) * dx * dy * dz;
        em += U[0] * dx * dy * dz;
        const double *G = g->gforce + 3 * (size_t)c;
        const double *v = u + 3 * (size_t)c;
        ef += U[0] * dx * dy * dz * fabs(v[0] * G[0] + v[1] * G[1] + v[2] * G[2]);
    }
    out[0] = mass; out[1] = momx; out[2] = momy; out[3] = momz; out[4] = amom; out[5] = etot;
    out[6] = ekin; out[7] = xcom; out[8] = ycom; out[9] = zcom; out[10] = em; out[11] = ef;
}

/* ---------------------------------------------------------- main loop */
/* main.f90:196-203, 211-555 with optional physics off (fluxangmom=.false.) */
int bxo_run(bxo_grid *g, int maxsteps, double starttime, double endtime, double *dts, double *times, int diag)
{
    double time = starttime, f1 = 1.0, dtold = 0.0, dt;
    int step = 0;
    double dg[12];
    for (;;) {
        if (time >= endtime || step >= maxsteps) break; /* :212 */
        step++;
        if (diag) bxo_diagnostics(g, dg); /* :222-223 */
        bxo_state(g);                     /* :232 */
        dt = bxo_courant(g);              /* :241 */
        dt = dt * f1;                     /* :267 */
        if (step > 1 && dtold > 0.0) dt = fmin2(dt, dtold * 1.1); /* :270 */
        dtold = dt;
        if (dt + time > endtime) dt = endtime - time; /* :279 */
        dt = dt * HALF;
        bxo_source(g, dt);        /* :289 */
        bxo_set_ghost_cells(g);   /* :312 */
        bxo_state(g);             /* :314 */
        bxo_velocity(g);          /* :318 */
        bxo_copy_old(g);          /* :327-333 */
        dt = dt * 2.0;            /* :342 */
        bxo_flux(g, 0, dt);       /* :346 */
        bxo_set_ghost_cells(g);   /* :351 */
        bxo_state(g);
        bxo_velocity(g);
        bxo_flux(g, 1, dt);       /* :368 */
        bxo_set_ghost_cells(g);   /* :374 */
        bxo_state(g);
        bxo_velocity(g);
        dt = HALF * dt;           /* :471 */
        bxo_source(g, dt);        /* :490 */
        bxo_set_ghost_cells(g);   /* :511 */
        bxo_state(g);
        bxo_velocity(g);
        dt = dt * 2.0;            /* :544 */
        time = time + dt;
        if (dts) dts[step - 1] = dt;
        if (times) times[step - 1] = time;
    }
    return step;
}
