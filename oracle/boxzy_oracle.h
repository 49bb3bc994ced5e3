/*
 * boxzy_oracle.h -- CPU restatement of the BOXZYHYDRO hydro step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the correctness oracle: plain C + OpenMP,
 * mirroring the reference's arithmetic order, array layout (Fortran-ordered
 * cons(5,ngrid), u(3,ngrid), ...) and pass structure.  Only tests, the smoke
 * check and the CPU-baseline leg of the benchmark may load it; the product
 * path (libboxzy_hip.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef BOXZY_ORACLE_H
#define BOXZY_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bxo_params {
    int nx, ny, nz;           /* cells per axis INCLUDING 2 ghost layers per side */
    double dx, dy, dz;
    double gammafix;          /* parameters.f90:52, namelist eos_input */
    double small_rho, small_eps, vlimit, cfl;  /* parameters.f90:69-74 */
    double xabun, yabun, zabun, mu_z;          /* eos.f90:144 (muc) */
    double avmag[3];          /* avmagx, avmagy, avmagz (parameters.f90:77) */
    int no_outflow[6];        /* xl, xr, yl, yr, zl, zr (parameters.f90:31-36) */
} bxo_params;

typedef struct bxo_grid bxo_grid;

bxo_grid *bxo_create(const bxo_params *prm);
void bxo_destroy(bxo_grid *g);

/* caller-owned arrays, Fortran order: cons(5,n) cons_old(5,n) u(3,n) p(n) adindx(n) qq(3,n) */
void bxo_bind(bxo_grid *g, double *cons, double *cons_old, double *u, double *p,
              double *adindx, double *qq);

int bxo_ngrid(const bxo_grid *g);
int bxo_nghost(const bxo_grid *g);
/* geometry export: boundary(n), ineigh(6,n) as 1-based ids like the reference */
void bxo_geometry(const bxo_grid *g, int *boundary, int *ineigh);
double bxo_muc(const bxo_grid *g);
double bxo_rgas(const bxo_grid *g);

void bxo_state(bxo_grid *g);                       /* state.f90:6-89      */
void bxo_velocity(bxo_grid *g);                    /* velocity.f90:4-64   */
double bxo_courant(bxo_grid *g);                   /* courant.f90:4-39, returns dt */
void bxo_source(bxo_grid *g, double dt);           /* source.f90:5-205    */
void bxo_set_ghost_cells(bxo_grid *g);             /* utils.f90:305-480   */
void bxo_flux(bxo_grid *g, int avg, double dt);    /* flux.f90:6-235      */
void bxo_copy_old(bxo_grid *g);                    /* main.f90:327-333    */
void bxo_avisc(bxo_grid *g);                       /* avisc.f90:4-26      */
void bxo_diagnostics(bxo_grid *g, double out[12]); /* utils.f90:1021-1098 */

/* main.f90:211-555: run until time>=endtime or maxsteps; dts/times receive
 * the full dt and the time after each step.  diag!=0 also runs the per-step
 * conservation diagnostics.  Returns the number of steps taken.            */
int bxo_run(bxo_grid *g, int maxsteps, double starttime, double endtime,
            double *dts, double *times, int diag);

/* scalar hooks for closed-form tests: the limiter (utils.f90:738-741) and
 * one HLL face (flux.f90:512-590), L/R = {rho, vx, vy, vz, p, eps, gamma} */
double bxo_slope(double c0, double c1, double c2);
void bxo_face(int idir, const double L[7], const double R[7], double F[5]);

/* set the OpenMP thread count used by the passes (0 = leave default) */
void bxo_set_threads(int n);
int bxo_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
