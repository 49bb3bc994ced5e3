/*
 * boxzy.h -- C-ABI of libboxzy_hip.so, the MI355X (gfx950) implementation of
 * the BOXZYHYDRO hydro step.
 *
 * The reference has no plugin or FFI layer: its hot path is a set of
 * argument-less Fortran subroutines acting on module globals
 * (grid_commons.f90:7-36) and called by main.f90.  Each entry point below
 * replaces one of those subroutines; the Fortran driver binds them with
 * ISO_C_BINDING and the Python package binds them with ctypes.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative code on failure;
 *    bx_last_error() returns the message of the last failure on the calling
 *    thread.  Nothing aborts.
 *  - Host arrays use the reference's Fortran layout: cons(5,ngrid),
 *    u(3,ngrid), p(ngrid), adindx(ngrid), qq(3,ngrid), cell id
 *    ix+(iy-1)nx+(iz-1)nx*ny.  Device state is SoA planes owned by the context.
 *  - A context is used by one host thread at a time; all work is ordered on
 *    the context's stream.  Entry points that return a host value
 *    synchronise that stream.
 */
#ifndef BOXZY_H
#define BOXZY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BX_OK 0
#define BX_ERR_ARG (-1)
#define BX_ERR_HIP (-2)
#define BX_ERR_STATE (-3)
#define BX_ERR_UNSUPPORTED (-4)

#define BX_PREC_FP64 0  /* plain FP64 drop-in */
#define BX_PREC_MIXED 1 /* FP32 storage/reconstruction, FP64 update and pressure recovery */

/* Namelist values that reach the hydro step (input.f90:7-16, parameters.f90). */
typedef struct bx_params {
    int32_t nx, ny, nz;       /* &grid_params: cells per axis incl. 2 ghost layers per side */
    double dx, dy, dz;        /* &grid_params */
    double gammafix;          /* &eos_input (H2STAT=-1) */
    double small_rho, small_eps, vlimit, cfl; /* &hydro */
    double xabun, yabun, zabun, mu_z;         /* &eos_input, enter muc (eos.f90:144) */
    double avmag[3];          /* &hydro avmagx, avmagy, avmagz */
    int32_t no_outflow[6];    /* &grid_params no_outflow_{xl,xr,yl,yr,zl,zr} */
    int32_t precision;        /* BX_PREC_FP64 or BX_PREC_MIXED */
    int32_t fluxangmom;       /* &grid_params: 1 = flux_ang (flux_ang.f90:6-618) and the
                                 source.f90:73-90 terms, exact call sequence only */
    double yoffset;           /* &grid_params: y shift of the flux_ang momentum transform */
} bx_params;

typedef struct bx_ctx bx_ctx;

const char *bx_last_error(void);

/* Replaces init_grid.f90:7-265 (allocation, boundary flags, neighbour and
 * ghost-donor tables) for the device state.  Selects device `device`. */
int bx_create(const bx_params *prm, int32_t device, bx_ctx **out);
int bx_destroy(bx_ctx *ctx);
int bx_ngrid(const bx_ctx *ctx, int64_t *ngrid, int64_t *nghost);

/* Host-only: the gather table that one set_ghost_cells call (utils.f90:305-480)
 * reduces to for this grid and wall set.  Ghost dst[k] receives cell src[k]
 * (src<0: slot -src-1 of scratch, i.e. the pre-call value of ghost
 * scratch_src[slot]) with momentum component j negated where bit j of sign[k]
 * is set.  Pass NULL arrays to query the sizes first.  Needs no GPU. */
int bx_ghost_table(const bx_params *prm, int64_t *nghost, int64_t *nscratch, int32_t *dst, int32_t *src,
                   int32_t *sign, int32_t *scratch_src);

/* Host <-> device copies of the reference's global arrays (Fortran layout).
 * Any pointer may be NULL to skip that array.  upload also sets
 * cons_old = cons, as init_conditions.f90:69-77 does.  The first transfer of
 * a multi-field array allocates a device staging buffer of at most 64 MiB,
 * which the context keeps until bx_destroy. */
int bx_upload(bx_ctx *ctx, const double *cons, const double *u, const double *p, const double *adindx);
int bx_download(bx_ctx *ctx, double *cons, double *u, double *p, double *adindx);
int bx_download_cons_old(bx_ctx *ctx, double *cons_old);

/* ---- one entry point per reference subroutine on the path ---- */
int bx_state(bx_ctx *ctx);                        /* state.f90:6-89 (H2STAT -1, 0, 1)   */
int bx_velocity(bx_ctx *ctx, double *maxv);       /* velocity.f90:4-64; maxv may be NULL */
int bx_courant(bx_ctx *ctx, double *dt);          /* courant.f90:4-39: dt = cfl/max rate */
int bx_source(bx_ctx *ctx, double tphase, int32_t siter, double dt); /* source.f90:5-205 */
int bx_set_ghost_cells(bx_ctx *ctx);              /* utils.f90:305-480                   */
int bx_flux(bx_ctx *ctx, int32_t avg, double dt); /* flux.f90:6-235; flux_ang.f90:6-281 when fluxangmom */
int bx_copy_old(bx_ctx *ctx);                     /* main.f90:327-333 cons_old = cons    */
int bx_avisc(bx_ctx *ctx, double *qq);            /* avisc.f90:4-26; qq(3,ngrid) host or NULL */

/* ---- whole steps: main.f90:211-555 ----
 * Advances up to `nsteps` steps or until *time >= endtime.  *time, *dtold and
 * *step carry the driver's scalars between calls (main.f90:196-198: start
 * with dtold=0, step=0).  dts (may be NULL) receives each step's full dt.
 * mode BX_STEP_EXACT runs the reference call sequence kernel by kernel;
 * BX_STEP_FUSED runs two fused stencil kernels per step and falls back to the
 * exact sequence for any step that meets a low-density cell (rho<1e-4), the
 * speed clamp, or a first step whose ghosts are not yet filled.
 * BX_STEP_GRAPH computes what BX_STEP_FUSED computes, bit for bit, with dt
 * formed on the device: runs of up to 32 steps that cannot reach endtime
 * (each dt is at most 1.1x the last) are replayed from a captured hipGraph of
 * two fused steps, with one host synchronisation per run instead of per step.
 * A run in which a step met an irregular branch is undone and its steps are
 * taken one at a time as in BX_STEP_FUSED; the first step and the steps near
 * endtime are always taken one at a time.
 * All modes give results bit-identical to the reference in FP64.
 * Returns the number of steps taken in *taken. */
#define BX_STEP_EXACT 0
#define BX_STEP_FUSED 1
#define BX_STEP_GRAPH 2
int bx_step(bx_ctx *ctx, int32_t mode, int32_t nsteps, double endtime, double *time, double *dtold,
            int32_t *step, double *dts, int32_t *taken);

/* One fused FP64 step that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It computes what bx_step(BX_STEP_FUSED) computes for one step of full dt
 * `dt` from a regular state (ghost cells filled with their donors' values,
 * u = m/rho, rho >= 1e-4, |u| <= vlimit), bit for bit, from the caller's
 * DEVICE planes cons_in (5 SoA planes of ngrid doubles: rho, mx, my, mz, E)
 * into cons_out (5 planes, ghosts filled; must not alias cons_in).  The
 * context supplies the grid, the parameters and the ghost table; its own
 * state is not read or written.  U*, p* and the other intermediates live in
 * a static device array of BX_STATIC_MAX_CELLS * BX_STATIC_BYTES_PER_CELL
 * bytes (plus 256) that each GPU holds until the process ends: ngrid above
 * BX_STATIC_MAX_CELLS is refused (BX_ERR_ARG).  Every call on one GPU shares
 * that array, so calls that overlap on one GPU (other streams, other
 * threads) race.  rate_out and flags_out are DEVICE words (either may be
 * NULL): the Courant max-rate of cons_out (the next dt is cfl/rate,
 * main.f90:266-281) and the irregular branches met (bit 1: rho < 1e-4,
 * bit 2: the vlimit clamp; the step is then not the reference's and must be
 * redone with bx_step; bit 8: a floor acted, informational).  Work is
 * enqueued on `stream` (a hipStream_t; NULL: the context's stream) and not
 * waited for.  FP64 contexts whose fused step is supported (every ghost
 * donor an interior cell, fixed gamma, plain flux, zero potential). */
#define BX_STATIC_MAX_CELLS 1048576
#define BX_STATIC_BYTES_PER_CELL 88
int bx_step_static(bx_ctx *ctx, const double *cons_in, double *cons_out, double dt, double *rate_out,
                   uint32_t *flags_out, void *stream);

/* Number of steps that bx_step(BX_STEP_FUSED) had to redo with the exact
 * sequence since the context was created. */
int bx_fallback_count(const bx_ctx *ctx, int64_t *count);

/* Number of steps that bx_step(BX_STEP_GRAPH) took from graph replays (runs
 * that were undone are not counted) since the context was created. */
int bx_graph_step_count(const bx_ctx *ctx, int64_t *count);

/* Steps bx_step has taken since the context was created: with the fused
 * kernels (including graph replays; steps redone are not counted) and with
 * the exact call sequence (the first step from a set-up, steps redone, and
 * every step of a grid the fused kernels do not cover).  Either pointer may
 * be NULL.  A diagnostic with no reference counterpart. */
int bx_step_counts(const bx_ctx *ctx, int64_t *fused, int64_t *exact);

/* Work split of the fused stage kernels.  variant 11 (the FP64 default) and
 * 10 (the mixed form's default): one kernel per stage over x-y tiles of
 * 64 x 8 (11) or 64 x 4 (10) cells marching z_chunk planes in z (y_chunk
 * unused).  Variants 0, 1, 4: two kernels per stage, a y-direction kernel of
 * y_chunk rows per wave and an x/z kernel of z_chunk planes per wave (1:
 * y-term loaded one plane early; 4: built for four waves per SIMD).  Used by
 * the autotuner; results do not change.  With gravity configured only
 * variant 11 is accepted (BX_ERR_UNSUPPORTED otherwise), and configuring
 * gravity returns another tuned variant to the default. */
int bx_tune(bx_ctx *ctx, int32_t y_chunk, int32_t z_chunk, int32_t variant);

/* Event timing of the fused stage kernels on the context's stream.
 * enable=1 starts (and clears), 0 stops, -1 only reads.  kernel_ms[4]
 * receives the summed durations of the y-direction and the x/z kernel of
 * stage 1, then of stage 2, over *steps steps; with one kernel per stage
 * (variants 10, 11) the y-direction entries are ~0 and the x/z entries hold
 * the stage kernels. */
int bx_profile(bx_ctx *ctx, int32_t enable, double *kernel_ms, int64_t *steps);

/* ------------------------------------------------------------------------
 * Batched shock tubes.  The reference runs one grid per process; this entry
 * family runs many independent tubes in one persistent launch (the queue of
 * tubes is ordered by estimated cost, largest first), optionally replayed
 * from a captured hipGraph.  There is no reference counterpart to cite: each
 * tube is what main.f90:211-555 computes for a grid that
 *   - has nz < 6 (flat z, init_grid.f90:241-247), outflow on every side;
 *   - holds the same set-up in every (y,z) row, except the frame cells
 *     (boundary=1), which hold rho=small_rho, m=0, E=small_eps with the row's
 *     u and p (init_conditions.f90:60-68).
 * A tube is passed as that row: nx cells of cons(5,nx), u(3,nx), p(nx)
 * (Fortran layout), rows of all tubes concatenated in order.  The result is
 * bit-identical to the reference run of the full grid (every row of which
 * equals the returned row, ghost cells holding their donors' values).
 */
typedef struct bx_batch bx_batch;

/* Host-only check that every tube meets the row form: flat z, outflow,
 * FP64, nx >= 8, set-up density >= 1e-4 outside the frame, ghost set-up
 * velocities equal to their donors'.  *bad receives the first failing tube
 * (or -1). */
int bx_batch_check(int32_t ntube, const bx_params *prm, const double *cons_rows, const double *u_rows,
                   int32_t *bad);

/* Upload the tubes; tube i runs until time >= endtime[i] (main.f90:212). */
int bx_batch_create(int32_t ntube, const bx_params *prm, const double *endtime, const double *cons_rows,
                    const double *u_rows, const double *p_rows, int32_t device, bx_batch **out);
int bx_batch_destroy(bx_batch *b);

/* Cells of all rows, and interior cells of the full grids the rows stand for. */
int bx_batch_cells(const bx_batch *b, int64_t *row_cells, int64_t *interior_cells);

/* Queue order (tube indices, largest estimated cost first). */
int bx_batch_order(const bx_batch *b, int32_t *order);

/* Work split: waves per workgroup (4, 8, 16) and workgroups per compute unit
 * (1..8) of the persistent launch.  Used by the autotuner; results do not
 * change. */
int bx_batch_tune(bx_batch *b, int32_t waves_per_block, int32_t blocks_per_cu);

/* Keep the first maxlog dt values of every tube (0: off). */
int bx_batch_log_dt(bx_batch *b, int32_t maxlog);

/* Run every tube from its set-up to its endtime (at most maxsteps steps,
 * 1000000 in the reference) with one persistent launch; use_graph replays a
 * captured hipGraph of the launch.  Asynchronous. */
int bx_batch_run(bx_batch *b, int32_t maxsteps, int32_t use_graph);

/* Wait for the run; *ms receives its duration on the device (HIP events).
 * When a gang could not run together (its members were not resident at the
 * same time) the run is repeated without gangs before this returns, and *ms
 * is that run's. */
int bx_batch_sync(bx_batch *b, double *ms);

/* Gangs: the ntubes costliest tubes (by the steps of the last finished run,
 * else by the set-up estimate) are each split over `members` workgroups
 * (contiguous ranges of cells; the two cells at either side of a range are
 * exchanged through device-coherent accesses, with two device-counter
 * barriers per step).  members x ntubes must not exceed the compute units.
 * members < 2 or ntubes = 0 turns gangs off.  Results do not change.  No
 * reference counterpart: the reference runs one tube per process. */
int bx_batch_gang(bx_batch *b, int32_t members, int32_t ntubes);

/* Runs repeated without gangs since the batch was created. */
int bx_batch_gang_fallbacks(const bx_batch *b, int32_t *n);

/* Final rows (cons(5,.), u(3,.), p(.); any may be NULL), time, steps and
 * status bits (1: the low-density EOS branch was met, 2: the vlimit clamp
 * was met, 4: a non-finite dt stopped the tube) per tube. */
int bx_batch_result(bx_batch *b, double *cons_rows, double *u_rows, double *p_rows, double *time, int32_t *steps,
                    uint32_t *status);

/* dt log, ntube x maxlog (entries past a tube's step count are unset). */
int bx_batch_dts(bx_batch *b, double *dts);

/* ------------------------------------------------------------------------
 * Batched grids.  Many independent grids of uneven size (3-D boxes, or flat-z
 * slabs with nz < 6) in one persistent launch: one workgroup steps one grid
 * as bx_step(BX_STEP_FUSED) does (the exact call sequence of main.f90:211-555
 * where required, fused steps on a regular state with shallow ghosts),
 * taking grids from a queue ordered by estimated cost, largest first.  No reference
 * counterpart to cite: the reference runs one grid per process, and each grid
 * here returns what that process returns, bit for bit.  Every grid has its
 * own bx_params (size, spacings, walls, gammafix, cfl, small_rho/eps,
 * vlimit, abundances) and endtime; the EOS is the fixed-gamma one
 * (H2STAT=-1), the flux the plain one (fluxangmom=0), the force that of a
 * zero potential, and the precision FP64.
 *
 * Host arrays are the grids' Fortran-layout arrays concatenated in order:
 * cons(5,.), u(3,.), p(.) over sum(nx*ny*nz) cells.  Suited to grids of up to
 * about 10^5 cells each (a grid is one workgroup's work). */
typedef struct bx_boxes bx_boxes;

/* Upload the grids' set-ups; grid i runs until time >= endtime[i]. */
int bx_boxes_create(int32_t nbox, const bx_params *prm, const double *endtime, const double *cons, const double *u,
                    const double *p, int32_t device, bx_boxes **out);
int bx_boxes_destroy(bx_boxes *b);

/* Cells of all grids (ghosts included), and their interior cells. */
int bx_boxes_cells(const bx_boxes *b, int64_t *cells, int64_t *interior_cells);

/* Queue order (grid indices, largest estimated cost first). */
int bx_boxes_order(const bx_boxes *b, int32_t *order);

/* Work split: waves per workgroup (4, 8 or 16) and workgroups per compute
 * unit (at most 16 waves per CU).  Used by the autotuner; results do not
 * change. */
int bx_boxes_tune(bx_boxes *b, int32_t waves_per_block, int32_t blocks_per_cu);

/* Keep the first maxlog dt values of every grid (0: off). */
int bx_boxes_log_dt(bx_boxes *b, int32_t maxlog);

/* Run every grid from its set-up to its endtime (at most maxsteps steps)
 * with one persistent launch; use_graph replays a captured hipGraph of the
 * launch.  Asynchronous; a later run starts again from the set-ups. */
int bx_boxes_run(bx_boxes *b, int32_t maxsteps, int32_t use_graph);

/* Wait for the run; *ms receives its duration on the device (HIP events). */
int bx_boxes_sync(bx_boxes *b, double *ms);

/* Final cons(5,.), u(3,.), p(.) (any may be NULL), time, steps and status
 * bits per grid (4: a NaN, zero or infinite dt stopped the grid before that
 * step). */
int bx_boxes_result(bx_boxes *b, double *cons, double *u, double *p, double *time, int32_t *steps,
                    uint32_t *status);

/* dt log, nbox x maxlog (entries past a grid's step count are unset). */
int bx_boxes_dts(bx_boxes *b, double *dts);

/* ------------------------------------------------------------------------
 * Adjoint of the FP64 step with a frozen dt sequence (no reference
 * counterpart: the reference has no derivatives).  The differentiated map is
 * the step of main.f90:211-555 on a regular state (ghost cells copies of their
 * donors, u = m/rho, rho >= 1e-4, |u| <= vlimit), as bx_step(BX_STEP_FUSED)
 * computes it, with dt given per step.  Branches are differentiated as taken
 * (limiter argument, HLL speed bounds, MUSCL lambda, internal-energy floors);
 * a run in which a floor of the stage update acts is refused.
 *
 * bx_adjoint_forward: fill the ghost cells of the current state, then take
 * nsteps steps, with dts_in[k] when dts_in is given, else with the Courant dt
 * of main.f90:266-281 (no endtime clamp); dts_out (may be NULL) receives the
 * dt of every step.  every > 0 keeps a checkpoint of the state every `every`
 * steps for bx_adjoint_backward; every == 0 keeps none.
 *
 * bx_adjoint_backward: lam_T and lam_0 are DEVICE pointers to 5 SoA planes of
 * ngrid doubles (rho, mx, my, mz, E): the adjoint of the final state of the
 * last checkpointed forward, and the resulting adjoint of its initial state
 * (ghost-cell adjoints folded into their donors).  A forward that was refused
 * or failed, took no steps or kept no checkpoints (every == 0) leaves nothing
 * to sweep: the backward then returns BX_ERR_STATE.  The forward keeps the
 * per-step states of its last segments in HBM, as many as the store budget
 * holds (bx_adjoint_store; at least one); every other segment is recomputed
 * from its checkpoint before its reverse sweep.  Both give the same bits.
 */
int bx_adjoint_forward(bx_ctx *ctx, int32_t nsteps, int32_t every, const double *dts_in, double *dts_out);
int bx_adjoint_backward(bx_ctx *ctx, const double *lam_T, double *lam_0);

/* Chunk lengths (cells) of the y- and z-direction reverse kernels; results do not change. */
int bx_adjoint_tune(bx_ctx *ctx, int32_t y_chunk, int32_t z_chunk);

/* Device memory (bytes) the next bx_adjoint_forward may spend on segment
 * stores, (5(every+1) + 6 every) planes of ngrid doubles each: -1 (default)
 * is half the free device memory, 0 keeps one store (every segment but the
 * last is recomputed).  *recomputed (may be NULL) receives the number of
 * segments backward sweeps have recomputed so far.  Results do not change. */
int bx_adjoint_store(bx_ctx *ctx, int64_t bytes, int64_t *recomputed);

/* Copy the FP64 conserved state between the context and a DEVICE buffer of
 * 5 SoA planes of ngrid doubles: to_ctx=1 sets the state (u and p are
 * recomputed by the next step), 0 reads it.  Synchronises. */
int bx_cons_device(bx_ctx *ctx, double *planes, int32_t to_ctx);

/* ------------------------------------------------------------------------
 * celldump files (write_files.f90:5-71 hydro part; read_hydro.f90:4-183
 * without DOUBLEGRID).  Binary: Fortran unformatted sequential records with
 * 4-byte record lengths -- 1234; time, step; the 8 units values; per cell
 * x, y, z, rho, p, u1, u2, u3, phi, E, muc (float64) and boundary (int32).
 * ascii != 0: the formatted form of write_files.f90:44-49.
 *
 * bx_write_celldump replaces write_files(step): it runs state() on the
 * context (as write_files.f90:31-33 does) and writes the file `path`.
 * bx_read_celldump replaces read_hydro(): it reads `path`, rebuilds the
 * momenta from u, resets adindx and the frame cells (read_hydro.f90:150-178),
 * uploads, and runs state(); *time and *step receive the file's values.
 * With gravity from a potential configured (bx_gravity_config), the phi
 * column is the context's potential and bx_read_celldump sets it from the
 * file (read_hydro.f90:129-131); otherwise phi is written as zero.
 * The encode/decode pair is the host-only part on host arrays (Fortran
 * layout): encode writes arrays that state() has already updated (muc_array
 * NULL: the fixed-gamma muc of eos.f90:144 in every cell; phi zero); decode
 * returns the arrays read_hydro holds before its final state() call.
 */
int bx_write_celldump(bx_ctx *ctx, const char *path, double time, int32_t step, int32_t ascii);
int bx_read_celldump(bx_ctx *ctx, const char *path, double *time, int32_t *step);
int bx_celldump_encode(const bx_params *prm, const char *path, double time, int32_t step, int32_t ascii,
                       const double *cons, const double *u, const double *p, const double *muc_array);
int bx_celldump_decode(const bx_params *prm, const char *path, double *time, int32_t *step, double *cons,
                       double *u, double *p, double *adindx);

/* ------------------------------------------------------------------------
 * Tabulated H2 equation of state (H2STAT = 0: fixed ortho:para ratio ac:bc;
 * 1: ortho/para equilibrium), eos.f90.  The keys of &eos_input
 * (input.f90:14-15) and the fixed table bounds of parameters.f90:44-58;
 * xabun, yabun, zabun, mu_z and gammafix come from bx_params.
 */
typedef struct bx_eos_params {
    int32_t H2STAT;    /* -1 fixed gamma (no tables), 0 or 1 */
    int32_t NEOS_RHO;  /* density rows of the tables */
    double dTk_eos, tk_bgrnd, brot, vib, ac, bc;                  /* &eos_input */
    double diss, ionize, rho_eos_low, rho_eos_high, tk_eos_cutoff; /* parameters.f90 */
} bx_eos_params;

/* The reference's defaults (parameters.f90:44-62).  Host-only. */
void bx_eos_defaults(bx_eos_params *e);
/* NEOS_T = int((tk_eos_cutoff - tk_bgrnd)/dTk_eos) + 1 (eos.f90:33).  Host-only. */
int bx_eos_size(const bx_eos_params *e, int32_t *neos_t);
/* Host-only: the tables initialize_eos() + calc_eos_table() (eos.f90:30-51,
 * 127-520) leave, bit-identical, in the reference's units after the call.
 * scalars[2] = log_rho_eos_low, drho_eos; rho_table and deng_eos_array hold
 * NEOS_RHO values, tk_table NEOS_T; the 2-D tables NEOS_T x NEOS_RHO in
 * Fortran order.  Any output may be NULL. */
int bx_eos_tables(const bx_params *prm, const bx_eos_params *e, double *scalars, double *rho_table,
                  double *deng_eos_array, double *tk_table, double *eng_table, double *gamma_table,
                  double *muc_table, double *eng_table2, double *tk_table2, double *muc_table2,
                  double *gamma_table2);
/* Replaces initialize_eos + calc_eos_table (main.f90:70-71): builds the tables
 * on the host, keeps them on the device, and makes bx_state() take the
 * tabulated branch (state.f90:58-79, get_gamma2 eos.f90:653-699), which also
 * sets muc_array.  H2STAT=-1 restores the fixed-gamma branch.  FP64 contexts,
 * before the first fused step; with tables, bx_step runs the exact sequence in
 * either mode (the fused kernels assume a fixed gamma) and the adjoint is
 * unavailable. */
int bx_eos_init(bx_ctx *ctx, const bx_eos_params *e);
/* muc_array(ngrid) (grid_commons.f90:12) as the last state() left it. */
int bx_download_muc(bx_ctx *ctx, double *muc_array);

/* ------------------------------------------------------------------------
 * Gravity from a given potential.  The reference computes gforce from the
 * potential phi (calc_gforce, utils.f90:549-571: centred differences on
 * interior cells) and adds an optional softened point mass placed at minus
 * the grid's mass moment (external_grav, gravity.f90:1477-1549); source()
 * applies gforce (source.f90:148-182).  phi itself comes from its
 * self-gravity solver, which is not part of this library: here it is given.
 * With flags set, bx_step makes the main-loop calls (main.f90:479-487:
 * calc_gforce, then external_grav, before source(tphase,1)).  The caller
 * makes the calls before the loop (main.f90:146-154) with bx_calc_gforce /
 * bx_external_grav.  With BX_GRAV_POTENTIAL alone, the fused and graph
 * modes take the fused kernels (walls on any sides), bit-identical
 * to the exact sequence, whenever gforce is the force of the current phi;
 * otherwise (after bx_set_phi, before the next calc_gforce) that step runs
 * the exact sequence, whose calc_gforce settles it.  With BX_GRAV_EXTERNAL
 * every step runs the exact sequence (the point mass's force needs the
 * corrector's mass moments before source(tphase,1)).  The allocation-free
 * step (bx_step_static) takes no gravity configuration.
 */
#define BX_GRAV_POTENTIAL 1 /* calc_gforce from phi (USEPOTENTIAL) */
#define BX_GRAV_EXTERNAL 2  /* external_grav point mass (EXTERNALGRAV) */
#define BX_GRAV_COM_SEQUENTIAL 4 /* mass moments summed in cell order, bit-identical to
                                   the reference's reduction on one OpenMP thread (one
                                   wave, about 1 s per call at 256^3); without it a
                                   fixed-order tree sum, equal to rounding */
/* flags 0 restores the zero-potential force.  object_mass and object_radius
 * are parameters.f90:83-86 (defaults 1.0 and 0.625).  FP64 contexts, before
 * the first fused step.  Sets phi = 0 and gforce = 0 (init_grid.f90:291-299). */
int bx_gravity_config(bx_ctx *ctx, int32_t flags, double object_mass, double object_radius);
int bx_set_phi(bx_ctx *ctx, const double *phi);   /* phi(ngrid), host */
int bx_calc_gforce(bx_ctx *ctx);                  /* utils.f90:549-571 */
/* gravity.f90:1477-1549; com (may be NULL) receives the mass moments
 * sum x rho dx dy dz, sum y ..., sum z ...  Synchronises when com is given. */
int bx_external_grav(bx_ctx *ctx, double *com);
int bx_download_gforce(bx_ctx *ctx, double *gforce); /* gforce(3,ngrid) */

/* Synchronise the context's stream. */
int bx_sync(bx_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
