// bx_internal.h -- context layout and launch helpers shared by the HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "boxzy.h"

namespace bx {

constexpr int kMaxD2 = 64;  // two-level ghost compositions a context may have (one wave)

// Scalars that live on the device so whole steps can run without host syncs.
struct DevScalars {
    unsigned long long rate_bits[2];  // Courant max-rate (non-negative doubles as bits), ping-pong
    unsigned long long maxv_bits;     // velocity.f90 maxv diagnostic
    unsigned int flags;               // irregular-branch flags raised by the fused kernels
    unsigned int pad;
    double dt;                        // full dt of the current step (main.f90:342)
    double dtold;                     // main.f90:271
    double time;                      // main.f90:545
    double endtime;
    int step;                         // main.f90:213
    int active;                       // 1 while the step loop runs (time<endtime)
    int logk;                         // steps of the current device-dt run (index into its dt log)
    int pad2;
};

// flags bits
enum : unsigned int {
    FLAG_LOWRHO = 1u,  // a state() call met rho<1e-4 (state.f90:52-56)
    FLAG_CLAMP = 2u,   // a velocity() call met |u_i|>vlimit (velocity.f90:33)
    FLAG_FLOOR = 8u,   // a floor of the stage update acted (flux.f90:225-228, the source re-round
                       // or the p* eps floor); informational: the step is exact either way
    FLAG_IRREGULAR = FLAG_LOWRHO | FLAG_CLAMP,
};

}  // namespace bx

struct bx_ctx {
    bx_params prm;
    int device;
    hipStream_t stream;
    int nx, ny, nz;
    int64_t n;        // ngrid
    int64_t nghost;
    double muc, rgas;
    double grav[3];   // uniform gravity acceleration; -0.0 from phi=0 (utils.f90:566-568)
    bool flatz;       // nz<6: z neighbours are the cell itself (init_grid.f90:241-247)

    // exact-path device state, SoA planes of n doubles
    double *cons;      // 5 planes: rho, mx, my, mz, E
    double *cons_old;  // 5 planes
    double *u;         // 3 planes
    double *p;
    double *adindx;
    double *qq;        // 3 planes
    double *states;    // exact-flux intermediate, state_planes(c) planes (face fluxes or interface states)

    // ghost table: ghost cell id, donor (>=0: cell id; <0: -(scratch slot+1)),
    // sign bits (bit k flips momentum k); scratch holds pre-fill values of
    // ghost donors (5 planes of nscratch)
    int32_t *ghost_dst;
    int32_t *ghost_src;
    int32_t *ghost_sign;
    int32_t *ghost_scratch_src;
    double *ghost_scratch;
    int64_t nscratch;
    // the same table composed over a pair of set_ghost_cells calls (every call
    // site of the step makes two in a row): ghost dst[k] <- +-(interior cell
    // fsrc[k]), exact when each scratch donor's own donor is an interior cell
    // (fghost_ok); equal to the table above when nscratch == 0
    int32_t *fghost_src;
    int32_t *fghost_sign;
    bool fghost_ok;
    // ghosts whose pair leaves the pre-pair value of another ghost h2 (whose
    // own pair gives an interior cell v): the corner ghosts of a box walled
    // in y and z.  After a step such a ghost holds +-(v's corrector result
    // before source(1)); a fix-up kernel after the corrector's ghost fill
    // writes it (d2_ubar: 6 doubles per entry, its scratch, the last one the
    // ghost's pressure for the materialised planes).  fghost_src/sign map
    // them to v for the other fills.
    int nd2;
    int32_t *d2_dst;   // ghost cell
    int32_t *d2_src;   // interior cell v
    int32_t *d2_sign;  // sign bits of the composition
    double *d2_ubar;

    // fused-path state
    double *ustar;     // 5 planes: stage-1 result U*
    double *pstar;     // stage-2 pressure p* (stale-velocity EOS, main.f90:351-355)
    double *unext;     // 5 planes: stage-2 result U^{n+1}
    bool fused_valid;  // device cons holds a regular state (see bx_step)
    bool fused_ready;  // fused buffers allocated, ghost table shallow
    bool fused_unsupported;
    bool ghosts_filled;  // ghost cells hold their donors' values (after any exact step)
    int fused_ly, fused_lz;  // y rows and z planes per wave of the stage kernels
    int fused_occ;           // stage kernel variant (bx_tune): 0, 1, 4 two kernels; 10, 11 one tiled kernel
    double *eflux;           // 5 planes: y-term of fluxtmp (FP64)
    bool prof_on;
    hipEvent_t ev[6];   // per stage: before the y kernel, between the two kernels, after the x/z kernel
    double prof_ms[4];  // summed durations: y kernel stage 1, x/z kernel stage 1, y kernel stage 2, x/z kernel stage 2
    int64_t prof_steps;
    bool aux_stale;    // u, p planes not materialised since the last fused step
    bool old_stale;    // cons_old not materialised since the last fused step
    int64_t fallbacks;
    int64_t nfused, nexact;  // steps bx_step took with the fused kernels / the exact sequence

    // captured pairs of fused steps (BX_STEP_GRAPH), keyed by the state buffer
    // they start from (the pair returns the state to it)
    hipGraph_t sgraph[2];
    hipGraphExec_t sexec[2];
    const void *skey[2];
    void *chunk0;          // 5 planes: the state at the start of a replayed chunk (rollback)
    int64_t graph_steps;   // steps taken from replays (kept, not undone)
    double *chunk_dts;     // dt of each step of a chunk
    double *stage;         // device staging of host transfers (Fortran layout), stage_len doubles
    int64_t stage_len;

    bx::DevScalars *dsc;   // device scalars
    bx::DevScalars *hsc;   // pinned host mirror

    // mixed-precision planes (FP32 storage)
    float *cons32, *ustar32, *unext32, *pstar32;
    bool state_in_32;  // the current state lives in cons32 (mixed form)

    struct bx_adj *adj;  // adjoint store (bx_adjoint.hip), or null
    struct bx_eos *eos;  // tabulated EOS (bx_eos.hip), or null for a fixed gamma

    // gravity (bx_gravity.hip): a given potential and per-cell gforce(3,n) planes
    int grav_flags;                // BX_GRAV_* bits, 0 = uniform -0.0 force of phi = 0
    double obj_mass, obj_radius;   // external_grav point mass (parameters.f90:83-86)
    double *phi;                   // n
    double *gforce;                // 3 planes
    double *grav_sums;             // centre-of-mass sums (3) and tree partials
    // The fused step with a potential (BX_GRAV_POTENTIAL alone): gforce is
    // settled when it was computed from the current phi (gforce_gen ==
    // phi_gen), so that source(0) and source(1) of a step see the same force;
    // gfused is gforce with the ghosts filled as their donors' (sign flips of
    // the momenta), built from generation gfused_gen.  p_fused: the pressure
    // plane holds what the last step's exact sequence would leave (written by
    // the fused corrector), so materialising keeps it.
    int64_t phi_gen, gforce_gen, gfused_gen;
    double *gfused;  // 3 planes
    bool p_fused;
};

namespace bx {

int set_error(int code, const std::string &msg);
int hip_check(hipError_t e, const char *what);

#define BX_HIP(call)                                             \
    do {                                                         \
        hipError_t _e = (call);                                  \
        if (_e != hipSuccess) return bx::hip_check(_e, #call);   \
    } while (0)

// planes of the exact flux's intermediate: the face fluxes, 5 per direction;
// with flux_ang the interface states, per direction 9 minus + 9 plus (rho, u,
// v, w, p, eps, gamma, r- and angular momentum)
inline int state_planes(const bx_ctx *c) { return c->prm.fluxangmom ? 54 : 15; }

inline unsigned int blocks_for(int64_t n, int threads)
{
    return (unsigned int)((n + threads - 1) / threads);
}

// kernels launched by the C-ABI (defined in bx_exact.hip / bx_fused.hip)
int launch_state(bx_ctx *c);
int launch_velocity(bx_ctx *c, bool want_maxv);
int launch_courant(bx_ctx *c);
int launch_source(bx_ctx *c, double dt);
int launch_ghosts(bx_ctx *c, double *planes5);
int launch_flux(bx_ctx *c, int avg, double dt);
int launch_avisc(bx_ctx *c);
int launch_materialize(bx_ctx *c);
int launch_regular_check(bx_ctx *c);
int fused_step(bx_ctx *c, double dt);
// one FP64 fused step U^n=u0 -> U*=us, p*=ps -> U^{n+1}=un (ghosts filled); the rate is zeroed
// first, the flags too when reset_flags (else they accumulate over steps)
int fused_step_into(bx_ctx *c, double dt, double *u0, double *us, double *ps, double *un, bool reset_flags);
// the same with dt formed on the device from the current Courant rate (the
// device scalars carry time, dtold, step and endtime); flags accumulate and
// each dt is logged into log[cap]
int fused_step_device_dt(bx_ctx *c, double *u0, double *us, double *ps, double *un, double *log, int cap);
void adjoint_free(bx_ctx *c);
void eos_free(bx_ctx *c);
void gravity_free(bx_ctx *c);
int launch_calc_gforce(bx_ctx *c);
int launch_external_grav(bx_ctx *c);
bool grav_settled(const bx_ctx *c);
int grav_fused_planes(bx_ctx *c);
int launch_state_tab(bx_ctx *c);
int fused_prepare(bx_ctx *c);
// BX_STEP_GRAPH: replay `pairs` captured pairs of fused steps from the current
// state (device dt, flags accumulating); the scalars must be on the device
int replay_step_pairs(bx_ctx *c, int pairs);
void drop_step_graphs(bx_ctx *c);
constexpr int kChunkSteps = 32;  // steps per replayed chunk, at most
int narrow_cons(bx_ctx *c);
int widen_cons(bx_ctx *c);

}  // namespace bx
