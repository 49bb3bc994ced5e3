// bx_tile.h -- one RK stage of the fused step as a single kernel: a workgroup
// holds an x-y tile of the grid with its +-2 halo in LDS and marches in z.
//
// Tile: 64 x-cells (one per lane) by TY rows of outputs.  The workgroup has
// TY+4 waves; wave w holds row r = w-2 of the tile (rows -2..TY+1, the two
// outer rows on each side being the y-halo).  The primitives of a plane are
// kept in a tile in LDS with one column of x-halo on each side; there are two
// such tiles, for planes k and k+1.  Every row writes its primitives one
// plane ahead, so that the tile of plane k is complete when plane k starts.
// Per z-plane k:
//
//   A  owner rows 0..TY-1 take the z face k+1/2 from their register queue in
//      z (the primitives of plane k from the tile, k+1 and k+2 in registers)
//      and park the z flux difference of plane k in LDS; every row writes its
//      primitives of plane k+1 into the other tile.  Wave r=TY+1 also takes
//      the x-edges of the owner rows on plane k+1: columns -2,-1,0 and
//      62..65, seven lanes per row, their primitives (columns -1 and 64 go
//      into the plane-(k+1) tile) and x-states.
//   B  owner rows form their x-states from the neighbours' primitives
//      (flux.f90:253-351, 391-479), the x-faces l-1/2 (the left plus state
//      from the next lower lane, or column -1's for lane 0) and with face
//      l+1/2 (next higher lane, or face 63+1/2) the x flux difference, parked
//      in LDS; then their y-states, publishing the y-plus state (row r-1's
//      is needed by row r's face r-1/2).  Row -1 publishes its y-plus state,
//      row TY keeps its y-minus state.                            -- barrier
//   C  owner rows and row TY evaluate their face y r-1/2 and publish it over
//      the y-plus state they read (no other row reads that slot).  Wave
//      r=TY+1 publishes the x-plus state of column -1 and evaluates face
//      x 63+1/2 of plane k+1 (the owners read plane k's in B).    -- barrier
//   D  owner rows read face y r+1/2 and apply the stage update and
//      everything pointwise up to the next stage (as the two-kernel form
//      does).
//
// A and B need no barrier between them: the tile they write in A is read
// from the next plane on, and the y-plus slot row r writes in B is read in
// D only by row r itself.  Every wave runs a loop of its own role with the
// same two barriers per plane (and one before the first), so that no
// register holds a value of another role.  Each face and each cell's
// primitives are evaluated once per workgroup, except at the tile's edges:
// the y-halo costs 4 rows of primitives, 2 rows of states and 1 row of faces
// per TY rows; the x-edges 7 cells of primitives, states and faces per row
// (one wave pass).
// Every arithmetic operation is the one the exact sequence performs on the
// same values, so the results are bit-identical to it.
#pragma once

#include "bx_stage.h"
#include "bx_waves.h"

namespace bx {

template <typename S, int TY>
struct TileLds {
    static constexpr int RS = 66;  // row stride of the primitive tile: columns -1..64
    S P[2][6][(TY + 4) * RS];      // rho, u0, u1, u2, p, eps of planes k, k+1 (tile k&1), rows -2..TY+1
    S Sy[6][(TY + 1) * 64];        // y-plus states of rows -1..TY-1 (index r+1), then y-face r-1/2 (index r)
    S SxL[6][TY];                  // x-plus state of column -1, per owner row
    S FxR[5][TY];                  // x-face flux 63+1/2, per owner row
    double Dx[5][TY * 64];         // x flux difference of each owner cell (phase B to D)
    double Dz[5][TY * 64];         // z flux difference of each owner cell (phase A to D)
};

template <typename S, int N>
__device__ __forceinline__ void store_w(S (&P)[6][N], int idx, const W6<S> &w)
{
    P[0][idx] = w.d;
    P[1][idx] = w.u0;
    P[2][idx] = w.u1;
    P[3][idx] = w.u2;
    P[4][idx] = w.p;
    P[5][idx] = w.e;
}

template <typename S, int N>
__device__ __forceinline__ W6<S> fetch_w(const S (&P)[6][N], int idx)
{
    W6<S> w;
    w.d = P[0][idx];
    w.u0 = P[1][idx];
    w.u1 = P[2][idx];
    w.u2 = P[3][idx];
    w.p = P[4][idx];
    w.e = P[5][idx];
    return w;
}

template <typename S, int N>
__device__ __forceinline__ void store_side(S (&B)[6][N], int idx, const Side<S> &s)
{
    B[0][idx] = s.d;
    B[1][idx] = s.v[0];
    B[2][idx] = s.v[1];
    B[3][idx] = s.v[2];
    B[4][idx] = s.p;
    B[5][idx] = s.e;
}

template <typename S, int N>
__device__ __forceinline__ Side<S> fetch_side(const S (&B)[6][N], int idx, S gam)
{
    Side<S> s;
    s.d = B[0][idx];
    s.v[0] = B[1][idx];
    s.v[1] = B[2][idx];
    s.v[2] = B[3][idx];
    s.p = B[4][idx];
    s.e = B[5][idx];
    s.g = gam;
    return s;
}

// value of the next lower lane; lane 0 takes `edge`
template <typename T>
__device__ __forceinline__ T from_lower(T v, T edge, int)
{
    return wave_shift_t<0x138>(v, edge);  // wave_shr:1
}

// value of the next higher lane; lane 63 takes `edge`
template <typename T>
__device__ __forceinline__ T from_upper(T v, T edge, int)
{
    return wave_shift_t<0x130>(v, edge);  // wave_shl:1
}

// Per-tile constants of one stage.
struct TileGeom {
    int nx, ny, nz;
    int64_t n, plane;
    int x0, j0, r, lane;
    int kb, ke;
};

// Owner rows 0..TY-1: the register queue in z and the stage update.  The
// primitives of plane k are read back from the tile where needed, and the x
// and z flux differences are parked in LDS until the update, to keep the
// registers live across the barriers few.
template <int STAGE, typename S, int TY, bool G>
__device__ __forceinline__ void tile_owner(const StageArgs &a, TileLds<S, TY> &L, const TileGeom &g,
                                           unsigned int &flg, double &rmax)
{
    constexpr int RS = TileLds<S, TY>::RS;
    const int lane = g.lane, r = g.r, nx = g.nx, ny = g.ny, nz = g.nz;
    const int64_t n = g.n;
    const int x = g.x0 + lane, y = g.j0 + r;
    const int xc = min(x, nx - 1), yc = min(y, ny - 1);
    const unsigned coff = (unsigned)(yc * nx + xc) * (unsigned)sizeof(S);
    const bool xy_ghost = x < 2 || x > nx - 3 || y < 2 || y > ny - 3;
    const bool is_out = !xy_ghost;
    const S gamS = S(a.gam);
    auto zplane = [&](int k) -> int64_t { return (int64_t)k * g.plane; };
    auto zghost = [&](int k) -> bool { return !a.flatz && (k < 2 || k > nz - 3); };
    const int c = (r + 2) * RS + lane + 1;  // this cell in the primitive tiles
    const int o = r * 64 + lane;            // this cell's slot in the per-owner buffers

    W6<S> q2{};  // primitives of plane k+1 (k+2 after the z face of plane k)
    S cs0 = S(0), cs1 = S(0);
    Side<S> zp_prev{};
    S Fz_prev[5] = {S(0), S(0), S(0), S(0), S(0)};
    Raw<S> nxt;
    const int kb = g.kb, ke = g.ke;
    if (!a.flatz) {
        const W6<S> t0 = load_prims<STAGE, S, G>(a, zplane(kb - 2), coff, flg);
        const W6<S> t1 = load_prims<STAGE, S, G>(a, zplane(kb - 1), coff, flg);
        const W6<S> t2 = load_prims<STAGE, S, G>(a, zplane(kb), coff, flg);
        const W6<S> t3 = load_prims<STAGE, S, G>(a, zplane(kb + 1), coff, flg);
        Side<S> zmi, zpl;
        dir_states<2>(t0, t1, t2, sound_of(t1, a.gam), xy_ghost || zghost(kb - 1), a, zmi, zpl);
        zp_prev = zpl;
        cs1 = sound_of(t2, a.gam);
        dir_states<2>(t1, t2, t3, cs1, xy_ghost || zghost(kb), a, zmi, zpl);
        face_flux<S, 2>(zp_prev, zmi, Fz_prev);
        zp_prev = zpl;
        store_w(L.P[kb & 1], c, t2);  // plane kb
        q2 = t3;
        nxt = load_raw<STAGE, S, G>(a, zplane(kb + 2), coff);
    } else {
        const W6<S> w = load_prims<STAGE, S, G>(a, zplane(kb), coff, flg);
        cs1 = sound_of(w, a.gam);
        store_w(L.P[kb & 1], c, w);
        if (kb + 1 < ke) nxt = load_raw<STAGE, S, G>(a, zplane(kb + 1), coff);
    }
    __syncthreads();  // every row's primitives of plane kb are in the tile

    for (int k = kb; k < ke; k++) {
        const bool ghost_k = xy_ghost || zghost(k);
        const int t = k & 1;
        // A: the z face k+1/2 and the z flux difference of plane k; plane
        // k+1 into the other tile
        if (!a.flatz) {
            const W6<S> q1 = q2;
            cs0 = cs1;
            const Raw<S> cur = nxt;
            // plane k+3 (the last iteration reloads a plane, so that the queue
            // needs no branch)
            nxt = load_raw<STAGE, S, G>(a, zplane(min(k + 3, nz - 1)), coff);
            q2 = prims<STAGE, S, G>(a, cur, flg);
            cs1 = sound_of(q1, a.gam);
            Side<S> zmi1, zpl1;
            dir_states<2>(fetch_w(L.P[t], c), q1, q2, cs1, xy_ghost || zghost(k + 1), a, zmi1, zpl1);
            S Fz1[5];
            face_flux<S, 2>(zp_prev, zmi1, Fz1);
            for (int i = 0; i < 5; i++) {
                L.Dz[i][o] = (double)Fz1[i] - (double)Fz_prev[i];
                Fz_prev[i] = Fz1[i];
            }
            zp_prev = zpl1;
            if (k + 1 < ke) store_w(L.P[t ^ 1], c, q1);
        } else {
            cs0 = cs1;
            if (k + 1 < ke) {
                const Raw<S> cur = nxt;
                nxt = load_raw<STAGE, S, G>(a, zplane(min(k + 2, ke - 1)), coff);
                const W6<S> w = prims<STAGE, S, G>(a, cur, flg);
                cs1 = sound_of(w, a.gam);
                store_w(L.P[t ^ 1], c, w);
            }
        }
        // B (no barrier: plane k's tile was complete before the last one): x-states, x-faces and the x flux difference; y-states
        {
            const W6<S> q0 = fetch_w(L.P[t], c);
            Side<S> xmi, xpl;
            dir_states<0>(fetch_w(L.P[t], c - 1), q0, fetch_w(L.P[t], c + 1), cs0, ghost_k, a, xmi, xpl);
            Side<S> xl;
            xl.d = from_lower(xpl.d, L.SxL[0][r], lane);
            xl.v[0] = from_lower(xpl.v[0], L.SxL[1][r], lane);
            xl.v[1] = from_lower(xpl.v[1], L.SxL[2][r], lane);
            xl.v[2] = from_lower(xpl.v[2], L.SxL[3][r], lane);
            xl.p = from_lower(xpl.p, L.SxL[4][r], lane);
            xl.e = from_lower(xpl.e, L.SxL[5][r], lane);
            xl.g = gamS;
            S Fx0[5];
            face_flux<S, 0>(xl, xmi, Fx0);  // face l-1/2
            for (int i = 0; i < 5; i++) {
                const S fx1 = from_upper(Fx0[i], L.FxR[i][r], lane);  // face l+1/2
                L.Dx[i][o] = (double)fx1 - (double)Fx0[i];
            }
        }
        Side<S> ymi;
        {
            Side<S> ypl;
            dir_states<1>(fetch_w(L.P[t], c - RS), fetch_w(L.P[t], c), fetch_w(L.P[t], c + RS), cs0, ghost_k, a,
                          ymi, ypl);
            store_side(L.Sy, (r + 1) * 64 + lane, ypl);
        }
        __syncthreads();
        // C: face y r-1/2, published over the y-plus state of row r-1
        S Fy0[5];
        face_flux<S, 1>(fetch_side(L.Sy, r * 64 + lane, gamS), ymi, Fy0);
        for (int i = 0; i < 5; i++) L.Sy[i][r * 64 + lane] = Fy0[i];
        __syncthreads();
        // D: face y r+1/2 and the update
        if (is_out) {
            const int64_t cz = zplane(k);
            double ui[5], uo[5], ft[5];
            for (int i = 0; i < 5; i++) {
                ui[i] = elem<S>(a.uin, (int64_t)i * n + cz, coff);
                uo[i] = STAGE == 2 ? (double)elem<S>(a.uold, (int64_t)i * n + cz, coff) : 0.0;
            }
            for (int i = 0; i < 5; i++) {
                // flux.f90:188-190 in the reference's order, accumulated in FP64
                const S fy1 = L.Sy[i][(r + 1) * 64 + lane];
                const double ey = (double)(S)(a.axz * ((double)fy1 - (double)Fy0[i]));
                const double dz = a.flatz ? 0.0 : L.Dz[i][o];
                const double s = (a.ayz * L.Dx[i][o] + ey) + a.axy * dz;
                ft[i] = -div_by(s * a.dt, a.vol, a.inv_vol, a.vol_pow2);
            }
            stage_update<STAGE, S, G>(a, ui, uo, ft, fetch_w(L.P[t], c), cz, coff, flg, rmax);
        }
    }
}

// Halo rows -2, -1 and TY: their primitives, the y-states of rows -1 and TY,
// and face y TY-1/2.
template <int STAGE, typename S, int TY, bool G>
__device__ __forceinline__ void tile_halo(const StageArgs &a, TileLds<S, TY> &L, const TileGeom &g,
                                          unsigned int &flg)
{
    constexpr int RS = TileLds<S, TY>::RS;
    const int lane = g.lane, r = g.r, nx = g.nx, ny = g.ny, nz = g.nz;
    const int x = g.x0 + lane, y = g.j0 + r;
    const int xc = min(x, nx - 1), yc = min(y, ny - 1);
    const unsigned coff = (unsigned)(yc * nx + xc) * (unsigned)sizeof(S);
    const bool xy_ghost = x < 2 || x > nx - 3 || y < 2 || y > ny - 3;
    const S gamS = S(a.gam);
    auto zplane = [&](int k) -> int64_t { return (int64_t)k * g.plane; };
    auto zghost = [&](int k) -> bool { return !a.flatz && (k < 2 || k > nz - 3); };
    const int c = (r + 2) * RS + lane + 1;
    const int kb = g.kb, ke = g.ke;
    Raw<S> nxt;
    store_w(L.P[kb & 1], c, load_prims<STAGE, S, G>(a, zplane(kb), coff, flg));
    if (kb + 1 < ke) nxt = load_raw<STAGE, S, G>(a, zplane(kb + 1), coff);
    __syncthreads();  // as the owners' first barrier
    for (int k = kb; k < ke; k++) {
        const int t = k & 1;
        if (k + 1 < ke) {
            // A: primitives of plane k+1 into the other tile
            const Raw<S> cur = nxt;
            nxt = load_raw<STAGE, S, G>(a, zplane(min(k + 2, ke - 1)), coff);
            store_w(L.P[t ^ 1], c, prims<STAGE, S, G>(a, cur, flg));
        }
        Side<S> ymi;
        if (r != -2) {
            // B: y-states of plane k
            const W6<S> w = fetch_w(L.P[t], c);
            Side<S> ypl;
            dir_states<1>(fetch_w(L.P[t], c - RS), w, fetch_w(L.P[t], c + RS), sound_of(w, a.gam),
                          xy_ghost || zghost(k), a, ymi, ypl);
            if (r == -1) store_side(L.Sy, lane, ypl);
        }
        __syncthreads();
        if (r == TY) {
            S F[5];
            face_flux<S, 1>(fetch_side(L.Sy, TY * 64 + lane, gamS), ymi, F);
            for (int i = 0; i < 5; i++) L.Sy[i][TY * 64 + lane] = F[i];
        }
        __syncthreads();
    }
}

// Wave r=TY+1: its own halo row, and the x-edges of the owner rows one plane
// ahead.  Lane 7*row+j holds column {-2,-1,0,62,63,64,65}[j] of owner row
// `row`.
template <int STAGE, typename S, int TY, bool G>
__device__ __forceinline__ void tile_edge(const StageArgs &a, TileLds<S, TY> &L, const TileGeom &g,
                                          unsigned int &flg)
{
    constexpr int RS = TileLds<S, TY>::RS;
    static_assert(7 * TY <= 64, "x-edge lanes of the owner rows must fit one wave");
    const int lane = g.lane, r = g.r, nx = g.nx, ny = g.ny, nz = g.nz;
    const int x = g.x0 + lane, y = g.j0 + r;
    const int xc = min(x, nx - 1), yc = min(y, ny - 1);
    const unsigned coff = (unsigned)(yc * nx + xc) * (unsigned)sizeof(S);
    const S gamS = S(a.gam);
    auto zplane = [&](int k) -> int64_t { return (int64_t)k * g.plane; };
    auto zghost = [&](int k) -> bool { return !a.flatz && (k < 2 || k > nz - 3); };
    const int c = (r + 2) * RS + lane + 1;

    const bool e_on = lane < 7 * TY;
    const int erow = lane / 7, ej = lane - 7 * (lane / 7);
    const int ecol = ej < 3 ? ej - 2 : ej + 59;
    const int ex = g.x0 + ecol, ey = g.j0 + erow;
    const unsigned eoff = (unsigned)(min(ey, ny - 1) * nx + min(max(ex, 0), nx - 1)) * (unsigned)sizeof(S);
    const bool e_xy_ghost = ex < 2 || ex > nx - 3 || ey < 2 || ey > ny - 3;

    // the three parts of the x-edge items of plane k
    W6<S> ew{};
    Side<S> emi{}, epl{};
    auto edge_prims = [&](int k, const Raw<S> &raw) {
        ew = prims<STAGE, S, G>(a, raw, flg);
        if (ej == 1 || ej == 5) store_w(L.P[k & 1], (erow + 2) * RS + ecol + 1, ew);
    };
    // neighbours within a row's seven lanes through whole-wave shifts (the
    // lanes whose neighbour lies in another row or outside the wave are not
    // used: columns -1, 63 and 64 take columns -2..0 and 62..65)
    auto edge_states = [&](int k) {
        dir_states<0>(lane_below_w(ew), ew, lane_above_w(ew), sound_of(ew, a.gam), e_xy_ghost || zghost(k), a, emi, epl);
    };
    auto edge_publish = [&]() {
        if (ej == 1) store_side(L.SxL, erow, epl);
        Side<S> xl;
        xl.d = lane_below(epl.d);
        xl.v[0] = lane_below(epl.v[0]);
        xl.v[1] = lane_below(epl.v[1]);
        xl.v[2] = lane_below(epl.v[2]);
        xl.p = lane_below(epl.p);
        xl.e = lane_below(epl.e);
        xl.g = gamS;
        if (ej == 5) {
            S F[5];
            face_flux<S, 0>(xl, emi, F);  // face 63+1/2
            for (int i = 0; i < 5; i++) L.FxR[i][erow] = F[i];
        }
    };

    const int kb = g.kb, ke = g.ke;
    Raw<S> nxt;
    store_w(L.P[kb & 1], c, load_prims<STAGE, S, G>(a, zplane(kb), coff, flg));
    if (kb + 1 < ke) nxt = load_raw<STAGE, S, G>(a, zplane(kb + 1), coff);
    Raw<S> enxt{};
    if (e_on) {
        edge_prims(kb, load_raw<STAGE, S, G>(a, zplane(kb), eoff));
        edge_states(kb);
        edge_publish();
        if (kb + 1 < ke) enxt = load_raw<STAGE, S, G>(a, zplane(kb + 1), eoff);
    }
    __syncthreads();  // as the owners' first barrier
    for (int k = kb; k < ke; k++) {
        const bool more = k + 1 < ke;
        if (more) {
            // A: primitives of plane k+1, its own row and the x-edges
            const Raw<S> cur = nxt;
            nxt = load_raw<STAGE, S, G>(a, zplane(min(k + 2, ke - 1)), coff);
            store_w(L.P[(k + 1) & 1], c, prims<STAGE, S, G>(a, cur, flg));
            if (e_on) {
                const Raw<S> ecur = enxt;
                enxt = load_raw<STAGE, S, G>(a, zplane(min(k + 2, ke - 1)), eoff);
                edge_prims(k + 1, ecur);
                edge_states(k + 1);  // B
            }
        }
        __syncthreads();
        if (e_on && more) edge_publish();  // C: the owners read plane k's in B
        __syncthreads();
    }
}

// One stage over the tile (bx, by) of 64 x TY outputs and the z planes of chunk bz.
// G: with gravity (StageArgs::gf, FP64).
template <int STAGE, typename S, int TY, bool G>
__device__ __forceinline__ void tile_stage(const StageArgs &a, TileLds<S, TY> &L, unsigned bx, unsigned by,
                                           unsigned bz)
{
    TileGeom g;
    g.nx = a.nx;
    g.ny = a.ny;
    g.nz = a.nz;
    g.n = a.n;
    g.plane = (int64_t)a.nx * a.ny;
    g.x0 = 2 + (int)bx * 64;
    g.j0 = 2 + (int)by * TY;
    g.r = wave_id() - 2;  // tile row of this wave, -2..TY+1 (uniform)
    g.lane = threadIdx.x & 63;
    g.kb = a.zlo + (int)bz * a.lz;
    g.ke = min(g.kb + a.lz, a.zhi);
    if (g.kb >= g.ke) return;  // uniform over the workgroup
    unsigned int flg = 0;
    double rmax = 0.0;
    if (g.r >= 0 && g.r < TY) {
        // The owner rows are the critical path between the barriers; each SIMD
        // holds two of them and one halo or edge wave, which is older and would
        // win the issue arbitration by age: the owners issue first
        // (2.06 -> 2.00 ms per step at 256^3).
        __builtin_amdgcn_s_setprio(1);
        tile_owner<STAGE, S, TY, G>(a, L, g, flg, rmax);
    }
    else if (g.r == TY + 1) tile_edge<STAGE, S, TY, G>(a, L, g, flg);
    else tile_halo<STAGE, S, TY, G>(a, L, g, flg);
    if (flg) atomicOr(a.flags, flg);
    if (STAGE == 2) wave_max_to(rmax, a.rate);
}

}  // namespace bx
