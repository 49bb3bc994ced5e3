// bx_eos.hip -- the tabulated H2 equation of state (H2STAT = 0 or 1).
//
// Host: the tables of initialize_eos() and calc_eos_table() (eos.f90:30-51,
// 127-520) for a fixed ortho:para ratio (H2STAT=0) or ortho/para equilibrium
// (H2STAT=1), including the re-gridding onto log-spaced energies
// (eos.f90:503-517 with get_gamma_norho, eos.f90:532-575) and the conversion
// to code units.  Every table entry is computed with the reference's formula
// and operation order using the C library's exp/log/log10/pow/sqrt, so the
// tables are the reference's bit for bit.  Quantities that depend on the
// temperature alone (the H2 partition sums of eos.f90:65-101, the Boltzmann
// factors, the translational prefactors) are computed once per temperature
// row instead of once per entry: they are pure functions of T, so the values
// are the same.  The entropy and grand-potential columns (sent, apot) and
// eostable.dat feed nothing on the path and are not produced.
//
// Device: state() with the tabulated branch (state.f90:58-79): get_gamma2
// (eos.f90:653-699) interpolates gamma, mu and T on the (log eps/rho, log rho)
// grid; p = rho T R / mu; the low-density override.  Each table entry carries
// its values and its interpolation slopes (64 bytes), stored beside the entry
// of the next density row at the same temperature index, so one lookup reads
// one 128-byte line.
#include <algorithm>
#include <climits>
#include <cmath>
#include <new>
#include <thread>
#include <vector>

#include "bx_internal.h"
#include "bx_math.h"
#include "bx_units.h"

namespace bx {

// parameters.f90:9-10 and eos.f90:16-19
constexpr double kRgasCGS = 8.254e7, kMp = 1.6726e-24, kKb = 1.38065e-16, kHplanck = 6.626e-27,
                 kMelec = 9.10934e-28, kPi = 3.1415926535897931;

// Fortran INT() of a real(8) as the reference's x86-64 build computes it
// (cvttsd2si): truncation, and INT_MIN for NaN and out-of-range values.
__host__ __device__ __forceinline__ int fort_int(double q)
{
    return (q > -2147483649.0 && q < 2147483648.0) ? (int)q : INT_MIN;
}

struct H2Row {  // per temperature row
    double tk, zp, zoprime, ze, dzpdt, dzedt, dzoprimedt;
    double ediss, eion, evib, s3_half, s3_e, zh2_int, zrot;
};

// eos.f90:65-101
static void h2_partition(double t, double brot, double &zp, double &zo, double &ze, double &zoprime,
                         double &dzpdt, double &dzodt, double &dzedt, double &dzoprimedt)
{
    zp = zo = ze = zoprime = 0.0;
    double sumo = 0.0, sump = 0.0;
    dzpdt = dzodt = dzedt = dzoprimedt = 0.0;
    for (int j = 0; j <= 100; j += 2) {
        const double f1 = (double)(2 * j + 1), f2 = (double)(j * (j + 1));
        const double f3 = (double)(2 * (j + 1) + 1), f4 = (double)((j + 1) * (j + 2));
        const double tt = t * t;
        zp = zp + f1 * std::exp(-((f2 * brot) / t));
        zo = zo + (3.0 * f3) * std::exp(-((f4 * brot) / t));
        zoprime = zoprime + (3.0 * f3) * std::exp(-(((f4 - 2.0) * brot) / t));
        dzpdt = dzpdt + (((f1 * f2) * brot) * std::exp(-((f2 * brot) / t))) / tt;
        dzodt = dzodt + ((((3.0 * f3) * f4) * brot) * std::exp(-((f4 * brot) / t))) / tt;
        dzoprimedt = dzoprimedt + ((((3.0 * f3) * (f4 - 2.0)) * brot) * std::exp(-(((f4 - 2.0) * brot) / t))) / tt;
        const double erro = std::fabs(sumo - zo) / zo, errp = std::fabs(sump - zp) / zp;
        if (erro < 1e-33 && errp < 1e-33) break;
        sumo = zo;
        sump = zp;
    }
    ze = zp + zo;
    dzedt = dzpdt + dzodt;
}

// eos.f90:107-111 without the 1/n: eul * (sqrt(2 pi m mp kB t / h^2))**3, eul = 1
static double trans_numer(double m, double t)
{
    const double s = std::sqrt((((((2.0 * kPi) * m) * kMp) * kKb) * t) / (kHplanck * kHplanck));
    return 1.0 * ((s * s) * s);
}

// Fortran max(int(jump*.75),1) with the default-real literal .75
static inline int next_jump(int jump) { return std::max((int)((float)jump * 0.75f), 1); }

// eos.f90:532-575 on one density column (tables in CGS at this point).  eng
// is the caller's variable: it is replaced by the end value when off-table.
static bool gamma_norho(double &eng, const double *E, const double *G, const double *M, const double *T,
                        int neos, double &tk, double &m, double &gam)
{
    int ientry = 1, jump = neos / 4, flag = 0;
    // the reference's search ends for any finite monotone column; bound it for NaN input
    for (long it = 0;; it++) {
        if (it > 4L * neos + 64) return false;
        const int inext = std::min(ientry + 1, neos);
        if (E[ientry - 1] <= eng && eng < E[inext - 1]) break;
        if (E[ientry - 1] > eng) {
            ientry = ientry - jump;
            jump = next_jump(jump);
            if (ientry < 1) {
                ientry = 1;
                if (eng <= E[0]) {
                    flag = 1;
                    break;
                }
            }
        } else {
            ientry = ientry + jump;
            jump = next_jump(jump);
            if (ientry > neos - 1) {
                ientry = neos - 1;
                if (eng >= E[ientry - 1]) {
                    flag = 2;
                    break;
                }
            }
        }
    }
    if (flag > 1) eng = E[neos - 1];
    else if (flag > 0) eng = E[0];
    const int a = ientry - 1, b = ientry;
    tk = T[a] + (T[b] - T[a]) / (E[b] - E[a]) * (eng - E[a]);
    gam = G[a] + (G[b] - G[a]) / (E[b] - E[a]) * (eng - E[a]);
    m = M[a] + (M[b] - M[a]) / (E[b] - E[a]) * (eng - E[a]);
    return true;
}

template <typename F>
static void parallel_for(int n, F f)
{
    const int nt = std::max(1, std::min<int>(std::min<int>((int)std::thread::hardware_concurrency(), 16), n));
    if (nt == 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int i = t; i < n; i += nt) f(i);
        });
    for (auto &x : th) x.join();
}

struct Tables {
    int neos = 0, nrho = 0;
    double log_rho_low = 0, drho = 0;
    std::vector<double> rho, deng, tk;                     // rho in code units at the end
    std::vector<double> eng, gam, muc, eng2, tk2, muc2, gam2;  // (irho, itk), itk fastest
};

static int table_size(const bx_eos_params *e, int *neos_t)
{
    if (!(e->dTk_eos > 0.0)) return set_error(BX_ERR_ARG, "EOS: dTk_eos must be > 0");
    const double q = (e->tk_eos_cutoff - e->tk_bgrnd) / e->dTk_eos;  // eos.f90:33
    if (!(q >= 1.0 && q < 1e8)) return set_error(BX_ERR_ARG, "EOS: (tk_eos_cutoff-tk_bgrnd)/dTk_eos out of range");
    if (e->NEOS_RHO < 2) return set_error(BX_ERR_ARG, "EOS: NEOS_RHO must be >= 2");
    *neos_t = (int)q + 1;
    return BX_OK;
}

// eos.f90:30-51 and 127-520
static int build_tables(const bx_params *p, const bx_eos_params *e, Tables &t)
{
    if (e->H2STAT < -1 || e->H2STAT > 1) return set_error(BX_ERR_ARG, "EOS: H2STAT must be -1, 0 or 1");
    int rc, neos;
    if ((rc = table_size(e, &neos))) return rc;
    const int nrho = e->NEOS_RHO;
    if ((double)neos * nrho > 4e8) return set_error(BX_ERR_ARG, "EOS: table larger than 4e8 entries");
    t.neos = neos;
    t.nrho = nrho;
    const size_t N = (size_t)neos * nrho;
    const double drho = (std::log10(e->rho_eos_high) - std::log10(e->rho_eos_low)) / (double)(nrho - 1);
    const double xabun = p->xabun, yabun = p->yabun, zabun = p->zabun, mu_z = p->mu_z;
    const double muc = 1.0 / (xabun * 0.5 + yabun * 0.25 + zabun / mu_z);  // eos.f90:144
    t.tk.resize(neos);
    for (int i = 1; i <= neos; i++) t.tk[i - 1] = e->tk_bgrnd + (double)(i - 1) * e->dTk_eos;
    t.log_rho_low = std::log10(e->rho_eos_low);
    t.drho = drho;
    t.rho.resize(nrho);
    for (int i = 1; i <= nrho; i++) t.rho[i - 1] = std::pow(10.0, t.log_rho_low + (double)(i - 1) * drho);
    for (auto *v : {&t.eng, &t.gam, &t.muc, &t.eng2, &t.tk2, &t.muc2, &t.gam2}) v->assign(N, 0.0);

    const double ac = e->ac, bc = e->bc, vib = e->vib, diss = e->diss, ionize = e->ionize;
    const double fa = ac / (ac + bc), fb = bc / (ac + bc);
    const double mebymp = kMelec / kMp;
    std::vector<H2Row> rows(neos);
    if (e->H2STAT >= 0) {
        parallel_for(neos, [&](int i) {
            H2Row &r = rows[i];
            const double tk = t.tk[i];
            double zo, dzodt;
            r.tk = tk;
            h2_partition(tk, e->brot, r.zp, zo, r.ze, r.zoprime, r.dzpdt, dzodt, r.dzedt, r.dzoprimedt);
            r.ediss = std::exp(-(diss / tk));
            r.eion = std::exp(-(ionize / tk));
            r.evib = std::exp(-(vib / tk));
            r.s3_half = trans_numer(0.5, tk);
            r.s3_e = trans_numer(mebymp, tk);
            const double vterm = ((vib / (tk * tk)) * r.evib) / (1.0 - r.evib);
            if (e->H2STAT == 0) {
                r.zh2_int = (std::pow(r.zp, fa) * std::pow(r.zoprime, fb)) / (1.0 - r.evib);
                r.zrot = ((fa * r.dzpdt) / r.zp + fb * (r.dzoprimedt / r.zoprime)) + vterm;
            } else {
                r.zh2_int = r.ze / (1.0 - r.evib);
                r.zrot = r.dzedt / r.ze + vterm;
            }
        });
    }
    parallel_for(nrho, [&](int ir) {
        const double den = t.rho[ir];
        const double nh = (xabun * den) / kMp;
        double *E = &t.eng[(size_t)ir * neos], *M = &t.muc[(size_t)ir * neos];
        for (int it = 0; it < neos; it++) {
            const double tk = t.tk[it];
            double xh1, xion;
            if (e->H2STAT >= 0) {
                const H2Row &r = rows[it];
                xh1 = ((0.5 * (r.s3_half / nh)) * r.ediss) / r.zh2_int;  // Tatum 1966
                xh1 = xh1 > 1e8 ? 1.0 : -(0.5 * xh1) + 0.5 * std::sqrt(xh1 * xh1 + 4.0 * xh1);
                xion = (r.s3_e / nh) * r.eion;
                xion = xion > 1e8 ? 1.0 : -(0.5 * xion) + 0.5 * std::sqrt(xion * xion + 4.0 * xion);
                double y;
                if (e->H2STAT == 0) {
                    y = xh1 * (1.5 + xion * 1.5) + (0.5 * (1.0 - xh1)) * (1.5 + tk * r.zrot);
                    E[it] = (kRgasCGS * tk) * ((xabun * y + (yabun * 0.25) * 1.5) + (zabun * 1.5) / mu_z) +
                            ((kRgasCGS * xabun) * xh1) * (xion * ionize + 0.5 * diss);
                } else {
                    y = xh1 * (((1.5 + (0.5 * diss) / tk) + (xion * ionize) / tk) + xion * 1.5) +
                        (0.5 * (1.0 - xh1)) * (1.5 + tk * r.zrot);
                    E[it] = (kRgasCGS * tk) * ((xabun * y + (yabun * 0.25) * 1.5) + (zabun * 1.5) / mu_z);
                }
            } else {
                // eos.f90:292-300: the single-gamma table, built for completeness
                xh1 = 1.0;
                xion = 0.0;
                E[it] = ((kRgasCGS * tk) / (p->gammafix - 1.0)) / muc;
                t.gam[(size_t)ir * neos + it] = p->gammafix;
            }
            M[it] = 1.0 / ((xabun * ((1.0 - 0.5 * (1.0 - xh1)) + xion) + yabun * 0.25) + zabun / mu_z);
        }
    });
    if (e->H2STAT >= 0) {
        // eos.f90:297-346: gamma1 from centred differences of mu and eps
        parallel_for(nrho, [&](int ir) {
            const int irm = ir == 0 ? 0 : ir - 1, irp = ir == nrho - 1 ? nrho - 1 : ir + 1;
            for (int it = 0; it < neos; it++) {
                const int itm = it == 0 ? 0 : it - 1, itp = it == neos - 1 ? neos - 1 : it + 1;
                const double *M = &t.muc[(size_t)ir * neos];
                const double mc = M[it];
                const double chiT = 1.0 - (t.tk[it] * (M[itp] - M[itm])) / (mc * (t.tk[itp] - t.tk[itm]));
                const double chiR = 1.0 - (t.rho[ir] * (t.muc[(size_t)irp * neos + it] - t.muc[(size_t)irm * neos + it])) /
                                              (mc * (t.rho[irp] - t.rho[irm]));
                const double *E = &t.eng[(size_t)ir * neos];
                const double cv = (E[itp] - E[itm]) / (t.tk[itp] - t.tk[itm]);
                t.gam[(size_t)ir * neos + it] = ((kRgasCGS * chiT) / (cv * mc)) * chiT + chiR;
            }
        });
    }
    // eos.f90:503-504
    std::copy_n(&t.gam[(size_t)1 * neos], neos, &t.gam[0]);
    std::copy_n(&t.gam[(size_t)(nrho - 2) * neos], neos, &t.gam[(size_t)(nrho - 1) * neos]);
    // eos.f90:505-514: energies log-spaced per column, and T, mu, gamma on them
    t.deng.assign(nrho, 0.0);
    std::vector<int> bad(nrho, 0);
    const Units scl = get_units();
    const double esc = (scl.time / scl.length) * (scl.time / scl.length);
    parallel_for(nrho, [&](int ir) {
        const size_t o = (size_t)ir * neos;
        const double *E = &t.eng[o];
        const double deng = (std::log10(E[neos - 1]) - std::log10(E[0])) / (double)(neos - 1);
        t.deng[ir] = deng;
        const double lo = std::log10(E[0]);
        for (int it = 0; it < neos; it++) {
            double eng = std::pow(10.0, lo + (double)it * deng);
            if (!gamma_norho(eng, E, &t.gam[o], &t.muc[o], t.tk.data(), neos, t.tk2[o + it], t.muc2[o + it],
                             t.gam2[o + it]))
                bad[ir] = 1;
            t.eng2[o + it] = eng;
        }
        // eos.f90:516-526: to code units
        for (int it = 0; it < neos; it++) {
            t.eng[o + it] = t.eng[o + it] * esc;
            t.eng2[o + it] = t.eng2[o + it] * esc;
        }
    });
    for (int ir = 0; ir < nrho; ir++) {
        if (bad[ir]) return set_error(BX_ERR_ARG, "EOS: energy search did not end (non-finite table)");
        t.rho[ir] = t.rho[ir] / scl.density;
    }
    return BX_OK;
}

// ------------------------------------------------------------------ device
// One entry of table 2 at (itk, irho): the values and, for itk < NEOS, the
// slopes (v(itk+1)-v(itk))/(eng2(itk+1)-eng2(itk)) that get_gamma2 forms at
// run time -- the same IEEE divisions of the same table values, done once.
struct alignas(64) TabEntry {
    double eng, gam, muc, tk, sg, sm, st, pad;
};
// The entries (itk, irho) and (itk, irho+1) side by side in one 128-byte
// line: a lookup, which needs both rows at the same itk, reads one line.
struct alignas(128) TabPair {
    TabEntry lo, hi;
};
// Per density row: rho_table(irho), 1/(rho_table(irho+1)-rho_table(irho))
// (NaN where the reference does not interpolate in rho), log10 of the first
// energy, deng_eos_array, and the first and last energies of the row.
struct alignas(64) RhoRow {
    double r0, inv_dr, leng1, deng, e_first, e_last, pad0, pad1;
};

struct EosDev {
    int neos, nrho;
    double log_rho_low, drho, density, tk_bgrnd;
    const TabPair *tab;  // (irho, itk), itk fastest (code units); hi = row irho+1
    const RhoRow *rows;
};

// eos.f90:653-699
__device__ __forceinline__ void get_gamma2(const EosDev &e, double eps, double rho, double &tk, double &m,
                                           double &gam)
{
    int irho = fort_int((log10(rho * e.density) - e.log_rho_low) / e.drho) + 1;
    irho = irho < e.nrho ? irho : e.nrho;
    if (irho < 1) irho = 1;
    const RhoRow R = e.rows[irho - 1];
    double eng = eps / rho;
    int ientry = fort_int((log10(eng) - R.leng1) / R.deng) + 1;
    if (ientry > e.neos - 1) {
        ientry = e.neos - 1;
        eng = R.e_last;
    }
    if (ientry < 1 || eng < R.e_first) {
        ientry = 1;
        eng = R.e_first;
    }
    const TabPair *pr = e.tab + (int64_t)(irho - 1) * e.neos + (ientry - 1);
    const TabEntry a = pr->lo, ap = pr->hi;  // hi: row irho+1 (the row itself on the last row)
    gam = a.gam + a.sg * (eng - a.eng);
    m = a.muc + a.sm * (eng - a.eng);
    tk = a.tk + a.st * (eng - a.eng);
    // rho_table(irho) /= rho_table(irhop): interpolate in rho; written as
    // selects so that no branch waits on the table load
    const double gamp = ap.gam + ap.sg * (eng - ap.eng);
    const double mp = ap.muc + ap.sm * (eng - ap.eng);
    const double tkp = ap.tk + ap.st * (eng - ap.eng);
    const double slope = R.inv_dr * (rho - R.r0);
    const bool interp = R.inv_dr == R.inv_dr;
    gam = interp ? gam + (gamp - gam) * slope : gam;
    m = interp ? m + (mp - m) * slope : m;
    tk = interp ? tk + (tkp - tk) * slope : tk;
}

// state.f90:30-79, tabulated branch (H2STAT >= 0).  Each thread takes CPT
// cells a block apart (coalesced), with indices clamped for the loads, so the
// table lookups of its cells are in flight together.
template <int CPT>
__global__ __launch_bounds__(256) void k_state_tab(int64_t n, double *__restrict__ cons,
                                                   const double *__restrict__ u, double *__restrict__ p,
                                                   double *__restrict__ ad, double *__restrict__ mucp,
                                                   EosDev e, double gamfix, double seps, double muc,
                                                   double rgas, int disk, int nx, int ny, double dx, double dy)
{
    const int64_t base = (int64_t)blockIdx.x * blockDim.x * CPT + threadIdx.x;
    double rho[CPT], ekin[CPT], tk[CPT], m[CPT], g[CPT];
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        const int64_t c0 = base + (int64_t)k * blockDim.x, c = c0 < n ? c0 : n - 1;
        rho[k] = cons[c];
        const double E = cons[4 * n + c];
        const double u0 = u[c], u1 = u[n + c], u2 = u[2 * n + c];
        ekin[k] = rho[k] * 0.5 * (u0 * u0 + u1 * u1 + u2 * u2);
        const double eps = vmax(E - ekin[k], seps);
        if (disk) {
            // nz == 1: the column is a disk of scale height hscale (state.f90:61-69)
            const int ix = (int)(c % nx) + 1, iy = (int)((c / nx) % ny) + 1;
            const double x = dx * ((double)ix - ((double)nx + 1.0) / 2.0);
            const double y = dy * ((double)iy - ((double)ny + 1.0) / 2.0);
            const double hscale = sqrt((ad[c] - 1.0) * eps / (2.0 * ekin[k]) * (x * x + y * y));
            get_gamma2(e, eps / (2.0 * hscale), rho[k] / (2.0 * hscale), tk[k], m[k], g[k]);
        } else {
            get_gamma2(e, eps, rho[k], tk[k], m[k], g[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        const int64_t c = base + (int64_t)k * blockDim.x;
        if (c >= n) break;
        ad[c] = g[k];
        mucp[c] = m[k];
        double pc = rho[k] * tk[k] * rgas / m[k];
        if (rho[k] < 1e-4) {
            pc = rgas * e.tk_bgrnd / muc * rho[k];
            cons[4 * n + c] = pc / (gamfix - 1.0) + ekin[k];
        }
        p[c] = pc;
    }
}

}  // namespace bx

struct bx_eos {
    bx::EosDev dev;
    bx::TabPair *tab;
    bx::RhoRow *rows;
    double *muc_plane;
};

namespace bx {

static void eos_release(bx_eos *q)
{
    if (!q) return;
    for (void *d : {(void *)q->tab, (void *)q->rows, (void *)q->muc_plane})
        if (d) (void)hipFree(d);
    delete q;
}

void eos_free(bx_ctx *c)
{
    if (!c) return;
    eos_release(c->eos);
    c->eos = nullptr;
}

int launch_state_tab(bx_ctx *c)
{
    const bx_eos *q = c->eos;
    constexpr int kCpt = 2;
    hipLaunchKernelGGL(k_state_tab<kCpt>, dim3(blocks_for(c->n, 256 * kCpt)), dim3(256), 0, c->stream, c->n, c->cons,
                       c->u, c->p,
                       c->adindx, q->muc_plane, q->dev, c->prm.gammafix, c->prm.small_eps, c->muc, c->rgas,
                       c->nz == 1 ? 1 : 0, c->nx, c->ny, c->prm.dx, c->prm.dy);
    BX_HIP(hipGetLastError());
    return BX_OK;
}

}  // namespace bx

using namespace bx;

extern "C" void bx_eos_defaults(bx_eos_params *e)
{
    if (!e) return;
    // parameters.f90:44-62
    e->H2STAT = 0;
    e->NEOS_RHO = 1000;
    e->dTk_eos = 5.0;
    e->tk_bgrnd = 1e-6;
    e->brot = 85.4;
    e->vib = 5987.0;
    e->ac = 1.0;
    e->bc = 3.0;
    e->diss = 52000.0;
    e->ionize = 157878.0;
    e->rho_eos_low = 1e-14;
    e->rho_eos_high = 1e2;
    e->tk_eos_cutoff = 5e4;
}

extern "C" int bx_eos_size(const bx_eos_params *e, int32_t *neos_t)
{
    if (!e || !neos_t) return set_error(BX_ERR_ARG, "bx_eos_size: null argument");
    int nt;
    int rc = table_size(e, &nt);
    if (rc) return rc;
    *neos_t = nt;
    return BX_OK;
}

extern "C" int bx_eos_tables(const bx_params *prm, const bx_eos_params *e, double *scalars, double *rho_table,
                             double *deng_eos_array, double *tk_table, double *eng_table, double *gamma_table,
                             double *muc_table, double *eng_table2, double *tk_table2, double *muc_table2,
                             double *gamma_table2)
{
    if (!prm || !e) return set_error(BX_ERR_ARG, "bx_eos_tables: null argument");
    Tables t;
    int rc = build_tables(prm, e, t);
    if (rc) return rc;
    if (scalars) {
        scalars[0] = t.log_rho_low;
        scalars[1] = t.drho;
    }
    auto put = [](double *dst, const std::vector<double> &v) {
        if (dst) std::copy(v.begin(), v.end(), dst);
    };
    put(rho_table, t.rho);
    put(deng_eos_array, t.deng);
    put(tk_table, t.tk);
    put(eng_table, t.eng);
    put(gamma_table, t.gam);
    put(muc_table, t.muc);
    put(eng_table2, t.eng2);
    put(tk_table2, t.tk2);
    put(muc_table2, t.muc2);
    put(gamma_table2, t.gam2);
    return BX_OK;
}

extern "C" int bx_eos_init(bx_ctx *c, const bx_eos_params *e)
{
    if (!c || !e) return set_error(BX_ERR_ARG, "bx_eos_init: null argument");
    if (c->prm.precision != BX_PREC_FP64)
        return set_error(BX_ERR_UNSUPPORTED, "bx_eos_init: the tabulated EOS runs on FP64 contexts");
    if (c->fused_ready || c->adj)
        return set_error(BX_ERR_STATE, "bx_eos_init: call before the first fused step or adjoint run");
    BX_HIP(hipSetDevice(c->device));
    if (e->H2STAT == -1) {  // fixed gamma: state() needs no table (state.f90:49-57)
        eos_free(c);
        c->fused_unsupported = false;  // fused_prepare decides again on the next fused step
        return BX_OK;
    }
    Tables t;
    int rc = build_tables(&c->prm, e, t);
    if (rc) return rc;
    const size_t N = (size_t)t.neos * t.nrho;
    auto entry = [&](size_t k, int it) {
        TabEntry q{t.eng2[k], t.gam2[k], t.muc2[k], t.tk2[k], 0.0, 0.0, 0.0, 0.0};
        if (it + 1 < t.neos) {  // eos.f90:684-689 slopes
            const double de = t.eng2[k + 1] - t.eng2[k];
            q.sg = (t.gam2[k + 1] - t.gam2[k]) / de;
            q.sm = (t.muc2[k + 1] - t.muc2[k]) / de;
            q.st = (t.tk2[k + 1] - t.tk2[k]) / de;
        }
        return q;
    };
    std::vector<TabPair> pk(N);
    parallel_for(t.nrho, [&](int ir) {
        const int irp = ir + 1 < t.nrho ? ir + 1 : ir;
        for (int it = 0; it < t.neos; it++)
            pk[(size_t)ir * t.neos + it] = TabPair{entry((size_t)ir * t.neos + it, it), entry((size_t)irp * t.neos + it, it)};
    });
    std::vector<RhoRow> rows(t.nrho);
    for (int ir = 0; ir < t.nrho; ir++) {
        const int irp = ir + 1 < t.nrho ? ir + 1 : ir;  // eos.f90:659-660
        const double r0 = t.rho[ir], r1 = t.rho[irp];
        const size_t o = (size_t)ir * t.neos;
        rows[ir] = RhoRow{r0, r0 == r1 ? std::nan("") : 1.0 / (r1 - r0), std::log10(t.eng2[o]), t.deng[ir],
                          t.eng2[o], t.eng2[o + t.neos - 1], 0.0, 0.0};
    }
    // built aside and attached only when complete, so a failed allocation leaves
    // the context on its previous EOS
    bx_eos *q = new (std::nothrow) bx_eos();
    if (!q) return set_error(BX_ERR_ARG, "bx_eos_init: out of host memory");
    hipError_t err = hipMalloc(&q->tab, N * sizeof(TabPair));
    if (err == hipSuccess) err = hipMalloc(&q->rows, t.nrho * sizeof(RhoRow));
    if (err == hipSuccess) err = hipMalloc(&q->muc_plane, c->n * sizeof(double));
    if (err == hipSuccess) err = hipMemcpy(q->tab, pk.data(), N * sizeof(TabPair), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(q->rows, rows.data(), t.nrho * sizeof(RhoRow), hipMemcpyHostToDevice);
    // init_grid.f90:294: muc_array starts at zero
    if (err == hipSuccess) err = hipMemset(q->muc_plane, 0, c->n * sizeof(double));
    if (err != hipSuccess) {
        eos_release(q);
        return hip_check(err, "bx_eos_init: table upload");
    }
    const Units scl = get_units();
    q->dev = EosDev{t.neos, t.nrho, t.log_rho_low, t.drho, scl.density, e->tk_bgrnd, q->tab, q->rows};
    eos_free(c);
    c->eos = q;
    // the fused kernels assume a fixed gamma: bx_step's fused mode runs exact steps
    c->fused_unsupported = true;
    return BX_OK;
}

extern "C" int bx_download_muc(bx_ctx *c, double *muc_array)
{
    if (!c || !muc_array) return set_error(BX_ERR_ARG, "bx_download_muc: null argument");
    BX_HIP(hipSetDevice(c->device));
    if (!c->eos) {  // state.f90:51: muc_array = muc everywhere
        std::fill(muc_array, muc_array + c->n, c->muc);
        return BX_OK;
    }
    BX_HIP(hipMemcpyAsync(muc_array, c->eos->muc_plane, c->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BX_HIP(hipStreamSynchronize(c->stream));
    return BX_OK;
}
