// bx_celldump.hip -- the reference's celldump files: writer (write_files.f90:5-71,
// hydro part) and restart reader (read_hydro.f90:4-183, without DOUBLEGRID).
//
// Binary form: Fortran unformatted sequential records, each framed by 4-byte
// record lengths:
//   1234                                       (int32, the read test)
//   time, step                                 (float64, int32)
//   units: length, time, mass, density, rgas, eps, kelvin, vel (8 float64,
//          derived_types.f90:13-16, values of units.f90:10-17)
//   per cell in id order: x, y, z, rho, p, u1, u2, u3, phi, E, muc (float64),
//          boundary (int32)
// ASCII form (write_ascii): the formats of write_files.f90:44-49.
//
// The encoder and decoder are host code on host arrays; bx_write_celldump and
// bx_read_celldump add the device side: the state() call the reference makes
// before writing and after reading, on the context's state, and the potential
// column when gravity from a potential is configured (written from and read
// into the context's phi; zero otherwise, as the reference's phi without a
// potential).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bx_internal.h"
#include "bx_units.h"

namespace {

void get_units(double s[8])
{
    const bx::Units u = bx::get_units();  // declaration order of type units
    const double v[8] = {u.length, u.time, u.mass, u.density, u.rgas, u.eps, u.kelvin, u.vel};
    for (int i = 0; i < 8; i++) s[i] = v[i];
}

double muc_of(const bx_params &p) { return 1.0 / (p.xabun * 0.5 + p.yabun * 0.25 + p.zabun / p.mu_z); }

// init_grid.f90:81-118: boundary flag and centred coordinates of cell (ix,iy,iz), 1-based
int boundary_of(int ix, int iy, int iz, int nx, int ny, int nz)
{
    const int f1 = (ix == 1 || ix == nx) + (iy == 1 || iy == ny) + (nz >= 6 && (iz == 1 || iz == nz));
    const int f2 = (ix == 2 || ix == nx - 1) + (iy == 2 || iy == ny - 1) + (nz >= 6 && (iz == 2 || iz == nz - 1));
    return f1 > 0 ? 1 : (f2 > 0 ? 2 : 0);
}

double centre(double d, int i, int n) { return d * ((double)i - ((double)n + 1.0) / 2.0); }

struct Rec {
    std::vector<unsigned char> b;
    void put(const void *p, size_t n)
    {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        b.insert(b.end(), c, c + n);
    }
    void d(double v) { put(&v, 8); }
    void i(int32_t v) { put(&v, 4); }
};

bool write_record(FILE *f, const Rec &r)
{
    const int32_t len = (int32_t)r.b.size();
    return fwrite(&len, 4, 1, f) == 1 && (r.b.empty() || fwrite(r.b.data(), 1, r.b.size(), f) == r.b.size()) &&
           fwrite(&len, 4, 1, f) == 1;
}

// Fortran 1pe<w>.<d>E<e> (e=0: the default exponent field) of one value
std::string fort_e(double v, int w, int d, int e)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*E", d, v);  // d.ddddE+xx, correctly rounded
    std::string s(buf);
    const size_t epos = s.find('E');
    std::string mant = s.substr(0, epos);
    const int ex = std::atoi(s.c_str() + epos + 1);
    char eb[16];
    if (e > 0) std::snprintf(eb, sizeof eb, "E%c%0*d", ex < 0 ? '-' : '+', e, std::abs(ex));
    else if (std::abs(ex) <= 99) std::snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
    else std::snprintf(eb, sizeof eb, "%c%03d", ex < 0 ? '-' : '+', std::abs(ex));
    std::string out = mant + eb;
    if ((int)out.size() < w) out = std::string(w - out.size(), ' ') + out;
    return out;
}

// The writer on host arrays; phi: the potential column (write_files.f90:52,
// 62), zero where null (no potential configured: phi stays at its set-up
// zero, init_grid.f90:291, 299).
int encode(const bx_params *prm, const char *path, double time, int32_t step, int32_t ascii, const double *cons,
           const double *u, const double *p, const double *muc_array, const double *phi_array)
{
    if (!prm || !path || !cons || !u || !p) return bx::set_error(BX_ERR_ARG, "bx_celldump_encode: null argument");
    const int nx = prm->nx, ny = prm->ny, nz = prm->nz;
    const int64_t n = (int64_t)nx * ny * nz;
    double scale[8];
    get_units(scale);
    const double muc = muc_of(*prm);
    FILE *f = std::fopen(path, ascii ? "w" : "wb");
    if (!f) return bx::set_error(BX_ERR_ARG, std::string("bx_celldump_encode: cannot open ") + path);
    bool ok = true;
    if (ascii) {
        // write_files.f90:44-49
        std::string line = std::string("#Time of snapshot at step:").substr(0, 25) + fort_e(time, 16, 8, 3);
        char ib[32];
        std::snprintf(ib, sizeof ib, "%8d", step);
        line += ib;
        ok = ok && std::fprintf(f, "%s\n", line.c_str()) > 0;
        line = "#";  // (A,1X,8(1pe15.8,1X)): the last 1X positions only, so no trailing blank
        for (int k = 0; k < 8; k++) line += " " + fort_e(scale[k], 15, 8, 0);
        ok = ok && std::fprintf(f, "%s\n", line.c_str()) > 0;
    } else {
        Rec r1, r2, r3;
        r1.i(1234);
        r2.d(time);
        r2.i(step);
        for (double v : scale) r3.d(v);
        ok = write_record(f, r1) && write_record(f, r2) && write_record(f, r3);
    }
    int64_t c = 0;
    for (int iz = 1; iz <= nz && ok; iz++)
        for (int iy = 1; iy <= ny && ok; iy++)
            for (int ix = 1; ix <= nx && ok; ix++, c++) {
                const double v[11] = {centre(prm->dx, ix, nx), centre(prm->dy, iy, ny), centre(prm->dz, iz, nz),
                                      cons[c * 5], p[c], u[c * 3], u[c * 3 + 1], u[c * 3 + 2],
                                      phi_array ? phi_array[c] : 0.0, cons[c * 5 + 4],
                                      muc_array ? muc_array[c] : muc};
                const int32_t bd = boundary_of(ix, iy, iz, nx, ny, nz);
                if (ascii) {
                    std::string line;
                    for (double x : v) line += fort_e(x, 16, 8, 3) + " ";
                    char bb[8];
                    std::snprintf(bb, sizeof bb, "%2d", bd);
                    line += bb;
                    ok = std::fprintf(f, "%s\n", line.c_str()) > 0;
                } else {
                    Rec r;
                    for (double x : v) r.d(x);
                    r.i(bd);
                    ok = write_record(f, r);
                }
            }
    if (std::fclose(f) != 0) ok = false;
    if (!ok || c != n) return bx::set_error(BX_ERR_ARG, std::string("bx_celldump_encode: write failed: ") + path);
    return 0;
}

// The reader on host arrays; phi_out (optional) receives the potential
// column, which read_hydro.f90:129-131 reads into phi.
int decode(const bx_params *prm, const char *path, double *time, int32_t *step, double *cons, double *u, double *p,
           double *adindx, double *phi_out)
{
    if (!prm || !path || !time || !step || !cons || !u || !p || !adindx)
        return bx::set_error(BX_ERR_ARG, "bx_celldump_decode: null argument");
    const int nx = prm->nx, ny = prm->ny, nz = prm->nz;
    const int64_t n = (int64_t)nx * ny * nz;
    FILE *f = std::fopen(path, "rb");
    if (!f) return bx::set_error(BX_ERR_ARG, std::string("bx_celldump_decode: cannot open ") + path);
    auto bad = [&](const std::string &why) {
        std::fclose(f);
        return bx::set_error(BX_ERR_ARG, "bx_celldump_decode: " + why + ": " + path);
    };
    std::vector<unsigned char> buf;
    auto record = [&](size_t want) -> bool {
        int32_t a = 0, b = 0;
        if (std::fread(&a, 4, 1, f) != 1 || a < 0) return false;
        buf.resize((size_t)a);
        if (a && std::fread(buf.data(), 1, (size_t)a, f) != (size_t)a) return false;
        if (std::fread(&b, 4, 1, f) != 1 || b != a) return false;
        return (size_t)a >= want;
    };
    // read_hydro.f90:119-131
    if (!record(4)) return bad("bad first record");
    if (!record(12)) return bad("bad time record");
    std::memcpy(time, buf.data(), 8);
    std::memcpy(step, buf.data() + 8, 4);
    if (!record(64)) return bad("bad units record");
    for (int64_t c = 0; c < n; c++) {
        if (!record(92)) return bad("short file (" + std::to_string(c) + " of " + std::to_string(n) + " cells)");
        double v[11];
        std::memcpy(v, buf.data(), sizeof v);
        cons[c * 5] = v[3];
        p[c] = v[4];
        u[c * 3] = v[5];
        u[c * 3 + 1] = v[6];
        u[c * 3 + 2] = v[7];
        cons[c * 5 + 4] = v[9];
        if (phi_out) phi_out[c] = v[8];
    }
    std::fclose(f);
    // read_hydro.f90:150-178: momenta from u, gamma, then the frame cells
    int64_t c = 0;
    for (int iz = 1; iz <= nz; iz++)
        for (int iy = 1; iy <= ny; iy++)
            for (int ix = 1; ix <= nx; ix++, c++) {
                for (int k = 0; k < 3; k++) cons[c * 5 + 1 + k] = u[c * 3 + k] * cons[c * 5];
                adindx[c] = prm->gammafix;
                if (boundary_of(ix, iy, iz, nx, ny, nz) == 1) {
                    cons[c * 5] = prm->small_rho;
                    cons[c * 5 + 1] = cons[c * 5 + 2] = cons[c * 5 + 3] = 0.0;
                    cons[c * 5 + 4] = prm->small_eps;
                }
            }
    return 0;
}

}  // namespace

extern "C" int bx_celldump_encode(const bx_params *prm, const char *path, double time, int32_t step, int32_t ascii,
                                  const double *cons, const double *u, const double *p, const double *muc_array)
{
    return encode(prm, path, time, step, ascii, cons, u, p, muc_array, nullptr);
}

extern "C" int bx_celldump_decode(const bx_params *prm, const char *path, double *time, int32_t *step, double *cons,
                                  double *u, double *p, double *adindx)
{
    return decode(prm, path, time, step, cons, u, p, adindx, nullptr);
}

extern "C" int bx_write_celldump(bx_ctx *c, const char *path, double time, int32_t step, int32_t ascii)
{
    if (!c || !path) return bx::set_error(BX_ERR_ARG, "bx_write_celldump: null argument");
    if (c->prm.precision != BX_PREC_FP64) return bx::set_error(BX_ERR_UNSUPPORTED, "bx_write_celldump: FP64 contexts only");
    // write_files.f90:31-33: state() first, so the pressure is current
    int rc;
    if ((rc = bx_state(c))) return rc;
    const size_t n = (size_t)c->n;
    std::vector<double> cons(5 * n), u(3 * n), p(n), muc;
    if ((rc = bx_download(c, cons.data(), u.data(), p.data(), nullptr))) return rc;
    if (c->eos) {  // per-cell muc_array of the tabulated EOS
        muc.resize(n);
        if ((rc = bx_download_muc(c, muc.data()))) return rc;
    }
    std::vector<double> phi;
    if (c->grav_flags && c->phi) {  // the configured potential
        phi.resize(n);
        BX_HIP(hipMemcpyAsync(phi.data(), c->phi, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        BX_HIP(hipStreamSynchronize(c->stream));
    }
    return encode(&c->prm, path, time, step, ascii, cons.data(), u.data(), p.data(), c->eos ? muc.data() : nullptr,
                  phi.empty() ? nullptr : phi.data());
}

extern "C" int bx_read_celldump(bx_ctx *c, const char *path, double *time, int32_t *step)
{
    if (!c || !path || !time || !step) return bx::set_error(BX_ERR_ARG, "bx_read_celldump: null argument");
    const size_t n = (size_t)c->n;
    std::vector<double> cons(5 * n), u(3 * n), p(n), ad(n), phi(n);
    int rc;
    if ((rc = decode(&c->prm, path, time, step, cons.data(), u.data(), p.data(), ad.data(), phi.data()))) return rc;
    if ((rc = bx_upload(c, cons.data(), u.data(), p.data(), ad.data()))) return rc;
    // the potential read with the state (read_hydro.f90:129-131), where one is configured
    if (c->grav_flags && (rc = bx_set_phi(c, phi.data()))) return rc;
    return bx_state(c);  // read_hydro.f90:180
}
