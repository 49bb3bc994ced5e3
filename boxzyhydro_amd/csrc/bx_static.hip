// bx_static.hip -- bx_step_static: one fused FP64 step between the caller's
// device planes that allocates nothing.  Its intermediates (U*, p*, the
// y-term planes) and its flag and rate words live in `scratch`, a file-scope
// device array of fixed size; every GPU holds its own copy of it until the
// process ends, and calls that overlap on one GPU share (and race on) it.
#include <cmath>
#include <mutex>

#include "bx_internal.h"

namespace bx {

int fused_supported(const bx_ctx *c);
void fused_defaults(bx_ctx *c);
int fused_step_words(bx_ctx *c, double dt, const double *u0, double *us, double *ps, double *un, double *ey,
                     unsigned int *flags, unsigned long long *rate);

constexpr size_t kStaticHeader = 256;  // flag and rate words, then the planes
constexpr size_t kStaticBytes = kStaticHeader + (size_t)BX_STATIC_MAX_CELLS * BX_STATIC_BYTES_PER_CELL;

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];  // device address of `scratch` on each GPU (null until resolved)
}  // namespace

// The device address of this GPU's `scratch` (device `dev` current).  The
// host-side name of the array is not a device address.
static int scratch_ptr(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= kMaxDevices) return set_error(BX_ERR_ARG, "bx_step_static: device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_check(e, "hipGetSymbolAddress(scratch)");
        if (!p) return set_error(BX_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_ptr[dev] = p;
    }
    *out = (unsigned char *)g_ptr[dev];
    return BX_OK;
}

__global__ void k_static_out(const unsigned int *flags, const unsigned long long *rate, uint32_t *flags_out,
                             double *rate_out)
{
    if (flags_out) *flags_out = *flags;
    if (rate_out) *rate_out = __longlong_as_double((long long)*rate);
}

}  // namespace bx

using namespace bx;

extern "C" int bx_step_static(bx_ctx *c, const double *cons_in, double *cons_out, double dt, double *rate_out,
                              uint32_t *flags_out, void *stream)
{
    if (!c || !cons_in || !cons_out) return set_error(BX_ERR_ARG, "bx_step_static: null argument");
    if (cons_in == cons_out) return set_error(BX_ERR_ARG, "bx_step_static: cons_in and cons_out must differ");
    if (c->prm.precision != BX_PREC_FP64) return set_error(BX_ERR_UNSUPPORTED, "bx_step_static: FP64 contexts only");
    if (c->n > BX_STATIC_MAX_CELLS)
        return set_error(BX_ERR_ARG, "bx_step_static: ngrid " + std::to_string(c->n) + " exceeds the static limit of " +
                                         std::to_string((long long)BX_STATIC_MAX_CELLS) + " cells");
    if (!(dt > 0.0) || !std::isfinite(dt)) return set_error(BX_ERR_ARG, "bx_step_static: dt must be positive and finite");
    if (int rc = fused_supported(c)) return rc;
    if (c->grav_flags) return set_error(BX_ERR_UNSUPPORTED, "bx_step_static: no gravity configuration (zero potential)");
    BX_HIP(hipSetDevice(c->device));
    unsigned char *base = nullptr;
    if (int rc = scratch_ptr(c->device, &base)) return rc;
    fused_defaults(c);
    unsigned int *flags = (unsigned int *)base;
    unsigned long long *rate = (unsigned long long *)(base + 64);
    double *us = (double *)(base + kStaticHeader);
    double *ps = us + 5 * c->n;
    double *ey = ps + c->n;
    const hipStream_t keep = c->stream;
    if (stream) c->stream = (hipStream_t)stream;
    int rc = fused_step_words(c, dt, cons_in, us, ps, cons_out, ey, flags, rate);
    if (!rc) {
        hipLaunchKernelGGL(k_static_out, dim3(1), dim3(1), 0, c->stream, flags, rate, flags_out, rate_out);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = hip_check(e, "k_static_out");
    }
    c->stream = keep;
    return rc;
}
