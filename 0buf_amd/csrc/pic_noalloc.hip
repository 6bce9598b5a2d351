/*
 * pic_noalloc.hip -- pic_moments_noalloc: the per-sort particle moments of
 * pic_moments, enqueued on the context stream without any device allocation
 * or host synchronisation, for callers that cannot allocate (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).
 *
 * The moment kernels reduce in two passes: nblocks workgroups each write 27
 * partial sums, then one small kernel adds them up.  Here the partial sums
 * live in `scratch`, a file-scope device array of PIC_NOALLOC_MAX_BLOCKS x 27
 * doubles (442,368 bytes) that every GPU the library runs on holds until the
 * process ends.  In host code the name `scratch` denotes a host-side stand-in,
 * not a device address: the device address is taken once per device with
 * hipGetSymbolAddress during pic_create (with that device current) and kept
 * in g_scratch[device]; kernels receive that pointer.
 *
 * Every call on one GPU shares that GPU's array, which keeps the last call's
 * contents.  Nothing needs clearing: workgroup b writes all 27 of its
 * partials before the final pass reads them, and the final pass reads only
 * the nblocks partials of its own call.  Calls that overlap on one GPU (two
 * contexts' streams, or two host threads) race on the array.
 */
#include <cstdio>
#include <mutex>

#include "pic_ctx.h"

#define PIC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[PIC_NOALLOC_MAX_BLOCKS * 27];

namespace {
std::mutex g_mu;
double *g_scratch[PIC_MAX_DEVICES] = {};

double *scratch_of(int dev) {
    if (dev < 0 || dev >= PIC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_scratch[dev];
}
}  // namespace

namespace picx {
/* set-up from pic_create, with device `dev` current: the device address of
 * this device's scratch array, once per device */
int noalloc_init(int dev) {
    if (dev < 0 || dev >= PIC_MAX_DEVICES) {
        snprintf(g_hip_err, sizeof(g_hip_err), "device %d beyond the %d the library tracks", dev,
                 PIC_MAX_DEVICES);
        return PIC_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[dev]) return PIC_SUCCESS;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return record_hip(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return record_hip(hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch[dev] = (double *)ptr;
    return PIC_SUCCESS;
}
}  // namespace picx

using namespace picx;

extern "C" int pic_moments_noalloc(pic_ctx *c, int nblocks, double *d_out27) {
    if (!c || !d_out27) return PIC_ERR_INVALID;
    if (nblocks < 1 || nblocks > PIC_NOALLOC_MAX_BLOCKS) {
        snprintf(g_hip_err, sizeof(g_hip_err),
                 "pic_moments_noalloc: nblocks %d outside 1..%d (the scratch array holds %d "
                 "blocks of partial sums)",
                 nblocks, PIC_NOALLOC_MAX_BLOCKS, PIC_NOALLOC_MAX_BLOCKS);
        return PIC_ERR_INVALID;
    }
    /* committing the staged particles would allocate the buckets */
    if (!c->committed) {
        snprintf(g_hip_err, sizeof(g_hip_err),
                 "pic_moments_noalloc: particles not committed (run a step or pic_moments first)");
        return PIC_ERR_STATE;
    }
    double *partials = scratch_of(c->dev);
    if (!partials) {
        snprintf(g_hip_err, sizeof(g_hip_err), "pic_moments_noalloc: no scratch for device %d",
                 c->dev);
        return PIC_ERR_STATE;
    }
    /* the FP32 form's packed particles are unpacked into the other set (no
     * allocation); a pending arrival's repair could allocate: refused */
    int rc = settle_noalloc(c);
    if (rc) return rc;
    if (c->fp32)
        launch_moments32(c->g, c->NC, c->p32, c->sp, c->d_slot_sort, partials, nblocks, d_out27,
                         c->stream);
    else
        launch_moments(c->NC, c->p, c->sp, c->d_slot_sort, partials, nblocks, d_out27, c->stream);
    HIPCHK(hipGetLastError());
    return PIC_SUCCESS;
}
