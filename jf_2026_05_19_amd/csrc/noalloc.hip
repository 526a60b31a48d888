// noalloc.hip -- K assembly for callers that cannot allocate device memory.
//
// sol105_assemble_K stages every element's 6x6 node blocks in a buffer the
// model allocates at upload, then sums them per BSR block in a fixed element
// order.  sol105_assemble_K_noalloc runs the same element kernels and the
// same gather, but stages the blocks in `scratch`, a file-scope device array
// of SOL105_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it runs
// on until the process ends.  The call itself allocates nothing: no
// hipMalloc, no hipFree, no allocating library routine.
//
// The array's name is a host-side stand-in in host code.  Its device address
// is looked up once per device with hipGetSymbolAddress (sol105_ctx_create
// calls noalloc_setup with that device current) and only that address is
// passed to kernels and copies.  Every call on one GPU shares that GPU's
// array: two calls that overlap on one GPU (two contexts, two threads) race.
// The element kernels overwrite every staged entry the gather reads, so
// nothing has to be cleared between calls.
#include "internal.h"

#include <mutex>

#define S105_MAX_DEVICES 64

namespace {

constexpr size_t kScratchDoubles = SOL105_NOALLOC_SCRATCH_BYTES / sizeof(double);

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double *g_ptr[S105_MAX_DEVICES] = {nullptr};
hipError_t g_status[S105_MAX_DEVICES] = {};
bool g_done[S105_MAX_DEVICES] = {false};

}  // namespace

namespace s105 {

// Looks up this device's copy of `scratch` (the caller has made `device`
// current).  Idempotent per device; the status is kept for the entry point.
hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= S105_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_status[device] = e;
        g_ptr[device] = (e == hipSuccess) ? static_cast<double *>(p) : nullptr;
        g_done[device] = true;
    }
    return g_status[device];
}

}  // namespace s105

extern "C" int sol105_assemble_K_noalloc(sol105_ctx *ctx, sol105_model *M, double *d_vals) {
    if (!ctx || !M) return s105::set_err(ctx, SOL105_E_INVALID, "null argument");
    const int dev = ctx->device;
    double *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (dev < 0 || dev >= S105_MAX_DEVICES || !g_done[dev])
            return s105::set_err(ctx, SOL105_E_STATE, "no-allocation scratch not set up for device %d", dev);
        if (g_status[dev] != hipSuccess || !g_ptr[dev])
            return s105::set_err(ctx, SOL105_E_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                                 hipGetErrorString(g_status[dev]));
        ptr = g_ptr[dev];
    }
    if (M->has_mpc())
        return s105::set_err(ctx, SOL105_E_UNSUPPORTED,
                             "models with RBE2/RBE3 constraints are assembled by sol105_assemble_K");
    const size_t need = ((size_t)M->n_q4 * 16 + (size_t)M->n_t3 * 9 + (size_t)M->n_bar * 4) * 36;
    if (need > kScratchDoubles)
        return s105::set_err(ctx, SOL105_E_INVALID,
                             "model needs %zu bytes of element-block staging; the no-allocation scratch holds %zu "
                             "(use sol105_assemble_K)",
                             need * sizeof(double), (size_t)SOL105_NOALLOC_SCRATCH_BYTES);
    (void)hipSetDevice(dev);
    hipStream_t s = ctx->stream;
    ElemArgs a = s105::elem_args(M);
    launch_q4_ke(a, ptr, s);
    launch_t3_ke(a, ptr, s);
    launch_bar_ke(a, ptr, s);
    launch_block_gather(M->pat.n_blocks(), M->d_cptr.p, M->d_cidx.p, ptr, M->d_K.p, s);
    HIPCHK(hipGetLastError());
    M->K_valid = true;
    M->factor_valid = M->u_valid = M->KG_valid = false;
    if (d_vals) {
        HIPCHK(hipMemcpyAsync(d_vals, M->d_K.p, M->pat.n_blocks() * 36 * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    return SOL105_OK;
}
