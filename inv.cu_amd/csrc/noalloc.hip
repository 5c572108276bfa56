// inv_misfit_traces_noalloc: the waveform misfit of a batch of traces with no
// device allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops).  Its scratch (per-trace misfits and their
// total) is a file-scope device array; every other buffer is the caller's or
// was allocated by inv_grid_set (the taper).
#include <mutex>

#include "internal.h"

namespace inv {

// [0, NOALLOC_MAX_TRACES): per-trace misfits; [NOALLOC_MAX_TRACES]: the total.
// 512 KiB per device, held until the process ends.
constexpr int NOALLOC_MAX_TRACES = 65536;
static __device__ __align__(256) double scratch[NOALLOC_MAX_TRACES + 32];

// Device address of `scratch` per device.  In host code the name `scratch`
// is a host-side stand-in, not a device address: kernels and copies take
// these pointers.
constexpr int NOALLOC_MAX_DEVICES = 64;
static double *g_scratch[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

void noalloc_init(int device) {
    INV_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, INV_ERR_ARG, "device id too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return;
    void *p = nullptr;
    INV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    INV_REQUIRE(p != nullptr, INV_ERR_HIP, "no device address for the misfit scratch");
    g_scratch[device] = static_cast<double *>(p);
}

static double *scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < NOALLOC_MAX_DEVICES) ? g_scratch[device] : nullptr;
}

// total = sum of chi[0 .. n), one block, fixed order (strided partial sums,
// then the 256 partials in lane order)
__global__ __launch_bounds__(256) void k_chi_total(const double *__restrict__ chi, int n,
                                                   double *__restrict__ total) {
    __shared__ double part[256];
    double s = 0;
    for (int k = threadIdx.x; k < n; k += 256) s += chi[k];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < 256; k++) t += part[k];
        *total = t;
    }
}

}  // namespace inv

using namespace inv;

extern "C" int inv_misfit_traces_noalloc(inv_ctx *ctx, const float *syn, const float *obs,
                                         float *adstf, int ntr, double *chi, double *total) {
    if (!ctx) return INV_ERR_ARG;
    try {
        Ctx &c = *ctx;
        INV_HIP(hipSetDevice(c.device));
        INV_REQUIRE(c.have_grid && c.taper.p, INV_ERR_STATE, "grid not set (inv_grid_set)");
        INV_REQUIRE(syn && obs && adstf && total, INV_ERR_ARG, "null argument");
        INV_REQUIRE(ntr >= 1 && ntr <= NOALLOC_MAX_TRACES, INV_ERR_ARG,
                    "ntr must be in [1, 65536] (the scratch of the no-allocation misfit)");
        double *sp = scratch_of(c.device);
        INV_REQUIRE(sp != nullptr, INV_ERR_STATE, "scratch not resolved for this device");
        launch_waveform_traces(syn, obs, adstf, c.taper.as<float>(), sp, c.g.nt, ntr, c.g.dt,
                               c.stream);
        k_chi_total<<<1, 256, 0, c.stream>>>(sp, ntr, sp + NOALLOC_MAX_TRACES);
        INV_HIP(hipGetLastError());
        if (chi)
            INV_HIP(hipMemcpyAsync(chi, sp, (size_t)ntr * sizeof(double), hipMemcpyDeviceToHost,
                                   c.stream));
        INV_HIP(hipMemcpyAsync(total, sp + NOALLOC_MAX_TRACES, sizeof(double),
                               hipMemcpyDeviceToHost, c.stream));
        INV_HIP(hipStreamSynchronize(c.stream));
        c.err.clear();
        return INV_OK;
    } catch (const inv::Error &e) {
        ctx->err = e.msg;
        return e.code;
    }
}
