// Trace misfits and adjoint sources on gfx950, batched over every trace of
// every shot slot in one launch (the reference runs one <<<nt,1>>> kernel and
// one cuBLAS norm with a host sync per trace, misfit/misfit.h:31-40).
#include "internal.h"

namespace inv {

// Block-wide FP64 sum: wavefront shuffles, then one LDS slot per wavefront.
template <int BLOCK>
__device__ double block_sum(double v) {
    __shared__ double part[BLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < BLOCK / 64; i++) s += part[i];
    return s;
}

// WaveformMisfit::run + diff (misfit/waveform.h:4-8,35-40):
//   r = (syn - obs) * tw;  adstf[nt-1-it] = r * tw * 2;  chi = ||r|| * sqrt(dt)
// One block per trace; blockIdx.x = comp*nrec + ir, blockIdx.y = slot.
constexpr int MISFIT_BLOCK = 256;
__global__ __launch_bounds__(MISFIT_BLOCK) void k_waveform(
    const float *__restrict__ syn_all, const float *__restrict__ obs_all,
    float *__restrict__ adstf_all, const float *__restrict__ tw, double *__restrict__ chi,
    int nt, int nrec, int ncomp, float sqrt_dt) {
    const int c = blockIdx.x / nrec, ir = blockIdx.x % nrec, b = blockIdx.y;
    const long base = ((long)b * ncomp + c) * nrec * nt + (long)ir * nt;
    const float *syn = syn_all + base, *obs = obs_all + base;
    float *ad = adstf_all + base;
    double ss = 0;
    for (int it = threadIdx.x; it < nt; it += MISFIT_BLOCK) {
        const float w = tw[it];
        const float r = (syn[it] - obs[it]) * w;
        ad[nt - 1 - it] = r * w * 2;
        ss += (double)r * (double)r;
    }
    ss = block_sum<MISFIT_BLOCK>(ss);
    if (threadIdx.x == 0) chi[((long)b * nrec + ir) * ncomp + c] = (double)((float)sqrt(ss) * sqrt_dt);
}

void misfit_waveform(Ctx &c, int count, int first) {
    const long tr = (long)c.ncomp * c.nrec * c.g.nt;
    const float sqrt_dt = sqrtf(c.g.dt);
    dim3 grid(c.ncomp * c.nrec, count);
    k_waveform<<<grid, MISFIT_BLOCK, 0, c.stream>>>(
        c.out.as<float>(), c.obs.as<float>() + (long)first * tr, c.adstf.as<float>(),
        c.taper.as<float>(), c.chi.as<double>(), c.g.nt, c.nrec, c.ncomp, sqrt_dt);
    INV_HIP(hipGetLastError());
}

void misfit_envelope(Ctx &c, int count, int first) {
    (void)c; (void)count; (void)first;
    throw Error{INV_ERR_UNSUPPORTED, "envelope misfit not built yet"};
}

}  // namespace inv
