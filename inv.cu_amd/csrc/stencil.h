// Device-side stencil helpers shared by the propagation kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace inv {

struct Geo {
    int nx, nz;
    long N;
    float dt, c8dx, c24dx, c8dz, c24dz;
    // fast-mode coefficients, RN(9/(8h)) and RN(1/(24h))
    float a1x, a2x, a1z, a2z;
};

// 4th-order staggered differences (solver/fdm.h:79,85,106,112):
//   D-f = 9 (f[k] - f[k-s]) / (8h) - (f[k+s] - f[k-2s]) / (24h)
//   D+f = 9 (f[k+s] - f[k]) / (8h) - (f[k+2s] - f[k-s]) / (24h)
// Exact mode (FAST = false) evaluates the reference's expression with IEEE
// divisions by c8 = 8h and c24 = 24h formed in FP32, so a cell's value is the
// reference formula's FP32 value bit for bit.  Fast mode multiplies by the
// FP32-rounded coefficients 9/(8h) and 1/(24h) instead: two multiplies in
// place of two ~11-instruction divides, the same FP32 accuracy against FP64
// but not bitwise equal to the exact mode.  f may point to global memory or
// to LDS; (c8, c24) and (a1, a2) are the x or z constants of Geo.
template <bool FAST>
__device__ __forceinline__ float d_minus(const float *__restrict__ f, long k, long s, float c8,
                                         float c24, float a1, float a2) {
    if (FAST) return a1 * (f[k] - f[k - s]) - a2 * (f[k + s] - f[k - 2 * s]);
    return 9.0f * (f[k] - f[k - s]) / c8 - (f[k + s] - f[k - 2 * s]) / c24;
}
template <bool FAST>
__device__ __forceinline__ float d_plus(const float *__restrict__ f, long k, long s, float c8,
                                        float c24, float a1, float a2) {
    if (FAST) return a1 * (f[k + s] - f[k]) - a2 * (f[k + 2 * s] - f[k - s]);
    return 9.0f * (f[k + s] - f[k]) / c8 - (f[k + 2 * s] - f[k - s]) / c24;
}

// velocity update v = ab (v + dt d / rho) (updateV*, fdm.h:152,157-158); fast
// mode reads the per-run field bdt = dt / rho in place of rho.
template <bool FAST>
__device__ __forceinline__ float vel_update(float ab, float v, float dt, float d, float rho_or_bdt) {
    if (FAST) return ab * (v + d * rho_or_bdt);
    return ab * (v + dt * d / rho_or_bdt);
}

}  // namespace inv
