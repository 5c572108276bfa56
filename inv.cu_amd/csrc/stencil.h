// Device-side stencil helpers shared by the propagation kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace inv {

// 4th-order staggered differences (solver/fdm.h:79,85,106,112):
//   D-f = 9 (f[k] - f[k-s]) / (8h) - (f[k+s] - f[k-2s]) / (24h)
//   D+f = 9 (f[k+s] - f[k]) / (8h) - (f[k+2s] - f[k-s]) / (24h)
// c8 = 8h and c24 = 24h are formed in FP32 on the host, as the reference forms
// them per use.  f may point to global memory or to LDS.
__device__ __forceinline__ float d_minus(const float *__restrict__ f, long k, long s, float c8,
                                         float c24) {
    return 9.0f * (f[k] - f[k - s]) / c8 - (f[k + s] - f[k - 2 * s]) / c24;
}
__device__ __forceinline__ float d_plus(const float *__restrict__ f, long k, long s, float c8,
                                        float c24) {
    return 9.0f * (f[k + s] - f[k]) / c8 - (f[k + 2 * s] - f[k - s]) / c24;
}

struct Geo {
    int nx, nz;
    long N;
    float dt, c8dx, c24dx, c8dz, c24dz;
};

}  // namespace inv
