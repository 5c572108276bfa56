// Device-side stencil helpers shared by the propagation kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/inv_hip.h"

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
// On values: dm(f[k], f[k-s], f[k+s], f[k-2s]) and dp(f[k+s], f[k], f[k+2s], f[k-s]).
template <bool FAST>
__device__ __forceinline__ float dm(float f0, float fm1, float fp1, float fm2, float c8, float c24,
                                    float a1, float a2) {
    if (FAST) return a1 * (f0 - fm1) - a2 * (fp1 - fm2);
    return 9.0f * (f0 - fm1) / c8 - (fp1 - fm2) / c24;
}
template <bool FAST>
__device__ __forceinline__ float dp(float fp1, float f0, float fp2, float fm1, float c8, float c24,
                                    float a1, float a2) {
    if (FAST) return a1 * (fp1 - f0) - a2 * (fp2 - fm1);
    return 9.0f * (fp1 - f0) / c8 - (fp2 - fm1) / c24;
}
template <bool FAST>
__device__ __forceinline__ float d_minus(const float *__restrict__ f, long k, long s, float c8,
                                         float c24, float a1, float a2) {
    return dm<FAST>(f[k], f[k - s], f[k + s], f[k - 2 * s], c8, c24, a1, a2);
}
template <bool FAST>
__device__ __forceinline__ float d_plus(const float *__restrict__ f, long k, long s, float c8,
                                        float c24, float a1, float a2) {
    return dp<FAST>(f[k + s], f[k], f[k + 2 * s], f[k - s], c8, c24, a1, a2);
}

// ---------------------------------------------------------------------------
// Reduced-precision snapshot storage.  A snapshot field x with max-abs m is
// stored as q = x * (SNAP_RANGE / m) in FP16 or BF16 (round to nearest even)
// and read back as q * (m / SNAP_RANGE).  SNAP_RANGE = 2^15 keeps the largest
// value below FP16's 65504 and leaves 2^-24 / 2^15 of m as the smallest
// normal magnitude, so values down to ~1e-9 m keep FP16's full 11-bit
// significand.  Arithmetic stays FP32.
// ---------------------------------------------------------------------------
constexpr float SNAP_RANGE = 32768.0f;

__device__ __forceinline__ unsigned short snap_enc16(float y, int fmt) {
    if (fmt == INV_SNAP_F16) {
        const _Float16 h = (_Float16)y;
        return __builtin_bit_cast(unsigned short, h);
    }
    unsigned u = __float_as_uint(y);  // BF16: round to nearest even (y finite)
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float snap_dec16(unsigned short q, int fmt) {
    if (fmt == INV_SNAP_F16) return (float)__builtin_bit_cast(_Float16, q);
    return __uint_as_float((unsigned)q << 16);
}

// encode factor SNAP_RANGE / m for a stored max-abs m (0 for an all-zero field)
__device__ __forceinline__ float snap_enc_scale(unsigned mbits) {
    const float m = __uint_as_float(mbits);
    return m > 0.0f ? SNAP_RANGE / m : 0.0f;
}

// store one snapshot value: FP32 as is, FP16/BF16 scaled by s.  The scaled
// value is rounded to FP32 first and then to 16 bits: the empty asm keeps the
// compiler from folding the multiply and the conversion into one mixed-
// precision instruction (a single rounding), which it does in some kernels
// and not in others, so that every kernel stores the same bits.
__device__ __forceinline__ unsigned short snap_enc_mul(float x, float s, int fmt) {
    float y = x * s;
    __asm__ volatile("" : "+v"(y));
    return snap_enc16(y, fmt);
}

__device__ __forceinline__ void snap_put(void *base, int fmt, long idx, float x, float s) {
    if (fmt == INV_SNAP_F32) static_cast<float *>(base)[idx] = x;
    else static_cast<unsigned short *>(base)[idx] = snap_enc_mul(x, s, fmt);
}

// read access to one 16-bit snapshot field, decoded to FP32
template <int FMT>
struct Snap16 {
    const unsigned short *q;
    float s;  // m / SNAP_RANGE
    __device__ __forceinline__ float operator[](long k) const { return snap_dec16(q[k], FMT) * s; }
};

// D+ of a field read through an accessor (the same expression as d_plus)
template <bool FAST, class A>
__device__ __forceinline__ float d_plus_acc(const A &f, long k, long s, float c8, float c24,
                                            float a1, float a2) {
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
