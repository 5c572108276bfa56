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
// to LDS; (c8, c24) and (a1, a2) are the x or z constants of Geo.  Fast
// mode fuses multiply-adds throughout (dm, dp and the updates below), the
// same expressions in every kernel, so the fast kernels agree bit for bit
// with each other as the exact ones do.
// On values: dm(f[k], f[k-s], f[k+s], f[k-2s]) and dp(f[k+s], f[k], f[k+2s], f[k-s]).
// Fast mode evaluates a1 (f0 - fm1) - a2 (fp1 - fm2) as one fused
// multiply-add on the product a2 (fp1 - fm2); with a1 = a2 = 0 it gives a
// zero, which the fast row kernels use to drop a difference outside its
// range (finite operands).
template <bool FAST>
__device__ __forceinline__ float dm(float f0, float fm1, float fp1, float fm2, float c8, float c24,
                                    float a1, float a2) {
    if (FAST) return __builtin_fmaf(a1, f0 - fm1, -(a2 * (fp1 - fm2)));
    return 9.0f * (f0 - fm1) / c8 - (fp1 - fm2) / c24;
}
template <bool FAST>
__device__ __forceinline__ float dp(float fp1, float f0, float fp2, float fm1, float c8, float c24,
                                    float a1, float a2) {
    if (FAST) return __builtin_fmaf(a1, fp1 - f0, -(a2 * (fp2 - fm1)));
    return 9.0f * (fp1 - f0) / c8 - (fp2 - fm1) / c24;
}
// The value of x in the previous / next lane of the wavefront (DPP
// wave_shr:1 / wave_shl:1; 0 in lane 0 / lane 63).  Every lane must be
// active.
__device__ __forceinline__ float lane_prev(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_next(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, true));
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
// Reduced-precision snapshot storage.  A snapshot field is scaled per row
// segment (the owned columns of one tile row of the kernel that writes it) by
// a power of two 2^e chosen from the segment's max-abs m so that m 2^e lies
// in [2^14, 2^15) (snap_exp); the product x 2^e is exact in FP32.  The
// formats store it as
//   FP16 / BF16  the nearest FP16 / BF16 value (round to nearest even),
//   I16          the nearest integer, |q| <= 32767 (16-bit fixed point
//                relative to the segment: 15 bits below its max-abs),
// and decode by multiplying with 2^-e.  Within a segment, each block of 4
// cells (counted from the segment's first column) refines the exponent by a
// 4-bit offset d in [0, 15], the one that brings the block's own max-abs
// toward [2^14, 2^15) (snap_blk_d): its cells are scaled by 2^(e + d).  The
// exponents of a slot live in its exponent table: a SNAPTAB_HDR-byte header
// holding the segment width (int) of the writing kernel, then one signed
// byte per [batch][field][row][seg], then the block offsets, nbb bytes per
// [batch][field][row]: two per byte, low nibble first, each segment starting
// on a byte (segment sg's block lb is nibble lb & 1 of byte
// sg * snap_seg_bytes(segw) + lb / 2 of its row).  All arithmetic stays FP32.
// ---------------------------------------------------------------------------
constexpr int SNAP_EMAX = 120;
constexpr int SNAPTAB_HDR = 64;
constexpr int SNAP_DMAX = 15;

// 2^e as a float, e in [-126, 127]
__device__ __host__ __forceinline__ float pow2f(int e) {
    const unsigned bits = (unsigned)(127 + e) << 23;
    float f;
    __builtin_memcpy(&f, &bits, 4);
    return f;
}

// e with m 2^e in [2^14, 32767.5) for normal m > 0 (one less when m 2^e
// would round to 2^15, so that |x 2^e| rounds to at most 32767 for every
// |x| <= m), 0 for m = 0, clamped to [-SNAP_EMAX, SNAP_EMAX]
__device__ __host__ __forceinline__ int snap_exp(float m) {
    if (!(m > 0.0f)) return 0;
    unsigned bits;
    __builtin_memcpy(&bits, &m, 4);
    const int k = (int)((bits >> 23) & 0xffu) - 126;  // m in [2^(k-1), 2^k)
    int e = 15 - k;
    e = e > SNAP_EMAX ? SNAP_EMAX : (e < -SNAP_EMAX ? -SNAP_EMAX : e);
    if (m * pow2f(e) >= 32767.5f) e -= 1;
    return e;
}

// block-offset bytes of one row segment of width segw
__device__ __host__ __forceinline__ int snap_seg_bytes(int segw) { return (segw + 7) >> 3; }

// offset d of a block of max-abs bm in a segment of exponent e: the block's
// own exponent minus e, in [0, SNAP_DMAX] (0 for an all-zero block)
__device__ __host__ __forceinline__ int snap_blk_d(float bm, int e) {
    if (!(bm > 0.0f)) return 0;
    const int d = snap_exp(bm) - e;
    return d < 0 ? 0 : (d > SNAP_DMAX ? SNAP_DMAX : d);
}

// the exponent of column j of a row: erow, the row's segment exponents
// (nseg of them); brow, its nbb block-offset bytes; segw, the writer's
// segment width.  Indexes are clamped to the row, so a table whose header
// holds no valid width cannot be read out of bounds.
__device__ __host__ __forceinline__ int snap_cell_exp(const signed char *erow,
                                                      const unsigned char *brow, int nseg,
                                                      int nbb, int segw, int j) {
    int sg = j / segw;
    sg = sg < nseg - 1 ? sg : nseg - 1;
    const int lb = (j - sg * segw) >> 2;
    int by = sg * snap_seg_bytes(segw) + (lb >> 1);
    by = by < nbb - 1 ? by : nbb - 1;
    return erow[sg] + ((brow[by] >> ((lb & 1) << 2)) & 15);
}

// one value of a 16-bit format: x 2^e encoded (e: its block's exponent)
__device__ __forceinline__ unsigned short snap_enc16(float x, int e, int fmt) {
    float y = x * pow2f(e);
    // keep the product a separate FP32 value (exact here) before the
    // conversion, in every kernel alike
    __asm__ volatile("" : "+v"(y));
    if (fmt == INV_SNAP_F16) {
        const _Float16 h = (_Float16)y;
        return __builtin_bit_cast(unsigned short, h);
    }
    if (fmt == INV_SNAP_I16) {
        float r = rintf(y);
        r = r > 32767.0f ? 32767.0f : (r < -32767.0f ? -32767.0f : r);
        return (unsigned short)(short)(int)r;
    }
    unsigned u = __float_as_uint(y);  // BF16: round to nearest even (y finite)
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float snap_dec16(unsigned short q, int e, int fmt) {
    float y;
    if (fmt == INV_SNAP_F16) y = (float)__builtin_bit_cast(_Float16, q);
    else if (fmt == INV_SNAP_I16) y = (float)(short)q;
    else y = __uint_as_float((unsigned)q << 16);
    return y * pow2f(-e);
}

// the exponents of row i of field f of batch slot b in a slot's table
__device__ __forceinline__ long snaptab_row(int b, int nsnapf, int f, int nx, int i, int nseg) {
    return SNAPTAB_HDR + (((long)b * nsnapf + f) * nx + i) * nseg;
}

// the block offsets of the same row (sblk: byte offset of the block-offset
// section in the slot's table, nbb: bytes per row)
__device__ __forceinline__ long snaptab_blk(long sblk, int b, int nsnapf, int f, int nx, int i,
                                            int nbb) {
    return sblk + (((long)b * nsnapf + f) * nx + i) * nbb;
}

// read access to one 16-bit snapshot field, decoded to FP32 (k a cell index;
// its exponent is that of its row segment plus its block's offset)
template <int FMT>
struct Snap16 {
    const unsigned short *q;
    const signed char *tab;    // the field's rows: tab[i * nseg + seg]
    const unsigned char *blk;  // the field's block offsets: blk[i * nbb + byte]
    int nz, nseg, nbb, segw;
    __device__ __forceinline__ float operator[](long k) const {
        const int i = (int)(k / nz), j = (int)(k - (long)i * nz);
        return snap_dec16(q[k],
                          snap_cell_exp(tab + (long)i * nseg, blk + (long)i * nbb, nseg, nbb,
                                        segw, j),
                          FMT);
    }
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
    if (FAST) return ab * __builtin_fmaf(d, rho_or_bdt, v);
    return ab * (v + dt * d / rho_or_bdt);
}

// stress updates (updateSY / updateSXZ, solver/fdm.h:138-149): the SH
// stresses s + dt mu d, the normal stresses s + dt ((lam + 2 mu) a + lam b)
// and the shear stress s + dt (mu (xz + zx)); and the displacement u + v dt
// (updateVY / updateVXZ, fdm.h:153,159-160), separately rounded in both
// modes: it is the recurrence by which the reference integrates velocity
// trace files into displacement (misfit/misfit.h:113-118), so displacement
// traces recorded here equal integrated velocity traces bit for bit
template <bool FAST>
__device__ __forceinline__ float s_sh(float s, float dt, float mu, float d) {
    if (FAST) return __builtin_fmaf(dt * mu, d, s);
    return s + dt * mu * d;
}
// The normal stresses as a pair: sxx from (xx, zz), szz from (zz, xx).  Fast
// mode writes them as s + dt (2 mu a + lam (xx + zz)) with the shared term
// formed once: with mu = 0 (a fluid) both increments are then the same
// number, as in exact arithmetic, where a fused lam a + lam b would round the
// two differently and leave a spurious deviatoric stress that the fluid
// cannot relax (traces in a water layer drift by 1e-4 over 4000 steps).
template <bool FAST>
__device__ __forceinline__ void s_normal(float sxx, float szz, float dt, float lam, float mu,
                                         float xx, float zz, float &oxx, float &ozz) {
    if (FAST) {
        const float c = lam * (xx + zz), m2 = 2.0f * mu;
        oxx = __builtin_fmaf(dt, __builtin_fmaf(m2, xx, c), sxx);
        ozz = __builtin_fmaf(dt, __builtin_fmaf(m2, zz, c), szz);
        return;
    }
    oxx = sxx + dt * ((lam + 2 * mu) * xx + lam * zz);
    ozz = szz + dt * ((lam + 2 * mu) * zz + lam * xx);
}
template <bool FAST>
__device__ __forceinline__ float s_shear(float s, float dt, float mu, float xz, float zx) {
    if (FAST) return __builtin_fmaf(dt * mu, xz + zx, s);
    return s + dt * (mu * (xz + zx));
}
template <bool FAST>
__device__ __forceinline__ float u_step(float u, float v, float dt) {
    return u + v * dt;
}

}  // namespace inv
