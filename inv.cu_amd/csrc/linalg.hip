// Field-wide operations on gfx950: the one-sided Gaussian smoother
// (filter/gaussian.h) and the BLAS-1 helpers the optimizers call
// (lib/device.h:157-192, optimizer/optimizer.h:49-107).  Dots and norms
// accumulate in FP64 with a fixed two-pass order, so results do not depend on
// the launch shape or on timing.
#include <cmath>
#include <vector>

#include "internal.h"

namespace inv {

constexpr int RBLOCK = 256;
constexpr int RGRID = 1024;  // partials of the first reduction pass

// ---------------------------------------------------------------------------
// GaussianFilter (filter/gaussian.h:4-56).  The reference evaluates
// gaussian(i - n) with int i and size_t n: for n > i the argument wraps to
// ~2^64, its square overflows FP32 and the weight is exactly 0, so only
// n <= i contributes.  Weights w[m], m = i - n, come from the host in FP32;
// terms whose weight underflowed to 0 (m > mcut) are skipped, which leaves
// every FP32 partial sum unchanged.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gauss_x(const float *__restrict__ data,
                                                 float *__restrict__ tmp,
                                                 const float *__restrict__ w, int nx, int nz,
                                                 int mcut) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nx || j >= nz) return;
    float acc = 0.0f;
    const int n0 = i - mcut > 0 ? i - mcut : 0;
    for (int n = n0; n <= i; n++) acc += w[i - n] * data[(long)n * nz + j];
    tmp[(long)i * nz + j] = acc;
}

__global__ __launch_bounds__(256) void k_gauss_z(float *__restrict__ data,
                                                 const float *__restrict__ tmp,
                                                 const float *__restrict__ w,
                                                 const float *__restrict__ sx,
                                                 const float *__restrict__ sz, int nx, int nz,
                                                 int mcut) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= nx || j >= nz) return;
    const float *row = tmp + (long)i * nz;
    float acc = 0.0f;
    const int n0 = j - mcut > 0 ? j - mcut : 0;
    for (int n = n0; n <= j; n++) acc += w[j - n] * row[n];
    data[(long)i * nz + j] = acc / (sx[i] * sz[j]);
}

static float gauss_weight(long m, int sigma) {
    // gaussian(), filter/gaussian.h:4-8, in FP32 with device::pi = 3.1415927
    const float pi = 3.1415927f;
    float xf = (float)m;
    float sigmaf = (float)sigma;
    return (1 / (sqrtf(2 * pi) * sigmaf)) * expf(-xf * xf / (2 * sigmaf * sigmaf));
}

void gauss_filter(Ctx &c, float *field, int nx, int nz, int sigma) {
    INV_REQUIRE(nx > 0 && nz > 0 && sigma > 0, INV_ERR_ARG, "filter: bad size or sigma");
    if (c.gauss_sigma != sigma || c.gauss_nx != nx || c.gauss_nz != nz) {
        const int nmax = nx > nz ? nx : nz;
        std::vector<float> w(nmax);
        int mcut = 0;
        for (int m = 0; m < nmax; m++) {
            w[m] = gauss_weight(m, sigma);
            if (w[m] != 0.0f) mcut = m;
        }
        // initialiseGaussian (gaussian.h:9-20): gsum = sumx(i) * sumz(j)
        std::vector<float> sx(nx), sz(nz);
        for (int i = 0; i < nx; i++) {
            float acc = 0;
            for (int n = 0; n <= i; n++) acc += w[i - n];
            sx[i] = acc;
        }
        for (int j = 0; j < nz; j++) {
            float acc = 0;
            for (int n = 0; n <= j; n++) acc += w[j - n];
            sz[j] = acc;
        }
        INV_HIP(hipStreamSynchronize(c.stream));
        c.gauss_w.ensure(nmax * sizeof(float));
        c.gauss_sx.ensure(nx * sizeof(float));
        c.gauss_sz.ensure(nz * sizeof(float));
        INV_HIP(hipMemcpy(c.gauss_w.p, w.data(), nmax * sizeof(float), hipMemcpyHostToDevice));
        INV_HIP(hipMemcpy(c.gauss_sx.p, sx.data(), nx * sizeof(float), hipMemcpyHostToDevice));
        INV_HIP(hipMemcpy(c.gauss_sz.p, sz.data(), nz * sizeof(float), hipMemcpyHostToDevice));
        c.gauss_sigma = sigma;
        c.gauss_nx = nx;
        c.gauss_nz = nz;
        c.gauss_m = mcut;
    }
    c.gauss_tmp.ensure((size_t)nx * nz * sizeof(float));
    dim3 grid((nz + 63) / 64, (nx + 3) / 4), block(64, 4);
    k_gauss_x<<<grid, block, 0, c.stream>>>(field, c.gauss_tmp.as<float>(), c.gauss_w.as<float>(),
                                            nx, nz, c.gauss_m);
    k_gauss_z<<<grid, block, 0, c.stream>>>(field, c.gauss_tmp.as<float>(), c.gauss_w.as<float>(),
                                            c.gauss_sx.as<float>(), c.gauss_sz.as<float>(), nx,
                                            nz, c.gauss_m);
    INV_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// element-wise ops (device::_calc / _copy / _init, lib/device.h:126-168)
// ---------------------------------------------------------------------------
__global__ void k_axpby(float *__restrict__ c, float ka, const float *__restrict__ a, float kb,
                        const float *__restrict__ b, long n) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        c[k] = ka * a[k] + kb * b[k];
}
__global__ void k_scale(float *__restrict__ c, float ka, const float *__restrict__ a, long n) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        c[k] = ka * a[k];
}
__global__ void k_mul(float *__restrict__ c, const float *__restrict__ a,
                      const float *__restrict__ b, long n) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        c[k] = a[k] * b[k];
}
// c = a: 16-byte vectors over the 16-byte-aligned part (device buffers are
// 256-byte aligned; an offset pointer copies its unaligned head and tail one
// float at a time), four vectors in flight per thread
__global__ void k_copy(float *__restrict__ c, const float *__restrict__ a, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x, nt = (long)gridDim.x * blockDim.x;
    const bool al = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(c)) & 15) == 0;
    const long nv = al ? n / 4 : 0;
    const float4 *a4 = reinterpret_cast<const float4 *>(a);
    float4 *c4 = reinterpret_cast<float4 *>(c);
    long k = t;
    for (; k + 3 * nt < nv; k += 4 * nt) {
        const float4 x0 = a4[k], x1 = a4[k + nt], x2 = a4[k + 2 * nt], x3 = a4[k + 3 * nt];
        c4[k] = x0;
        c4[k + nt] = x1;
        c4[k + 2 * nt] = x2;
        c4[k + 3 * nt] = x3;
    }
    for (; k < nv; k += nt) c4[k] = a4[k];
    for (long e = 4 * nv + t; e < n; e += nt) c[e] = a[e];
}
__global__ void k_fill1(float *__restrict__ c, float v, long n) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        c[k] = v;
}

static unsigned ew_grid(size_t n) {
    size_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return (unsigned)(g ? g : 1);
}

void vec_axpby(Ctx &c, float *out, float ka, const float *a, float kb, const float *b, size_t n) {
    k_axpby<<<ew_grid(n), 256, 0, c.stream>>>(out, ka, a, kb, b, (long)n);
    INV_HIP(hipGetLastError());
}
void vec_scale(Ctx &c, float *out, float ka, const float *a, size_t n) {
    k_scale<<<ew_grid(n), 256, 0, c.stream>>>(out, ka, a, (long)n);
    INV_HIP(hipGetLastError());
}
void vec_mul(Ctx &c, float *out, const float *a, const float *b, size_t n) {
    k_mul<<<ew_grid(n), 256, 0, c.stream>>>(out, a, b, (long)n);
    INV_HIP(hipGetLastError());
}
void vec_copy(Ctx &c, float *out, const float *a, size_t n) {
    k_copy<<<ew_grid((n + 3) / 4), 256, 0, c.stream>>>(out, a, (long)n);
    INV_HIP(hipGetLastError());
}
void vec_fill(Ctx &c, float *out, float v, size_t n) {
    k_fill1<<<ew_grid(n), 256, 0, c.stream>>>(out, v, (long)n);
    INV_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// reductions: pass 1 writes RGRID partials (block b covers a fixed contiguous
// range), pass 2 combines them in index order.
// ---------------------------------------------------------------------------
template <typename T, typename Op>
__device__ T block_reduce(T v, Op op, T ident) {
    __shared__ T part[RBLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    T s = ident;
    if (threadIdx.x == 0)
        for (int i = 0; i < RBLOCK / 64; i++) s = op(s, part[i]);
    return s;
}

struct AddD {
    __device__ double operator()(double a, double b) const { return a + b; }
};
struct MaxF {
    __device__ float operator()(float a, float b) const { return a > b ? a : b; }
};

__global__ __launch_bounds__(RBLOCK) void k_dot_p1(const float *__restrict__ a,
                                                   const float *__restrict__ b, long n,
                                                   double *__restrict__ part) {
    const long per = (n + gridDim.x - 1) / gridDim.x;
    const long lo = (long)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
    double s = 0;
    for (long k = lo + threadIdx.x; k < hi; k += RBLOCK) s += (double)a[k] * (double)b[k];
    s = block_reduce<double>(s, AddD(), 0.0);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(RBLOCK) void k_sum_p1(const double *__restrict__ x, long n,
                                                   double *__restrict__ part) {
    const long per = (n + gridDim.x - 1) / gridDim.x;
    const long lo = (long)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
    double s = 0;
    for (long k = lo + threadIdx.x; k < hi; k += RBLOCK) s += x[k];
    s = block_reduce<double>(s, AddD(), 0.0);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(RBLOCK) void k_sum_p2(const double *__restrict__ part, int m,
                                                   double *__restrict__ out) {
    double s = 0;
    for (int k = threadIdx.x; k < m; k += RBLOCK) s += part[k];
    s = block_reduce<double>(s, AddD(), 0.0);
    if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(RBLOCK) void k_amax_p1(const float *__restrict__ a, long n,
                                                    float *__restrict__ part) {
    const long per = (n + gridDim.x - 1) / gridDim.x;
    const long lo = (long)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
    float s = 0;
    for (long k = lo + threadIdx.x; k < hi; k += RBLOCK) s = fmaxf(s, fabsf(a[k]));
    s = block_reduce<float>(s, MaxF(), 0.0f);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(RBLOCK) void k_amax_p2(const float *__restrict__ part, int m,
                                                    float *__restrict__ out) {
    float s = 0;
    for (int k = threadIdx.x; k < m; k += RBLOCK) s = fmaxf(s, part[k]);
    s = block_reduce<float>(s, MaxF(), 0.0f);
    if (threadIdx.x == 0) *out = s;
}

static void ensure_red(Ctx &c) { c.red.ensure((RGRID + 8) * sizeof(double)); }

double vec_dot(Ctx &c, const float *a, const float *b, size_t n) {
    ensure_red(c);
    double *part = c.red.as<double>();
    k_dot_p1<<<RGRID, RBLOCK, 0, c.stream>>>(a, b, (long)n, part);
    k_sum_p2<<<1, RBLOCK, 0, c.stream>>>(part, RGRID, part + RGRID);
    INV_HIP(hipGetLastError());
    double out = 0;
    INV_HIP(hipMemcpyAsync(&out, part + RGRID, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    INV_HIP(hipStreamSynchronize(c.stream));
    return out;
}

void sum_f64(Ctx &c, const double *x, size_t n, double *dev_out) {
    ensure_red(c);
    double *part = c.red.as<double>();
    k_sum_p1<<<RGRID, RBLOCK, 0, c.stream>>>(x, (long)n, part);
    k_sum_p2<<<1, RBLOCK, 0, c.stream>>>(part, RGRID, dev_out);
    INV_HIP(hipGetLastError());
}

float vec_amax(Ctx &c, const float *a, size_t n) {
    ensure_red(c);
    float *part = reinterpret_cast<float *>(c.red.as<double>());
    k_amax_p1<<<RGRID, RBLOCK, 0, c.stream>>>(a, (long)n, part);
    k_amax_p2<<<1, RBLOCK, 0, c.stream>>>(part, RGRID, part + RGRID);
    INV_HIP(hipGetLastError());
    float out = 0;
    INV_HIP(hipMemcpyAsync(&out, part + RGRID, sizeof(float), hipMemcpyDeviceToHost, c.stream));
    INV_HIP(hipStreamSynchronize(c.stream));
    return out;
}

}  // namespace inv
