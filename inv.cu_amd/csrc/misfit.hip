// Trace misfits and adjoint sources on gfx950, batched over every trace of
// every shot slot in one launch (the reference runs one <<<nt,1>>> kernel and
// one cuBLAS norm with a host sync per trace, misfit/misfit.h:31-40).
#include <complex>

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

// The same misfit for slots whose traces are shorter than the run (the
// request server pads a batch to its longest member): slot b uses the first
// nt_b samples of its rows and its own taper row, and writes its adjoint
// source into the last nt_b samples of an nt-sample row (zeros before), so
// that an adjoint of nt steps injects nothing until step nt - nt_b and then
// exactly what an adjoint of nt_b steps would.  The per-sample arithmetic and
// the FP64 summation order are those of k_waveform at nt = nt_b.
__global__ __launch_bounds__(MISFIT_BLOCK) void k_waveform_slots(
    const float *__restrict__ syn_all, const float *__restrict__ obs_all,
    float *__restrict__ adstf_all, const float *__restrict__ tw_all,
    const int *__restrict__ slot_nt, double *__restrict__ chi, int nt, int nrec, int ncomp,
    float sqrt_dt) {
    const int c = blockIdx.x / nrec, ir = blockIdx.x % nrec, b = blockIdx.y;
    const long base = ((long)b * ncomp + c) * nrec * nt + (long)ir * nt;
    const float *syn = syn_all + base, *obs = obs_all + base, *tw = tw_all + (long)b * nt;
    const int nb = slot_nt[b], pad = nt - nb;
    float *ad = adstf_all + base + pad;
    for (int it = threadIdx.x; it < pad; it += MISFIT_BLOCK) adstf_all[base + it] = 0.0f;
    double ss = 0;
    for (int it = threadIdx.x; it < nb; it += MISFIT_BLOCK) {
        const float w = tw[it];
        const float r = (syn[it] - obs[it]) * w;
        ad[nb - 1 - it] = r * w * 2;
        ss += (double)r * (double)r;
    }
    ss = block_sum<MISFIT_BLOCK>(ss);
    if (threadIdx.x == 0) chi[((long)b * nrec + ir) * ncomp + c] = (double)((float)sqrt(ss) * sqrt_dt);
}

void misfit_waveform_slots(Ctx &c, int count, const int *slot_nt, const float *taper_tab) {
    INV_REQUIRE(count >= 1 && count <= c.g.batch, INV_ERR_ARG, "slot count out of range");
    const float sqrt_dt = sqrtf(c.g.dt);
    dim3 grid(c.ncomp * c.nrec, count);
    k_waveform_slots<<<grid, MISFIT_BLOCK, 0, c.stream>>>(
        c.out.as<float>(), c.obs.as<float>(), c.adstf.as<float>(), taper_tab, slot_nt,
        c.chi.as<double>(), c.g.nt, c.nrec, c.ncomp, sqrt_dt);
    INV_HIP(hipGetLastError());
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

void launch_waveform_traces(const float *syn, const float *obs, float *adstf, const float *tw,
                            double *chi, int nt, int ntr, float dt, hipStream_t stream) {
    // ntr traces as one component of ntr receivers of one slot: trace r is
    // row r of [ntr][nt], its misfit chi[r]
    k_waveform<<<dim3(ntr, 1), MISFIT_BLOCK, 0, stream>>>(syn, obs, adstf, tw, chi, nt, ntr, 1,
                                                         sqrtf(dt));
    INV_HIP(hipGetLastError());
}

double misfit_trace(Ctx &c, int type, const float *syn, const float *obs, float *adstf) {
    c.scal.ensure(8 * sizeof(double));
    if (type == 0) {
        k_waveform<<<dim3(1, 1), MISFIT_BLOCK, 0, c.stream>>>(syn, obs, adstf, c.taper.as<float>(),
                                                             c.scal.as<double>(), c.g.nt, 1, 1,
                                                             sqrtf(c.g.dt));
        INV_HIP(hipGetLastError());
    } else {
        envelope_trace(c, syn, obs, adstf, c.scal.as<double>());
    }
    double chi = 0;
    INV_HIP(hipMemcpyAsync(&chi, c.scal.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    INV_HIP(hipStreamSynchronize(c.stream));
    return chi;
}

// ---------------------------------------------------------------------------
// Envelope misfit (misfit/envelope.h:4-106).  The reference runs ~14 kernels
// and three cuFFT C2C transforms per trace; here one 1024-thread block per
// trace keeps the trace in LDS and runs the three analytic-signal transforms
// A = IFFT(mask * FFT(x)) (unnormalised) with a mixed-radix Stockham FFT
// (radix 4, 2, 3, 5, 7; twiddles from a host FP64 table rounded to FP32).
// A length with another prime factor (cuFFT takes any length) goes through
// Bluestein's algorithm: the DFT of length n as a convolution with a chirp,
// done by power-of-two transforms of length M >= 2n - 1.  A transform longer
// than the LDS holds (n or M above ENV_MAXN) runs the same passes out of place
// between two per-trace buffers in global memory (the G variant), traces in
// chunks of at most ENV_CHUNK_BYTES of buffers.
// ---------------------------------------------------------------------------
constexpr int EB = 1024;               // threads per trace
constexpr int ENV_MAXN = 16 * EB;      // longest LDS transform (elements per thread <= 16)
constexpr int ENV_MAXN_G = 1 << 20;    // longest trace
constexpr size_t ENV_CHUNK_BYTES = size_t(512) << 20;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// One Stockham pass of radix R over the LDS buffer: every thread first loads
// all inputs of its butterflies, then (after a barrier) writes the outputs,
// so the pass is out-of-place in effect while using one buffer.
template <int R>
__device__ __noinline__ void fft_pass(float2 *buf, int n, int ns, const float2 *__restrict__ W, bool inv) {
    constexpr int BM = (ENV_MAXN + R * EB - 1) / (R * EB);
    const int nb = n / R;
    float2 v[BM][R];
#pragma unroll
    for (int c = 0; c < BM; c++) {
        const int j = threadIdx.x + c * EB;
        if (j < nb)
#pragma unroll
            for (int r = 0; r < R; r++) v[c][r] = buf[j + r * nb];
    }
    __syncthreads();
    const int span = n / (ns * R);
#pragma unroll
    for (int c = 0; c < BM; c++) {
        const int j = threadIdx.x + c * EB;
        if (j < nb) {
            const int k = j % ns;
#pragma unroll
            for (int r = 1; r < R; r++) {
                float2 w = W[(k * r * span) % n];
                if (inv) w.y = -w.y;
                v[c][r] = cmul(v[c][r], w);
            }
            const int base = (j / ns) * ns * R + k;
#pragma unroll
            for (int q = 0; q < R; q++) {
                float2 acc = v[c][0];
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = W[((q * r) % R) * (n / R)];
                    if (inv) w.y = -w.y;
                    const float2 t = cmul(v[c][r], w);
                    acc.x += t.x;
                    acc.y += t.y;
                }
                buf[base + q * ns] = acc;
            }
        }
    }
    __syncthreads();
}

// The same pass from src to dst (global memory: no register staging, any n).
template <int R>
__device__ __noinline__ void fft_pass_g(const float2 *src, float2 *dst, int n, int ns,
                                        const float2 *__restrict__ W, bool inv) {
    const int nb = n / R;
    const int span = n / (ns * R);
    for (int j = threadIdx.x; j < nb; j += EB) {
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = src[j + r * nb];
        const int k = j % ns;
#pragma unroll
        for (int r = 1; r < R; r++) {
            float2 w = W[(k * r * span) % n];
            if (inv) w.y = -w.y;
            v[r] = cmul(v[r], w);
        }
        const int base = (j / ns) * ns * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) {
            float2 acc = v[0];
#pragma unroll
            for (int r = 1; r < R; r++) {
                float2 w = W[((q * r) % R) * (n / R)];
                if (inv) w.y = -w.y;
                const float2 t = cmul(v[r], w);
                acc.x += t.x;
                acc.y += t.y;
            }
            dst[base + q * ns] = acc;
        }
    }
    __syncthreads();
}

// The FFT of buf in place through tmp (both n elements, global memory).
__device__ void fft_g(float2 *buf, float2 *tmp, int n, const int *__restrict__ fac, int nfac,
                      const float2 *__restrict__ W, bool inv) {
    int ns = 1;
    float2 *src = buf, *dst = tmp;
    for (int f = 0; f < nfac; f++) {
        const int R = fac[f];
        switch (R) {
            case 4: fft_pass_g<4>(src, dst, n, ns, W, inv); break;
            case 2: fft_pass_g<2>(src, dst, n, ns, W, inv); break;
            case 3: fft_pass_g<3>(src, dst, n, ns, W, inv); break;
            case 5: fft_pass_g<5>(src, dst, n, ns, W, inv); break;
            default: fft_pass_g<7>(src, dst, n, ns, W, inv); break;
        }
        ns *= R;
        float2 *t = src;
        src = dst;
        dst = t;
    }
    if (src != buf) {
        for (int i = threadIdx.x; i < n; i += EB) buf[i] = src[i];
        __syncthreads();
    }
}

__device__ void fft_lds(float2 *buf, int n, const int *__restrict__ fac, int nfac,
                        const float2 *__restrict__ W, bool inv) {
    int ns = 1;
    for (int f = 0; f < nfac; f++) {
        const int R = fac[f];
        switch (R) {
            case 4: fft_pass<4>(buf, n, ns, W, inv); break;
            case 2: fft_pass<2>(buf, n, ns, W, inv); break;
            case 3: fft_pass<3>(buf, n, ns, W, inv); break;
            case 5: fft_pass<5>(buf, n, ns, W, inv); break;
            default: fft_pass<7>(buf, n, ns, W, inv); break;
        }
        ns *= R;
    }
}

// The transform plan of one trace length n: a mixed-radix FFT of length n
// (m = 0), or Bluestein with power-of-two transforms of length m: with the
// chirp w[k] = exp(-i pi k^2 / n),
//   X[j] = w[j] (a * b)[j],  a[k] = x[k] w[k],  b[k] = conj(w[|k|]),
// the cyclic convolution of length m taken as IFFT(FFT(a) B) with
// B = FFT(b) / m precomputed in FP64 (bf, and bi = FFT(conj b) / m for the
// inverse transform, whose chirp is conj(w)).  fac / W: factors and twiddles
// of the FFT length (n or m).
struct EnvPlan {
    const float2 *W;
    const int *fac;
    int nfac, m;
    const float2 *chirp, *bf, *bi;
};

// an FFT of length len in LDS (G false) or through tmp (G true)
template <bool G>
__device__ void fft_any(float2 *buf, float2 *tmp, int len, const EnvPlan &p, bool inv) {
    if (G) fft_g(buf, tmp, len, p.fac, p.nfac, p.W, inv);
    else fft_lds(buf, len, p.fac, p.nfac, p.W, inv);
}

// the DFT of buf[0 .. n-1] in place (buf holds m elements for Bluestein;
// tmp as many, for the G variant)
template <bool G>
__device__ void dft_any(float2 *buf, float2 *tmp, int n, const EnvPlan &p, bool inv) {
    if (p.m == 0) {
        fft_any<G>(buf, tmp, n, p, inv);
        return;
    }
    for (int i = threadIdx.x; i < p.m; i += EB) {
        float2 a = make_float2(0.0f, 0.0f);
        if (i < n) {
            float2 w = p.chirp[i];
            if (inv) w.y = -w.y;
            a = cmul(buf[i], w);
        }
        buf[i] = a;
    }
    __syncthreads();
    fft_any<G>(buf, tmp, p.m, p, false);
    const float2 *B = inv ? p.bi : p.bf;
    for (int i = threadIdx.x; i < p.m; i += EB) buf[i] = cmul(buf[i], B[i]);
    __syncthreads();
    fft_any<G>(buf, tmp, p.m, p, true);
    for (int i = threadIdx.x; i < n; i += EB) {
        float2 w = p.chirp[i];
        if (inv) w.y = -w.y;
        buf[i] = cmul(buf[i], w);
    }
    __syncthreads();
}

// analytic signal in place (EnvelopeMisfit::hilbert + _hilbert mask,
// misfit/envelope.h:4-28,87-93): bin 0 kept, positive bins doubled, the rest
// zeroed, except that bin n/2+1 of an even length is left untouched.
template <bool G>
__device__ void analytic(float2 *buf, float2 *tmp, int n, const EnvPlan &p) {
    dft_any<G>(buf, tmp, n, p, false);
    for (int i = threadIdx.x; i < n; i += EB) {
        if (i == 0) continue;
        float2 h = buf[i];
        if (n % 2 == 0) {
            if (i < n / 2 + 1) h = make_float2(h.x * 2, h.y * 2);
            else if (i > n / 2 + 1) h = make_float2(0.0f, 0.0f);
        } else {
            if (i < (n + 1) / 2) h = make_float2(h.x * 2, h.y * 2);
            else h = make_float2(0.0f, 0.0f);
        }
        buf[i] = h;
    }
    __syncthreads();
    dft_any<G>(buf, tmp, n, p, true);
}

template <typename T, typename Op>
__device__ T env_reduce(T v, Op op) {
    __shared__ T part[EB / 64];
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) part[w] = v;
    __syncthreads();
    T s = part[0];
    for (int i = 1; i < EB / 64; i++) s = op(s, part[i]);
    return s;  // every thread gets the result
}

// One workgroup per trace t = t0 + blockIdx.y * gridDim.x + blockIdx.x of
// the [slot][component][receiver] traces.  G false: the transform buffer in
// LDS, per-sample intermediates at 3 nt t in scratch.  G true: trace t's
// buffers at (t - t0) gstride floats in scratch: the transform buffer and
// its partner (len elements each, len = m or nt), then the intermediates.
template <bool G>
__global__ __launch_bounds__(EB) void k_envelope(const float *__restrict__ syn_all,
                                                 const float *__restrict__ obs_all,
                                                 float *__restrict__ adstf_all,
                                                 const EnvPlan plan,
                                                 double *__restrict__ chi,
                                                 float *__restrict__ scratch, int nt, int nrec,
                                                 int ncomp, float sqrt_dt, long t0, long gstride) {
    extern __shared__ float2 lbuf[];
    const long t = t0 + (long)blockIdx.y * gridDim.x + blockIdx.x;
    const int b = (int)(t / ((long)ncomp * nrec));
    const int rem = (int)(t - (long)b * ncomp * nrec);
    const int c = rem / nrec, ir = rem % nrec;
    const long base = t * nt;
    const float *syn = syn_all + base, *obs = obs_all + base;
    float *ad = adstf_all + base;
    const int n = nt;
    const float fn = (float)n;
    const int len = plan.m ? plan.m : n;
    float2 *ebuf = lbuf, *tmp = nullptr;
    // per-sample intermediates; element i is only touched by thread i % EB
    float *esyn = scratch + 3 * base;
    if (G) {
        ebuf = reinterpret_cast<float2 *>(scratch + (t - t0) * gstride);
        tmp = ebuf + len;
        esyn = reinterpret_cast<float *>(tmp + len);
    }
    float *hsyn = esyn + n, *etmp = esyn + 2 * n;

    for (int i = threadIdx.x; i < n; i += EB) ebuf[i] = make_float2(syn[i], 0.0f);  // copyR2C
    __syncthreads();
    analytic<G>(ebuf, tmp, n, plan);
    float amax = 0.0f;
    for (int i = threadIdx.x; i < n; i += EB) {
        const float2 a = ebuf[i];
        const float e = sqrtf(a.x * a.x + a.y * a.y) / fn;  // copyC2Abs
        esyn[i] = e;
        hsyn[i] = a.y / fn;                                  // copyC2Imag
        amax = fmaxf(amax, fabsf(e));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += EB) ebuf[i] = make_float2(obs[i], 0.0f);
    __syncthreads();
    analytic<G>(ebuf, tmp, n, plan);
    amax = env_reduce<float>(amax, [](float x, float y) { return x > y ? x : y; });
    // max = max(0.05 * amax(E_syn), 0.05) with double literals (envelope.h:70-73)
    float mx = (float)((double)amax * 0.05);
    if ((double)mx < 0.05) mx = (float)0.05;
    double ss = 0.0;
    for (int i = threadIdx.x; i < n; i += EB) {
        const float2 a = ebuf[i];
        const float eobs = sqrtf(a.x * a.x + a.y * a.y) / fn;
        const float es = esyn[i];
        const float et = (es - eobs) / (es + mx);  // envelopeTmp
        etmp[i] = et;
        hsyn[i] = hsyn[i] * et;                     // calc(ersd, ersd, etmp)
        const float r = es - eobs;                  // calc(ersd, 1, esyn, -1, eobs)
        ss += (double)r * (double)r;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += EB) ebuf[i] = make_float2(hsyn[i], 0.0f);
    __syncthreads();
    analytic<G>(ebuf, tmp, n, plan);
    for (int i = threadIdx.x; i < n; i += EB) {
        const float ersd = ebuf[i].y / fn;
        ad[n - 1 - i] = etmp[i] * syn[i] - ersd;  // prepareEnvelopeSTF
    }
    ss = env_reduce<double>(ss, [](double x, double y) { return x + y; });
    if (threadIdx.x == 0) chi[((long)b * nrec + ir) * ncomp + c] = (double)((float)sqrt(ss) * sqrt_dt);
}

// radix-2 FFT in FP64 on the host (sign -1 forward), for the Bluestein
// filter spectra
static void fft_host(std::vector<std::complex<double>> &a) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = -2.0 * M_PI / (double)len;
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; k++) {
                const std::complex<double> w(std::cos(ang * (double)k), std::sin(ang * (double)k));
                const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

// elements of one trace's transform (LDS elements of the LDS variant)
static int env_len(const Ctx &c) { return c.env_m ? c.env_m : c.env_n; }
// floats of one trace's buffers in the global variant: two transform buffers
// and three nt-sample intermediates, rounded to an even count (float2 bases)
static long env_gstride(const Ctx &c) {
    const long f = 4L * env_len(c) + 3L * c.env_n;
    return (f + 1) & ~1L;
}

static void envelope_setup(Ctx &c) {
    const int n = c.g.nt;
    if (c.env_n == n) return;
    INV_REQUIRE(n >= 2 && n <= ENV_MAXN_G, INV_ERR_UNSUPPORTED,
                "envelope misfit supports 2 <= nt <= " + std::to_string(ENV_MAXN_G));
    auto factor = [](int len) {
        std::vector<int> f;
        for (int p : {4, 2, 3, 5, 7})
            while (len % p == 0) {
                f.push_back(p);
                len /= p;
            }
        return std::make_pair(f, len);
    };
    auto [fac, rest] = factor(n);
    int m = 0;  // Bluestein transform length, 0 for a direct one
    if (rest != 1) {
        m = 1;
        while (m < 2 * n - 1) m <<= 1;
        fac = factor(m).first;
    }
    const int len = m ? m : n;
    const bool global = len > ENV_MAXN;
    const size_t lds = (size_t)len * sizeof(float2);
    if (!global && lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_envelope<false>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error{INV_ERR_UNSUPPORTED, "device cannot give " + std::to_string(lds) +
                                                 " bytes of LDS to one workgroup (nt too long)"};
        }
    }
    std::vector<float2> W(len);
    for (int k = 0; k < len; k++) {
        const double ang = -2.0 * M_PI * (double)k / (double)len;
        W[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    std::vector<float2> blue;  // chirp [n], bf [m], bi [m]
    if (m) {
        std::vector<std::complex<double>> w(n), bf(m), bi(m);
        for (int k = 0; k < n; k++) {
            // k^2 mod 2n keeps the angle small: exp(-i pi k^2 / n) has period 2n in k^2
            const long q = (long)k * k % (2L * n);
            const double ang = -M_PI * (double)q / (double)n;
            w[k] = std::complex<double>(std::cos(ang), std::sin(ang));
        }
        for (int k = 0; k < n; k++) {
            bf[k] = std::conj(w[k]);
            bi[k] = w[k];
            if (k) {
                bf[m - k] = std::conj(w[k]);
                bi[m - k] = w[k];
            }
        }
        fft_host(bf);
        fft_host(bi);
        blue.resize((size_t)n + 2 * m);
        for (int k = 0; k < n; k++) blue[k] = make_float2((float)w[k].real(), (float)w[k].imag());
        for (int k = 0; k < m; k++) {
            blue[n + k] = make_float2((float)(bf[k].real() / m), (float)(bf[k].imag() / m));
            blue[n + m + k] = make_float2((float)(bi[k].real() / m), (float)(bi[k].imag() / m));
        }
    }
    INV_HIP(hipStreamSynchronize(c.stream));
    c.env_w.ensure(len * sizeof(float2));
    c.env_fac.ensure(fac.size() * sizeof(int) + 4);
    INV_HIP(hipMemcpy(c.env_w.p, W.data(), len * sizeof(float2), hipMemcpyHostToDevice));
    INV_HIP(hipMemcpy(c.env_fac.p, fac.data(), fac.size() * sizeof(int), hipMemcpyHostToDevice));
    if (m) {
        c.env_blue.ensure(blue.size() * sizeof(float2));
        INV_HIP(hipMemcpy(c.env_blue.p, blue.data(), blue.size() * sizeof(float2),
                          hipMemcpyHostToDevice));
    }
    c.env_nfac = (int)fac.size();
    c.env_m = m;
    c.env_n = n;
    c.env_global = global ? 1 : 0;
}

static EnvPlan env_plan(Ctx &c) {
    EnvPlan p{c.env_w.as<float2>(), c.env_fac.as<int>(), c.env_nfac, c.env_m, nullptr, nullptr,
              nullptr};
    if (c.env_m) {
        const float2 *b = c.env_blue.as<float2>();
        p.chirp = b;
        p.bf = b + c.env_n;
        p.bi = b + c.env_n + c.env_m;
    }
    return p;
}

// k_envelope over `traces` traces of the [slot][component][receiver] layout:
// one launch (LDS variant) or launches of at most ENV_CHUNK_BYTES of
// per-trace buffers each (global variant)
static void envelope_launch(Ctx &c, const float *syn, const float *obs, float *adstf,
                            double *chi, long traces, int nrec, int ncomp) {
    const EnvPlan plan = env_plan(c);
    const float sdt = sqrtf(c.g.dt);
    if (!c.env_global) {
        c.env_scratch.ensure(3 * (size_t)traces * c.g.nt * sizeof(float));
        dim3 grid(ncomp * nrec, (unsigned)(traces / ((long)ncomp * nrec)));
        k_envelope<false><<<grid, EB, (size_t)env_len(c) * sizeof(float2), c.stream>>>(
            syn, obs, adstf, plan, chi, c.env_scratch.as<float>(), c.g.nt, nrec, ncomp, sdt, 0,
            0);
        INV_HIP(hipGetLastError());
        return;
    }
    const long gs = env_gstride(c);
    const long per = gs * (long)sizeof(float);
    long chunk = (long)(ENV_CHUNK_BYTES / per);
    if (chunk < 1) chunk = 1;
    if (chunk > traces) chunk = traces;
    c.env_gbuf.ensure((size_t)chunk * per);
    for (long t0 = 0; t0 < traces; t0 += chunk) {
        const long k = traces - t0 < chunk ? traces - t0 : chunk;
        k_envelope<true><<<dim3((unsigned)k, 1), EB, 0, c.stream>>>(
            syn, obs, adstf, plan, chi, c.env_gbuf.as<float>(), c.g.nt, nrec, ncomp, sdt, t0, gs);
        INV_HIP(hipGetLastError());
    }
}

void envelope_trace(Ctx &c, const float *syn, const float *obs, float *adstf, double *chi_dev) {
    envelope_setup(c);
    envelope_launch(c, syn, obs, adstf, chi_dev, 1, 1, 1);
}

void misfit_envelope(Ctx &c, int count, int first) {
    envelope_setup(c);
    const long tr = (long)c.ncomp * c.nrec * c.g.nt;
    envelope_launch(c, c.out.as<float>(), c.obs.as<float>() + (long)first * tr,
                    c.adstf.as<float>(), c.chi.as<double>(), (long)count * c.ncomp * c.nrec,
                    c.nrec, c.ncomp);
}

}  // namespace inv
