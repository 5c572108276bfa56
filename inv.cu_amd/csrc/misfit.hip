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
// Lengths with another prime factor are rejected.
// ---------------------------------------------------------------------------
constexpr int EB = 1024;               // threads per trace
constexpr int ENV_MAXN = 16 * EB;      // longest trace (elements per thread <= 16)

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

// analytic signal in place (EnvelopeMisfit::hilbert + _hilbert mask,
// misfit/envelope.h:4-28,87-93): bin 0 kept, positive bins doubled, the rest
// zeroed, except that bin n/2+1 of an even length is left untouched.
__device__ void analytic_lds(float2 *buf, int n, const int *fac, int nfac, const float2 *W) {
    fft_lds(buf, n, fac, nfac, W, false);
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
    fft_lds(buf, n, fac, nfac, W, true);
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

__global__ __launch_bounds__(EB) void k_envelope(const float *__restrict__ syn_all,
                                                 const float *__restrict__ obs_all,
                                                 float *__restrict__ adstf_all,
                                                 const float2 *__restrict__ W,
                                                 const int *__restrict__ fac, int nfac,
                                                 double *__restrict__ chi,
                                                 float *__restrict__ scratch, int nt, int nrec,
                                                 int ncomp, float sqrt_dt) {
    extern __shared__ float2 ebuf[];
    const int c = blockIdx.x / nrec, ir = blockIdx.x % nrec, b = blockIdx.y;
    const long base = ((long)b * ncomp + c) * nrec * nt + (long)ir * nt;
    const float *syn = syn_all + base, *obs = obs_all + base;
    float *ad = adstf_all + base;
    const int n = nt;
    const float fn = (float)n;
    // per-sample intermediates; element i is only touched by thread i % EB
    float *esyn = scratch + 3 * base, *hsyn = esyn + n, *etmp = esyn + 2 * n;

    for (int i = threadIdx.x; i < n; i += EB) ebuf[i] = make_float2(syn[i], 0.0f);  // copyR2C
    __syncthreads();
    analytic_lds(ebuf, n, fac, nfac, W);
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
    analytic_lds(ebuf, n, fac, nfac, W);
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
    analytic_lds(ebuf, n, fac, nfac, W);
    for (int i = threadIdx.x; i < n; i += EB) {
        const float ersd = ebuf[i].y / fn;
        ad[n - 1 - i] = etmp[i] * syn[i] - ersd;  // prepareEnvelopeSTF
    }
    ss = env_reduce<double>(ss, [](double x, double y) { return x + y; });
    if (threadIdx.x == 0) chi[((long)b * nrec + ir) * ncomp + c] = (double)((float)sqrt(ss) * sqrt_dt);
}

static void envelope_setup(Ctx &c) {
    const int n = c.g.nt;
    if (c.env_n == n) return;
    INV_REQUIRE(n >= 2 && n <= ENV_MAXN, INV_ERR_UNSUPPORTED,
                "envelope misfit supports 2 <= nt <= " + std::to_string(ENV_MAXN));
    std::vector<int> fac;
    int m = n;
    for (int p : {4, 2, 3, 5, 7})
        while (m % p == 0) {
            fac.push_back(p);
            m /= p;
        }
    INV_REQUIRE(m == 1, INV_ERR_UNSUPPORTED,
                "envelope misfit needs nt with prime factors 2, 3, 5, 7 only (nt = " +
                    std::to_string(n) + ")");
    const size_t lds = (size_t)n * sizeof(float2);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_envelope),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error{INV_ERR_UNSUPPORTED, "device cannot give " + std::to_string(lds) +
                                                 " bytes of LDS to one workgroup (nt too long)"};
        }
    }
    std::vector<float2> W(n);
    for (int k = 0; k < n; k++) {
        const double ang = -2.0 * M_PI * (double)k / (double)n;
        W[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    INV_HIP(hipStreamSynchronize(c.stream));
    c.env_w.ensure(n * sizeof(float2));
    c.env_fac.ensure(fac.size() * sizeof(int) + 4);
    INV_HIP(hipMemcpy(c.env_w.p, W.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    INV_HIP(hipMemcpy(c.env_fac.p, fac.data(), fac.size() * sizeof(int), hipMemcpyHostToDevice));
    c.env_nfac = (int)fac.size();
    c.env_n = n;
}

void envelope_trace(Ctx &c, const float *syn, const float *obs, float *adstf, double *chi_dev) {
    envelope_setup(c);
    c.env_scratch.ensure(3 * (size_t)c.g.nt * sizeof(float));
    k_envelope<<<dim3(1, 1), EB, (size_t)c.g.nt * sizeof(float2), c.stream>>>(
        syn, obs, adstf, c.env_w.as<float2>(), c.env_fac.as<int>(), c.env_nfac, chi_dev,
        c.env_scratch.as<float>(), c.g.nt, 1, 1, sqrtf(c.g.dt));
    INV_HIP(hipGetLastError());
}

void misfit_envelope(Ctx &c, int count, int first) {
    envelope_setup(c);
    const long tr = (long)c.ncomp * c.nrec * c.g.nt;
    c.env_scratch.ensure(3 * (size_t)count * tr * sizeof(float));
    dim3 grid(c.ncomp * c.nrec, count);
    k_envelope<<<grid, EB, (size_t)c.g.nt * sizeof(float2), c.stream>>>(
        c.out.as<float>(), c.obs.as<float>() + (long)first * tr, c.adstf.as<float>(),
        c.env_w.as<float2>(), c.env_fac.as<int>(), c.env_nfac, c.chi.as<double>(),
        c.env_scratch.as<float>(), c.g.nt, c.nrec, c.ncomp, sqrtf(c.g.dt));
    INV_HIP(hipGetLastError());
}

}  // namespace inv
