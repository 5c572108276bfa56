// ORACLE — CPU restatement of libccy/inv.cu's forward+adjoint gradient path.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker.  The product never links it.
//
// Templated on the arithmetic type T (float = the reference's FP32 path,
// double = the FP64 oracle).  Expression order follows the reference line by
// line (cited per function); built with -ffp-contract=off so nothing fuses.
// The loops mirror the reference's kernel sequence (divergence into scratch,
// then updates) so that the CPU timing of this file is the timing of the
// reference's algorithm on host cores.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

int g_threads = 0;

inline int threads() {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

const float kPiF = 3.1415927f;  // host::pi and device::pi, lib/device.h:68,96

// ---------------------------------------------------------------------------
// stencils, solver/fdm.h:76-136
// D-  (stress divergence, backward staggered), valid 2 <= i < n-2
// D+  (velocity gradient, forward staggered), valid 1 <= i < n-2
// ---------------------------------------------------------------------------
template <typename T>
inline T dminus(const T *f, long k, long s, T h) {
    return (T)9 * (f[k] - f[k - s]) / ((T)8 * h) - (f[k + s] - f[k - 2 * s]) / ((T)24 * h);
}
template <typename T>
inline T dplus(const T *f, long k, long s, T h) {
    return (T)9 * (f[k + s] - f[k]) / ((T)8 * h) - (f[k + 2 * s] - f[k - s]) / ((T)24 * h);
}

template <typename T>
struct Grid {
    int nx, nz;
    long n;
    T dx, dz, dt;
};

// divSY, solver/fdm.h:76-87
template <typename T>
void div_sy(const Grid<T> &g, T *dsy, const T *sxy, const T *szy) {
    const int nx = g.nx, nz = g.nz;
#pragma omp parallel for num_threads(threads()) schedule(static)
    for (int i = 0; i < nx; i++) {
        for (int j = 0; j < nz; j++) {
            long k = (long)i * nz + j;
            T v;
            if (i >= 2 && i < nx - 2) v = dminus(sxy, k, nz, g.dx);
            else v = 0;
            if (j >= 2 && j < nz - 2) v += dminus(szy, k, 1, g.dz);
            dsy[k] = v;
        }
    }
}

// divSXZ, solver/fdm.h:88-102
template <typename T>
void div_sxz(const Grid<T> &g, T *dsx, T *dsz, const T *sxx, const T *szz, const T *sxz) {
    const int nx = g.nx, nz = g.nz;
#pragma omp parallel for num_threads(threads()) schedule(static)
    for (int i = 0; i < nx; i++) {
        for (int j = 0; j < nz; j++) {
            long k = (long)i * nz + j;
            T a, b;
            if (i >= 2 && i < nx - 2) {
                a = dminus(sxx, k, nz, g.dx);
                b = dminus(sxz, k, nz, g.dx);
            } else {
                a = 0;
                b = 0;
            }
            if (j >= 2 && j < nz - 2) {
                a += dminus(sxz, k, 1, g.dz);
                b += dminus(szz, k, 1, g.dz);
            }
            dsx[k] = a;
            dsz[k] = b;
        }
    }
}

// divVY, solver/fdm.h:103-117
template <typename T>
void div_vy(const Grid<T> &g, T *dx_, T *dz_, const T *vy) {
    const int nx = g.nx, nz = g.nz;
#pragma omp parallel for num_threads(threads()) schedule(static)
    for (int i = 0; i < nx; i++) {
        for (int j = 0; j < nz; j++) {
            long k = (long)i * nz + j;
            dx_[k] = (i >= 1 && i < nx - 2) ? dplus(vy, k, nz, g.dx) : (T)0;
            dz_[k] = (j >= 1 && j < nz - 2) ? dplus(vy, k, 1, g.dz) : (T)0;
        }
    }
}

// divVXZ, solver/fdm.h:118-136
template <typename T>
void div_vxz(const Grid<T> &g, T *dvxdx, T *dvxdz, T *dvzdx, T *dvzdz, const T *vx,
             const T *vz) {
    const int nx = g.nx, nz = g.nz;
#pragma omp parallel for num_threads(threads()) schedule(static)
    for (int i = 0; i < nx; i++) {
        for (int j = 0; j < nz; j++) {
            long k = (long)i * nz + j;
            if (i >= 1 && i < nx - 2) {
                dvxdx[k] = dplus(vx, k, nz, g.dx);
                dvzdx[k] = dplus(vz, k, nz, g.dx);
            } else {
                dvxdx[k] = 0;
                dvzdx[k] = 0;
            }
            if (j >= 1 && j < nz - 2) {
                dvxdz[k] = dplus(vx, k, 1, g.dz);
                dvzdz[k] = dplus(vz, k, 1, g.dz);
            } else {
                dvxdz[k] = 0;
                dvzdz[k] = 0;
            }
        }
    }
}

// updateSY, solver/fdm.h:138-142
template <typename T>
void update_sy(const Grid<T> &g, T *sxy, T *szy, const T *dvydx, const T *dvydz, const T *mu) {
    const T dt = g.dt;
#pragma omp parallel for num_threads(threads()) schedule(static)
    for (long k = 0; k < g.n; k++) {
        sxy[k] += dt * mu[k] * dvydx[k];
        szy[k] += dt * mu[k] * dvydz[k];
    }
}

// updateSXZ, solver/fdm.h:143-149
template <typename T>
void update_sxz(const Grid<T> &g, T *sxx, T *szz, T *sxz, const T *dvxdx, const T *dvxdz,
                const T *dvzdx, const T *dvzdz, const T *lam, const T *mu) {
    const T dt = g.dt;
#pragma omp parallel for num_threads(threads()) schedule(static)
    for (long k = 0; k < g.n; k++) {
        sxx[k] += dt * ((lam[k] + 2 * mu[k]) * dvxdx[k] + lam[k] * dvzdz[k]);
        szz[k] += dt * ((lam[k] + 2 * mu[k]) * dvzdz[k] + lam[k] * dvxdx[k]);
        sxz[k] += dt * (mu[k] * (dvxdz[k] + dvzdx[k]));
    }
}

// updateVY, solver/fdm.h:150-154
template <typename T>
void update_vy(const Grid<T> &g, T *vy, T *uy, const T *dsy, const T *rho, const T *ab) {
    const T dt = g.dt;
#pragma omp parallel for num_threads(threads()) schedule(static)
    for (long k = 0; k < g.n; k++) {
        vy[k] = ab[k] * (vy[k] + dt * dsy[k] / rho[k]);
        uy[k] += vy[k] * dt;
    }
}

// updateVXZ, solver/fdm.h:155-161
template <typename T>
void update_vxz(const Grid<T> &g, T *vx, T *vz, T *ux, T *uz, const T *dsx, const T *dsz,
                const T *rho, const T *ab) {
    const T dt = g.dt;
#pragma omp parallel for num_threads(threads()) schedule(static)
    for (long k = 0; k < g.n; k++) {
        vx[k] = ab[k] * (vx[k] + dt * dsx[k] / rho[k]);
        vz[k] = ab[k] * (vz[k] + dt * dsz[k] / rho[k]);
        ux[k] += vx[k] * dt;
        uz[k] += vz[k] * dt;
    }
}

// Working state of one propagation (FdmSolver members, solver/fdm.h:222-271).
template <typename T>
struct Prop {
    Grid<T> g;
    const orc_sim *s;
    std::vector<T> lam, mu, rho, ab;
    std::vector<T> vx, vy, vz, ux, uy, uz, sxx, szz, sxz, sxy, szy;
    std::vector<T> dsx, dsy, dsz, dvxdx, dvxdz, dvydx, dvydz, dvzdx, dvzdz;
    std::vector<T> fw1, fw2, fw3, fw4;  // *_fw derivative scratch (fdm.h:257-263)

    explicit Prop(const orc_sim *sim) : s(sim) {
        g.nx = sim->nx;
        g.nz = sim->nz;
        g.n = (long)sim->nx * sim->nz;
        g.dx = (T)sim->dx;
        g.dz = (T)sim->dz;
        g.dt = (T)sim->dt;
        const long n = g.n;
        lam.assign(sim->lambda, sim->lambda + n);
        mu.assign(sim->mu, sim->mu + n);
        rho.assign(sim->rho, sim->rho + n);
        ab.assign(sim->ab, sim->ab + n);
        auto z = [n](std::vector<T> &v) { v.assign(n, (T)0); };
        if (sim->sh) {
            z(vy); z(uy); z(sxy); z(szy); z(dsy); z(dvydx); z(dvydz);
        }
        if (sim->psv) {
            z(vx); z(vz); z(ux); z(uz); z(sxx); z(szz); z(sxz);
            z(dsx); z(dsz); z(dvxdx); z(dvxdz); z(dvzdx); z(dvzdz);
        }
        z(fw1); z(fw2); z(fw3); z(fw4);
    }
    // initWavefields, solver/fdm.h:295-311
    void zero() {
        auto z = [](std::vector<T> &v) { std::fill(v.begin(), v.end(), (T)0); };
        if (s->sh) { z(vy); z(uy); z(sxy); z(szy); }
        if (s->psv) { z(vx); z(vz); z(ux); z(uz); z(sxx); z(szz); z(sxz); }
    }
    // divS, solver/fdm.h:273-281
    void divS() {
        if (s->sh) div_sy(g, dsy.data(), sxy.data(), szy.data());
        if (s->psv) div_sxz(g, dsx.data(), dsz.data(), sxx.data(), szz.data(), sxz.data());
    }
    // divV, solver/fdm.h:282-294
    void divV() {
        if (s->sh) {
            update_vy(g, vy.data(), uy.data(), dsy.data(), rho.data(), ab.data());
            div_vy(g, dvydx.data(), dvydz.data(), vy.data());
            update_sy(g, sxy.data(), szy.data(), dvydx.data(), dvydz.data(), mu.data());
        }
        if (s->psv) {
            update_vxz(g, vx.data(), vz.data(), ux.data(), uz.data(), dsx.data(), dsz.data(),
                       rho.data(), ab.data());
            div_vxz(g, dvxdx.data(), dvxdz.data(), dvzdx.data(), dvzdz.data(), vx.data(),
                    vz.data());
            update_sxz(g, sxx.data(), szz.data(), sxz.data(), dvxdx.data(), dvxdz.data(),
                       dvzdx.data(), dvzdz.data(), lam.data(), mu.data());
        }
    }
    int nsnap() const { return (s->sh ? 2 : 0) + (s->psv ? 4 : 0); }
    int nstate() const { return (s->sh ? 4 : 0) + (s->psv ? 7 : 0); }
    void dump_state(T *out) const {
        if (!out) return;
        const long n = g.n;
        int f = 0;
        auto put = [&](const std::vector<T> &v) { std::memcpy(out + (long)f++ * n, v.data(), n * sizeof(T)); };
        if (s->sh) { put(vy); put(uy); put(sxy); put(szy); }
        if (s->psv) { put(vx); put(vz); put(ux); put(uz); put(sxx); put(szz); put(sxz); }
    }
};

// runForward, solver/fdm.h:515-570 (snapshot slots fdm.h:521-533, 552-560)
template <typename T>
int forward(const orc_sim *s, int isrc, int nsteps, T *out_x, T *out_y, T *out_z, T *snap,
            T *state) {
    if (!s || s->nx < 1 || s->nz < 1 || s->nt < 1) return -1;
    Prop<T> p(s);
    const int nt = s->nt;
    const long n = p.g.n;
    const int nz = s->nz;
    if (nsteps < 0 || nsteps > nt) nsteps = nt;
    const int nsfe = s->sfe > 0 ? nt / s->sfe : 0;
    const int nf = p.nsnap();
    p.zero();
    for (int it = 0; it < nsteps; it++) {
        int isfe = -1;
        if (s->sfe && s->adj && (it + 1) % s->sfe == 0) isfe = nsfe - (it + 1) / s->sfe;
        if (isfe >= 0 && snap) {
            T *slot = snap + (long)isfe * nf * n;
            int f = 0;
            if (s->sh) { std::memcpy(slot + (long)f * n, p.uy.data(), n * sizeof(T)); f += 2; }
            if (s->psv) {
                std::memcpy(slot + (long)f * n, p.ux.data(), n * sizeof(T));
                std::memcpy(slot + (long)(f + 1) * n, p.uz.data(), n * sizeof(T));
            }
        }
        p.divS();
        // addSTF, solver/fdm.h:163-180 (point-force sources)
        for (int is = 0; is < s->nsrc; is++) {
            if ((isrc < 0 || isrc == is) && !(s->src_type && s->src_type[is] == 1)) {
                long km = (long)s->src_ix[is] * nz + s->src_iz[is];
                long ks = (long)is * nt + it;
                if (s->sh) p.dsy[km] += (T)s->stf_y[ks];
                if (s->psv) {
                    p.dsx[km] += (T)s->stf_x[ks];
                    p.dsz[km] += (T)s->stf_z[ks];
                }
            }
        }
        p.divV();
        // explosive sources: equal stf into sxx and szz after the stress update
        if (s->psv && s->src_type)
            for (int is = 0; is < s->nsrc; is++)
                if ((isrc < 0 || isrc == is) && s->src_type[is] == 1) {
                    const long km = (long)s->src_ix[is] * nz + s->src_iz[is];
                    const T f = (T)s->stf_y[(long)is * nt + it];
                    p.sxx[km] += f;
                    p.szz[km] += f;
                }
        // saveRec, solver/fdm.h:181-196 with the choice of fdm.h:540-551
        const bool disp = s->adj || s->trace_type == 1;
        for (int ir = 0; ir < s->nrec; ir++) {
            long km = (long)s->rec_ix[ir] * nz + s->rec_iz[ir];
            long kr = (long)ir * nt + it;
            if (s->sh && out_y) out_y[kr] = disp ? p.uy[km] : p.vy[km];
            if (s->psv) {
                if (out_x) out_x[kr] = disp ? p.ux[km] : p.vx[km];
                if (out_z) out_z[kr] = disp ? p.uz[km] : p.vz[km];
            }
        }
        if (isfe >= 0 && snap) {
            T *slot = snap + (long)isfe * nf * n;
            int f = 0;
            if (s->sh) { std::memcpy(slot + (long)(f + 1) * n, p.vy.data(), n * sizeof(T)); f += 2; }
            if (s->psv) {
                std::memcpy(slot + (long)(f + 2) * n, p.vx.data(), n * sizeof(T));
                std::memcpy(slot + (long)(f + 3) * n, p.vz.data(), n * sizeof(T));
            }
        }
    }
    p.dump_state(state);
    return 0;
}

// runAdjoint, solver/fdm.h:571-612 with interaction* of fdm.h:198-219
template <typename T>
int adjoint(const orc_sim *s, const T *adstf_x, const T *adstf_y, const T *adstf_z,
            const T *snap, T *k_lambda, T *k_mu, T *k_rho, T *state) {
    if (!s || s->sfe <= 0 || s->nt % s->sfe != 0) return -1;
    Prop<T> p(s);
    const int nt = s->nt, nz = s->nz;
    const long n = p.g.n;
    const int nf = p.nsnap();
    const T ndt = (T)s->sfe * p.g.dt;  // sfe * dt, fdm.h:585
    p.zero();
    for (int it = 0; it < nt; it++) {
        p.divS();
        for (int ir = 0; ir < s->nrec; ir++) {
            long km = (long)s->rec_ix[ir] * nz + s->rec_iz[ir];
            long ks = (long)ir * nt + it;
            if (s->sh) p.dsy[km] += adstf_y[ks];
            if (s->psv) {
                p.dsx[km] += adstf_x[ks];
                p.dsz[km] += adstf_z[ks];
            }
        }
        p.divV();
        if ((it + s->sfe) % s->sfe == 0) {
            const int isfe = (it + s->sfe) / s->sfe - 1;
            const T *slot = snap + (long)isfe * nf * n;
            int f = 0;
            if (s->sh) {
                const T *uy_fw = slot + (long)f * n;
                const T *vy_fw = slot + (long)(f + 1) * n;
                f += 2;
                div_vy(p.g, p.dvydx.data(), p.dvydz.data(), p.uy.data());
                div_vy(p.g, p.fw1.data(), p.fw2.data(), uy_fw);
                const T *vy = p.vy.data();
                const T *a = p.dvydx.data(), *b = p.dvydz.data();
                const T *afw = p.fw1.data(), *bfw = p.fw2.data();
                const bool ir = s->inv_rho, im = s->inv_mu;
#pragma omp parallel for num_threads(threads()) schedule(static)
                for (long k = 0; k < n; k++) {
                    if (ir) k_rho[k] -= vy_fw[k] * vy[k] * ndt;
                    if (im) k_mu[k] -= (a[k] * afw[k] + b[k] * bfw[k]) * ndt;
                }
            }
            if (s->psv) {
                const T *ux_fw = slot + (long)f * n;
                const T *uz_fw = slot + (long)(f + 1) * n;
                const T *vx_fw = slot + (long)(f + 2) * n;
                const T *vz_fw = slot + (long)(f + 3) * n;
                div_vxz(p.g, p.dvxdx.data(), p.dvxdz.data(), p.dvzdx.data(), p.dvzdz.data(),
                        p.ux.data(), p.uz.data());
                div_vxz(p.g, p.fw1.data(), p.fw2.data(), p.fw3.data(), p.fw4.data(), ux_fw, uz_fw);
                const T *vx = p.vx.data(), *vz = p.vz.data();
                const T *xx = p.dvxdx.data(), *xz = p.dvxdz.data();
                const T *zx = p.dvzdx.data(), *zz = p.dvzdz.data();
                const T *xxf = p.fw1.data(), *xzf = p.fw2.data();
                const T *zxf = p.fw3.data(), *zzf = p.fw4.data();
                const bool ir = s->inv_rho, im = s->inv_mu, il = s->inv_lambda;
#pragma omp parallel for num_threads(threads()) schedule(static)
                for (long k = 0; k < n; k++) {
                    if (ir) k_rho[k] -= (vx_fw[k] * vx[k] + vz_fw[k] * vz[k]) * ndt;
                    if (im)
                        k_mu[k] -= (2 * xx[k] * xxf[k] + 2 * zz[k] * zzf[k] +
                                    (xz[k] + zx[k]) * (zxf[k] + xzf[k])) * ndt;
                    if (il) k_lambda[k] -= ((xx[k] + zz[k]) * (xxf[k] + zzf[k])) * ndt;
                }
            }
        }
    }
    p.dump_state(state);
    return 0;
}

// ---------------------------------------------------------------------------
// misfits
// ---------------------------------------------------------------------------
// taper, misfit/waveform.h:9-27 (float, with the double literals 0.5)
void taper_f32(int nt, float dt, float *tw) {
    float t_end = (nt - 1) * dt;
    float taper_width = t_end / 10;
    float t_min = taper_width;
    float t_max = t_end - taper_width;
    for (int it = 0; it < nt; it++) {
        float t = it * dt;
        if (t <= t_min) tw[it] = (float)(0.5 + 0.5 * cosf(kPiF * (t_min - t) / (taper_width)));
        else if (t >= t_max) tw[it] = (float)(0.5 + 0.5 * cosf(kPiF * (t_max - t) / (taper_width)));
        else tw[it] = 1;
    }
}

// WaveformMisfit::run + diff, misfit/waveform.h:4-8, 35-40
template <typename T>
double waveform(int nt, T dt, const T *syn, const T *obs, T *adstf) {
    std::vector<float> twf(nt);
    taper_f32(nt, (float)dt, twf.data());
    double ss = 0;
    for (int it = 0; it < nt; it++) {
        T tw = (T)twf[it];
        T r = (syn[it] - obs[it]) * tw;
        adstf[nt - it - 1] = r * tw * 2;
        ss += (double)r * (double)r;
    }
    // cublasSnrm2 * sqrt(dt); the norm is accumulated in double here.
    if (sizeof(T) == sizeof(float)) return (double)((float)std::sqrt(ss) * std::sqrt((float)dt));
    return std::sqrt(ss) * std::sqrt((double)dt);
}

// Mixed-radix recursive DFT (unnormalised, sign -1 forward / +1 inverse) used
// in place of cuFFT's C2C transforms (misfit/envelope.h:87-92).
void fft_rec(std::complex<double> *x, int n, int stride, int sign, std::complex<double> *out,
             std::complex<double> *tmp) {
    if (n == 1) { out[0] = x[0]; return; }
    int p = 0;
    for (int f : {4, 2, 3, 5, 7}) if (n % f == 0) { p = f; break; }
    if (p == 0) {
        // prime length: direct DFT
        for (int k = 0; k < n; k++) {
            std::complex<double> acc = 0;
            for (int t = 0; t < n; t++) {
                double ang = sign * 2.0 * M_PI * (double)((long)k * t % n) / n;
                acc += x[(long)t * stride] * std::complex<double>(std::cos(ang), std::sin(ang));
            }
            out[k] = acc;
        }
        return;
    }
    int m = n / p;
    // sub-transforms of the p decimated sequences
    for (int r = 0; r < p; r++) fft_rec(x + (long)r * stride, m, stride * p, sign, tmp + (long)r * m, out);
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            std::complex<double> acc = 0;
            for (int r = 0; r < p; r++) {
                int kk = k + q * m;
                double ang = sign * 2.0 * M_PI * (double)(((long)r * kk) % n) / n;
                acc += tmp[(long)r * m + k] * std::complex<double>(std::cos(ang), std::sin(ang));
            }
            out[k + q * m] = acc;
        }
    }
    // copy back through tmp is not needed: out holds the result
}

void dft(std::vector<std::complex<double>> &a, int sign) {
    const int n = (int)a.size();
    std::vector<std::complex<double>> out(n), tmp(n), src(a);
    fft_rec(src.data(), n, 1, sign, out.data(), tmp.data());
    a.swap(out);
}

// _hilbert mask, misfit/envelope.h:4-28 (bin n/2+1 untouched for even n)
void hilbert_mask(std::vector<std::complex<double>> &h) {
    const size_t n = h.size();
    for (size_t i = 1; i < n; i++) {
        if (n % 2 == 0) {
            if (i < n / 2 + 1) h[i] *= 2.0;
            else if (i > n / 2 + 1) h[i] = 0;
        } else {
            if (i < (n + 1) / 2) h[i] *= 2.0;
            else h[i] = 0;
        }
    }
}

// EnvelopeMisfit::hilbert, misfit/envelope.h:87-93: A = IFFT(mask * FFT(x))
void analytic(int n, const double *x, std::vector<std::complex<double>> &a) {
    a.assign(n, 0);
    for (int i = 0; i < n; i++) a[i] = x[i];
    dft(a, -1);
    hilbert_mask(a);
    dft(a, +1);
}

// EnvelopeMisfit::run, misfit/envelope.h:62-86 (transform in double, the
// element-wise steps in T)
template <typename T>
double envelope(int nt, T dt, const T *syn, const T *obs, T *adstf) {
    std::vector<double> xs(nt), xo(nt);
    for (int i = 0; i < nt; i++) { xs[i] = syn[i]; xo[i] = obs[i]; }
    std::vector<std::complex<double>> cs, co;
    analytic(nt, xs.data(), cs);
    analytic(nt, xo.data(), co);
    std::vector<T> esyn(nt), eobs(nt), etmp(nt), ersd(nt);
    for (int i = 0; i < nt; i++) {
        esyn[i] = (T)(std::abs(cs[i]) / nt);  // copyC2Abs, envelope.h:42-45
        eobs[i] = (T)(std::abs(co[i]) / nt);
    }
    T amax = 0;
    for (int i = 0; i < nt; i++) amax = std::max(amax, (T)std::fabs(esyn[i]));
    T mx = (T)((double)amax * 0.05);  // envelope.h:70-73 (double literal 0.05)
    if ((double)mx < 0.05) mx = (T)0.05;
    for (int i = 0; i < nt; i++) etmp[i] = (esyn[i] - eobs[i]) / (esyn[i] + mx);  // envelopeTmp
    for (int i = 0; i < nt; i++) ersd[i] = (T)(cs[i].imag() / nt);               // copyC2Imag
    for (int i = 0; i < nt; i++) ersd[i] = ersd[i] * etmp[i];                    // calc(c=a*b)
    std::vector<double> xr(nt);
    for (int i = 0; i < nt; i++) xr[i] = ersd[i];
    std::vector<std::complex<double>> cr;
    analytic(nt, xr.data(), cr);
    for (int i = 0; i < nt; i++) ersd[i] = (T)(cr[i].imag() / nt);
    for (int it = 0; it < nt; it++) adstf[nt - it - 1] = etmp[it] * syn[it] - ersd[it];  // prepareEnvelopeSTF
    double ss = 0;
    for (int i = 0; i < nt; i++) {
        T r = (T)1 * esyn[i] + (T)-1 * eobs[i];
        ss += (double)r * (double)r;
    }
    if (sizeof(T) == sizeof(float)) return (double)((float)std::sqrt(ss) * std::sqrt((float)dt));
    return std::sqrt(ss) * std::sqrt((double)dt);
}

// ---------------------------------------------------------------------------
// GaussianFilter, filter/gaussian.h:4-56.  gaussian(i - n) with int i and
// size_t n wraps for n > i; (float)2^64 squared overflows and the weight is
// exactly 0, so only n <= i contributes.  Weights are formed in float.
// ---------------------------------------------------------------------------
float gauss_weight(long m, int sigma) {
    float xf = (float)m;
    float sigmaf = (float)sigma;
    return (1 / (sqrtf(2 * kPiF) * sigmaf)) * expf(-xf * xf / (2 * sigmaf * sigmaf));
}

template <typename T>
void gauss(int nx, int nz, int sigma, T *data) {
    const int nmax = nx > nz ? nx : nz;
    std::vector<T> w(nmax);
    for (int m = 0; m < nmax; m++) w[m] = (T)gauss_weight(m, sigma);
    std::vector<T> sx(nx), sz(nz);
    for (int i = 0; i < nx; i++) {
        T acc = 0;
        for (int n = 0; n <= i; n++) acc += w[i - n];
        sx[i] = acc;
    }
    for (int j = 0; j < nz; j++) {
        T acc = 0;
        for (int n = 0; n <= j; n++) acc += w[j - n];
        sz[j] = acc;
    }
    // Terms whose weight underflowed to exactly 0 (m > mcut) add +0 to the
    // partial sum and are skipped; every sum keeps the reference's order
    // (n ascending).
    int mcut = 0;
    for (int m = 0; m < nmax; m++)
        if (w[m] != (T)0) mcut = m;
    const long N = (long)nx * nz;
    std::vector<T> tmp(N);
#pragma omp parallel for num_threads(threads()) schedule(static)
    for (int i = 0; i < nx; i++) {  // filterKernelX, the j loop innermost
        T *out = tmp.data() + (long)i * nz;
        for (int j = 0; j < nz; j++) out[j] = 0;
        for (int n = std::max(0, i - mcut); n <= i; n++) {
            const T wn = w[i - n];
            const T *row = data + (long)n * nz;
            for (int j = 0; j < nz; j++) out[j] += wn * row[j];
        }
    }
#pragma omp parallel for num_threads(threads()) schedule(static)
    for (int i = 0; i < nx; i++) {  // filterKernelZ
        for (int j = 0; j < nz; j++) {
            T acc = 0;
            for (int n = std::max(0, j - mcut); n <= j; n++) acc += w[j - n] * tmp[(long)i * nz + n];
            T gs = sx[i] * sz[j];  // initialiseGaussian
            data[(long)i * nz + j] = acc / gs;
        }
    }
}

// Misfit::run, misfit/misfit.h:15-51
template <typename T>
double gradient(const orc_sim *s, const T *obs, int misfit_type, int sigma, int kernel,
                T *k_lambda, T *k_mu, T *k_rho) {
    const int nt = s->nt, nrec = s->nrec;
    const long n = (long)s->nx * s->nz;
    const long tr = (long)nrec * nt;
    const int ncomp = (s->sh ? 1 : 0) + (s->psv ? 2 : 0);
    if (kernel) {
        std::fill(k_lambda, k_lambda + n, (T)0);
        std::fill(k_mu, k_mu + n, (T)0);
        std::fill(k_rho, k_rho + n, (T)0);
    }
    orc_sim sim = *s;
    sim.adj = 1;
    Prop<T> tmp(&sim);
    const int nsfe = s->sfe > 0 ? nt / s->sfe : 0;
    std::vector<T> snap(kernel ? (size_t)nsfe * tmp.nsnap() * n : 0);
    std::vector<T> ox(tr), oy(tr), oz(tr), ax(tr, 0), ay(tr, 0), az(tr, 0);
    T misfit = 0;
    for (int isrc = 0; isrc < s->nsrc; isrc++) {
        forward<T>(&sim, isrc, -1, ox.data(), oy.data(), oz.data(), kernel ? snap.data() : nullptr,
                   nullptr);
        const T *ob = obs + (long)isrc * ncomp * tr;
        for (int ir = 0; ir < nrec; ir++) {
            long irt = (long)ir * nt;
            int c = 0;
            auto run = [&](const T *syn, const T *o, T *ad) -> T {
                double chi = misfit_type == 1 ? envelope<T>(nt, (T)s->dt, syn, o, ad)
                                              : waveform<T>(nt, (T)s->dt, syn, o, ad);
                return (T)chi;
            };
            if (s->sh) { misfit += run(oy.data() + irt, ob + c * tr + irt, ay.data() + irt); c++; }
            if (s->psv) {
                misfit += run(ox.data() + irt, ob + c * tr + irt, ax.data() + irt);
                misfit += run(oz.data() + irt, ob + (c + 1) * tr + irt, az.data() + irt);
            }
        }
        if (kernel)
            adjoint<T>(&sim, ax.data(), ay.data(), az.data(), snap.data(), k_lambda, k_mu, k_rho,
                       nullptr);
    }
    if (kernel && sigma > 0) {
        gauss<T>(s->nx, s->nz, sigma, k_lambda);
        gauss<T>(s->nx, s->nz, sigma, k_mu);
        gauss<T>(s->nx, s->nz, sigma, k_rho);
    }
    return (double)misfit;
}

}  // namespace

extern "C" {

void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void) { return threads(); }

int orc_grid_dims(int npt, float xmax, float zmax, int *nx, int *nz, float *dx, float *dz) {
    if (npt < 4 || !(xmax > 0) || !(zmax > 0)) return -1;
    float a = zmax / xmax;
    float b = a + 1;
    float c = 1 - npt;
    size_t gx = (size_t)std::round((-b + std::sqrt(b * b - 4 * a * c)) / (2 * a)) + 1;
    size_t gz = (size_t)npt / gx;
    *nx = (int)gx;
    *nz = (int)gz;
    *dx = xmax / (float)(gx - 1);
    *dz = zmax / (float)(gz - 1);
    return 0;
}

void orc_absbound(int nx, int nz, int width, float alpha, int left, int right, int bottom,
                  int top, float *ab) {
    const long w = width;
    for (int i = 0; i < nx; i++) {
        for (int j = 0; j < nz; j++) {
            float v = 1;
            auto f = [&](long d) {
                float x = alpha * (float)d;
                return expf(-(x * x));
            };
            // size_t comparisons as in the reference: nx - w wraps when nx < w
            const size_t uw = (size_t)w, ui = (size_t)i, uj = (size_t)j;
            if (left && ui + 1 < uw) v *= f(w - i - 1);
            if (right && ui > (size_t)nx - uw) v *= f(w + i - nx);
            if (bottom && uj > (size_t)nz - uw) v *= f(w + j - nz);
            if (top && uj + 1 < uw) v *= f(w - j - 1);
            ab[(long)i * nz + j] = v;
        }
    }
}

int orc_calc_idx(float coord, float L, int n) {
    return (int)((double)(coord / L * (float)(n - 1)) + 0.5);
}

void orc_ricker(int nt, float dt, float f0, float t0, float *stf) {
    float amax = 0;
    for (int it = 0; it < nt; it++) {
        float t = (float)it * dt - t0;
        float pft = kPiF * f0 * t;
        stf[it] = -t * expf(-pft * pft);
        if (fabsf(stf[it]) > amax) amax = fabsf(stf[it]);
    }
    if (amax > 0)
        for (int it = 0; it < nt; it++) stf[it] /= amax;
}

void orc_source(int nt, float dt, const float *cfg, float *stf_x, float *stf_y, float *stf_z) {
    std::vector<float> stf(nt);
    orc_ricker(nt, dt, cfg[0], cfg[1], stf.data());
    const float angle = cfg[2], amp = cfg[3];
    for (int it = 0; it < nt; it++) {
        stf_x[it] = amp * stf[it] * cosf(angle);
        stf_y[it] = amp * stf[it];
        stf_z[it] = amp * stf[it] * sinf(angle);
    }
}

void orc_mesh2grid(int nx, int nz, float dx, float dz, float dmax, int npt, const float *px,
                   const float *pz, const float *pvp, const float *pvs, const float *prho,
                   float *lam, float *mu, float *rho) {
#pragma omp parallel for num_threads(threads()) schedule(static)
    for (int i = 0; i < nx; i++) {
        for (int j = 0; j < nz; j++) {
            long k = (long)i * nz + j;
            float x = i * dx;
            float z = j * dz;
            float dmin = dmax;
            for (int l = 0; l < npt; l++) {
                float ddx = x - px[l];
                float ddz = z - pz[l];
                float d = ddx * ddx + ddz * ddz;
                if (d < dmin) {
                    dmin = d;
                    lam[k] = pvp[l];
                    mu[k] = pvs[l];
                    rho[k] = prho[l];
                }
            }
        }
    }
}

void orc_vps2lm(int n, float *lam, float *mu, const float *rho) {
    for (int k = 0; k < n; k++) {
        float vp = lam[k], vs = mu[k];
        if (vp > vs) lam[k] = rho[k] * (vp * vp - 2 * vs * vs);
        else lam[k] = 0;
        mu[k] = rho[k] * vs * vs;
    }
}

void orc_lm2vps(int n, float *vp, float *vs, const float *rho) {
    for (int k = 0; k < n; k++) {
        float l = vp[k], m = vs[k];
        vp[k] = sqrtf((l + 2 * m) / rho[k]);
        vs[k] = sqrtf(m / rho[k]);
    }
}

int orc_snap_fields(const orc_sim *s) { return (s->sh ? 2 : 0) + (s->psv ? 4 : 0); }
int orc_state_fields(const orc_sim *s) { return (s->sh ? 4 : 0) + (s->psv ? 7 : 0); }

int orc_forward_f32(const orc_sim *s, int isrc, int nsteps, float *ox, float *oy, float *oz,
                    float *snap, float *state) {
    return forward<float>(s, isrc, nsteps, ox, oy, oz, snap, state);
}
int orc_forward_f64(const orc_sim *s, int isrc, int nsteps, double *ox, double *oy, double *oz,
                    double *snap, double *state) {
    return forward<double>(s, isrc, nsteps, ox, oy, oz, snap, state);
}
int orc_adjoint_f32(const orc_sim *s, const float *ax, const float *ay, const float *az,
                    const float *snap, float *kl, float *km, float *kr, float *state) {
    return adjoint<float>(s, ax, ay, az, snap, kl, km, kr, state);
}
int orc_adjoint_f64(const orc_sim *s, const double *ax, const double *ay, const double *az,
                    const double *snap, double *kl, double *km, double *kr, double *state) {
    return adjoint<double>(s, ax, ay, az, snap, kl, km, kr, state);
}

double orc_waveform_f32(int nt, float dt, const float *syn, const float *obs, float *adstf) {
    return waveform<float>(nt, dt, syn, obs, adstf);
}
double orc_waveform_f64(int nt, double dt, const double *syn, const double *obs, double *adstf) {
    return waveform<double>(nt, dt, syn, obs, adstf);
}
double orc_envelope_f32(int nt, float dt, const float *syn, const float *obs, float *adstf) {
    return envelope<float>(nt, dt, syn, obs, adstf);
}
double orc_envelope_f64(int nt, double dt, const double *syn, const double *obs, double *adstf) {
    return envelope<double>(nt, dt, syn, obs, adstf);
}
void orc_hilbert_f64(int n, const double *x, double *re, double *im) {
    std::vector<std::complex<double>> a;
    analytic(n, x, a);
    for (int i = 0; i < n; i++) { re[i] = a[i].real(); im[i] = a[i].imag(); }
}

void orc_gauss_f32(int nx, int nz, int sigma, float *data) { gauss<float>(nx, nz, sigma, data); }
void orc_gauss_f64(int nx, int nz, int sigma, double *data) { gauss<double>(nx, nz, sigma, data); }

double orc_gradient_f32(const orc_sim *s, const float *obs, int misfit_type, int sigma, int kernel,
                        float *kl, float *km, float *kr) {
    return gradient<float>(s, obs, misfit_type, sigma, kernel, kl, km, kr);
}
double orc_gradient_f64(const orc_sim *s, const double *obs, int misfit_type, int sigma,
                        int kernel, double *kl, double *km, double *kr) {
    return gradient<double>(s, obs, misfit_type, sigma, kernel, kl, km, kr);
}

}  // extern "C"
