// Sanitizer run of the CPU oracle (oracle/oracle.cpp): every entry point on
// small problems, built with AddressSanitizer and UndefinedBehaviorSanitizer
// by tests/test_sanitizers.py.  A memory or UB error aborts the run; the
// checks below only assert that the results are finite and that the FP32 and
// FP64 instantiations agree loosely (the tight checks live in the Python
// oracle tests).  Prints "ALL OK" at the end.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../oracle/oracle.h"

using std::vector;

static int failures = 0;
#define CHECK(cond)                                                                   \
    do {                                                                              \
        if (cond) std::printf("ok   %s\n", #cond);                                     \
        else {                                                                        \
            std::printf("FAIL %s  (line %d)\n", #cond, __LINE__);                      \
            failures++;                                                               \
        }                                                                             \
    } while (0)

template <class T>
static bool finite(const vector<T> &v) {
    for (T x : v)
        if (!std::isfinite((double)x)) return false;
    return true;
}

template <class T>
static double rel(const vector<T> &a, const vector<double> &b) {
    double n = 0, d = 0;
    for (size_t i = 0; i < b.size(); i++) {
        n += ((double)a[i] - b[i]) * ((double)a[i] - b[i]);
        d += b[i] * b[i];
    }
    return d > 0 ? std::sqrt(n / d) : std::sqrt(n);
}

// one problem: SH and/or P-SV on a 44 x 38 grid (nz not a multiple of
// anything the GPU tiles use), 2 sources (one explosive when P-SV), 5
// receivers, nt 80
static void run_case(int sh, int psv, int misfit_type, int sigma) {
    const int nx = 44, nz = 38, nt = 80, N = nx * nz;
    const float dx = 10, dz = 10, dt = 1e-3f;
    vector<float> lam(N), mu(N), rho(N), ab(N);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < nz; j++) {
            const int k = i * nz + j;
            const float vp = 3000 + 4.0f * j + ((i / 7 + j / 5) % 2 ? 150.0f : 0.0f);
            const float vs = j < 3 ? 0.0f : vp / 1.73f;  // a fluid top layer
            rho[k] = j < 3 ? 1000.0f : 2200.0f;
            lam[k] = vp;
            mu[k] = vs;
        }
    orc_vps2lm(N, lam.data(), mu.data(), rho.data());
    orc_absbound(nx, nz, 8, 0.015f, 1, 1, 1, 0, ab.data());
    const int nsrc = 2, nrec = 5;
    int src_ix[nsrc] = {11, 30}, src_iz[nsrc] = {6, 9}, src_type[nsrc] = {0, psv ? 1 : 0};
    int rec_ix[nrec] = {4, 12, 20, 28, 39}, rec_iz[nrec] = {5, 5, 5, 5, 37};
    vector<float> sx(nsrc * nt), sy(nsrc * nt), sz(nsrc * nt);
    for (int s = 0; s < nsrc; s++) {
        const float cfg[4] = {12.0f, 0.06f, 30.0f * s + 10.0f, 1e9f};
        orc_source(nt, dt, cfg, &sx[s * nt], &sy[s * nt], &sz[s * nt]);
    }
    orc_sim S{};
    S.nx = nx, S.nz = nz, S.dx = dx, S.dz = dz, S.dt = dt, S.nt = nt;
    S.sh = sh, S.psv = psv, S.sfe = 10, S.adj = 1, S.trace_type = 0;
    S.inv_lambda = S.inv_mu = S.inv_rho = 1;
    S.lambda = lam.data(), S.mu = mu.data(), S.rho = rho.data(), S.ab = ab.data();
    S.nsrc = nsrc, S.src_ix = src_ix, S.src_iz = src_iz;
    S.stf_x = sx.data(), S.stf_y = sy.data(), S.stf_z = sz.data();
    S.nrec = nrec, S.rec_ix = rec_ix, S.rec_iz = rec_iz;
    S.src_type = src_type;

    const int nf = orc_snap_fields(&S), ns = orc_state_fields(&S), nsfe = nt / S.sfe;
    // forward of source 0, FP32 and FP64, with snapshots and final state
    vector<float> tx(nrec * nt), ty(nrec * nt), tz(nrec * nt), snap((size_t)nsfe * nf * N),
        st((size_t)ns * N);
    vector<double> dtx(nrec * nt), dty(nrec * nt), dtz(nrec * nt), dsnap((size_t)nsfe * nf * N),
        dst((size_t)ns * N);
    CHECK(orc_forward_f32(&S, 0, -1, tx.data(), ty.data(), tz.data(), snap.data(), st.data()) == 0);
    CHECK(orc_forward_f64(&S, 0, -1, dtx.data(), dty.data(), dtz.data(), dsnap.data(),
                          dst.data()) == 0);
    CHECK(finite(snap) && finite(st));
    CHECK(rel(snap, dsnap) < 1e-4);
    // a partial run (every source) without snapshots or state
    CHECK(orc_forward_f32(&S, -1, 17, tx.data(), ty.data(), tz.data(), nullptr, nullptr) == 0);

    // adjoint from the recorded traces used as adjoint sources
    vector<float> kl(N), km(N), kr(N);
    vector<double> dkl(N), dkm(N), dkr(N);
    CHECK(orc_forward_f32(&S, 0, -1, tx.data(), ty.data(), tz.data(), snap.data(), nullptr) == 0);
    CHECK(orc_adjoint_f32(&S, tx.data(), ty.data(), tz.data(), snap.data(), kl.data(), km.data(),
                          kr.data(), st.data()) == 0);
    CHECK(orc_adjoint_f64(&S, dtx.data(), dty.data(), dtz.data(), dsnap.data(), dkl.data(),
                          dkm.data(), dkr.data(), dst.data()) == 0);
    CHECK(finite(kl) && finite(km) && finite(kr));
    CHECK(rel(kr, dkr) < 1e-3);

    // gradient against observed data from a perturbed model
    const int ncomp = (sh ? 1 : 0) + (psv ? 2 : 0);
    vector<float> obs((size_t)nsrc * ncomp * nrec * nt);
    vector<double> dobs(obs.size());
    vector<float> mu2 = mu;
    for (int k = 0; k < N; k++) mu2[k] *= 1.02f;
    orc_sim T = S;
    T.mu = mu2.data();
    for (int s = 0; s < nsrc; s++) {
        CHECK(orc_forward_f32(&T, s, -1, tx.data(), ty.data(), tz.data(), nullptr, nullptr) == 0);
        float *o = &obs[(size_t)s * ncomp * nrec * nt];
        int c = 0;
        if (sh) std::copy(ty.begin(), ty.end(), o + (size_t)(c++) * nrec * nt);
        if (psv) {
            std::copy(tx.begin(), tx.end(), o + (size_t)(c++) * nrec * nt);
            std::copy(tz.begin(), tz.end(), o + (size_t)(c++) * nrec * nt);
        }
    }
    for (size_t i = 0; i < obs.size(); i++) dobs[i] = obs[i];
    const double f32 = orc_gradient_f32(&S, obs.data(), misfit_type, sigma, 1, kl.data(),
                                        km.data(), kr.data());
    const double f64 = orc_gradient_f64(&S, dobs.data(), misfit_type, sigma, 1, dkl.data(),
                                        dkm.data(), dkr.data());
    CHECK(std::isfinite(f32) && f64 > 0 && std::fabs(f32 - f64) <= 1e-3 * f64);
    CHECK(finite(kl) && finite(km) && finite(kr));
    CHECK(rel(kr, dkr) < 1e-2);
}

static void run_setup() {
    int nx, nz;
    float dx, dz;
    CHECK(orc_grid_dims(35 * 17, 340.0f, 160.0f, &nx, &nz, &dx, &dz) == 0);
    CHECK(orc_calc_idx(55.0f, 340.0f, nx) >= 0);
    // mesh2grid on scattered points (off the grid nodes)
    const int npt = 97;
    vector<float> px(npt), pz(npt), pvp(npt), pvs(npt), prho(npt);
    for (int p = 0; p < npt; p++) {
        px[p] = 3.5f * (p % 11) + 0.25f;
        pz[p] = 2.25f * (p / 11) + 0.5f;
        pvp[p] = 2000.0f + p, pvs[p] = 1000.0f + p, prho[p] = 2000.0f;
    }
    const int gx = 12, gz = 9;
    vector<float> l(gx * gz), m(gx * gz), r(gx * gz);
    orc_mesh2grid(gx, gz, 3.0f, 2.0f, 40.0f * 40.0f + 30.0f * 30.0f, npt, px.data(), pz.data(),
                  pvp.data(), pvs.data(), prho.data(), l.data(), m.data(), r.data());
    CHECK(finite(l) && finite(m) && finite(r));
    vector<float> vp = l, vs = m;
    orc_lm2vps(gx * gz, vp.data(), vs.data(), r.data());
    CHECK(finite(vp) && finite(vs));
    // misfits on every length from 2 to 70 (FFT sizes of every factorisation)
    for (int n = 2; n <= 70; n++) {
        vector<float> a(n), b(n), ad(n);
        vector<double> da(n), db(n), dad(n), re(n), im(n);
        for (int t = 0; t < n; t++) {
            a[t] = std::sin(0.3f * t) + 0.1f * t, b[t] = std::cos(0.2f * t);
            da[t] = a[t], db[t] = b[t];
        }
        const double w = orc_waveform_f32(n, 1e-3f, a.data(), b.data(), ad.data());
        const double e = orc_envelope_f32(n, 1e-3f, a.data(), b.data(), ad.data());
        const double e64 = orc_envelope_f64(n, 1e-3, da.data(), db.data(), dad.data());
        orc_hilbert_f64(n, da.data(), re.data(), im.data());
        if (!(std::isfinite(w) && std::isfinite(e) && std::isfinite(e64) && finite(ad) &&
              finite(dad) && finite(re) && finite(im))) {
            std::printf("FAIL misfits at n = %d\n", n);
            failures++;
        }
    }
    // the smoother at sigma 1..9 on a grid smaller than its support (sigma 0
    // means no filter and is never passed)
    for (int sg = 1; sg < 10; sg++) {
        vector<float> f(13 * 7);
        vector<double> d(13 * 7);
        for (size_t k = 0; k < f.size(); k++) f[k] = (float)(k % 5), d[k] = (double)(k % 5);
        orc_gauss_f32(13, 7, sg, f.data());
        orc_gauss_f64(13, 7, sg, d.data());
        if (!(finite(f) && finite(d))) {
            std::printf("FAIL gauss at sigma %d\n", sg);
            failures++;
        }
    }
    vector<float> stf(64);
    orc_ricker(64, 1e-3f, 20.0f, 0.05f, stf.data());
    CHECK(finite(stf));
}

int main() {
    orc_set_threads(2);
    run_setup();
    run_case(1, 0, 0, 0);
    run_case(0, 1, 1, 3);
    run_case(1, 1, 0, 5);
    if (failures) {
        std::printf("%d FAILED\n", failures);
        return 1;
    }
    std::printf("ALL OK\n");
    return 0;
}
