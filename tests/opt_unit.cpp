// GPU checks of the optimizers of inv.cu_amd/host/inv.h (NLCGOptimizer,
// LBFGSOptimizer, the line search) on a quadratic misfit
//   f(m) = 1/2 sum over the three parameter fields of a (m - t)^2,
// gradient a (m - t), with the search vectors on the device (the library's
// vector operations, FP64 dot products).  Built by the library's Makefile,
// run by tests/test_gpu_optimizer.py; prints one line per check and "ALL OK".
//
//  * NLCG (optimizer/nlcg.h:8-52): the second direction is -g + beta p_old
//    with the closed forms of Fletcher-Reeves, beta = g.g / g0.g0, and
//    Polak-Ribiere, beta = g.(g - g0) / g0.g0 (no clamp); periodic restart
//    (inv_cycle), loss of conjugacy (angle(g, g0) < nlcg_thresh) and
//    ascent (p.g > 0) return -1.
//  * L-BFGS (optimizer/lbfgs.h:18-91): each direction equals the textbook
//    two-loop recursion (Nocedal & Wright, Algorithm 7.4, gamma from the
//    newest pair) computed here in FP64 on the host over the stored pairs,
//    newest first: while fewer than mem pairs are stored those are the
//    pairs; after that the reference's slot rotation keeps mem - 1 distinct
//    pairs with the newest one also in the last slot (lbfgs.h:29-36), and
//    lbfgs_fix = 1 gives the plain ring of the mem newest pairs.
//  * end to end: iterations of gradient, direction and line search lower the
//    misfit at every iteration for both optimizers.
#include <cmath>
#include <cstdio>
#include <random>

#include "../inv.cu_amd/host/inv.h"

using namespace invhost;

static int failures = 0;
#define CHECK(cond, ...)                                                            \
    do {                                                                            \
        if (cond) {                                                                 \
            std::printf("ok   %s\n", #cond);                                        \
        } else {                                                                    \
            std::printf("FAIL %s (line %d) ", #cond, __LINE__);                     \
            std::printf(__VA_ARGS__);                                               \
            std::printf("\n");                                                      \
            failures++;                                                             \
        }                                                                           \
    } while (0)

using Vec = std::vector<double>;  // the three fields one after another

// A solver whose only role is to own the parameter and gradient fields of a
// small grid (no propagation); exports are no-ops.
struct FieldSolver : FdmSolver {
    void setup(int nx, int nz) {
        INVH(nullptr, inv_ctx_create(0, &ctx));
        inv_grid_desc d{};
        d.nx = nx;
        d.nz = nz;
        d.dx = d.dz = 10.0f;
        d.dt = 1e-3f;
        d.nt = 10;
        d.psv = 1;
        d.sfe = 10;
        d.abs_width = 2;
        d.abs_alpha = 0.015f;
        d.inv_lambda = d.inv_mu = d.inv_rho = 1;
        d.batch = 1;
        INVH(ctx, inv_grid_set(ctx, &d));
        dim.nx = nx;
        dim.nz = nz;
        inv_lambda = inv_mu = inv_rho = 1;
        INVH(ctx, inv_field_ptr(ctx, INV_LAMBDA, &lambda));
        INVH(ctx, inv_field_ptr(ctx, INV_MU, &mu));
        INVH(ctx, inv_field_ptr(ctx, INV_RHO, &rho));
        INVH(ctx, inv_field_ptr(ctx, INV_K_LAMBDA, &k_lambda));
        INVH(ctx, inv_field_ptr(ctx, INV_K_MU, &k_mu));
        INVH(ctx, inv_field_ptr(ctx, INV_K_RHO, &k_rho));
    }
    void exportAxis() override {}
    void exportKernel(size_t) override {}
    void exportModel(size_t) override {}
    size_t n() const { return (size_t)dim.nx * dim.nz; }
    float **m() { return fm; }
    float **g() { return fg; }
    float *fm[3], *fg[3];
    void bind() {
        fm[0] = lambda; fm[1] = mu; fm[2] = rho;
        fg[0] = k_lambda; fg[1] = k_mu; fg[2] = k_rho;
    }
};

static void put(inv_ctx *ctx, float **dst, const Vec &v, size_t n) {
    std::vector<float> t(n);
    for (int ip = 0; ip < 3; ip++) {
        for (size_t k = 0; k < n; k++) t[k] = (float)v[ip * n + k];
        INVH(ctx, inv_memcpy_h2d(ctx, dst[ip], t.data(), n * sizeof(float)));
    }
}
static Vec get(inv_ctx *ctx, float **src, size_t n) {
    std::vector<float> t(n);
    Vec v(3 * n);
    for (int ip = 0; ip < 3; ip++) {
        INVH(ctx, inv_memcpy_d2h(ctx, t.data(), src[ip], n * sizeof(float)));
        for (size_t k = 0; k < n; k++) v[ip * n + k] = t[k];
    }
    return v;
}
// values as the device holds them (FP32)
static Vec f32(const Vec &v) {
    Vec r(v.size());
    for (size_t k = 0; k < v.size(); k++) r[k] = (float)v[k];
    return r;
}
static double dot(const Vec &a, const Vec &b) {
    double s = 0;
    for (size_t k = 0; k < a.size(); k++) s += a[k] * b[k];
    return s;
}
static Vec axpy(double ka, const Vec &a, double kb, const Vec &b) {
    Vec r(a.size());
    for (size_t k = 0; k < a.size(); k++) r[k] = ka * a[k] + kb * b[k];
    return r;
}
static double relerr(const Vec &a, const Vec &b) {
    const Vec d = axpy(1, a, -1, b);
    return std::sqrt(dot(d, d) / dot(b, b));
}

// the quadratic misfit on device fields: a (curvature) and t (minimiser)
struct QuadMisfit : Misfit {
    FieldSolver *fs = nullptr;
    float *a[3], *t[3], *tmp = nullptr, *w = nullptr;
    Vec ha, ht;
    void setup(FieldSolver *s, std::mt19937 &rng) {
        fs = s;
        const size_t n = s->n();
        std::uniform_real_distribution<double> ua(1.0, 20.0), ut(-1.0, 1.0);
        ha.resize(3 * n);
        ht.resize(3 * n);
        for (auto &x : ha) x = (float)ua(rng);
        for (auto &x : ht) x = (float)ut(rng);
        for (int ip = 0; ip < 3; ip++) {
            void *p;
            INVH(s->ctx, inv_alloc(s->ctx, n * sizeof(float), &p));
            a[ip] = (float *)p;
            INVH(s->ctx, inv_alloc(s->ctx, n * sizeof(float), &p));
            t[ip] = (float *)p;
        }
        void *p;
        INVH(s->ctx, inv_alloc(s->ctx, n * sizeof(float), &p));
        tmp = (float *)p;
        INVH(s->ctx, inv_alloc(s->ctx, n * sizeof(float), &p));
        w = (float *)p;
        put(s->ctx, a, ha, n);
        put(s->ctx, t, ht, n);
        ref = 1;
    }
    double run(bool kernel = false, bool snapshot = false) override {
        (void)snapshot;
        inv_ctx *ctx = fs->ctx;
        const size_t n = fs->n();
        double f = 0;
        for (int ip = 0; ip < 3; ip++) {
            INVH(ctx, inv_vec_axpby(ctx, tmp, 1.0f, fs->m()[ip], -1.0f, t[ip], n));
            float *out = kernel ? fs->g()[ip] : w;
            INVH(ctx, inv_vec_mul(ctx, out, a[ip], tmp, n));
            double d = 0;
            INVH(ctx, inv_vec_dot_f64(ctx, tmp, out, n, &d));
            f += 0.5 * d;
        }
        return f;
    }
    Vec grad(const Vec &m) const {  // in FP32 as the device computes it
        Vec g(m.size());
        for (size_t k = 0; k < m.size(); k++)
            g[k] = (float)((float)ha[k] * (float)((float)m[k] - (float)ht[k]));
        return g;
    }
};

static map<string, string> base_cfg(const string &dir) {
    return {{"config", dir},     {"inv_iteration", "8"}, {"inv_cycle", "0"}, {"ls_step", "10"},
            {"ls_step_max", "0.5"}, {"ls_step_init", "0.05"}, {"ls_thresh", "1.2"},
            {"nlcg_type", "1"},  {"nlcg_thresh", "0"},   {"lbfgs_mem", "3"},  {"lbfgs_thresh", "0"}};
}

struct NLCG : NLCGOptimizer {
    using Optimizer::computeDirection;
    using Optimizer::g_old;
    using Optimizer::inv_count;
    using Optimizer::lineSearch;
    using Optimizer::m_old;
    using Optimizer::p_new;
    using Optimizer::p_old;
    using Optimizer::p_copy;
    using Optimizer::restartSearch;
};
struct LBFGS : LBFGSOptimizer {
    using Optimizer::computeDirection;
    using Optimizer::g_old;
    using Optimizer::lineSearch;
    using Optimizer::m_old;
    using Optimizer::p_new;
    using Optimizer::p_old;
    using Optimizer::p_copy;
    using Optimizer::restartSearch;
};

static void test_nlcg(const string &dir, FieldSolver &fs, QuadMisfit &mf, std::mt19937 &rng) {
    const size_t n = fs.n();
    std::normal_distribution<double> nd;
    Vec g1(3 * n), g2(3 * n);
    for (auto &x : g1) x = (float)nd(rng);
    for (auto &x : g2) x = (float)nd(rng);
    for (int type : {0, 1}) {
        auto cfg = base_cfg(dir);
        cfg["nlcg_type"] = std::to_string(type);
        Config config(cfg);
        NLCG o;
        o.init(&config, &fs, &mf);
        put(fs.ctx, fs.g(), g1, n);
        CHECK(o.computeDirection() == 0, "first direction");
        Vec p1 = get(fs.ctx, o.p_new, n);
        CHECK(relerr(p1, axpy(-1, g1, 0, g1)) == 0.0, "p1 != -g1");
        o.p_copy(o.p_old, o.p_new);
        o.p_copy(o.g_old, fs.g());
        put(fs.ctx, fs.g(), g2, n);
        const int rc = o.computeDirection();
        const double beta = type == 0 ? dot(g2, g2) / dot(g1, g1)
                                      : dot(g2, axpy(1, g2, -1, g1)) / dot(g1, g1);
        const Vec want = axpy(-1, g2, beta, p1);
        const double e = relerr(get(fs.ctx, o.p_new, n), want);
        CHECK(rc == 1 && e < 1e-6, "type %d: rc %d, rel err %.3e", type, rc, e);
    }
    {   // ascent: g2 = -g1/2 with Polak-Ribiere gives beta 0.75 and p = -g1/4
        auto cfg = base_cfg(dir);
        Config config(cfg);
        NLCG o;
        o.init(&config, &fs, &mf);
        put(fs.ctx, fs.g(), g1, n);
        o.computeDirection();
        o.p_copy(o.p_old, o.p_new);
        o.p_copy(o.g_old, fs.g());
        put(fs.ctx, fs.g(), axpy(-0.5, g1, 0, g1), n);
        CHECK(o.computeDirection() == -1, "ascent direction not restarted");
    }
    {   // loss of conjugacy: g2 almost parallel to g1, threshold 0.5 rad
        auto cfg = base_cfg(dir);
        cfg["nlcg_thresh"] = "0.5";
        Config config(cfg);
        NLCG o;
        o.init(&config, &fs, &mf);
        put(fs.ctx, fs.g(), g1, n);
        o.computeDirection();
        o.p_copy(o.p_old, o.p_new);
        o.p_copy(o.g_old, fs.g());
        put(fs.ctx, fs.g(), axpy(0.9, g1, 0.01, g2), n);
        CHECK(o.computeDirection() == -1, "conjugacy loss not restarted");
    }
    {   // periodic restart after inv_cycle directions
        auto cfg = base_cfg(dir);
        cfg["inv_cycle"] = "1";
        Config config(cfg);
        NLCG o;
        o.init(&config, &fs, &mf);
        put(fs.ctx, fs.g(), g1, n);
        o.computeDirection();
        o.p_copy(o.p_old, o.p_new);
        o.p_copy(o.g_old, fs.g());
        put(fs.ctx, fs.g(), g2, n);
        CHECK(o.computeDirection() == -1, "periodic restart missing");
    }
}

// textbook two-loop recursion over pairs (newest first)
static Vec two_loop(const Vec &g, const std::vector<std::pair<Vec, Vec>> &pairs) {
    Vec q = g;
    std::vector<double> al(pairs.size()), rh(pairs.size());
    for (size_t i = 0; i < pairs.size(); i++) {
        const Vec &s = pairs[i].first, &y = pairs[i].second;
        rh[i] = 1 / dot(y, s);
        al[i] = rh[i] * dot(s, q);
        q = axpy(1, q, -al[i], y);
    }
    const double gamma = dot(pairs[0].first, pairs[0].second) /
                         dot(pairs[0].second, pairs[0].second);
    Vec r = axpy(gamma, q, 0, q);
    for (size_t i = pairs.size(); i-- > 0;) {
        const double be = rh[i] * dot(pairs[i].second, r);
        r = axpy(1, r, al[i] - be, pairs[i].first);
    }
    return axpy(-1, r, 0, r);
}

static void test_lbfgs(const string &dir, FieldSolver &fs, QuadMisfit &mf, std::mt19937 &rng) {
    const size_t n = fs.n();
    const int mem = 3;
    std::normal_distribution<double> nd;
    for (int fix : {0, 1}) {
        auto cfg = base_cfg(dir);
        cfg["lbfgs_fix"] = std::to_string(fix);
        Config config(cfg);
        LBFGS o;
        o.init(&config, &fs, &mf);
        // iterates of a descent walk on the quadratic, gradients exact (FP32)
        Vec m(3 * n);
        for (auto &x : m) x = (float)nd(rng);
        std::vector<std::pair<Vec, Vec>> hist;  // all pairs, oldest first
        Vec mprev, gprev;
        double worst = 0;
        int bad_rc = 0;
        for (int k = 0; k < 7; k++) {
            const Vec g = mf.grad(m);
            put(fs.ctx, fs.m(), m, n);
            put(fs.ctx, fs.g(), g, n);
            if (k > 0) {
                put(fs.ctx, o.m_old, mprev, n);
                put(fs.ctx, o.g_old, gprev, n);
                // the pair as the device forms it (FP32 differences)
                hist.push_back({f32(axpy(1, f32(m), -1, f32(mprev))),
                                f32(axpy(1, f32(g), -1, f32(gprev)))});
            }
            const int rc = o.computeDirection();
            const Vec p = get(fs.ctx, o.p_new, n);
            if (k == 0) {
                worst = std::max(worst, relerr(p, axpy(-1, g, 0, g)));
            } else {
                std::vector<std::pair<Vec, Vec>> use;
                const int h = (int)hist.size();
                if (fix || h < mem) {
                    for (int i = h - 1; i >= std::max(0, h - mem); i--) use.push_back(hist[i]);
                } else {
                    // mem - 1 newest pairs, then the newest again (aliased slot)
                    for (int i = h - 1; i >= h - (mem - 1); i--) use.push_back(hist[i]);
                    use.push_back(hist[h - 1]);
                }
                worst = std::max(worst, relerr(p, two_loop(g, use)));
                bad_rc += rc != 1;
            }
            // next iterate: a step along the direction with some noise
            mprev = m;
            gprev = g;
            for (size_t q = 0; q < m.size(); q++)
                m[q] = (float)(m[q] + 0.04 * p[q] + 1e-3 * nd(rng));
        }
        CHECK(worst < 2e-5 && bad_rc == 0, "fix %d: worst rel err %.3e, bad rc %d", fix, worst,
              bad_rc);
    }
}

template <class O>
static void test_descent(const string &dir, FieldSolver &fs, QuadMisfit &mf, const char *name) {
    const size_t n = fs.n();
    auto cfg = base_cfg(dir);
    Config config(cfg);
    O o;
    o.init(&config, &fs, &mf);
    // start away from the minimiser (and from 0: the first trial step is
    // ls_step_init * max|m| / max|p|)
    std::mt19937 rng(7);
    std::normal_distribution<double> nd;
    Vec m0 = mf.ht;
    for (auto &x : m0) x = (float)(x + 0.5 * nd(rng));
    put(fs.ctx, fs.m(), m0, n);
    double prev = 0, first = 0;
    bool down = true;
    for (int it = 0; it < 8; it++) {
        const double f = mf.run(true);
        if (it == 0) first = f;
        else down = down && f < prev;
        prev = f;
        if (o.computeDirection() < 0) o.restartSearch();
        o.p_copy(o.m_old, fs.m());
        o.p_copy(o.p_old, o.p_new);
        o.p_copy(o.g_old, fs.g());
        if (o.lineSearch((float)f) < 0) break;
    }
    const double last = mf.run(false);
    CHECK(down && last < prev && last < 0.05 * first, "%s: misfit %.4e -> %.4e", name, first,
          last);
}

int main(int argc, char **argv) {
    const string dir = argc > 1 ? argv[1] : "/tmp/inv_opt_unit";
    createDirectory(dir);
    std::mt19937 rng(12345);
    FieldSolver fs;
    fs.setup(48, 40);
    fs.bind();
    QuadMisfit mf;
    mf.setup(&fs, rng);
    test_nlcg(dir, fs, mf, rng);
    test_lbfgs(dir, fs, mf, rng);
    test_descent<NLCG>(dir, fs, mf, "NLCG");
    test_descent<LBFGS>(dir, fs, mf, "L-BFGS");
    if (failures) {
        std::printf("%d FAILED\n", failures);
        return 1;
    }
    std::printf("ALL OK\n");
    return 0;
}
