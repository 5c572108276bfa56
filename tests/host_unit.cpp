// Host-only checks of the plugin layer (inv.cu_amd/host/inv.h): config
// parsing rules, the line-search arithmetic and step rules, and the L-BFGS
// slot rotation.  No device call is made.  Built and run by
// tests/test_host_unit.py; prints one line per check and "ALL OK" at the end.
//
// Expected values: the line-search numbers are worked by hand from the rules of
// optimizer/optimizer.h:235-357 on small histories (the parabola vertices and
// the least-squares fit come from numpy.polyfit on the same points).
#include <cassert>
#include <cstdlib>

#include "../inv.cu_amd/host/inv.h"

using namespace invhost;

static int failures = 0;
#define CHECK(cond)                                                                   \
    do {                                                                              \
        if (cond) std::cout << "ok   " << #cond << "\n";                              \
        else {                                                                        \
            std::cout << "FAIL " << #cond << "  (line " << __LINE__ << ")\n";         \
            failures++;                                                               \
        }                                                                             \
    } while (0)

static bool near(double a, double b, double tol = 1e-6) {
    return std::fabs(a - b) <= tol * std::max(1.0, std::fabs(b));
}

struct TestOpt : Optimizer {
    using Optimizer::backtrack;
    using Optimizer::bracket;
    using Optimizer::polyfit;
    int rule = 0;  // 0 bracket (NLCG), 1 backtrack (L-BFGS)
    float calcStep(size_t sc, float sm, int &st) override {
        return (float)(rule ? backtrack(sc, sm, st) : bracket(sc, sm, st));
    }
    int computeDirection() override { return 0; }
    void history(vector<double> lens, vector<double> vals, vector<double> gtg,
                 vector<double> gtp, size_t step, double thresh) {
        ls_lens = lens;
        ls_vals = vals;
        ls_count = lens.size();
        ls_gtg = gtg;
        ls_gtp = gtp;
        inv_count = gtg.size();
        ls_step = step;
        ls_thresh = thresh;
    }
};

static void write(const string &path, const string &text) {
    std::ofstream(path) << text;
}

static void test_config(const string &dir) {
    createDirectory(dir + "/sub");
    // first definition wins: command line, then config.ini, then inherit
    write(dir + "/base.ini", "nt = 99\ndt = 0.5\nextra = 7 # from base\n");
    write(dir + "/config.ini",
          "[global]\n"
          "mode = 2                 # comment after the value\n"
          "  nt = 5000\n"
          "#sh = 1\n"
          "psv=1\n"
          "name = abc\n"
          "half = 2.5\n"
          "neg = -1.5\n"
          "inherit = " + dir + "/base.ini\n"
          "nt = 1\n");
    write(dir + "/sources.dat",
          "# x | z | type | f0 | t0 | angle | amp\n"
          "1 2 0 0.02 0 0 1e10\n"
          "   3 4 0 0.02 0 0 1e10\n"
          "5 6 0 0.02 0 0\n"
          "7 8 0 0.02 0 0 1e10 \n");
    write(dir + "/stations.dat", "10 20\n30 40\n50\n");
    map<string, string> cfg;
    cfg["config"] = dir;
    cfg["mode"] = "3";
    Config c(cfg);
    CHECK(c.i["mode"] == 3);                    // the command line wins
    CHECK(c.i["nt"] == 5000);                   // first line of the file wins
    CHECK(near(c.f["dt"], 0.5));                // from the inherited file
    CHECK(c.i["extra"] == 7);
    CHECK(c.i.count("sh") == 0);                // '#' keys are skipped
    CHECK(c.i["psv"] == 1);
    CHECK(c.i.count("name") == 0);              // not a number: no f/i entry
    CHECK(c.i["half"] == 3 && near(c.f["half"], 2.5));  // std::round
    CHECK(c.i["neg"] == -2);
    // rows: full rows only, and a trailing blank leaves the stream short of eof
    CHECK(c.src.size() == 2);
    CHECK(c.src.size() == 2 && c.src[1][0] == 3.0f && near(c.src[1][6], 1e10));
    CHECK(c.rec.size() == 2 && c.rec[1][1] == 40.0f);
    struct stat st;
    CHECK(stat((dir + "/output").c_str(), &st) == 0);
}

static void test_polyfit() {
    // exact parabola through three points
    double x[3] = {0, 0.25, 0.4045085}, y[3], p[3];
    for (int i = 0; i < 3; i++) y[i] = 3 * x[i] * x[i] - 2 * x[i] + 1;
    const double rss = TestOpt::polyfit(x, y, p, 3);
    CHECK(near(p[0], 3) && near(p[1], -2) && near(p[2], 1) && rss < 1e-20);
    // least squares through four points (numpy.polyfit)
    double x4[4] = {0, 0.25, 0.4, 0.6}, y4[4] = {10, 8, 7.5, 9.0};
    TestOpt::polyfit(x4, y4, p, 4);
    CHECK(near(p[0], 20.14873506, 1e-8) && near(p[1], -13.91541562, 1e-8) &&
          near(p[2], 10.03940784, 1e-8));
}

static void test_steps() {
    TestOpt o;
    int st = 9;
    double a;
    // first trial of the first search: 1 / g.g
    o.history({0}, {10}, {4}, {-2}, 10, 1.2);
    a = o.bracket(0, 100, st);
    CHECK(near(a, 0.25) && st == 0);
    a = o.bracket(0, 0.1, st);
    CHECK(near(a, 0.0618034) && st == 0);  // clamped to 0.618034 step_max
    // lower misfit, no bracket yet: expand by the golden ratio
    o.history({0, 0.25}, {10, 8}, {4}, {-2}, 10, 1.2);
    a = o.bracket(1, 100, st);
    CHECK(near(a, 1.618034 * 0.25) && st == 0);
    a = o.bracket(1, 0.3, st);
    CHECK(near(a, 0.3) && st == 1);  // later trials clamp to step_max and stop
    // bracketed, vertex within the threshold of a trial: take the best trial
    o.history({0, 0.25, 0.4045085}, {10, 8, 9}, {4}, {-2}, 10, 1.2);
    a = o.bracket(2, 100, st);
    CHECK(near(a, 0.25) && st == 1);
    // bracketed, vertex far from every trial: try the vertex
    o.history({0, 0.25, 0.30}, {10, 8, 9}, {4}, {-2}, 10, 1.2);
    a = o.bracket(2, 100, st);
    CHECK(near(a, 0.16785714285714282, 1e-9) && st == 0);
    // bracketed but the least-squares parabola through the four points from
    // the one before the best opens downward (numpy.polyfit: -0.725 x^2 ...):
    // the best trial is accepted rather than evaluated again
    o.history({0, 1, 2, 3, 4}, {10, 0.1, 0, 3, 0}, {4}, {-2}, 10, 1.2);
    a = o.bracket(4, 100, st);
    CHECK(near(a, 2) && st == 1);
    // bracketed, vertex far from every trial, but ls_step trials made: accept
    o.history({0, 0.25, 0.30}, {10, 8, 9}, {4}, {-2}, 2, 1.2);
    a = o.bracket(2, 100, st);
    CHECK(near(a, 0.25) && st == 1);
    // misfit went up: quadratic backtrack, clamped to [0.1, 0.5] x1
    o.history({0, 0.25}, {10, 12}, {4}, {-2}, 10, 1.2);
    a = o.bracket(1, 100, st);
    CHECK(near(a, 0.025) && st == 0);
    // out of trials without a bracket: fail
    o.history({0, 0.25, 0.5}, {10, 12, 13}, {4}, {-2}, 1, 1.2);
    a = o.bracket(2, 100, st);
    CHECK(a == 0 && st == -1);
    // second search, first trial: previous best length x g.p ratio
    o.history({0, 0.25, 0}, {10, 8, 7.9}, {4, 3}, {-2, -1.5}, 10, 1.2);
    a = o.bracket(0, 100, st);
    CHECK(near(a, 0.25 * 2 / 1.5) && st == 0);
    // L-BFGS rule after the first search: unit step, accept any decrease
    a = o.backtrack(0, 0.5, st);
    CHECK(near(a, 0.5) && st == 0);
    a = o.backtrack(0, 3, st);
    CHECK(near(a, 1) && st == 0);
    o.history({0, 0.25, 0, 0.5}, {10, 8, 7.9, 7.0}, {4, 3}, {-2, -1.5}, 10, 1.2);
    a = o.backtrack(1, 3, st);
    CHECK(near(a, 0.5) && st == 1);
    o.history({0, 0.25, 0, 0.5}, {10, 8, 7.9, 8.5}, {4, 3}, {-2, -1.5}, 10, 1.2);
    a = o.backtrack(1, 3, st);
    CHECK(near(a, 0.125 / 1.7) && st == 0);
    o.history({0, 0.25, 0, 0.5, 0.07}, {10, 8, 7.9, 8.5, 8.0}, {4, 3}, {-2, -1.5}, 1, 1.2);
    a = o.backtrack(2, 3, st);
    CHECK(a == 0 && st == -1);
    // first search of L-BFGS follows the NLCG rule
    o.history({0}, {10}, {4}, {-2}, 10, 1.2);
    a = o.backtrack(0, 100, st);
    CHECK(near(a, 0.25) && st == 0);
}

// Plugins written against the reference's interfaces (source/source.h:3-43,
// optimizer/optimizer.h:359-360): the same virtual signatures override here.
class BoxSource : public Source {
  public:
    void create(float *stf, float f0_, float t0_) override {
        for (size_t it = 0; it < nt; it++) {
            const float t = it * dt - t0_;
            stf[it] = (t >= 0 && t < 1 / f0_) ? 1.0f : 0.0f;
        }
    }
};

class FixedStepOptimizer : public Optimizer {
  public:
    int directions = 0;
    int computeDirection() override {
        directions++;
        return 0;
    }
    float calcStep(size_t step_count, float step_max, int &status) override {
        status = step_count > 0 ? 1 : 0;
        return 0.5f * step_max;
    }
};

static void test_reference_style_plugins() {
    BoxSource src;
    float cfg[4] = {8.0f, 0.0f, 0.5f, 3.0f};  // f0, t0, angle, amp
    src.init(cfg, 40, 0.01f);
    CHECK(src.nt == 40 && src.stf_x && src.stf_y && src.stf_z);
    CHECK(src.stf_y[0] == 3.0f && src.stf_y[12] == 3.0f && src.stf_y[13] == 0.0f);
    CHECK(src.stf_x[5] == 3.0f * 1.0f * cosf(0.5f) && src.stf_z[5] == 3.0f * 1.0f * sinf(0.5f));
    FixedStepOptimizer opt;  // instantiable: every pure virtual is overridden
    int st = -9;
    CHECK(opt.calcStep(0, 2.0f, st) == 1.0f && st == 0);
    CHECK(opt.calcStep(3, 2.0f, st) == 1.0f && st == 1);
    // the library's own Ricker source through the same interface equals the
    // C ABI's wavelet bit for bit
    RickerSource rk;
    float rcfg[4] = {12.0f, 0.08f, 0.7f, 1e9f};
    rk.init(rcfg, 300, 1e-3f);
    vector<float> x(300), y(300), z(300);
    inv_source_ricker(300, 1e-3f, 12.0f, 0.08f, 0.7f, 1e9f, x.data(), y.data(), z.data());
    bool same = true;
    for (int i = 0; i < 300; i++)
        same = same && rk.stf_x[i] == x[i] && rk.stf_y[i] == y[i] && rk.stf_z[i] == z[i];
    CHECK(same);
}

static void test_rotation() {
    // mem = 3, buffers A=0 B=1 C=2: the reference's rotation yields
    // [B,A,B], [A,B,A], [B,A,B] ... (slots 0 and 2 alias, C is never used)
    vector<int> si = {0, 1, 2};
    LBFGSOptimizer::rotateSlots(si, false);
    CHECK((si == vector<int>{1, 0, 1}));
    LBFGSOptimizer::rotateSlots(si, false);
    CHECK((si == vector<int>{0, 1, 0}));
    LBFGSOptimizer::rotateSlots(si, false);
    CHECK((si == vector<int>{1, 0, 1}));
    vector<int> ring = {0, 1, 2};
    LBFGSOptimizer::rotateSlots(ring, true);
    CHECK((ring == vector<int>{2, 0, 1}));
    LBFGSOptimizer::rotateSlots(ring, true);
    CHECK((ring == vector<int>{1, 2, 0}));
    vector<int> one = {0};
    LBFGSOptimizer::rotateSlots(one, false);
    CHECK((one == vector<int>{0}));
}

int main(int argc, char **argv) {
    const string dir = argc > 1 ? argv[1] : "/tmp/inv_host_unit";
    test_config(dir);
    test_polyfit();
    test_steps();
    test_rotation();
    test_reference_style_plugins();
    if (failures) {
        std::cout << failures << " FAILED\n";
        return 1;
    }
    std::cout << "ALL OK\n";
    return 0;
}
