// Host side of the inv.cu gradient path on MI355X: the reference's plugin
// classes (Config, Source, Solver, Filter, Misfit, Optimizer and the
// module:: factories of lib/index.h:32-65) with the same names, ids, config
// keys and file layouts, running every array operation through the C ABI of
// libinv_hip.so (include/inv_hip.h).  Device arrays are owned by the library
// context; the classes hold device pointers to them, as the reference's do.
#pragma once

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <dirent.h>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/inv_hip.h"

namespace invhost {

using std::map;
using std::string;
using std::vector;

// Library calls: a failed call is an error with the library's message (the
// reference ignores CUDA errors; silently continuing on a device error is not
// kept).
inline void check(inv_ctx *ctx, int rc, const char *what) {
    if (rc != INV_OK) {
        throw std::runtime_error(string(what) + " failed (" + std::to_string(rc) +
                                 "): " + (ctx ? inv_last_error(ctx) : ""));
    }
}
#define INVH(ctx, call) ::invhost::check((ctx), (call), #call)

// ---------------------------------------------------------------------------
// filesystem helpers (lib/mkdir.h): mkdir -p and "delete the files of a
// directory, then the directory"
// ---------------------------------------------------------------------------
inline void createDirectory(const string &path) {
    string cur;
    std::stringstream ss(path);
    string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        mkdir(cur.c_str(), 0755);
    }
}

inline void removeDirectory(const string &path) {
    DIR *d = opendir(path.c_str());
    if (!d) return;
    while (dirent *e = readdir(d)) {
        string n = e->d_name;
        if (n == "." || n == "..") continue;
        std::remove((path + "/" + n).c_str());
    }
    closedir(d);
    rmdir(path.c_str());
}

// ---------------------------------------------------------------------------
// Config (lib/config.h:7-109): `key = value` lines, sections and text after
// the first value token ignored, keys starting with '#' skipped, the first
// definition wins (command line, then the file, then its `inherit` file);
// every value that parses fully as a float is kept as f[key] and rounded as
// i[key].  sources.dat rows are 7 floats (x z type f0 t0 angle amp),
// stations.dat rows 2 floats (x z); other lines are skipped.
// ---------------------------------------------------------------------------
class Config {
  public:
    vector<vector<float>> src;
    vector<vector<float>> rec;
    map<string, float> f;
    map<string, int> i;
    string path;

    explicit Config(map<string, string> cfg) {
        path = cfg["config"];
        loadConfig(path + "/config.ini", cfg);
        if (cfg["inherit"].size()) loadConfig(cfg["inherit"], cfg);
        parseConfig(cfg);
        loadRows(path + "/sources.dat", 7, src);
        loadRows(path + "/stations.dat", 2, rec);
        if (i["clean"]) removeDirectory(path + "/output");
        createDirectory(path + "/output");
    }

  private:
    static void loadConfig(const string &file, map<string, string> &cfg) {
        std::ifstream in(file);
        string line;
        while (std::getline(in, line)) {
            size_t pos = line.find_first_not_of(" \t");
            if (pos != string::npos) line = line.substr(pos);
            pos = line.find('=');
            if (pos == string::npos) continue;
            std::istringstream ks(line.substr(0, pos)), vs(line.substr(pos + 1));
            string key, value;
            ks >> key;
            vs >> value;
            if (key.size() && value.size() && key[0] != '#' && !cfg[key].size()) cfg[key] = value;
        }
    }
    void parseConfig(const map<string, string> &cfg) {
        for (auto const &kv : cfg) {
            if (kv.first == "config" || kv.first == "inherit") continue;
            std::istringstream vs(kv.second);
            float v;
            vs >> v;
            if (vs.eof() && !vs.fail()) {
                f[kv.first] = v;
                i[kv.first] = (int)std::round(v);
            }
        }
    }
    static void loadRows(const string &file, size_t n, vector<vector<float>> &rows) {
        std::ifstream in(file);
        string line;
        while (std::getline(in, line)) {
            size_t pos = line.find_first_not_of(" \t");
            if (pos != string::npos) line = line.substr(pos);
            std::istringstream ss(line);
            vector<float> r(n);
            for (size_t k = 0; k < n; k++) ss >> r[k];
            if (ss.eof() && !ss.fail()) rows.push_back(r);
        }
    }
};

// ---------------------------------------------------------------------------
// Source / RickerSource (source/source.h:3-43, source/ricker.h:3-22)
// ---------------------------------------------------------------------------
class Source {
  protected:
    float f0 = 0, t0 = 0, angle = 0, amp = 0;

  public:
    size_t nt = 0;
    float dt = 0;
    float *stf_x = nullptr, *stf_y = nullptr, *stf_z = nullptr;  // host [nt]
    // the normalised wavelet of the source type into stf[nt]
    virtual void create(float *stf, float f0, float t0) = 0;
    virtual void init(float *cfg, size_t nt0, float dt0) {
        f0 = cfg[0];
        t0 = cfg[1];
        angle = cfg[2];
        amp = cfg[3];
        nt = nt0;
        dt = dt0;
        release();
        stf_x = new float[nt]();
        stf_y = new float[nt]();
        stf_z = new float[nt]();
        vector<float> stf(nt, 0.0f);
        create(stf.data(), f0, t0);
        // the components in the FP32 expression order of inv_source_ricker
        for (size_t it = 0; it < nt; it++) {
            stf_x[it] = amp * stf[it] * cosf(angle);
            stf_y[it] = amp * stf[it];
            stf_z[it] = amp * stf[it] * sinf(angle);
        }
    }
    // injection mechanism (inv_sources_type): 0 point force, 1 explosive
    virtual int mechanism() const { return 0; }
    virtual ~Source() { release(); }

  private:
    void release() {
        delete[] stf_x;
        delete[] stf_y;
        delete[] stf_z;
        stf_x = stf_y = stf_z = nullptr;
    }
};

class RickerSource : public Source {
  public:
    // the library's Ricker wavelet (amp 1, angle 0: stf_y is the wavelet)
    void create(float *stf, float f0_, float t0_) override {
        vector<float> sx(nt), sz(nt);
        int rc = inv_source_ricker((int)nt, dt, f0_, t0_, 0.0f, 1.0f, sx.data(), stf, sz.data());
        if (rc != INV_OK) throw std::runtime_error("invalid Ricker parameters");
    }
};

// Source id 1, no counterpart in the reference (whose only source is the
// point force of id 0): the same Ricker wavelet, injected as an explosion
// (amp times the wavelet added to sxx and szz after each stress update).
class ExplosiveRickerSource : public RickerSource {
  public:
    int mechanism() const override { return 1; }
};

namespace module {
Source *source(size_t id);
}

// ---------------------------------------------------------------------------
// Solver / FdmSolver (solver/solver.h:7-268, solver/fdm.h:222-613)
// ---------------------------------------------------------------------------
struct Dim {
    int nx = 0, nz = 0;
    operator int() const { return nx * nz; }
    int hk(int i, int j) const { return i * nz + j; }
};

class Solver {
  protected:
    string path;
    size_t sfe = 0, wfe = 0, wae = 0;
    size_t abs_width = 0;
    float abs_alpha = 0;
    bool abs_left = false, abs_top = false, abs_right = false, abs_bottom = false;
    vector<float> src_x, src_z, rec_x, rec_z;

  public:
    bool sh = false, psv = false;
    bool inv_lambda = false, inv_mu = false, inv_rho = false;
    bool adj = false;
    size_t trace_type = 0, model_type = 0;
    Dim dim;
    size_t nt = 0, nsrc = 0, nrec = 0;
    float dx = 0, dz = 0, dt = 0;
    // device arrays (owned by the library context)
    float *adstf_x = nullptr, *adstf_y = nullptr, *adstf_z = nullptr;
    float *out_x = nullptr, *out_y = nullptr, *out_z = nullptr;
    float *lambda = nullptr, *mu = nullptr, *rho = nullptr;
    float *k_lambda = nullptr, *k_mu = nullptr, *k_rho = nullptr;

    virtual void init(Config *config) = 0;
    virtual void importModel(bool true_model) = 0;
    virtual void initKernels() = 0;
    virtual void exportAxis() = 0;
    virtual void exportKernel(size_t n = 0) = 0;
    virtual void exportModel(size_t n = 0) = 0;
    virtual void runForward(int isrc, bool trace = false, bool snapshot = false) = 0;
    virtual void runAdjoint(int isrc, bool snapshot = false) = 0;
    virtual ~Solver() {}
};

class FdmSolver : public Solver {
  public:
    inv_ctx *ctx = nullptr;
    int batch = 1;
    int ncomp = 0;
    bool tune_pending = false;  // extension key autotune
    float xmax = 0, zmax = 0;

    ~FdmSolver() override {
        if (ctx) inv_ctx_destroy(ctx);
    }

    // Solver::init + FdmSolver::init (solver.h:158-224, fdm.h:350-462)
    void init(Config *config) override {
        auto &f = config->f;
        auto &i = config->i;
        dt = f["dt"];
        nt = i["nt"];
        sfe = i["sfe"];
        wfe = i["wfe"];
        wae = i["wae"];
        sh = i["sh"];
        psv = i["psv"];
        adj = i["mode"] == 0 || i["mode"] == 2;
        trace_type = i["trace_type"];
        model_type = i["model_type"];
        abs_alpha = f["abs_alpha"];
        abs_width = i["abs_width"];
        abs_left = i["abs_left"];
        abs_top = i["abs_top"];
        abs_right = i["abs_right"];
        abs_bottom = i["abs_bottom"];
        inv_lambda = i["inv_lambda"];
        inv_mu = i["inv_mu"];
        inv_rho = i["inv_rho"];
        path = config->path;
        nsrc = config->src.size();
        nrec = config->rec.size();
        if (!nsrc || !nrec) throw std::runtime_error("no sources or no stations in " + path);
        // grid from the true model's point cloud (fdm.h:354-389)
        int npt = 0;
        auto read_max = [&](const string &comp) {
            vector<float> v = readBin(path + "/model_true/proc000000_" + comp + ".bin", npt);
            return *std::max_element(v.begin(), v.end());
        };
        xmax = read_max("x");
        zmax = read_max("z");
        int nx, nz;
        INVH(nullptr, inv_grid_dims(npt, xmax, zmax, &nx, &nz, &dx, &dz));
        dim.nx = nx;
        dim.nz = nz;
        // extension keys: shots per launch, arithmetic mode, graph capture
        batch = std::max(1, std::min<int>(i.count("batch") ? i["batch"] : 16, (int)nsrc));
        INVH(nullptr, inv_ctx_create(i.count("device") ? i["device"] : 0, &ctx));
        inv_grid_desc d{};
        d.nx = nx;
        d.nz = nz;
        d.dx = dx;
        d.dz = dz;
        d.dt = dt;
        d.nt = (int)nt;
        d.sh = sh;
        d.psv = psv;
        d.adj = adj;
        d.trace_type = (int)trace_type;
        d.sfe = (int)sfe;
        d.abs_width = (int)abs_width;
        d.abs_alpha = abs_alpha;
        d.abs_left = abs_left;
        d.abs_right = abs_right;
        d.abs_top = abs_top;
        d.abs_bottom = abs_bottom;
        d.inv_lambda = inv_lambda;
        d.inv_mu = inv_mu;
        d.inv_rho = inv_rho;
        d.batch = batch;
        INVH(ctx, inv_grid_set(ctx, &d));
        if (i.count("math")) INVH(ctx, inv_set_math(ctx, i["math"]));
        if (i.count("graphs")) INVH(ctx, inv_set_graphs(ctx, i["graphs"]));
        // forward-snapshot storage (inv_set_snapshots: 0 FP32, 1 FP16, 2 BF16,
        // 3 I16; checkpoint interval) and tile autotuning at the first model
        // load
        if (i.count("snap_format") || i.count("ckpt"))
            INVH(ctx, inv_set_snapshots(ctx, i.count("snap_format") ? i["snap_format"] : 0,
                                        i.count("ckpt") ? i["ckpt"] : 0));
        tune_pending = i.count("autotune") && i["autotune"];
        // sources: positions (calcIdx, fdm.h:454-455) and wavelets (solver.h:191-200)
        vector<int> sx(nsrc), szi(nsrc);
        vector<float> stfx(nsrc * nt), stfy(nsrc * nt), stfz(nsrc * nt);
        vector<int> mech(nsrc, 0);
        for (size_t is = 0; is < nsrc; is++) {
            const vector<float> &s = config->src[is];
            src_x.push_back(s[0]);
            src_z.push_back(s[1]);
            sx[is] = inv_calc_idx(s[0], xmax, nx);
            szi[is] = inv_calc_idx(s[1], zmax, nz);
            std::unique_ptr<Source> src(module::source((size_t)s[2]));
            if (!src) throw std::runtime_error("unknown source type");
            float cfg[4] = {s[3], s[4], s[5], s[6]};
            src->init(cfg, nt, dt);
            mech[is] = src->mechanism();
            std::copy(src->stf_x, src->stf_x + nt, stfx.begin() + is * nt);
            std::copy(src->stf_y, src->stf_y + nt, stfy.begin() + is * nt);
            std::copy(src->stf_z, src->stf_z + nt, stfz.begin() + is * nt);
        }
        INVH(ctx, inv_sources_set(ctx, (int)nsrc, sx.data(), szi.data(), stfx.data(),
                                  stfy.data(), stfz.data()));
        if (std::count(mech.begin(), mech.end(), 1)) INVH(ctx, inv_sources_type(ctx, mech.data()));
        vector<int> rx(nrec), rz(nrec);
        for (size_t ir = 0; ir < nrec; ir++) {
            rec_x.push_back(config->rec[ir][0]);
            rec_z.push_back(config->rec[ir][1]);
            rx[ir] = inv_calc_idx(config->rec[ir][0], xmax, nx);
            rz[ir] = inv_calc_idx(config->rec[ir][1], zmax, nz);
        }
        INVH(ctx, inv_receivers_set(ctx, (int)nrec, rx.data(), rz.data()));
        ncomp = (sh ? 1 : 0) + (psv ? 2 : 0);
        auto ptr = [&](int field) {
            float *p = nullptr;
            INVH(ctx, inv_field_ptr(ctx, field, &p));
            return p;
        };
        lambda = ptr(INV_LAMBDA);
        mu = ptr(INV_MU);
        rho = ptr(INV_RHO);
        k_lambda = ptr(INV_K_LAMBDA);
        k_mu = ptr(INV_K_MU);
        k_rho = ptr(INV_K_RHO);
        float *out = ptr(INV_OUT), *ad = ptr(INV_ADSTF);
        const size_t tr = nrec * nt;
        int c = 0;
        if (sh) {
            out_y = out + c * tr;
            adstf_y = ad + c * tr;
            c++;
        }
        if (psv) {
            out_x = out + c * tr;
            adstf_x = ad + c * tr;
            out_z = out + (c + 1) * tr;
            adstf_z = ad + (c + 1) * tr;
        }
    }

    // Solver::importModel + FdmSolver::importModel (solver.h:225-252,
    // fdm.h:463-476): model_{true,init}/proc000000_{x,z,vp,vs,rho} or
    // {lambda,mu} when model_type = 1
    void importModel(bool true_model) override {
        const string dir = path + (true_model ? "/model_true" : "/model_init") + "/proc000000_";
        int n0 = 0, n = 0;
        vector<float> x = readBin(dir + "x.bin", n0);
        vector<float> z = readBin(dir + "z.bin", n);
        check_npt(n0, n);
        vector<float> p1 = readBin(dir + (model_type ? "lambda.bin" : "vp.bin"), n);
        check_npt(n0, n);
        vector<float> p2 = readBin(dir + (model_type ? "mu.bin" : "vs.bin"), n);
        check_npt(n0, n);
        vector<float> pr = readBin(dir + "rho.bin", n);
        check_npt(n0, n);
        INVH(ctx, inv_model_from_points(ctx, n0, x.data(), z.data(), p1.data(), p2.data(),
                                        pr.data(), xmax, zmax, (int)model_type));
        if (tune_pending) {
            tune_pending = false;
            INVH(ctx, inv_autotune(ctx, 0, batch, (int)std::min<size_t>(nt, 10), nullptr, nullptr,
                                   nullptr));
        }
    }

    void initKernels() override { INVH(ctx, inv_kernels_zero(ctx)); }
    bool exportsAdjoint() const { return wae != 0; }

    // exportAxis (fdm.h:500-514)
    void exportAxis() override {
        vector<float> x(dim), z(dim);
        for (int i = 0; i < dim.nx; i++)
            for (int j = 0; j < dim.nz; j++) {
                x[dim.hk(i, j)] = i * dx;
                z[dim.hk(i, j)] = j * dz;
            }
        exportData("x", x);
        exportData("z", z);
    }

    // Solver::exportKernel (solver.h:255-259)
    void exportKernel(size_t n = 0) override {
        vector<float> kl(dim), km(dim), kr(dim);
        INVH(ctx, inv_kernels_get(ctx, kl.data(), km.data(), kr.data()));
        if (inv_lambda) exportData("k_lambda", kl, n);
        if (inv_mu) exportData("k_mu", km, n);
        if (inv_rho) exportData("k_rho", kr, n);
    }

    // FdmSolver::exportModel (fdm.h:477-494) / Solver::exportModel (solver.h:260-264)
    void exportModel(size_t n = 0) override {
        vector<float> a(dim), b(dim), r(dim);
        if (model_type == 0) {
            INVH(ctx, inv_model_to_vps(ctx, a.data(), b.data(), r.data()));
            if (inv_lambda && inv_mu) exportData("vp", a, n);
            if (inv_mu) exportData("vs", b, n);
            if (inv_rho) exportData("rho", r, n);
        } else {
            INVH(ctx, inv_model_download_lmr(ctx, a.data(), b.data(), r.data()));
            if (inv_lambda) exportData("lambda", a, n);
            if (inv_mu) exportData("mu", b, n);
            if (inv_rho) exportData("rho", r, n);
        }
    }

    // runForward (fdm.h:515-570): isrc < 0 fires every source at once
    void runForward(int isrc, bool trace = false, bool snapshot = false) override {
        const int flags = adj ? INV_FWD_SNAPSHOTS : 0;
        if (snapshot && wfe) {
            INVH(ctx, inv_forward_begin(ctx, isrc, 1, flags));
            steppedExport(wfe, -1);
        } else {
            INVH(ctx, inv_forward(ctx, isrc, 1, flags));
        }
        if (trace) exportTraces((size_t)std::max(isrc, 0));
    }

    // runAdjoint (fdm.h:571-612), driven by adstf_*
    void runAdjoint(int isrc, bool snapshot = false) override {
        if (snapshot && wae) {
            INVH(ctx, inv_adjoint_begin(ctx, isrc, 1));
            steppedExport(wae, isrc);
        } else {
            INVH(ctx, inv_adjoint(ctx, isrc, 1));
        }
    }

    // Batched extensions: shots first..first+count-1 in one launch per step.
    void runForwardBatch(int first, int count) {
        INVH(ctx, inv_forward(ctx, first, count, adj ? INV_FWD_SNAPSHOTS : 0));
    }
    void runAdjointBatch(int first, int count) { INVH(ctx, inv_adjoint(ctx, first, count)); }

    // exportTraces (solver.h:38-108): SU, 240-byte header + nt float32 per receiver
    void exportTraces(size_t isrc) {
        vector<float> tr((size_t)ncomp * nrec * nt);
        INVH(ctx, inv_traces_get(ctx, 0, tr.data()));
        int32_t h1[28] = {0};
        int16_t h2[2] = {0, 0}, h3[2] = {0, 0};
        float h4[30] = {0};
        const float xs = src_x[isrc], zs = src_z[isrc];
        int16_t dt_int2 = 0;
        if (!(dt * 1e6 > std::pow(2, 15))) dt_int2 = (int16_t)(dt * 1e6);
        h1[18] = (int32_t)std::round(xs);
        h1[19] = (int32_t)std::round(zs);
        h2[1] = (int16_t)nt;
        h3[0] = dt_int2;
        char istr[16];
        std::snprintf(istr, sizeof istr, "%06zu", isrc);
        auto write = [&](const string &comp, const float *data) {
            std::ofstream out(path + "/output/" + (trace_type ? "u" : "v") + comp + "_" + istr +
                                  ".su",
                              std::ofstream::binary);
            for (size_t ir = 0; ir < nrec; ir++) {
                h1[0] = (int32_t)ir + 1;
                h1[9] = (int32_t)std::round(rec_x[ir] - xs);
                h1[20] = (int32_t)std::round(rec_x[ir]);
                h1[21] = (int32_t)std::round(rec_z[ir]);
                if (nrec > 1) h4[1] = rec_x[1] - rec_x[0];
                out.write(reinterpret_cast<const char *>(h1), sizeof h1);
                out.write(reinterpret_cast<const char *>(h2), sizeof h2);
                out.write(reinterpret_cast<const char *>(h3), sizeof h3);
                out.write(reinterpret_cast<const char *>(h4), sizeof h4);
                out.write(reinterpret_cast<const char *>(data + ir * nt), nt * sizeof(float));
            }
        };
        const size_t trn = nrec * nt;
        int c = 0;
        if (sh) write("y", tr.data() + trn * c++);
        if (psv) {
            write("x", tr.data() + trn * c);
            write("z", tr.data() + trn * (c + 1));
        }
    }

    // exportData (solver.h:109-120): int32 npt, then npt float32
    void exportData(const string &comp, const vector<float> &data, size_t n = 0) {
        char num[16];
        std::snprintf(num, sizeof num, "%06zu", n);
        std::ofstream out(path + "/output/proc" + num + "_" + comp + ".bin", std::ofstream::binary);
        const int32_t npt = (int32_t)dim;
        out.write(reinterpret_cast<const char *>(&npt), sizeof npt);
        out.write(reinterpret_cast<const char *>(data.data()), (size_t)npt * sizeof(float));
    }

    static vector<float> readBin(const string &file, int &npt) {
        std::ifstream in(file, std::ifstream::binary);
        if (!in) throw std::runtime_error("cannot open " + file);
        int32_t n = 0;
        in.read(reinterpret_cast<char *>(&n), sizeof n);
        if (n <= 0) throw std::runtime_error("bad point count in " + file);
        vector<float> v((size_t)n);
        in.read(reinterpret_cast<char *>(v.data()), (size_t)n * sizeof(float));
        npt = n;
        return v;
    }

  private:
    static void check_npt(int a, int b) {
        // the reference prints "error: invalid model" and goes on (solver.h:231-233)
        if (a != b) throw std::runtime_error("error: invalid model");
    }

    // exportSnapshot (fdm.h:312-347) every `we` steps of a begun run
    void steppedExport(size_t we, int isrc) {
        string istr;
        if (isrc >= 0) {
            istr = std::to_string(isrc);
            const size_t digits = std::to_string(nsrc).size();
            while (istr.size() < digits) istr = "0" + istr;
        }
        const int nf = (sh ? 4 : 0) + (psv ? 7 : 0);
        vector<float> st((size_t)nf * dim);
        int pos = 0;
        while (pos < (int)nt) {
            const int prev = pos;
            INVH(ctx, inv_run_advance(ctx, (int)we - (pos % (int)we), &pos));
            if (pos == prev) break;
            if (pos % (int)we != 0) continue;
            INVH(ctx, inv_state_get(ctx, 0, st.data()));
            auto field = [&](int k) {
                return vector<float>(st.begin() + (size_t)k * dim, st.begin() + (size_t)(k + 1) * dim);
            };
            const int p0 = sh ? 4 : 0;  // state order SH {vy,uy,sxy,szy}, P-SV {vx,vz,ux,uz,...}
            if (trace_type == 0) {
                if (sh) exportData("vy" + istr, field(0), pos);
                if (psv) {
                    exportData("vx" + istr, field(p0), pos);
                    exportData("vz" + istr, field(p0 + 1), pos);
                }
            } else if (trace_type == 1) {
                if (sh) exportData("uy" + istr, field(1), pos);
                if (psv) {
                    exportData("ux" + istr, field(p0 + 2), pos);
                    exportData("uz" + istr, field(p0 + 3), pos);
                }
            }
        }
    }
};

// ---------------------------------------------------------------------------
// Filter / GaussianFilter (filter/filter.h:3-15, filter/gaussian.h:3-61)
// ---------------------------------------------------------------------------
class Filter {
  protected:
    Dim dim;
    size_t sigma = 0;

  public:
    inv_ctx *ctx = nullptr;
    virtual void init(size_t nx, size_t nz, size_t s) {
        dim.nx = (int)nx;
        dim.nz = (int)nz;
        sigma = s;
    }
    virtual void apply(float *dev) = 0;
    virtual ~Filter() {}
};

class GaussianFilter : public Filter {
  public:
    void apply(float *dev) override {
        if (sigma == 0) return;
        INVH(ctx, inv_filter_gauss1s(ctx, dev, dim.nx, dim.nz, (int)sigma));
    }
};

// ---------------------------------------------------------------------------
// Misfit (misfit/misfit.h:3-153), WaveformMisfit (misfit/waveform.h),
// EnvelopeMisfit (misfit/envelope.h)
// ---------------------------------------------------------------------------
class Misfit {
  protected:
    FdmSolver *solver = nullptr;
    Filter *filter = nullptr;
    int type = 0;  // 0 waveform, 1 envelope
    bool trace_file = false;

    // per-trace misfit on device pointers (misfit.h:11)
    virtual float run(float *syn, float *obs, float *adstf) {
        double chi = 0;
        INVH(solver->ctx, inv_misfit_trace(solver->ctx, type, syn, obs, adstf, &chi));
        return (float)chi;
    }

  public:
    float ref = 1;
    int progress = 0;  // extension key: a line on stderr every `progress` batches

    // Misfit::run (misfit.h:15-51).  Shots run in batches of `batch` per
    // launch; when adjoint wavefield exports are due (snapshot and wae > 0)
    // they run one at a time, as in the reference, so that the exports are
    // per shot.  The per-trace misfits of a shot are one launch
    // (inv_misfit) rather than one call per trace.  Virtual (the
    // reference's is not) so that a misfit of another kind can drive the
    // optimizers; code written against the reference is unaffected.
    virtual double run(bool kernel = false, bool snapshot = false) {
        inv_ctx *ctx = solver->ctx;
        if (kernel) solver->initKernels();
        double misfit = 0;
        if (snapshot && kernel && solver->exportsAdjoint()) {
            for (size_t isrc = 0; isrc < solver->nsrc; isrc++) {
                solver->runForward((int)isrc);
                double f = 0;
                INVH(ctx, inv_misfit(ctx, (int)isrc, 1, type, nullptr, &f));
                misfit += f;
                solver->runAdjoint((int)isrc, snapshot);
            }
        } else {
            for (size_t first = 0; first < solver->nsrc; first += solver->batch) {
                const int count = (int)std::min<size_t>(solver->batch, solver->nsrc - first);
                solver->runForwardBatch((int)first, count);
                double f = 0;
                INVH(ctx, inv_misfit(ctx, (int)first, count, type, nullptr, &f));
                misfit += f;
                if (kernel) solver->runAdjointBatch((int)first, count);
                if (progress && (first / solver->batch) % progress == 0)
                    std::cerr << (kernel ? "gradient" : "misfit") << ": shots " << first + count
                              << " of " << solver->nsrc << std::endl;
            }
        }
        if (kernel && filter) {
            filter->apply(solver->k_lambda);
            filter->apply(solver->k_mu);
            filter->apply(solver->k_rho);
        }
        return misfit;
    }

    // Misfit::init (misfit.h:91-151): observed data from SU files
    // (trace_file = 1, velocity integrated to displacement when trace_type
    // is 0) or modelled on model_true, then model_init is loaded.
    virtual void init(Config *config, Solver *s, Filter *fl) {
        progress = config->i.count("progress") ? config->i["progress"] : 0;
        solver = dynamic_cast<FdmSolver *>(s);
        if (!solver) throw std::runtime_error("Misfit needs the FDM solver");
        filter = fl;
        solver->init(config);
        if (filter) {
            filter->ctx = solver->ctx;
            filter->init(solver->dim.nx, solver->dim.nz, config->i["filter_sigma"]);
        }
        ref = 1;
        inv_ctx *ctx = solver->ctx;
        const size_t nt = solver->nt, nrec = solver->nrec;
        if (config->i["trace_file"]) {
            vector<float> buf((size_t)solver->ncomp * nrec * nt);
            for (size_t is = 0; is < solver->nsrc; is++) {
                int c = 0;
                auto read = [&](const string &comp) {
                    char istr[16];
                    std::snprintf(istr, sizeof istr, "%06zu", is);
                    const string file = config->path + "/traces/" +
                                        (solver->trace_type ? "u" : "v") + comp + "_" + istr + ".su";
                    std::ifstream in(file, std::ifstream::binary);
                    if (!in) throw std::runtime_error("cannot open " + file);
                    float *dst = buf.data() + (size_t)c++ * nrec * nt;
                    for (size_t ir = 0; ir < nrec; ir++) {
                        char hdr[240];
                        in.read(hdr, 240);
                        in.read(reinterpret_cast<char *>(dst + ir * nt), nt * sizeof(float));
                        if (solver->trace_type == 0) {
                            float *b = dst + ir * nt;  // v2u, misfit.h:113-120
                            b[0] *= solver->dt;
                            for (size_t it = 1; it < nt; it++) b[it] = b[it - 1] + b[it] * solver->dt;
                        }
                    }
                };
                if (solver->sh) read("y");
                if (solver->psv) {
                    read("x");
                    read("z");
                }
                INVH(ctx, inv_obs_set(ctx, (int)is, buf.data()));
            }
        } else {
            std::cout << "Generating traces" << std::endl;
            solver->importModel(true);
            for (size_t first = 0; first < solver->nsrc; first += solver->batch) {
                const int count = (int)std::min<size_t>(solver->batch, solver->nsrc - first);
                INVH(ctx, inv_forward(ctx, (int)first, count, 0));
                INVH(ctx, inv_obs_from_traces(ctx, (int)first, count));
                if (progress && (first / solver->batch) % progress == 0)
                    std::cerr << "traces: shots " << first + count << " of " << solver->nsrc
                              << std::endl;
            }
        }
        solver->importModel(false);
    }
    virtual ~Misfit() {}
};

class WaveformMisfit : public Misfit {
  public:
    WaveformMisfit() { type = 0; }
};

class EnvelopeMisfit : public Misfit {
  public:
    EnvelopeMisfit() { type = 1; }
};

// ---------------------------------------------------------------------------
// Optimizer (optimizer/optimizer.h:15-514), NLCG (nlcg.h), L-BFGS (lbfgs.h).
// The search vectors are device arrays; dot products and the line-search
// arithmetic are FP64.
// ---------------------------------------------------------------------------
class Optimizer {
  protected:
    Misfit *misfit = nullptr;
    FdmSolver *solver = nullptr;
    inv_ctx *ctx = nullptr;
    size_t N = 0;
    float *m_new[3] = {}, *g_new[3] = {};
    float *m_old[3] = {}, *g_old[3] = {}, *p_new[3] = {}, *p_old[3] = {};
    vector<double> ls_lens, ls_vals, ls_gtg, ls_gtp;
    bool inv_parameter[3] = {};
    size_t inv_cycle = 0, inv_iteration = 0;
    size_t ls_step = 0;
    double ls_step_max = 0, ls_step_init = 0, ls_thresh = 0;
    size_t eval_count = 0, inv_count = 0, ls_count = 0;
    std::chrono::steady_clock::time_point time_start;

    float *alloc_field() {
        void *p = nullptr;
        INVH(ctx, inv_alloc(ctx, N * sizeof(float), &p));
        return static_cast<float *>(p);
    }
    double p_amax(float **a) {
        double m = 0;
        for (int ip = 0; ip < 3; ip++)
            if (inv_parameter[ip]) {
                float v = 0;
                INVH(ctx, inv_vec_amax(ctx, a[ip], N, &v));
                m = std::max(m, (double)v);
            }
        return m;
    }
    double p_dot(float **a, float **b) {
        double s = 0;
        for (int ip = 0; ip < 3; ip++)
            if (inv_parameter[ip]) {
                double d = 0;
                INVH(ctx, inv_vec_dot_f64(ctx, a[ip], b[ip], N, &d));
                s += d;
            }
        return s;
    }
    void p_calc(float **c, double ka, float **a) {
        for (int ip = 0; ip < 3; ip++)
            if (inv_parameter[ip]) INVH(ctx, inv_vec_scale(ctx, c[ip], (float)ka, a[ip], N));
    }
    void p_calc(float **c, double ka, float **a, double kb, float **b) {
        for (int ip = 0; ip < 3; ip++)
            if (inv_parameter[ip])
                INVH(ctx, inv_vec_axpby(ctx, c[ip], (float)ka, a[ip], (float)kb, b[ip], N));
    }
    void p_copy(float **dst, float **src) {
        for (int ip = 0; ip < 3; ip++)
            if (inv_parameter[ip]) INVH(ctx, inv_vec_copy(ctx, dst[ip], src[ip], N));
    }
    double p_angle(float **p, float **g, double k = 1) {
        const double xx = p_dot(p, p), yy = p_dot(g, g), xy = k * p_dot(p, g);
        return std::acos(xy / std::sqrt(xx * yy));
    }

    // least-squares parabola through n <= 4 points, Householder QR in FP64
    // (polyfit + solveQR, optimizer.h:109-182)
    static double polyfit(const double *x, const double *y, double *p, size_t n) {
        double A[4][3], b[4];
        for (size_t i = 0; i < n; i++) {
            A[i][0] = x[i] * x[i];
            A[i][1] = x[i];
            A[i][2] = 1;
            b[i] = y[i];
        }
        const size_t nc = std::min<size_t>(3, n);
        for (size_t k = 0; k < nc; k++) {
            double norm = 0;
            for (size_t i = k; i < n; i++) norm += A[i][k] * A[i][k];
            norm = std::sqrt(norm);
            if (norm == 0) continue;
            const double alpha = A[k][k] > 0 ? -norm : norm;
            double v[4] = {0, 0, 0, 0};
            for (size_t i = k; i < n; i++) v[i] = A[i][k];
            v[k] -= alpha;
            double vv = 0;
            for (size_t i = k; i < n; i++) vv += v[i] * v[i];
            if (vv == 0) continue;
            for (size_t j = k; j < 3; j++) {
                double s = 0;
                for (size_t i = k; i < n; i++) s += v[i] * A[i][j];
                for (size_t i = k; i < n; i++) A[i][j] -= 2 * s / vv * v[i];
            }
            double s = 0;
            for (size_t i = k; i < n; i++) s += v[i] * b[i];
            for (size_t i = k; i < n; i++) b[i] -= 2 * s / vv * v[i];
        }
        p[0] = p[1] = p[2] = 0;
        for (int k = (int)nc - 1; k >= 0; k--) {
            double s = b[k];
            for (size_t j = k + 1; j < 3; j++) s -= A[k][j] * p[j];
            p[k] = A[k][k] != 0 ? s / A[k][k] : 0;
        }
        double rss = 0;
        for (size_t i = 0; i < n; i++) {
            const double e = p[0] * x[i] * x[i] + p[1] * x[i] + p[2];
            rss += (y[i] - e) * (y[i] - e);
        }
        return rss;
    }

    // getHistory (optimizer.h:184-211): the last step_count+1 trials sorted
    // by step length; returns the number of earlier line searches
    size_t getHistory(vector<double> &x, vector<double> &f, size_t step_count) {
        x.assign(step_count + 1, 0);
        f.assign(step_count + 1, 0);
        for (size_t i = 0; i < step_count + 1; i++) {
            const size_t j = ls_count - 1 - step_count + i;
            x[i] = ls_lens[j];
            f[i] = ls_vals[j];
        }
        for (size_t i = 0; i < step_count + 1; i++)
            for (size_t j = i + 1; j < step_count + 1; j++)
                if (x[j] < x[i]) {
                    std::swap(x[i], x[j]);
                    std::swap(f[i], f[j]);
                }
        size_t update_count = 0;
        for (size_t i = 0; i < ls_count; i++)
            if (std::fabs(ls_lens[i]) < 1e-6) update_count++;
        if (update_count > 0) update_count--;
        return update_count;
    }
    static size_t argmin(const double *f, size_t n) {
        size_t idx = 0;
        for (size_t i = 1; i < n; i++)
            if (f[i] < f[idx]) idx = i;
        return idx;
    }
    static bool checkBracket(const double *x, const double *f, size_t n) {
        (void)x;
        const size_t imin = argmin(f, n);
        const double fmin = f[imin];
        if (fmin < f[0])
            for (size_t i = imin; i < n; i++)
                if (f[i] > fmin) return true;
        return false;
    }
    // goodEnough (optimizer.h:235-256); x0 is set from the parabola when it
    // opens upward (convex is then true)
    bool goodEnough(const double *x, const double *f, size_t n, double &x0, bool &convex) {
        const double thresh = std::log10(ls_thresh);
        convex = false;
        if (!checkBracket(x, f, n)) return false;
        double p[3];
        const size_t idx = argmin(f, n) - 1;
        const size_t fitlen = std::min<size_t>(4, n - idx);
        polyfit(x + idx, f + idx, p, fitlen);
        if (p[0] <= 0) {
            std::cout << "  line search error" << std::endl;
        } else {
            convex = true;
            x0 = -p[1] / (2 * p[0]);
            for (size_t i = 1; i < n; i++)
                if (std::fabs(std::log10(x[i] / x0)) < thresh) return true;
        }
        return false;
    }
    static double backtrack2(double f0, double g0, double x1, double f1, double b1, double b2) {
        double x2 = -g0 * x1 * x1 / (2 * (f1 - f0 - g0 * x1));
        if (x2 > b2 * x1) x2 = b2 * x1;
        else if (x2 < b1 * x1) x2 = b1 * x1;
        return x2;
    }
    // bracket (optimizer.h:267-326), the NLCG step rule
    double bracket(size_t step_count, double step_max, int &status) {
        vector<double> x, f;
        const size_t update_count = getHistory(x, f, step_count);
        double alpha = 0;
        if (step_count == 0) {
            if (update_count == 0) {
                alpha = 1 / ls_gtg[inv_count - 1];
            } else {
                const size_t idx = argmin(ls_vals.data(), ls_count - 1);
                alpha = ls_lens[idx] * ls_gtp[inv_count - 2] / ls_gtp[inv_count - 1];
            }
            status = 0;
        } else if (checkBracket(x.data(), f.data(), step_count + 1)) {
            // A bracket: accept the best trial when a trial lies near the
            // parabola's vertex, else try the vertex.  Two deviations, so the
            // loop always ends: when the fitted parabola does not open upward
            // (the reference then leaves the step unset) the best trial is
            // accepted, and so it is once the search has made ls_step trials.
            const double best = x[argmin(f.data(), step_count + 1)];
            double x0 = best;
            bool convex = false;
            if (goodEnough(x.data(), f.data(), step_count + 1, x0, convex) || !convex ||
                step_count >= ls_step) {
                alpha = best;
                status = 1;
            } else {
                alpha = x0;
                status = 0;
            }
        } else if (step_count <= ls_step) {
            size_t i;
            for (i = 1; i < step_count + 1; i++)
                if (f[i] > f[0]) break;
            if (i == step_count + 1) {
                alpha = 1.618034 * x[step_count];
            } else {
                const double slope = ls_gtp[inv_count - 1] / ls_gtg[inv_count - 1];
                alpha = backtrack2(f[0], slope, x[1], f[1], 0.1, 0.5);
            }
            status = 0;
        } else {
            alpha = 0;
            status = -1;
        }
        if (alpha > step_max) {
            if (step_count == 0) {
                alpha = 0.618034 * step_max;
                status = 0;
            } else {
                alpha = step_max;
                status = 1;
            }
        }
        return alpha;
    }
    // backtrack (optimizer.h:327-357), the L-BFGS step rule
    double backtrack(size_t step_count, double step_max, int &status) {
        vector<double> x, f;
        const size_t update_count = getHistory(x, f, step_count);
        const size_t idx = argmin(f.data(), step_count + 1);
        double alpha = 0;
        if (update_count == 0) {
            alpha = bracket(step_count, step_max, status);
        } else if (step_count == 0) {
            alpha = std::min(1.0, step_max);
            status = 0;
        } else if (f[idx] < f[0]) {
            alpha = x[idx];
            status = 1;
        } else if (step_count <= ls_step) {
            const double slope = ls_gtp[inv_count - 1] / ls_gtg[inv_count - 1];
            alpha = backtrack2(f[0], slope, x[1], f[1], 0.1, 0.5);
            status = 0;
        } else {
            alpha = 0;
            status = -1;
        }
        return alpha;
    }

    void ls_push(double len, double val) {
        if (ls_lens.size() <= ls_count) {
            ls_lens.resize(ls_count + 1);
            ls_vals.resize(ls_count + 1);
        }
        ls_lens[ls_count] = len;
        ls_vals[ls_count] = val;
        ls_count++;
    }

    // the step rule (optimizer.h:359, the reference's float signature; the
    // rules below compute in FP64 and round the step once)
    virtual float calcStep(size_t step_count, float step_max, int &status) = 0;

    // lineSearch (optimizer.h:360-419)
    virtual int lineSearch(float f) {
        int status = 0;
        double alpha = 0;
        const double norm_m = p_amax(m_new), norm_p = p_amax(p_new);
        const double gtg = p_dot(g_new, g_new), gtp = p_dot(g_new, p_new);
        const double step_max = ls_step_max * norm_m / norm_p;
        size_t step_count = 0;
        ls_push(0, f);
        if (ls_gtg.size() < inv_count) {
            ls_gtg.resize(inv_count);
            ls_gtp.resize(inv_count);
        }
        ls_gtg[inv_count - 1] = gtg;
        ls_gtp[inv_count - 1] = gtp;
        double alpha_current = 0;
        if (ls_step_init && ls_count <= 1) alpha = ls_step_init * norm_m / norm_p;
        else alpha = calcStep(step_count, step_max, status);
        while (true) {
            p_calc(m_new, 1, m_new, alpha - alpha_current, p_new);
            alpha_current = alpha;
            ls_push(alpha, misfit->run(false));
            eval_count++;
            step_count++;
            alpha = calcStep(step_count, step_max, status);
            std::cout << "  step " << (step_count < 10 ? "0" : "") << step_count
                      << "  misfit = " << ls_vals[ls_count - 1] / misfit->ref << std::endl;
            if (status > 0) {
                std::cout << "  alpha = " << alpha << std::endl;
                p_calc(m_new, 1, m_new, alpha - alpha_current, p_new);
                return status;
            } else if (status < 0) {
                p_calc(m_new, 1, m_new, -alpha_current, p_new);
                if (p_angle(p_new, g_new, -1) < 1e-3) {
                    std::cout << "  line search failed" << std::endl;
                    return status;
                }
                std::cout << "  restarting line search..." << std::endl;
                restartSearch();
                return lineSearch(f);
            }
        }
    }

    virtual int computeDirection() = 0;
    virtual void restartSearch() {
        p_calc(p_new, -1, g_new);
        ls_count = 0;
        inv_count = 1;
    }

  public:
    // Optimizer::run (optimizer.h:428-467)
    void run() {
        solver->exportAxis();
        solver->exportModel();
        time_start = std::chrono::steady_clock::now();
        for (size_t iter = 0; iter < inv_iteration; iter++) {
            std::cout << "\nStarting iteration " << iter + 1 << " / " << inv_iteration << std::endl;
            const double f = misfit->run(true);
            if (iter == 0) misfit->ref = (float)f;
            eval_count += 2;
            std::cout << "  step 00  misfit = " << f / misfit->ref << std::endl;
            if (computeDirection() < 0) restartSearch();
            p_copy(m_old, m_new);
            p_copy(p_old, p_new);
            p_copy(g_old, g_new);
            const int ls = lineSearch(f);
            if (misfit->progress)
                std::cerr << "iteration " << iter + 1 << ": "
                          << std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                           time_start)
                                 .count()
                          << " s since the start, " << eval_count << " evaluations" << std::endl;
            if (ls < 0) break;
            solver->exportKernel(iter + 1);
            solver->exportModel(iter + 1);
        }
        const double elapsed =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - time_start).count();
        // evaluated before anything is printed, so that progress lines on
        // stderr do not split the line
        const double final_misfit = misfit->run(false) / misfit->ref;
        std::cout << "\nFinal misfit: " << final_misfit << std::endl;
        std::cout << "Elapsed time: ";
        if (elapsed > 60) {
            const size_t t_min = (size_t)(elapsed / 60);
            const size_t t_sec = (size_t)(elapsed - t_min * 60);
            std::cout << t_min << "min " << t_sec << "s" << std::endl;
        } else {
            std::cout << std::round(elapsed) << std::endl;
        }
        // not in the reference: the same time in seconds, and the forward
        // evaluations of the iterations (gradients count two)
        std::cout << "Elapsed seconds: " << elapsed << "  evaluations: " << eval_count
                  << std::endl;
    }

    // Optimizer::init (optimizer.h:468-510)
    virtual void init(Config *config, Solver *s, Misfit *m) {
        misfit = m;
        solver = dynamic_cast<FdmSolver *>(s);
        ctx = solver->ctx;
        N = (size_t)solver->dim;
        inv_parameter[0] = solver->inv_lambda;
        inv_parameter[1] = solver->inv_mu;
        inv_parameter[2] = solver->inv_rho;
        inv_iteration = config->i["inv_iteration"];
        inv_cycle = config->i["inv_cycle"];
        ls_step = config->i["ls_step"];
        ls_step_max = config->f["ls_step_max"];
        ls_step_init = config->f["ls_step_init"];
        ls_thresh = config->f["ls_thresh"];
        m_new[0] = solver->lambda;
        m_new[1] = solver->mu;
        m_new[2] = solver->rho;
        g_new[0] = solver->k_lambda;
        g_new[1] = solver->k_mu;
        g_new[2] = solver->k_rho;
        for (int ip = 0; ip < 3; ip++) {
            m_old[ip] = alloc_field();
            g_old[ip] = alloc_field();
            p_new[ip] = alloc_field();
            p_old[ip] = alloc_field();
        }
        eval_count = inv_count = ls_count = 0;
    }
    virtual ~Optimizer() {}
};

class NLCGOptimizer : public Optimizer {
  protected:
    size_t nlcg_type = 0;
    double nlcg_thresh = 0;

    // Polak-Ribiere and Fletcher-Reeves beta (nlcg.h:8-19); p_new is scratch
    double pollak() {
        p_calc(p_new, 1, g_new, -1, g_old);
        return p_dot(g_new, p_new) / p_dot(g_old, g_old);
    }
    double fletcher() { return p_dot(g_new, g_new) / p_dot(g_old, g_old); }

    float calcStep(size_t step_count, float step_max, int &status) override {
        return (float)bracket(step_count, step_max, status);
    }
    // computeDirection (nlcg.h:24-52)
    int computeDirection() override {
        inv_count++;
        if (inv_count == 1) {
            p_calc(p_new, -1, g_new);
            return 0;
        } else if (inv_cycle && inv_cycle < inv_count) {
            std::cout << "  restarting NLCG... [periodic restart]" << std::endl;
            return -1;
        }
        double beta = 0;
        switch (nlcg_type) {
            case 0: beta = fletcher(); break;
            case 1: beta = pollak(); break;
            default: beta = 0;
        }
        p_calc(p_new, -1, g_new, beta, p_old);
        if (p_angle(g_new, g_old) < nlcg_thresh) {
            std::cout << "  restarting NLCG... [loss of conjugacy]" << std::endl;
            return -1;
        }
        if (p_dot(p_new, g_new) > 0) {
            std::cout << "  restarting NLCG... [not a descent direction]" << std::endl;
            return -1;
        }
        return 1;
    }

  public:
    void init(Config *config, Solver *s, Misfit *m) override {
        Optimizer::init(config, s, m);
        nlcg_type = config->i["nlcg_type"];
        nlcg_thresh = config->f["nlcg_thresh"];
    }
};

class LBFGSOptimizer : public Optimizer {
  protected:
    size_t lbfgs_mem = 0, lbfgs_used = 0;
    double lbfgs_thresh = 0;
    bool lbfgs_fix = false;
    vector<float *> S, Y;  // mem buffers of 3 fields each, device resident
    vector<int> si;        // slot -> buffer index

    float calcStep(size_t step_count, float step_max, int &status) override {
        return (float)backtrack(step_count, step_max, status);
    }
    void restartSearch() override {
        Optimizer::restartSearch();
        lbfgs_used = 0;
    }
    float **buf(vector<float *> &v, int b) { return &v[(size_t)b * 3]; }

    // computeDirection (lbfgs.h:18-91), two-loop recursion on device vectors
    int computeDirection() override {
        inv_count++;
        if (inv_count == 1) {
            p_calc(p_new, -1, g_new);
            return 0;
        } else if (inv_cycle && inv_cycle < inv_count) {
            std::cout << "  restarting LBFGS... [periodic restart]" << std::endl;
            return -1;
        }
        rotateSlots(si, lbfgs_fix);
        p_calc(buf(S, si[0]), 1, m_new, -1, m_old);
        p_calc(buf(Y, si[0]), 1, g_new, -1, g_old);
        if (lbfgs_used < lbfgs_mem) lbfgs_used++;
        vector<double> rh(lbfgs_used), al(lbfgs_used);
        p_copy(p_new, g_new);
        double sty = 1, yty = 1;
        for (size_t i = 0; i < lbfgs_used; i++) {
            float **s = buf(S, si[i]), **y = buf(Y, si[i]);
            rh[i] = 1 / p_dot(y, s);
            al[i] = rh[i] * p_dot(s, p_new);
            p_calc(p_new, 1, p_new, -al[i], y);
            if (i == 0) {
                sty = 1 / rh[i];
                yty = p_dot(y, y);
            }
        }
        p_calc(p_new, sty / yty, p_new);
        for (int i = (int)lbfgs_used - 1; i >= 0; i--) {
            float **s = buf(S, si[i]), **y = buf(Y, si[i]);
            const double be = rh[i] * p_dot(y, p_new);
            p_calc(p_new, 1, p_new, al[i] - be, s);
        }
        const double angle = p_angle(g_new, p_new);
        const double pi = 3.1415927;
        if (angle >= pi / 2) {
            std::cout << "  restarting LBFGS... [not a descent direction]" << std::endl;
            return -1;
        } else if (angle > pi / 2 - lbfgs_thresh) {
            std::cout << "  restarting LBFGS... [practical safeguard]" << std::endl;
            return -1;
        }
        p_calc(p_new, -1, p_new);
        return 1;
    }

  public:
    // Slot rotation before storing a new (s, y) pair.  The reference binds a
    // reference to slot mem-1, shifts the slots up and then stores that
    // reference (by then slot mem-2's buffer) into slot 0, so that slots 0 and
    // mem-1 share one buffer and the memory holds mem-1 distinct pairs
    // (lbfgs.h:29-36).  That behaviour is kept; fix = true gives a plain ring.
    static void rotateSlots(vector<int> &si, bool fix) {
        const int mem = (int)si.size();
        const int last = si[mem - 1];
        for (int i = mem - 1; i > 0; i--) si[i] = si[i - 1];
        si[0] = fix ? last : si[mem - 1];
    }

    void init(Config *config, Solver *s, Misfit *m) override {
        Optimizer::init(config, s, m);
        lbfgs_mem = std::max(1, config->i["lbfgs_mem"]);
        lbfgs_thresh = config->f["lbfgs_thresh"];
        lbfgs_fix = config->i.count("lbfgs_fix") && config->i["lbfgs_fix"];
        lbfgs_used = 0;
        for (size_t b = 0; b < lbfgs_mem; b++)
            for (int ip = 0; ip < 3; ip++) {
                S.push_back(alloc_field());
                Y.push_back(alloc_field());
            }
        si.resize(lbfgs_mem);
        for (size_t b = 0; b < lbfgs_mem; b++) si[b] = (int)b;
    }
};

// ---------------------------------------------------------------------------
// module:: factories (lib/index.h:32-65).  Filter id 1 ("none" in the
// reference's config comment) gives no filter; the reference returns nullptr
// there and then dereferences it.
// ---------------------------------------------------------------------------
namespace module {
inline Solver *solver(size_t id) { return id == 0 ? new FdmSolver() : nullptr; }
inline Source *source(size_t id) {
    switch (id) {
        case 0: return new RickerSource();
        case 1: return new ExplosiveRickerSource();
        default: return nullptr;
    }
}
inline Filter *filter(size_t id) { return id == 0 ? new GaussianFilter() : nullptr; }
inline Misfit *misfit(size_t id) {
    switch (id) {
        case 0: return new WaveformMisfit();
        case 1: return new EnvelopeMisfit();
        default: return nullptr;
    }
}
inline Optimizer *optimizer(size_t id) {
    switch (id) {
        case 0: return new NLCGOptimizer();
        case 1: return new LBFGSOptimizer();
        default: return nullptr;
    }
}
}  // namespace module

}  // namespace invhost
