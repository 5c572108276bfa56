// Request server: continuous batching of shot-solve requests on one device.
//
// Threads: callers submit requests into a FIFO queue; one scheduler thread
// turns the queue into batches (policy in serve_policy.h) and hands each to an
// idle lane; every lane has its own host thread and its own context (own HIP
// stream, own buffers), so the batches of different lanes run concurrently
// on the device.  A lane keeps its context's buffers between batches of one
// class, so a stream of like-sized requests does not reallocate.
//
// A batch puts one shot per slot of one fused-kernel launch.  Slots read their
// own request's resident model through a per-slot pointer table, and a batch
// runs for its longest member's nt: shorter members' wavelets are zero-padded
// at the end, their traces are the first nt_b samples of their rows, and
// their adjoint sources are written into the last nt_b samples of the run
// (zeros before), so their adjoint field stays exactly zero until step
// nt - nt_b and then pairs adjoint step nt - nt_b + k with forward step
// nt_b - 1 - k, as a run of nt_b steps would.  A slot's arithmetic does not
// depend on the batch around it: every result equals the request run alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "internal.h"
#include "serve_policy.h"

namespace inv {
namespace serve {

using Clock = std::chrono::steady_clock;

struct Model {
    int id = 0, nx = 0, nz = 0;
    DevBuf lam, mu, rho;
    std::map<unsigned, DevBuf> bdt;  // dt / rho per time step (keyed by the bits of dt)
    int refs = 0;                    // queued or running requests using it
};

struct ClassInfo {
    inv_grid_desc desc{};  // nt and batch are set per batch
    std::vector<int> rec_ix, rec_iz;
};

struct Request {
    long long id = 0;
    inv_request r{};
    std::string cls;
    const ClassInfo *ci = nullptr;
    Model *model = nullptr;
    int status = INV_REQ_QUEUED, error = INV_OK;
    std::string err;
    double misfit = 0, t_submit = 0, t_start = 0, t_done = 0;
    int lane = -1, batch_shots = 0, batch_requests = 0;
};

struct Lane {
    int idx = 0;
    std::thread th;
    inv_ctx *ctx = nullptr;
    std::string ctx_cls;  // class the context is configured for (worker side)
    bool rec_set = false;
    std::string cls;      // class of the buffers, as the policy sees it
    size_t held = 0;      // bytes reserved (policy view)
    size_t actual = 0;    // bytes the lane held after its last batch
    int state = LANE_IDLE;
    bool has_work = false;
    std::vector<Request *> batch;
    std::condition_variable cv;
    DevBuf slot_nt, taper, ksum;  // per-slot trace lengths, tapers, kernel sums
};

struct Server {
    inv_server_desc d{};
    Policy pol;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<Pending> queue;
    std::map<long long, std::unique_ptr<Request>> reqs;
    std::map<int, std::unique_ptr<Model>> models;
    std::map<std::string, ClassInfo> classes;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::thread sched;
    bool stop = false, stop_workers = false;
    long long next_id = 1;
    int next_model = 1;
    size_t models_bytes = 0, peak = 0;
    long long submitted = 0, completed = 0, failed = 0, batches = 0, shots = 0;
    double cell_updates = 0;
    std::string err;
    hipStream_t ctl = nullptr;  // model uploads

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    size_t held_total() const {
        size_t n = models_bytes;
        for (auto &l : lanes) n += l->held;
        return n;
    }
};

static void release_lane(Lane &L) {
    if (L.ctx) inv_ctx_destroy(L.ctx);
    L.ctx = nullptr;
    L.ctx_cls.clear();
    L.rec_set = false;
    L.slot_nt.release();
    L.taper.release();
    L.ksum.release();
}

static size_t lane_bytes(Lane &L) {
    size_t n = L.slot_nt.bytes + L.taper.bytes + L.ksum.bytes;
    if (L.ctx) n += ctx_bytes_held(*L.ctx);
    return n;
}

// Bytes a lane needs for a batch of `shots` shots of class ci run for nt steps.
static size_t need_bytes(const Server &S, const ClassInfo &ci, int shots, int nt) {
    inv_grid_desc g = ci.desc;
    g.nt = nt;
    g.batch = S.d.plain ? 1 : shots;
    const size_t N = (size_t)g.nx * g.nz;
    size_t n = ctx_bytes_estimate(g, shots, (int)ci.rec_ix.size(),
                                  g.adj ? S.d.snap_format : INV_SNAP_F32, 0);
    // lane buffers and the slot-model table, with allocator rounding
    n += (size_t)shots * sizeof(int) + (size_t)shots * nt * sizeof(float) + 3 * N * sizeof(float) +
         (size_t)shots * 3 * sizeof(void *) + 4 * (2u << 20);
    return n;
}

static void check(inv_ctx *ctx, int code) {
    if (code != INV_OK) throw Error{code, inv_last_error(ctx)};
}

// dt / rho of a model for the fast math mode, built once per dt on first use.
static const float *model_bdt(Server &S, Model &m, float dt, hipStream_t stream) {
    unsigned key;
    std::memcpy(&key, &dt, sizeof key);
    std::lock_guard<std::mutex> lk(S.mu);
    auto f = m.bdt.find(key);
    if (f != m.bdt.end()) return f->second.as<float>();
    const long N = (long)m.nx * m.nz;
    DevBuf b;
    b.ensure(N * sizeof(float));
    try {
        launch_bdt(b.as<float>(), m.rho.as<float>(), dt, N, stream);
        INV_HIP(hipStreamSynchronize(stream));
    } catch (...) {
        b.release();
        throw;
    }
    m.bdt[key] = b;
    S.models_bytes += b.bytes;
    return b.as<float>();
}

static void ensure_ctx(Server &S, Lane &L, const std::string &cls) {
    if (L.ctx && L.ctx_cls == cls) return;
    release_lane(L);
    const int code = inv_ctx_create(S.d.device, &L.ctx);
    if (code != INV_OK) {
        L.ctx = nullptr;
        throw Error{code, "context creation failed"};
    }
    L.ctx->retain = true;
    // the lanes run one- and two-step launches only: a serving load whose
    // lanes timed and ran three-step tiles stopped making progress (cause not
    // found), so that path stays off here
    L.ctx->tune_trip = false;
    L.ctx_cls = cls;
}

static void configure(Server &S, Lane &L, const ClassInfo &ci, int nt, int batch) {
    inv_grid_desc g = ci.desc;
    g.nt = nt;
    g.batch = batch;
    check(L.ctx, inv_grid_set(L.ctx, &g));
    if (S.d.tile_x) check(L.ctx, inv_set_variant(L.ctx, INV_VARIANT_FUSED, S.d.tile_x));
    check(L.ctx, inv_set_math(L.ctx, S.d.math));
    check(L.ctx, inv_set_graphs(L.ctx, 0));
    if (g.adj) check(L.ctx, inv_set_snapshots(L.ctx, S.d.snap_format, 0));
    if (!L.rec_set) {
        check(L.ctx, inv_receivers_set(L.ctx, (int)ci.rec_ix.size(), ci.rec_ix.data(),
                                       ci.rec_iz.data()));
        L.rec_set = true;
    }
}

static void set_slot_models(Server &S, Lane &L, const std::vector<Request *> &batch, int slots) {
    Ctx &c = *L.ctx;
    std::vector<const float *> ptrs;
    ptrs.reserve(3 * (size_t)slots);
    for (Request *q : batch) {
        Model &m = *q->model;
        const float *r0 = S.d.math == INV_MATH_FAST ? model_bdt(S, m, q->r.dt, c.stream)
                                                    : m.rho.as<float>();
        const int n = S.d.plain ? 1 : q->r.nshots;
        for (int s = 0; s < n; s++) {
            ptrs.push_back(r0);
            ptrs.push_back(m.lam.as<float>());
            ptrs.push_back(m.mu.as<float>());
        }
    }
    INV_REQUIRE((int)ptrs.size() == 3 * slots, INV_ERR_STATE, "slot model table size");
    c.slot_model.ensure(ptrs.size() * sizeof(void *));
    INV_HIP(hipMemcpyAsync(c.slot_model.p, ptrs.data(), ptrs.size() * sizeof(void *),
                           hipMemcpyHostToDevice, c.stream));
    INV_HIP(hipStreamSynchronize(c.stream));
    c.slot_models = true;
    c.have_model = true;
}

// One batch of like-class requests, one shot per slot, padded to the longest.
static void run_batched(Server &S, Lane &L, const std::vector<Request *> &batch) {
    const ClassInfo &ci = *batch[0]->ci;
    int shots = 0, nt = 0;
    for (Request *q : batch) {
        shots += q->r.nshots;
        nt = std::max(nt, q->r.nt);
    }
    ensure_ctx(S, L, batch[0]->cls);
    configure(S, L, ci, nt, shots);
    Ctx &c = *L.ctx;
    const bool adj = ci.desc.adj != 0;
    const long N = c.N, rows = (long)c.ncomp * c.nrec;

    std::vector<int> ix(shots), iz(shots), slot_nt(shots);
    std::vector<float> sx((size_t)shots * nt, 0.0f), sy(sx.size(), 0.0f), sz(sx.size(), 0.0f);
    std::vector<float> tw(adj ? (size_t)shots * nt : 0, 0.0f);
    int b = 0;
    for (Request *q : batch) {
        const inv_request &r = q->r;
        std::vector<float> t1(adj ? r.nt : 0);
        if (adj) waveform_taper(r.nt, r.dt, t1.data());
        for (int s = 0; s < r.nshots; s++, b++) {
            ix[b] = r.src_ix[s];
            iz[b] = r.src_iz[s];
            slot_nt[b] = r.nt;
            const size_t o = (size_t)b * nt, i = (size_t)s * r.nt;
            std::memcpy(&sx[o], r.stf_x + i, r.nt * sizeof(float));
            std::memcpy(&sy[o], r.stf_y + i, r.nt * sizeof(float));
            std::memcpy(&sz[o], r.stf_z + i, r.nt * sizeof(float));
            if (adj) std::memcpy(&tw[o], t1.data(), r.nt * sizeof(float));
        }
    }
    check(L.ctx, inv_sources_set(L.ctx, shots, ix.data(), iz.data(), sx.data(), sy.data(),
                                 sz.data()));
    set_slot_models(S, L, batch, shots);
    // tile of the step kernel: timed once per grid class, cached after that
    if (!S.d.tile_x)
        check(L.ctx, inv_autotune(L.ctx, 0, shots, std::min(nt, 10), nullptr, nullptr, nullptr));
    check(L.ctx, inv_forward(L.ctx, 0, shots, adj ? INV_FWD_SNAPSHOTS : 0));

    // traces: slot b's rows hold nt samples, the request's the first nt_b
    b = 0;
    for (Request *q : batch) {
        const inv_request &r = q->r;
        for (int s = 0; s < r.nshots; s++, b++)
            if (r.traces)
                INV_HIP(hipMemcpy2DAsync(r.traces + (size_t)s * rows * r.nt, r.nt * sizeof(float),
                                         c.out.as<float>() + (size_t)b * rows * nt,
                                         nt * sizeof(float), r.nt * sizeof(float), rows,
                                         hipMemcpyDeviceToHost, c.stream));
    }
    if (!adj) {
        INV_HIP(hipStreamSynchronize(c.stream));
        return;
    }
    // observed traces into the slots' rows of the resident obs buffer
    b = 0;
    for (Request *q : batch) {
        const inv_request &r = q->r;
        for (int s = 0; s < r.nshots; s++, b++)
            INV_HIP(hipMemcpy2DAsync(c.obs.as<float>() + (size_t)b * rows * nt, nt * sizeof(float),
                                     r.obs + (size_t)s * rows * r.nt, r.nt * sizeof(float),
                                     r.nt * sizeof(float), rows, hipMemcpyHostToDevice, c.stream));
    }
    L.slot_nt.ensure((size_t)shots * sizeof(int));
    L.taper.ensure(tw.size() * sizeof(float));
    INV_HIP(hipMemcpyAsync(L.slot_nt.p, slot_nt.data(), (size_t)shots * sizeof(int),
                           hipMemcpyHostToDevice, c.stream));
    INV_HIP(hipMemcpyAsync(L.taper.p, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice,
                           c.stream));
    misfit_waveform_slots(c, shots, L.slot_nt.as<int>(), L.taper.as<float>());
    std::vector<double> chi((size_t)shots * c.nrec * c.ncomp);
    INV_HIP(hipMemcpyAsync(chi.data(), c.chi.p, chi.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c.stream));
    if (shots == 1) check(L.ctx, inv_kernels_zero(L.ctx));
    check(L.ctx, inv_adjoint(L.ctx, 0, shots));
    // each request's kernels: the sum of its slots' accumulators in slot order
    L.ksum.ensure(3 * (size_t)N * sizeof(float));
    b = 0;
    for (Request *q : batch) {
        const inv_request &r = q->r;
        const float *k[3] = {c.k_lambda.as<float>(), c.k_mu.as<float>(), c.k_rho.as<float>()};
        if (shots > 1 && (r.k_lambda || r.k_mu || r.k_rho)) {
            kslot_sum(c, L.ksum.as<float>(), b, r.nshots);
            for (int f = 0; f < 3; f++) k[f] = L.ksum.as<float>() + f * N;
        }
        float *dst[3] = {r.k_lambda, r.k_mu, r.k_rho};
        for (int f = 0; f < 3; f++)
            if (dst[f])
                INV_HIP(hipMemcpyAsync(dst[f], k[f], N * sizeof(float), hipMemcpyDeviceToHost,
                                       c.stream));
        b += r.nshots;
    }
    INV_HIP(hipStreamSynchronize(c.stream));
    // misfit: per shot, the trace misfits in [receiver][component] order
    // (misfit/misfit.h:31-40), then over the request's shots
    b = 0;
    const long per = (long)c.nrec * c.ncomp;
    for (Request *q : batch) {
        double total = 0;
        for (int s = 0; s < q->r.nshots; s++, b++) {
            double part = 0;
            for (long t = 0; t < per; t++) part += chi[(size_t)b * per + t];
            total += part;
        }
        q->misfit = total;
    }
}

// The plain baseline: one request, its shots one per launch in shot order,
// kernels accumulated shot after shot as Misfit::run does
// (misfit/misfit.h:25-45), no graph.
static void run_plain(Server &S, Lane &L, Request *q) {
    const ClassInfo &ci = *q->ci;
    const inv_request &r = q->r;
    ensure_ctx(S, L, q->cls);
    configure(S, L, ci, r.nt, 1);
    Ctx &c = *L.ctx;
    const bool adj = ci.desc.adj != 0;
    const long rows = (long)c.ncomp * c.nrec;
    check(L.ctx, inv_sources_set(L.ctx, r.nshots, r.src_ix, r.src_iz, r.stf_x, r.stf_y, r.stf_z));
    set_slot_models(S, L, {q}, 1);
    if (!S.d.tile_x)
        check(L.ctx, inv_autotune(L.ctx, 0, 1, std::min(r.nt, 10), nullptr, nullptr, nullptr));
    if (adj) check(L.ctx, inv_kernels_zero(L.ctx));
    double total = 0;
    for (int s = 0; s < r.nshots; s++) {
        check(L.ctx, inv_forward(L.ctx, s, 1, adj ? INV_FWD_SNAPSHOTS : 0));
        if (r.traces)
            INV_HIP(hipMemcpyAsync(r.traces + (size_t)s * rows * r.nt, c.out.p,
                                   (size_t)rows * r.nt * sizeof(float), hipMemcpyDeviceToHost,
                                   c.stream));
        if (!adj) continue;
        check(L.ctx, inv_obs_set(L.ctx, s, r.obs + (size_t)s * rows * r.nt));
        double chi = 0;
        check(L.ctx, inv_misfit(L.ctx, s, 1, 0, nullptr, &chi));
        total += chi;
        check(L.ctx, inv_adjoint(L.ctx, s, 1));
    }
    if (adj) {
        const float *k[3] = {c.k_lambda.as<float>(), c.k_mu.as<float>(), c.k_rho.as<float>()};
        float *dst[3] = {r.k_lambda, r.k_mu, r.k_rho};
        for (int f = 0; f < 3; f++)
            if (dst[f])
                INV_HIP(hipMemcpyAsync(dst[f], k[f], c.N * sizeof(float), hipMemcpyDeviceToHost,
                                       c.stream));
    }
    INV_HIP(hipStreamSynchronize(c.stream));
    q->misfit = total;
}

static void worker(Server *S, Lane *L) {
    (void)hipSetDevice(S->d.device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        L->cv.wait(lk, [&] { return L->has_work || S->stop_workers; });
        if (!L->has_work) break;
        const std::vector<Request *> batch = L->batch;
        lk.unlock();
        int code = INV_OK;
        std::string msg;
        try {
            if (S->d.plain) run_plain(*S, *L, batch[0]);
            else run_batched(*S, *L, batch);
        } catch (const Error &e) {
            code = e.code;
            msg = e.msg;
        } catch (const std::bad_alloc &) {
            code = INV_ERR_NOMEM;
            msg = "host allocation failed";
        } catch (const std::exception &e) {
            code = INV_ERR_ARG;
            msg = e.what();
        }
        size_t actual = 0;
        if (code != INV_OK) release_lane(*L);  // start the next batch from a fresh context
        else actual = lane_bytes(*L);
        lk.lock();
        const double t = S->now();
        for (Request *q : batch) {
            q->t_done = t;
            q->model->refs--;
            if (code != INV_OK) {
                q->status = INV_REQ_FAILED;
                q->error = code;
                q->err = msg;
                S->failed++;
            } else {
                q->status = INV_REQ_DONE;
                S->completed++;
                S->shots += q->r.nshots;
                S->cell_updates += (double)q->r.nshots * q->r.nx * q->r.nz * q->r.nt *
                                   (q->r.adjoint ? 2 : 1);
            }
        }
        S->batches++;
        L->actual = actual;
        // buffers only grow within a class, so what a lane keeps can exceed
        // the estimate of any single batch: reserve what it holds
        L->held = std::max(L->held, actual);
        if (code != INV_OK) {
            L->held = 0;
            L->cls.clear();
        }
        L->has_work = false;
        L->batch.clear();
        L->state = LANE_IDLE;
        S->cv_done.notify_all();
        S->cv_sched.notify_all();
    }
}

static void scheduler(Server *S) {
    (void)hipSetDevice(S->d.device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        if (S->stop && S->queue.empty()) break;
        std::vector<LaneView> views;
        for (auto &l : S->lanes) views.push_back({l->state, l->cls, l->held});
        Plan p;
        const NeedFn need = [&](const std::string &cls, int shots, int nt) {
            return need_bytes(*S, S->classes.at(cls), shots, nt);
        };
        if (!plan_next(S->queue, views, S->models_bytes, S->pol, need, p)) {
            S->cv_sched.wait(lk);
            continue;
        }
        if (!p.trim.empty()) {
            // release idle lanes' contexts to make room (outside the lock: a
            // free may wait for the device)
            for (int t : p.trim) S->lanes[t]->state = LANE_BUSY;
            lk.unlock();
            for (int t : p.trim) release_lane(*S->lanes[t]);
            lk.lock();
            for (int t : p.trim) {
                Lane &T = *S->lanes[t];
                T.held = 0;
                T.actual = 0;
                T.cls.clear();
                T.state = LANE_IDLE;
            }
        }
        Lane &L = *S->lanes[p.lane];
        for (size_t j = 0; j < p.members[0]; j++) S->queue[j].bypassed++;
        std::vector<Request *> batch;
        for (size_t m : p.members) batch.push_back(S->reqs.at(S->queue[m].id).get());
        for (auto it = p.members.rbegin(); it != p.members.rend(); ++it)
            S->queue.erase(S->queue.begin() + (long)*it);
        L.held = (L.cls == batch[0]->cls) ? std::max(L.held, p.need) : p.need;
        L.cls = batch[0]->cls;
        L.state = LANE_BUSY;
        L.batch = batch;
        L.has_work = true;
        const double t = S->now();
        for (Request *q : batch) {
            q->status = INV_REQ_RUNNING;
            q->t_start = t;
            q->lane = L.idx;
            q->batch_shots = p.shots;
            q->batch_requests = (int)batch.size();
        }
        S->peak = std::max(S->peak, S->held_total());
        L.cv.notify_one();
    }
}

static std::string class_key(const inv_request &r) {
    std::string k;
    auto put = [&](const void *p, size_t n) { k.append(static_cast<const char *>(p), n); };
    const int ints[] = {r.nx, r.nz, r.sh, r.psv, r.adjoint, r.adjoint ? 0 : r.trace_type,
                        r.adjoint ? r.sfe : 0, r.abs_width, r.abs_left, r.abs_right, r.abs_top,
                        r.abs_bottom, r.nrec};
    const float fl[] = {r.dx, r.dz, r.dt, r.abs_alpha};
    put(ints, sizeof ints);
    put(fl, sizeof fl);
    put(r.rec_ix, (size_t)r.nrec * sizeof(int));
    put(r.rec_iz, (size_t)r.nrec * sizeof(int));
    return k;
}

static void validate(const Server &S, const inv_request &r, const Model *m) {
    INV_REQUIRE(m, INV_ERR_ARG, "unknown model id");
    INV_REQUIRE(m->nx == r.nx && m->nz == r.nz, INV_ERR_ARG, "model grid differs from the request's");
    INV_REQUIRE(r.nx >= 5 && r.nz >= 5 && (long)r.nx * r.nz < (1L << 31), INV_ERR_ARG, "bad grid");
    INV_REQUIRE((r.sh != 0) != (r.psv != 0), INV_ERR_ARG, "a request is SH or P-SV");
    INV_REQUIRE(r.nt >= 1 && r.dt > 0 && r.dx > 0 && r.dz > 0, INV_ERR_ARG, "bad nt/dt/dx/dz");
    INV_REQUIRE(r.nshots >= 1 && r.nshots <= S.pol.max_slots, INV_ERR_ARG,
                "nshots must be in [1, max_slots]");
    INV_REQUIRE(r.src_ix && r.src_iz && r.stf_x && r.stf_y && r.stf_z, INV_ERR_ARG, "null sources");
    INV_REQUIRE(r.nrec >= 1 && r.rec_ix && r.rec_iz, INV_ERR_ARG, "bad receivers");
    for (int s = 0; s < r.nshots; s++)
        INV_REQUIRE(r.src_ix[s] >= 0 && r.src_ix[s] < r.nx && r.src_iz[s] >= 0 && r.src_iz[s] < r.nz,
                    INV_ERR_ARG, "source outside the grid");
    for (int q = 0; q < r.nrec; q++)
        INV_REQUIRE(r.rec_ix[q] >= 0 && r.rec_ix[q] < r.nx && r.rec_iz[q] >= 0 && r.rec_iz[q] < r.nz,
                    INV_ERR_ARG, "receiver outside the grid");
    if (r.adjoint) {
        INV_REQUIRE(r.sfe >= 1 && r.nt % r.sfe == 0, INV_ERR_ARG, "adjoint needs nt % sfe == 0");
        INV_REQUIRE(r.obs, INV_ERR_ARG, "adjoint needs observed traces");
    }
}

}  // namespace serve
}  // namespace inv

using namespace inv;
using namespace inv::serve;

struct inv_server : public inv::serve::Server {};

template <typename F>
static int sguard(inv_server *srv, F &&f) {
    if (!srv) return INV_ERR_ARG;
    try {
        INV_HIP(hipSetDevice(srv->d.device));
        f(*srv);
        return INV_OK;
    } catch (const inv::Error &e) {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->err = "host allocation failed";
        return INV_ERR_NOMEM;
    }
}

static void fill_info(const Request &q, inv_request_info *info) {
    if (!info) return;
    info->status = q.status;
    info->error = q.error;
    info->misfit = q.misfit;
    info->t_submit = q.t_submit;
    info->t_start = q.t_start;
    info->t_done = q.t_done;
    info->lane = q.lane;
    info->batch_shots = q.batch_shots;
    info->batch_requests = q.batch_requests;
}

extern "C" {

int inv_server_create(const inv_server_desc *desc, inv_server **out) {
    if (!desc || !out) return INV_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || desc->device < 0 || desc->device >= n) {
        (void)hipGetLastError();
        return INV_ERR_ARG;
    }
    if (desc->max_lanes < 1 || desc->max_lanes > 16 || desc->max_slots < 1 ||
        desc->max_bypass < 0 || desc->snap_format < INV_SNAP_F32 ||
        desc->snap_format > INV_SNAP_I16 ||
        (desc->math != INV_MATH_EXACT && desc->math != INV_MATH_FAST) ||
        (desc->tile_x != 0 && desc->tile_x != 8 && desc->tile_x != 16 && desc->tile_x != 32))
        return INV_ERR_ARG;
    std::unique_ptr<inv_server> S(new (std::nothrow) inv_server());
    if (!S) return INV_ERR_NOMEM;
    S->d = *desc;
    if (hipSetDevice(desc->device) != hipSuccess ||
        hipStreamCreateWithFlags(&S->ctl, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        return INV_ERR_HIP;
    }
    size_t budget = desc->hbm_budget;
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipStreamDestroy(S->ctl);
            return INV_ERR_HIP;
        }
        budget = fr / 10 * 9;
    }
    S->pol.budget = budget;
    S->pol.max_slots = desc->max_slots;
    S->pol.max_bypass = desc->plain ? 0 : desc->max_bypass;
    S->pol.batching = !desc->plain;
    S->t0 = Clock::now();
    const int nl = desc->plain ? 1 : desc->max_lanes;
    for (int l = 0; l < nl; l++) {
        S->lanes.emplace_back(new Lane());
        S->lanes.back()->idx = l;
    }
    Server *raw = S.get();
    for (auto &l : S->lanes) l->th = std::thread(worker, raw, l.get());
    S->sched = std::thread(scheduler, raw);
    *out = S.release();
    return INV_OK;
}

int inv_server_destroy(inv_server *srv) {
    if (!srv) return INV_ERR_ARG;
    (void)hipSetDevice(srv->d.device);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_sched.notify_all();
    srv->sched.join();
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->cv_done.wait(lk, [&] {
            for (auto &l : srv->lanes)
                if (l->has_work) return false;
            return true;
        });
        srv->stop_workers = true;
        for (auto &l : srv->lanes) l->cv.notify_one();
    }
    for (auto &l : srv->lanes) l->th.join();
    for (auto &l : srv->lanes) release_lane(*l);
    for (auto &kv : srv->models) {
        kv.second->lam.release();
        kv.second->mu.release();
        kv.second->rho.release();
        for (auto &b : kv.second->bdt) b.second.release();
    }
    (void)hipStreamDestroy(srv->ctl);
    delete srv;
    return INV_OK;
}

const char *inv_server_last_error(const inv_server *srv) {
    return srv ? srv->err.c_str() : "null server";
}

int inv_server_model_add(inv_server *srv, int nx, int nz, const float *lambda, const float *mu,
                         const float *rho, int *id) {
    return sguard(srv, [&](Server &S) {
        INV_REQUIRE(id && lambda && mu && rho && nx >= 5 && nz >= 5 &&
                        (long)nx * nz < (1L << 31),
                    INV_ERR_ARG, "bad model");
        const size_t b = (size_t)nx * nz * sizeof(float);
        {
            std::lock_guard<std::mutex> lk(S.mu);
            INV_REQUIRE(S.held_total() + 4 * b <= S.pol.budget, INV_ERR_NOMEM,
                        "model does not fit the device-memory budget");
        }
        std::unique_ptr<Model> m(new Model());
        m->nx = nx;
        m->nz = nz;
        m->lam.ensure(b);
        m->mu.ensure(b);
        m->rho.ensure(b);
        INV_HIP(hipMemcpyAsync(m->lam.p, lambda, b, hipMemcpyHostToDevice, S.ctl));
        INV_HIP(hipMemcpyAsync(m->mu.p, mu, b, hipMemcpyHostToDevice, S.ctl));
        INV_HIP(hipMemcpyAsync(m->rho.p, rho, b, hipMemcpyHostToDevice, S.ctl));
        INV_HIP(hipStreamSynchronize(S.ctl));
        std::lock_guard<std::mutex> lk(S.mu);
        m->id = S.next_model++;
        *id = m->id;
        S.models_bytes += m->lam.bytes + m->mu.bytes + m->rho.bytes;
        S.models[m->id] = std::move(m);
        S.cv_sched.notify_all();
    });
}

int inv_server_model_remove(inv_server *srv, int id) {
    return sguard(srv, [&](Server &S) {
        std::unique_ptr<Model> m;
        {
            std::lock_guard<std::mutex> lk(S.mu);
            auto f = S.models.find(id);
            INV_REQUIRE(f != S.models.end(), INV_ERR_ARG, "unknown model id");
            INV_REQUIRE(f->second->refs == 0, INV_ERR_STATE, "model in use by a request");
            m = std::move(f->second);
            S.models.erase(f);
            S.models_bytes -= m->lam.bytes + m->mu.bytes + m->rho.bytes;
            for (auto &b : m->bdt) S.models_bytes -= b.second.bytes;
        }
        m->lam.release();
        m->mu.release();
        m->rho.release();
        for (auto &b : m->bdt) b.second.release();
        std::lock_guard<std::mutex> lk(S.mu);
        S.cv_sched.notify_all();
    });
}

int inv_server_submit(inv_server *srv, const inv_request *req, long long *id) {
    return sguard(srv, [&](Server &S) {
        INV_REQUIRE(req && id, INV_ERR_ARG, "null request");
        std::unique_ptr<Request> q(new Request());
        q->r = *req;
        q->cls = class_key(*req);
        std::lock_guard<std::mutex> lk(S.mu);
        INV_REQUIRE(!S.stop, INV_ERR_STATE, "server is shutting down");
        auto fm = S.models.find(req->model);
        validate(S, *req, fm == S.models.end() ? nullptr : fm->second.get());
        q->model = fm->second.get();
        auto fc = S.classes.find(q->cls);
        if (fc == S.classes.end()) {
            ClassInfo ci;
            inv_grid_desc &g = ci.desc;
            g.nx = req->nx;
            g.nz = req->nz;
            g.dx = req->dx;
            g.dz = req->dz;
            g.dt = req->dt;
            g.nt = req->nt;
            g.sh = req->sh;
            g.psv = req->psv;
            g.adj = req->adjoint ? 1 : 0;
            g.trace_type = req->adjoint ? 0 : req->trace_type;
            g.sfe = req->adjoint ? req->sfe : 0;
            g.abs_width = req->abs_width;
            g.abs_alpha = req->abs_alpha;
            g.abs_left = req->abs_left;
            g.abs_right = req->abs_right;
            g.abs_top = req->abs_top;
            g.abs_bottom = req->abs_bottom;
            g.inv_lambda = g.inv_mu = g.inv_rho = 1;
            g.batch = 1;
            ci.rec_ix.assign(req->rec_ix, req->rec_ix + req->nrec);
            ci.rec_iz.assign(req->rec_iz, req->rec_iz + req->nrec);
            fc = S.classes.emplace(q->cls, std::move(ci)).first;
        }
        q->ci = &fc->second;
        INV_REQUIRE(S.models_bytes + need_bytes(S, fc->second, req->nshots, req->nt) <=
                        S.pol.budget,
                    INV_ERR_NOMEM, "request does not fit the device-memory budget");
        q->id = S.next_id++;
        q->t_submit = S.now();
        q->model->refs++;
        Pending p;
        p.id = q->id;
        p.cls = q->cls;
        p.shots = req->nshots;
        p.nt = req->nt;
        S.queue.push_back(p);
        *id = q->id;
        S.reqs[q->id] = std::move(q);
        S.submitted++;
        S.cv_sched.notify_all();
    });
}

int inv_server_wait(inv_server *srv, long long id, double timeout_s, inv_request_info *info) {
    if (!srv) return INV_ERR_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto f = srv->reqs.find(id);
    if (f == srv->reqs.end()) {
        srv->err = "unknown request id";
        return INV_ERR_ARG;
    }
    Request *q = f->second.get();
    auto done = [&] { return q->status == INV_REQ_DONE || q->status == INV_REQ_FAILED; };
    if (timeout_s < 0) {
        srv->cv_done.wait(lk, done);
    } else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        fill_info(*q, info);
        srv->err = "timed out waiting for the request";
        return INV_ERR_STATE;
    }
    fill_info(*q, info);
    const int code = q->status == INV_REQ_DONE ? INV_OK : q->error;
    if (code != INV_OK) srv->err = q->err;
    srv->reqs.erase(f);
    return code;
}

int inv_server_poll(inv_server *srv, long long id, inv_request_info *info) {
    if (!srv) return INV_ERR_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto f = srv->reqs.find(id);
    if (f == srv->reqs.end()) {
        srv->err = "unknown request id";
        return INV_ERR_ARG;
    }
    fill_info(*f->second, info);
    return INV_OK;
}

int inv_server_stats_get(inv_server *srv, inv_server_stats *st) {
    if (!srv || !st) return INV_ERR_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    st->submitted = srv->submitted;
    st->completed = srv->completed;
    st->failed = srv->failed;
    st->batches = srv->batches;
    st->shots = srv->shots;
    st->cell_updates = srv->cell_updates;
    st->budget = srv->pol.budget;
    st->models_bytes = srv->models_bytes;
    st->bytes_held = srv->held_total();
    st->bytes_peak = std::max(srv->peak, st->bytes_held);
    st->bytes_actual = 0;
    st->lanes_busy = 0;
    for (auto &l : srv->lanes) {
        st->bytes_actual += l->actual;
        if (l->state == LANE_BUSY) st->lanes_busy++;
    }
    st->queued = (int)srv->queue.size();
    st->now = srv->now();
    return INV_OK;
}

}  // extern "C"
