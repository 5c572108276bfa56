// extern "C" boundary of libinv_hip.so (declared in include/inv_hip.h) and the
// per-context resource management behind it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <mutex>
#include <new>

#include "internal.h"
#include "step.h"

namespace inv {

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

void DevBuf::ensure(size_t nbytes) {
    if (nbytes <= bytes && p) return;
    // queued kernels may still use the old buffer
    if (p) (void)hipDeviceSynchronize();
    release();
    if (nbytes == 0) nbytes = 16;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        throw Error{INV_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(nbytes) +
                                       " bytes failed: " + hipGetErrorString(e)};
    }
    bytes = nbytes;
}

// Absorbing profile of initAbsbound (solver/fdm.h:48-74) as 1-D factor
// tables: ab(i,j) = ((L(i) * R(i)) * B(j)) * T(j), the reference's product
// order, with factor exp(-(alpha d)^2) in FP32 (absent sides give 1).
static void build_absbound(Ctx &c) {
    const inv_grid_desc &g = c.g;
    const size_t w = (size_t)g.abs_width;
    auto f = [&](long d) {
        const float x = g.abs_alpha * (float)d;
        return expf(-(x * x));
    };
    std::vector<float> ax(g.nx), bz(g.nz), tz(g.nz);
    for (int i = 0; i < g.nx; i++) {
        float v = 1;
        if (g.abs_left && (size_t)i + 1 < w) v *= f((long)w - i - 1);
        if (g.abs_right && (size_t)i > (size_t)g.nx - w) v *= f((long)w + i - g.nx);
        ax[i] = v;
    }
    for (int j = 0; j < g.nz; j++) {
        bz[j] = (g.abs_bottom && (size_t)j > (size_t)g.nz - w) ? f((long)w + j - g.nz) : 1.0f;
        tz[j] = (g.abs_top && (size_t)j + 1 < w) ? f((long)w - j - 1) : 1.0f;
    }
    c.abx.ensure(g.nx * sizeof(float));
    c.abzb.ensure(g.nz * sizeof(float));
    c.abzt.ensure(g.nz * sizeof(float));
    INV_HIP(hipMemcpy(c.abx.p, ax.data(), g.nx * sizeof(float), hipMemcpyHostToDevice));
    INV_HIP(hipMemcpy(c.abzb.p, bz.data(), g.nz * sizeof(float), hipMemcpyHostToDevice));
    INV_HIP(hipMemcpy(c.abzt.p, tz.data(), g.nz * sizeof(float), hipMemcpyHostToDevice));
}

// taper of misfit/waveform.h:9-27 in FP32 (with its double literals 0.5)
void waveform_taper(int nt, float dt, float *tw) {
    const float pi = 3.1415927f;
    float t_end = (nt - 1) * dt;
    float taper_width = t_end / 10;
    float t_min = taper_width;
    float t_max = t_end - taper_width;
    for (int it = 0; it < nt; it++) {
        float t = it * dt;
        if (t <= t_min) tw[it] = (float)(0.5 + 0.5 * cosf(pi * (t_min - t) / (taper_width)));
        else if (t >= t_max) tw[it] = (float)(0.5 + 0.5 * cosf(pi * (t_max - t) / (taper_width)));
        else tw[it] = 1;
    }
}

static void build_taper(Ctx &c) {
    const int nt = c.g.nt;
    std::vector<float> tw(nt);
    waveform_taper(nt, c.g.dt, tw.data());
    c.taper.ensure(nt * sizeof(float));
    INV_HIP(hipMemcpy(c.taper.p, tw.data(), nt * sizeof(float), hipMemcpyHostToDevice));
}

static void alloc_slots(Ctx &c) {
    const size_t B = (size_t)c.g.batch, N = (size_t)c.N, f = sizeof(float);
    auto st = [&](DevBuf &b, bool on) {
        if (on) b.ensure(B * N * f);
        else b.release();
    };
    st(c.vy, c.g.sh); st(c.uy, c.g.sh); st(c.sxy, c.g.sh); st(c.szy, c.g.sh);
    st(c.vx, c.g.psv); st(c.vz, c.g.psv); st(c.ux, c.g.psv); st(c.uz, c.g.psv);
    st(c.sxx, c.g.psv); st(c.szz, c.g.psv); st(c.sxz, c.g.psv);
    c.lam.ensure(N * f);
    c.mu.ensure(N * f);
    c.rho.ensure(N * f);
    // dt / rho of the fast math mode: allocated here, with the model, so that
    // launch arguments built before a run never see a null pointer
    c.bdt.ensure(N * f);
    c.k_lambda.ensure(N * f);
    c.k_mu.ensure(N * f);
    c.k_rho.ensure(N * f);
    INV_HIP(hipMemset(c.k_lambda.p, 0, N * f));
    INV_HIP(hipMemset(c.k_mu.p, 0, N * f));
    INV_HIP(hipMemset(c.k_rho.p, 0, N * f));
    if (c.g.adj && c.g.batch > 1) c.kslot.ensure(B * 3 * N * f);
    else if (!c.retain) c.kslot.release();
    // snapshot storage is sized before the first run that needs it, for the
    // format and checkpoint interval in force then (inv_set_snapshots)
    if (!c.retain) {
        c.snap.release();
        c.sexp.release();
        c.ckbuf.release();
        c.rstate.release();
    }
    c.snap_format = INV_SNAP_F32;
    c.ckpt = 0;
    c.nslots = c.nsfe;
}

static void alloc_traces(Ctx &c) {
    if (!c.have_grid) return;
    const size_t tr = (size_t)c.ncomp * std::max(c.nrec, 1) * c.g.nt;
    c.out.ensure((size_t)c.g.batch * tr * sizeof(float));
    c.adstf.ensure((size_t)c.g.batch * tr * sizeof(float));
    INV_HIP(hipMemset(c.out.p, 0, (size_t)c.g.batch * tr * sizeof(float)));
    INV_HIP(hipMemset(c.adstf.p, 0, (size_t)c.g.batch * tr * sizeof(float)));
    c.chi.ensure((size_t)c.g.batch * std::max(c.nrec, 1) * c.ncomp * sizeof(double));
    if (c.nsrc > 0) {
        c.obs.ensure((size_t)c.nsrc * tr * sizeof(float));
        INV_HIP(hipMemset(c.obs.p, 0, (size_t)c.nsrc * tr * sizeof(float)));
    }
}

static void build_receivers(Ctx &c) {
    std::vector<std::vector<int>> cells(1), rows(1);
    for (int r = 0; r < c.nrec; r++) {
        cells[0].push_back(c.rec_ix[r] * c.g.nz + c.rec_iz[r]);
        rows[0].push_back(r);
    }
    build_point_list(c, c.rec_list, 1, cells, rows);
}

static void require_ready(Ctx &c) {
    INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set (inv_grid_set)");
    INV_REQUIRE(c.have_model, INV_ERR_STATE, "model not set");
    INV_REQUIRE(c.have_src && c.have_rec, INV_ERR_STATE, "sources/receivers not set");
}

static void check_slots(Ctx &c, int first, int count) {
    INV_REQUIRE(count >= 1 && count <= c.g.batch, INV_ERR_ARG, "count must be in [1, batch]");
    if (first == -1) {
        INV_REQUIRE(count == 1, INV_ERR_ARG, "first = -1 (all sources) needs count = 1");
    } else {
        INV_REQUIRE(first >= 0 && first + count <= c.nsrc, INV_ERR_ARG, "shot range out of bounds");
    }
}

// Waits for the context's stream; a persistent-kernel launch that gave up
// waiting for a neighbouring tile (its tiles were not all resident) left its
// error word set: reported here, and cleared.
static void sync(Ctx &c) {
    INV_HIP(hipStreamSynchronize(c.stream));
    if (c.res_ctl.p) {
        unsigned w[2] = {0, 0};
        INV_HIP(hipMemcpy(w, c.res_ctl.p, sizeof w, hipMemcpyDeviceToHost));
        if (w[1]) {
            w[1] = 0;
            INV_HIP(hipMemcpy(c.res_ctl.p, w, sizeof w, hipMemcpyHostToDevice));
            throw Error{INV_ERR_HIP,
                        "persistent kernel: a tile gave up waiting for a neighbouring tile's "
                        "halo rows; the run's wavefields are invalid"};
        }
    }
}

static void drop_graphs(Ctx &c) {
    (void)hipStreamSynchronize(c.stream);
    for (auto &kv : c.graphs) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    c.graphs.clear();
}

// Everything a captured loop bakes in changes: forget the graphs.
static void invalidate(Ctx &c) {
    drop_graphs(c);
    c.gen++;
}

std::vector<DevBuf *> ctx_buffers(Ctx &c) {
    return {&c.lam, &c.mu, &c.rho, &c.abx, &c.abzb, &c.abzt, &c.k_lambda, &c.k_mu, &c.k_rho,
            &c.kslot, &c.vx, &c.vz, &c.ux, &c.uz, &c.sxx, &c.szz, &c.sxz, &c.vy, &c.uy,
            &c.sxy, &c.szy, &c.snap, &c.sexp, &c.ckbuf, &c.rstate, &c.out, &c.adstf, &c.obs,
            &c.chi, &c.taper, &c.red, &c.scal, &c.stf, &c.rec_list.ptr, &c.rec_list.cell,
            &c.rec_list.row, &c.gauss_w, &c.gauss_tmp, &c.gauss_sx, &c.gauss_sz, &c.vx_b,
            &c.vz_b, &c.sxx_b, &c.szz_b, &c.sxz_b, &c.vy_b, &c.sxy_b, &c.szy_b, &c.f_rec_v.ptr,
            &c.f_rec_v.cell, &c.f_rec_v.row, &c.f_rec_t.ptr, &c.f_rec_t.cell, &c.f_rec_t.row,
            &c.env_w, &c.env_fac, &c.env_scratch, &c.env_blue, &c.env_gbuf, &c.bdt,
            &c.slot_model, &c.xbuf, &c.res_ctl};
}

size_t ctx_bytes_held(Ctx &c) {
    size_t n = 0;
    for (DevBuf *b : ctx_buffers(c)) n += b->bytes;
    for (auto &kv : c.src_lists) n += kv.second.ptr.bytes + kv.second.cell.bytes + kv.second.row.bytes;
    return n;
}

// Upper bound of what a context holds after a forward (+ adjoint) run with
// the fused kernel, the waveform misfit and no filter: every buffer of
// alloc_slots, alloc_traces, the second state set and the snapshot storage,
// plus allocator rounding of 2 MiB per buffer and 16 MiB for the small ones
// (point lists, tapers, absorbing tables, reduction scratch).
size_t ctx_bytes_estimate(const inv_grid_desc &d, int nsrc, int nrec, int snap_format, int ckpt) {
    const size_t B = (size_t)d.batch, N = (size_t)d.nx * d.nz, f = sizeof(float);
    const int ncomp = (d.sh ? 1 : 0) + (d.psv ? 2 : 0), nsnapf = (d.sh ? 2 : 0) + (d.psv ? 4 : 0);
    const int nstate = (d.sh ? 4 : 0) + (d.psv ? 7 : 0), nsecond = (d.sh ? 3 : 0) + (d.psv ? 5 : 0);
    size_t n = 0;
    int nbuf = 0;
    auto add = [&](size_t b) {
        n += b;
        nbuf++;
    };
    for (int i = 0; i < nstate + nsecond; i++) add(B * N * f);
    for (int i = 0; i < 7; i++) add(N * f);  // lam, mu, rho, dt/rho, three kernels
    if (d.adj && B > 1) add(3 * B * N * f);
    const size_t tr = (size_t)ncomp * std::max(nrec, 1) * d.nt * f;
    add(B * tr);
    add(B * tr);
    add((size_t)std::max(nsrc, 1) * tr);
    add(3 * (size_t)std::max(nsrc, 1) * d.nt * f);
    add(B * std::max(nrec, 1) * ncomp * sizeof(double));
    if (d.adj && d.sfe > 0) {
        const int nsfe = d.nt / d.sfe;
        const int nslots = ckpt > 0 ? ckpt / d.sfe : nsfe;
        const size_t es = snap_format == INV_SNAP_F32 ? 4 : 2;
        add(B * (size_t)nslots * nsnapf * N * es);
        if (snap_format != INV_SNAP_F32)
            add((size_t)nslots * snap_tab_bytes(d.batch, nsnapf, d.nx, d.nz));
        if (ckpt > 0) {
            add((size_t)(d.nt / ckpt - 1) * nstate * B * N * f);
            add((size_t)((d.sh ? 7 : 0) + (d.psv ? 12 : 0)) * B * N * f);
        }
    }
    return n + (size_t)nbuf * (2u << 20) + (16u << 20);
}

}  // namespace inv

using namespace inv;

template <typename F>
static int guard(inv_ctx *ctx, F &&f) {
    if (!ctx) return INV_ERR_ARG;
    try {
        INV_HIP(hipSetDevice(ctx->device));
        f(*ctx);
        ctx->err.clear();
        return INV_OK;
    } catch (const inv::Error &e) {
        ctx->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        ctx->err = "host allocation failed";
        return INV_ERR_NOMEM;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return INV_ERR_ARG;
    }
}

extern "C" {

int inv_grid_dims(int npt, float xmax, float zmax, int *nx, int *nz, float *dx, float *dz) {
    if (!nx || !nz || !dx || !dz || npt < 4 || !(xmax > 0) || !(zmax > 0)) return INV_ERR_ARG;
    // solver/fdm.h:382-389, in FP32 as the reference evaluates it
    float a = zmax / xmax;
    float b = a + 1;
    float c = 1 - npt;
    const size_t gx = (size_t)std::round((-b + std::sqrt(b * b - 4 * a * c)) / (2 * a)) + 1;
    const size_t gz = (size_t)npt / gx;
    if (gx < 2 || gz < 2) return INV_ERR_ARG;
    *nx = (int)gx;
    *nz = (int)gz;
    *dx = xmax / (float)(gx - 1);
    *dz = zmax / (float)(gz - 1);
    return INV_OK;
}

int inv_calc_idx(float coord, float length, int n) {
    // (int)(coord / L * (n - 1) + 0.5): FP32 product, double-literal rounding
    return (int)((double)(coord / length * (float)(n - 1)) + 0.5);
}

int inv_source_ricker(int nt, float dt, float f0, float t0, float angle, float amp,
                      float *stf_x, float *stf_y, float *stf_z) {
    if (nt < 1 || !stf_x || !stf_y || !stf_z) return INV_ERR_ARG;
    const float pi = 3.1415927f;  // host::pi, lib/device.h:68
    std::vector<float> stf(nt);
    float amax = 0;
    for (int it = 0; it < nt; it++) {
        const float t = (float)it * dt - t0;
        const float pft = pi * f0 * t;
        stf[it] = -t * expf(-pft * pft);
        if (fabsf(stf[it]) > amax) amax = fabsf(stf[it]);
    }
    if (amax > 0)
        for (int it = 0; it < nt; it++) stf[it] /= amax;
    for (int it = 0; it < nt; it++) {
        stf_x[it] = amp * stf[it] * cosf(angle);
        stf_y[it] = amp * stf[it];
        stf_z[it] = amp * stf[it] * sinf(angle);
    }
    return INV_OK;
}

int inv_device_count(int *count) {
    if (!count) return INV_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return INV_OK;
}

int inv_ctx_create(int device, inv_ctx **out) {
    if (!out) return INV_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
        (void)hipGetLastError();
        return INV_ERR_ARG;
    }
    inv_ctx *c = new (std::nothrow) inv_ctx();
    if (!c) return INV_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return INV_ERR_HIP;
    }
    try {
        noalloc_init(device);  // the no-allocation misfit's scratch on this device
    } catch (const inv::Error &e) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return e.code;
    }
    *out = c;
    return INV_OK;
}

int inv_ctx_destroy(inv_ctx *ctx) {
    if (!ctx) return INV_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    Ctx &c = *ctx;
    for (DevBuf *b : ctx_buffers(c)) b->release();
    for (auto &kv : c.src_lists) release_list(kv.second);
    drop_graphs(c);
    for (void *p : c.user_allocs) (void)hipFree(p);
    prof_release(c);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return INV_OK;
}
const char *inv_last_error(const inv_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int inv_sync(inv_ctx *ctx) {
    return guard(ctx, [](Ctx &c) { sync(c); });
}

int inv_mem_info(inv_ctx *ctx, size_t *free_bytes, size_t *total_bytes) {
    return guard(ctx, [&](Ctx &) {
        INV_REQUIRE(free_bytes && total_bytes, INV_ERR_ARG, "null output");
        INV_HIP(hipMemGetInfo(free_bytes, total_bytes));
    });
}

int inv_grid_set(inv_ctx *ctx, const inv_grid_desc *d) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(d, INV_ERR_ARG, "null descriptor");
        INV_REQUIRE(d->nx >= 5 && d->nz >= 5, INV_ERR_ARG, "grid must be at least 5x5");
        INV_REQUIRE(d->sh || d->psv, INV_ERR_ARG, "enable sh and/or psv");
        INV_REQUIRE(d->nt >= 1 && d->dt > 0 && d->dx > 0 && d->dz > 0, INV_ERR_ARG,
                    "bad nt/dt/dx/dz");
        INV_REQUIRE(d->batch >= 1, INV_ERR_ARG, "batch must be >= 1");
        // cell indices (point lists, kernels' row spans, which reach 20 rows
        // past the grid) are 32-bit
        INV_REQUIRE((long)(d->nx + 32) * d->nz < (1L << 31), INV_ERR_ARG, "grid too large");
        if (d->adj) {
            // nt % sfe != 0 makes the reference's adjoint read past the last
            // snapshot slot (solver/fdm.h:583-585); reject it here.
            INV_REQUIRE(d->sfe >= 1 && d->nt % d->sfe == 0, INV_ERR_ARG,
                        "adj needs sfe >= 1 with nt % sfe == 0");
        }
        sync(c);
        c.g = *d;
        c.N = (long)d->nx * d->nz;
        c.nsfe = (d->adj && d->sfe > 0) ? d->nt / d->sfe : 0;
        c.ncomp = (d->sh ? 1 : 0) + (d->psv ? 2 : 0);
        c.nsnapf = (d->sh ? 2 : 0) + (d->psv ? 4 : 0);
        c.have_grid = true;
        c.have_model = false;
        c.slot_models = false;
        build_absbound(c);
        build_taper(c);
        alloc_slots(c);
        alloc_traces(c);
        c.parity = 0;
        invalidate(c);
        fused_invalidate_lists(c);
        if (c.have_rec) build_receivers(c);
    });
}

int inv_model_upload_lmr(inv_ctx *ctx, const float *lambda, const float *mu, const float *rho) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set");
        INV_REQUIRE(lambda && mu && rho, INV_ERR_ARG, "null model array");
        // hipMemcpy does not order against the non-blocking context stream:
        // wait for kernels that may still read the model
        sync(c);
        const size_t b = c.N * sizeof(float);
        INV_HIP(hipMemcpy(c.lam.p, lambda, b, hipMemcpyHostToDevice));
        INV_HIP(hipMemcpy(c.mu.p, mu, b, hipMemcpyHostToDevice));
        INV_HIP(hipMemcpy(c.rho.p, rho, b, hipMemcpyHostToDevice));
        c.have_model = true;
        c.slot_models = false;
    });
}

int inv_model_download_lmr(inv_ctx *ctx, float *lambda, float *mu, float *rho) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set");
        sync(c);
        const size_t b = c.N * sizeof(float);
        if (lambda) INV_HIP(hipMemcpy(lambda, c.lam.p, b, hipMemcpyDeviceToHost));
        if (mu) INV_HIP(hipMemcpy(mu, c.mu.p, b, hipMemcpyDeviceToHost));
        if (rho) INV_HIP(hipMemcpy(rho, c.rho.p, b, hipMemcpyDeviceToHost));
    });
}

int inv_model_from_points(inv_ctx *ctx, int npt, const float *px, const float *pz,
                          const float *p1, const float *p2, const float *prho, float xmax,
                          float zmax, int model_type) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set");
        INV_REQUIRE(npt > 0 && px && pz && p1 && p2 && prho, INV_ERR_ARG, "bad model points");
        const int nx = c.g.nx, nz = c.g.nz;
        const float dx = c.g.dx, dz = c.g.dz;
        std::vector<float> l(c.N), m(c.N), r(c.N);
        // Fast path: every grid node has a model point exactly on it.  Then
        // the brute-force nearest search of mesh2grid (strict <, first wins)
        // picks the first point at distance 0, which a lookup finds directly.
        std::vector<int> hit(c.N, -1);
        for (int p = 0; p < npt; p++) {
            const float fi = px[p] / dx, fj = pz[p] / dz;
            const long i = std::lround(fi), j = std::lround(fj);
            if (i < 0 || i >= nx || j < 0 || j >= nz) continue;
            if ((float)i * dx != px[p] || (float)j * dz != pz[p]) continue;
            const long k = i * nz + j;
            if (hit[k] < 0) hit[k] = p;
        }
        bool all = true;
        for (long k = 0; k < c.N; k++) all = all && hit[k] >= 0;
        if (all) {
            for (long k = 0; k < c.N; k++) {
                l[k] = p1[hit[k]];
                m[k] = p2[hit[k]];
                r[k] = prho[hit[k]];
            }
        } else {
            // general case: the nearest point of each node, on the device
            // (the same choice as mesh2grid's scan, model.hip)
            sync(c);
            model_nearest(c, npt, px, pz, p1, p2, prho, xmax * xmax + zmax * zmax);
            const size_t b = c.N * sizeof(float);
            INV_HIP(hipMemcpy(l.data(), c.lam.p, b, hipMemcpyDeviceToHost));
            INV_HIP(hipMemcpy(m.data(), c.mu.p, b, hipMemcpyDeviceToHost));
            INV_HIP(hipMemcpy(r.data(), c.rho.p, b, hipMemcpyDeviceToHost));
        }
        sync(c);
        if (model_type == 0) {
            // vps2lm, solver/fdm.h:25-36
            for (long k = 0; k < c.N; k++) {
                const float vp = l[k], vs = m[k];
                l[k] = vp > vs ? r[k] * (vp * vp - 2 * vs * vs) : 0.0f;
                m[k] = r[k] * vs * vs;
            }
        }
        const size_t b = c.N * sizeof(float);
        INV_HIP(hipMemcpy(c.lam.p, l.data(), b, hipMemcpyHostToDevice));
        INV_HIP(hipMemcpy(c.mu.p, m.data(), b, hipMemcpyHostToDevice));
        INV_HIP(hipMemcpy(c.rho.p, r.data(), b, hipMemcpyHostToDevice));
        c.have_model = true;
        c.slot_models = false;
    });
}

int inv_model_to_vps(inv_ctx *ctx, float *vp, float *vs, float *rho) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid && vp && vs && rho, INV_ERR_ARG, "bad arguments");
        sync(c);
        const size_t b = c.N * sizeof(float);
        INV_HIP(hipMemcpy(vp, c.lam.p, b, hipMemcpyDeviceToHost));
        INV_HIP(hipMemcpy(vs, c.mu.p, b, hipMemcpyDeviceToHost));
        INV_HIP(hipMemcpy(rho, c.rho.p, b, hipMemcpyDeviceToHost));
        // lm2vps, solver/fdm.h:37-43
        for (long k = 0; k < c.N; k++) {
            const float l = vp[k], m = vs[k];
            vp[k] = sqrtf((l + 2 * m) / rho[k]);
            vs[k] = sqrtf(m / rho[k]);
        }
    });
}

int inv_sources_set(inv_ctx *ctx, int nsrc, const int *ix, const int *iz, const float *stf_x,
                    const float *stf_y, const float *stf_z) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set");
        INV_REQUIRE(nsrc >= 1 && ix && iz && stf_x && stf_y && stf_z, INV_ERR_ARG, "bad sources");
        for (int s = 0; s < nsrc; s++)
            INV_REQUIRE(ix[s] >= 0 && ix[s] < c.g.nx && iz[s] >= 0 && iz[s] < c.g.nz,
                        INV_ERR_ARG, "source outside the grid");
        sync(c);
        c.nsrc = nsrc;
        c.src_ix.assign(ix, ix + nsrc);
        c.src_iz.assign(iz, iz + nsrc);
        c.src_type.assign(nsrc, 0);
        const size_t b = (size_t)nsrc * c.g.nt * sizeof(float);
        c.stf.ensure(3 * b);
        INV_HIP(hipMemcpy(c.stf.as<char>(), stf_x, b, hipMemcpyHostToDevice));
        INV_HIP(hipMemcpy(c.stf.as<char>() + b, stf_y, b, hipMemcpyHostToDevice));
        INV_HIP(hipMemcpy(c.stf.as<char>() + 2 * b, stf_z, b, hipMemcpyHostToDevice));
        c.have_src = true;
        invalidate(c);
        fused_invalidate_lists(c);
        alloc_traces(c);
    });
}

int inv_sources_type(inv_ctx *ctx, const int *type) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_src, INV_ERR_STATE, "sources not set");
        std::vector<int> t(c.nsrc, 0);
        if (type)
            for (int s = 0; s < c.nsrc; s++) {
                INV_REQUIRE(type[s] == 0 || type[s] == 1, INV_ERR_ARG,
                            "source type must be 0 (point force) or 1 (explosive)");
                t[s] = type[s];
            }
        sync(c);
        c.src_type = t;
        invalidate(c);
        fused_invalidate_lists(c);
    });
}

int inv_receivers_set(inv_ctx *ctx, int nrec, const int *ix, const int *iz) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set");
        INV_REQUIRE(nrec >= 1 && ix && iz, INV_ERR_ARG, "bad receivers");
        for (int r = 0; r < nrec; r++)
            INV_REQUIRE(ix[r] >= 0 && ix[r] < c.g.nx && iz[r] >= 0 && iz[r] < c.g.nz,
                        INV_ERR_ARG, "receiver outside the grid");
        sync(c);
        c.nrec = nrec;
        c.rec_ix.assign(ix, ix + nrec);
        c.rec_iz.assign(iz, iz + nrec);
        build_receivers(c);
        invalidate(c);
        fused_invalidate_lists(c);
        c.have_rec = true;
        alloc_traces(c);
    });
}

static void run_steps_direct(Ctx &c, int kind, int first, int count, int flags, int it0,
                             int it1) {
    const bool split = c.variant == INV_VARIANT_SPLIT;
    if (kind == 1) {
        if (split) fdm_forward(c, first, count, flags, it0, it1);
        else fused_forward(c, first, count, flags, it0, it1);
    } else {
        if (split) fdm_adjoint(c, first, count, it0, it1);
        else fused_adjoint(c, first, count, it0, it1);
    }
}

static int forward_end(const Ctx &c) {
    return (c.step_limit > 0 && c.step_limit < c.g.nt) ? c.step_limit : c.g.nt;
}

// A whole forward or adjoint run is captured once into a HIP graph (every
// launch of the time loop, nt to nt + nt/sfe kernels) and replayed after
// that; runs with event sampling, partial runs and stepping go direct.
static void run_steps(Ctx &c, int kind, int first, int count, int flags, int it0, int it1) {
    const int end = kind == 1 ? forward_end(c) : c.g.nt;
    INV_REQUIRE(c.variant == INV_VARIANT_FUSED || (c.snap_format == INV_SNAP_F32 && c.ckpt == 0),
                INV_ERR_UNSUPPORTED,
                "16-bit snapshots and checkpoints need the fused column kernel (variant 0)");
    INV_REQUIRE(c.variant != INV_VARIANT_FUSED || c.g.nz % c.col_v == 0, INV_ERR_UNSUPPORTED,
                "the column kernel's columns per lane must divide nz");
    INV_REQUIRE(!c.slot_models || c.variant != INV_VARIANT_SPLIT, INV_ERR_UNSUPPORTED,
                "per-slot models need a fused step kernel");
    INV_REQUIRE(!c.any_explosive() || c.variant == INV_VARIANT_FUSED, INV_ERR_UNSUPPORTED,
                "explosive sources need the fused column kernel (variant 0)");
    INV_REQUIRE(c.variant != INV_VARIANT_FUSED || c.pair_tx == 0 || c.g.nz % c.pair_v == 0,
                INV_ERR_UNSUPPORTED, "the two-step kernel's columns per lane must divide nz");
    if (kind == 2 || (flags & INV_FWD_SNAPSHOTS)) snap_storage(c);
    if (!c.use_graphs || it0 != 0 || it1 != end) {
        run_steps_direct(c, kind, first, count, flags, it0, it1);
        return;
    }
    const std::string key = std::to_string(kind) + "/" + std::to_string(first) + "/" +
                            std::to_string(count) + "/" + std::to_string(flags) + "/" +
                            std::to_string(end);
    auto found = c.graphs.find(key);
    if (found == c.graphs.end() || found->second.gen != c.gen) {
        if (found != c.graphs.end()) {
            sync(c);
            if (found->second.exec) (void)hipGraphExecDestroy(found->second.exec);
            if (found->second.graph) (void)hipGraphDestroy(found->second.graph);
            c.graphs.erase(found);
        }
        // everything that allocates, uploads or synchronises happens before
        // the capture; inside it the run only launches kernels (no event
        // sampling inside a capture)
        if (c.variant == INV_VARIANT_SPLIT) fdm_prepare(c, kind, first, count);
        else fused_prepare(c, kind, first, count, flags);
        sync(c);
        Ctx::GraphEntry e;
        const int every = c.prof.every;
        c.prof.every = 0;
        INV_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        try {
            run_steps_direct(c, kind, first, count, flags, it0, it1);
        } catch (...) {
            c.prof.every = every;
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        c.prof.every = every;
        e.parity = c.parity;  // the host loop ran: the set the replay ends on
        INV_HIP(hipStreamEndCapture(c.stream, &e.graph));
        INV_HIP(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
        e.gen = c.gen;
        found = c.graphs.emplace(key, e).first;
    }
    // with sampling on, each replay is bracketed by events (units = steps)
    hipEvent_t ev;
    const bool p = prof_begin_any(c, &ev);
    INV_HIP(hipGraphLaunch(found->second.exec, c.stream));
    if (p) prof_end(c, kind == 1 ? PROF_FWD_RUN : PROF_ADJ_RUN, ev, end);
    // the replay does not run the host loop: set the state-set parity it ends
    // on (the fused kernels alternate sets each launch, one or two steps per
    // launch, as the captured loop recorded; the kernel pair works in place)
    c.parity = c.variant == INV_VARIANT_SPLIT ? 0 : found->second.parity;
}

int inv_forward(inv_ctx *ctx, int first, int count, int flags) {
    return guard(ctx, [&](Ctx &c) {
        require_ready(c);
        check_slots(c, first, count);
        run_steps(c, 1, first, count, flags, 0, forward_end(c));
        c.run_kind = 0;
    });
}

int inv_adjoint(inv_ctx *ctx, int first, int count) {
    return guard(ctx, [&](Ctx &c) {
        require_ready(c);
        check_slots(c, first < 0 ? 0 : first, count);
        INV_REQUIRE(c.g.adj, INV_ERR_STATE, "adjoint needs adj = 1 (snapshots)");
        run_steps(c, 2, first, count, 0, 0, c.g.nt);
        c.run_kind = 0;
    });
}

int inv_forward_begin(inv_ctx *ctx, int first, int count, int flags) {
    return guard(ctx, [&](Ctx &c) {
        require_ready(c);
        check_slots(c, first, count);
        c.run_kind = 1;
        c.run_first = first;
        c.run_count = count;
        c.run_flags = flags;
        c.run_it = 0;
    });
}

int inv_adjoint_begin(inv_ctx *ctx, int first, int count) {
    return guard(ctx, [&](Ctx &c) {
        require_ready(c);
        check_slots(c, first < 0 ? 0 : first, count);
        INV_REQUIRE(c.g.adj, INV_ERR_STATE, "adjoint needs adj = 1 (snapshots)");
        c.run_kind = 2;
        c.run_first = first;
        c.run_count = count;
        c.run_flags = 0;
        c.run_it = 0;
    });
}

int inv_run_advance(inv_ctx *ctx, int nsteps, int *position) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.run_kind != 0, INV_ERR_STATE, "no run begun (inv_forward_begin)");
        INV_REQUIRE(nsteps >= 0, INV_ERR_ARG, "nsteps must be >= 0");
        const int end = c.run_kind == 1 ? forward_end(c) : c.g.nt;
        const int it1 = std::min(end, c.run_it + nsteps);
        if (it1 > c.run_it)
            run_steps(c, c.run_kind, c.run_first, c.run_count, c.run_flags, c.run_it, it1);
        c.run_it = it1;
        if (position) *position = it1;
        if (it1 == end) c.run_kind = 0;
    });
}

int inv_kernels_zero(inv_ctx *ctx) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set");
        vec_fill(c, c.k_lambda.as<float>(), 0.0f, c.N);
        vec_fill(c, c.k_mu.as<float>(), 0.0f, c.N);
        vec_fill(c, c.k_rho.as<float>(), 0.0f, c.N);
    });
}

int inv_kernels_get(inv_ctx *ctx, float *kl, float *km, float *kr) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set");
        sync(c);
        const size_t b = c.N * sizeof(float);
        if (kl) INV_HIP(hipMemcpy(kl, c.k_lambda.p, b, hipMemcpyDeviceToHost));
        if (km) INV_HIP(hipMemcpy(km, c.k_mu.p, b, hipMemcpyDeviceToHost));
        if (kr) INV_HIP(hipMemcpy(kr, c.k_rho.p, b, hipMemcpyDeviceToHost));
    });
}

int inv_kernels_set(inv_ctx *ctx, const float *kl, const float *km, const float *kr) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set");
        sync(c);
        const size_t b = c.N * sizeof(float);
        if (kl) INV_HIP(hipMemcpy(c.k_lambda.p, kl, b, hipMemcpyHostToDevice));
        if (km) INV_HIP(hipMemcpy(c.k_mu.p, km, b, hipMemcpyHostToDevice));
        if (kr) INV_HIP(hipMemcpy(c.k_rho.p, kr, b, hipMemcpyHostToDevice));
    });
}

static size_t trace_bytes(const Ctx &c) { return (size_t)c.ncomp * c.nrec * c.g.nt * sizeof(float); }

int inv_traces_get(inv_ctx *ctx, int slot, float *host) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_rec && host && slot >= 0 && slot < c.g.batch, INV_ERR_ARG, "bad slot");
        sync(c);
        INV_HIP(hipMemcpy(host, c.out.as<char>() + slot * trace_bytes(c), trace_bytes(c),
                          hipMemcpyDeviceToHost));
    });
}

int inv_adstf_set(inv_ctx *ctx, int slot, const float *host) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_rec && host && slot >= 0 && slot < c.g.batch, INV_ERR_ARG, "bad slot");
        sync(c);
        INV_HIP(hipMemcpy(c.adstf.as<char>() + slot * trace_bytes(c), host, trace_bytes(c),
                          hipMemcpyHostToDevice));
    });
}

int inv_adstf_get(inv_ctx *ctx, int slot, float *host) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_rec && host && slot >= 0 && slot < c.g.batch, INV_ERR_ARG, "bad slot");
        sync(c);
        INV_HIP(hipMemcpy(host, c.adstf.as<char>() + slot * trace_bytes(c), trace_bytes(c),
                          hipMemcpyDeviceToHost));
    });
}

int inv_obs_set(inv_ctx *ctx, int isrc, const float *host) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_src && c.have_rec && host && isrc >= 0 && isrc < c.nsrc, INV_ERR_ARG,
                    "bad shot");
        sync(c);
        INV_HIP(hipMemcpy(c.obs.as<char>() + isrc * trace_bytes(c), host, trace_bytes(c),
                          hipMemcpyHostToDevice));
    });
}

int inv_obs_get(inv_ctx *ctx, int isrc, float *host) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_src && c.have_rec && host && isrc >= 0 && isrc < c.nsrc, INV_ERR_ARG,
                    "bad shot");
        sync(c);
        INV_HIP(hipMemcpy(host, c.obs.as<char>() + isrc * trace_bytes(c), trace_bytes(c),
                          hipMemcpyDeviceToHost));
    });
}

int inv_obs_from_traces(inv_ctx *ctx, int first, int count) {
    return guard(ctx, [&](Ctx &c) {
        require_ready(c);
        check_slots(c, first, count);
        INV_REQUIRE(first >= 0, INV_ERR_ARG, "first must name a shot");
        INV_HIP(hipMemcpyAsync(c.obs.as<char>() + first * trace_bytes(c), c.out.p,
                               count * trace_bytes(c), hipMemcpyDeviceToDevice, c.stream));
    });
}

static double run_misfit(Ctx &c, int first, int count, int type, double *chi_host) {
    if (type == 0) misfit_waveform(c, count, first);
    else if (type == 1) misfit_envelope(c, count, first);
    else throw Error{INV_ERR_ARG, "misfit type must be 0 (waveform) or 1 (envelope)"};
    const size_t ntr = (size_t)count * c.nrec * c.ncomp;
    c.scal.ensure(8 * sizeof(double));
    sum_f64(c, c.chi.as<double>(), ntr, c.scal.as<double>());
    double total = 0;
    INV_HIP(hipMemcpyAsync(&total, c.scal.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (chi_host)
        INV_HIP(hipMemcpyAsync(chi_host, c.chi.p, ntr * sizeof(double), hipMemcpyDeviceToHost,
                               c.stream));
    sync(c);
    return total;
}

int inv_misfit(inv_ctx *ctx, int first, int count, int type, double *chi, double *total) {
    return guard(ctx, [&](Ctx &c) {
        require_ready(c);
        check_slots(c, first, count);
        INV_REQUIRE(first >= 0, INV_ERR_ARG, "first must name a shot");
        const double t = run_misfit(c, first, count, type, chi);
        if (total) *total = t;
    });
}

int inv_misfit_trace(inv_ctx *ctx, int type, const float *syn, const float *obs, float *adstf,
                     double *chi) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set");
        INV_REQUIRE(syn && obs && adstf && chi, INV_ERR_ARG, "null argument");
        INV_REQUIRE(type == 0 || type == 1, INV_ERR_ARG, "misfit type must be 0 or 1");
        *chi = misfit_trace(c, type, syn, obs, adstf);
    });
}

int inv_gradient(inv_ctx *ctx, int misfit_type, int sigma, int kernel, double *misfit) {
    return guard(ctx, [&](Ctx &c) {
        require_ready(c);
        INV_REQUIRE(c.g.adj, INV_ERR_STATE, "gradient needs adj = 1");
        if (kernel) {
            vec_fill(c, c.k_lambda.as<float>(), 0.0f, c.N);
            vec_fill(c, c.k_mu.as<float>(), 0.0f, c.N);
            vec_fill(c, c.k_rho.as<float>(), 0.0f, c.N);
        }
        double total = 0;
        for (int first = 0; first < c.nsrc; first += c.g.batch) {
            const int count = std::min(c.g.batch, c.nsrc - first);
            run_steps(c, 1, first, count, kernel ? INV_FWD_SNAPSHOTS : 0, 0, c.g.nt);
            total += run_misfit(c, first, count, misfit_type, nullptr);
            if (kernel) run_steps(c, 2, first, count, 0, 0, c.g.nt);
        }
        if (kernel && sigma > 0) {
            gauss_filter(c, c.k_lambda.as<float>(), c.g.nx, c.g.nz, sigma);
            gauss_filter(c, c.k_mu.as<float>(), c.g.nx, c.g.nz, sigma);
            gauss_filter(c, c.k_rho.as<float>(), c.g.nx, c.g.nz, sigma);
        }
        sync(c);
        if (misfit) *misfit = total;
    });
}

int inv_filter_gauss1s(inv_ctx *ctx, float *dev_field, int nx, int nz, int sigma) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(dev_field, INV_ERR_ARG, "null field");
        gauss_filter(c, dev_field, nx, nz, sigma);
    });
}

int inv_vec_axpby(inv_ctx *ctx, float *cc, float ka, const float *a, float kb, const float *b,
                  size_t n) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(cc && a && b, INV_ERR_ARG, "null vector");
        vec_axpby(c, cc, ka, a, kb, b, n);
    });
}
int inv_vec_scale(inv_ctx *ctx, float *cc, float ka, const float *a, size_t n) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(cc && a, INV_ERR_ARG, "null vector");
        vec_scale(c, cc, ka, a, n);
    });
}
int inv_vec_mul(inv_ctx *ctx, float *cc, const float *a, const float *b, size_t n) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(cc && a && b, INV_ERR_ARG, "null vector");
        vec_mul(c, cc, a, b, n);
    });
}
int inv_vec_copy(inv_ctx *ctx, float *dst, const float *src, size_t n) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(dst && src, INV_ERR_ARG, "null vector");
        vec_copy(c, dst, src, n);
    });
}
int inv_vec_fill(inv_ctx *ctx, float *dst, float value, size_t n) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(dst, INV_ERR_ARG, "null vector");
        vec_fill(c, dst, value, n);
    });
}
int inv_vec_dot_f64(inv_ctx *ctx, const float *a, const float *b, size_t n, double *out) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(a && b && out, INV_ERR_ARG, "null argument");
        *out = vec_dot(c, a, b, n);
    });
}
int inv_vec_amax(inv_ctx *ctx, const float *a, size_t n, float *out) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(a && out, INV_ERR_ARG, "null argument");
        *out = vec_amax(c, a, n);
    });
}

int inv_field_ptr(inv_ctx *ctx, int field, float **out) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(out, INV_ERR_ARG, "null output");
        INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set");
        float *p = nullptr;
        switch (field) {
            case INV_LAMBDA: p = c.lam.as<float>(); break;
            case INV_MU: p = c.mu.as<float>(); break;
            case INV_RHO: p = c.rho.as<float>(); break;
            case INV_K_LAMBDA: p = c.k_lambda.as<float>(); break;
            case INV_K_MU: p = c.k_mu.as<float>(); break;
            case INV_K_RHO: p = c.k_rho.as<float>(); break;
            case INV_OUT: p = c.out.as<float>(); break;
            case INV_ADSTF: p = c.adstf.as<float>(); break;
            case INV_OBS: p = c.obs.as<float>(); break;
            default: throw Error{INV_ERR_ARG, "unknown field id"};
        }
        INV_REQUIRE(p, INV_ERR_STATE, "field not allocated");
        *out = p;
    });
}

int inv_alloc(inv_ctx *ctx, size_t bytes, void **out) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(out && bytes > 0, INV_ERR_ARG, "bad allocation request");
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error{INV_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e)};
        }
        c.user_allocs.push_back(p);
        *out = p;
    });
}

int inv_free(inv_ctx *ctx, void *dev) {
    return guard(ctx, [&](Ctx &c) {
        auto it = std::find(c.user_allocs.begin(), c.user_allocs.end(), dev);
        INV_REQUIRE(it != c.user_allocs.end(), INV_ERR_ARG, "pointer not from inv_alloc");
        sync(c);
        INV_HIP(hipFree(dev));
        c.user_allocs.erase(it);
    });
}

int inv_memcpy_h2d(inv_ctx *ctx, void *dev, const void *host, size_t bytes) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(dev && host, INV_ERR_ARG, "null pointer");
        sync(c);
        INV_HIP(hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice));
    });
}

int inv_memcpy_d2h(inv_ctx *ctx, void *host, const void *dev, size_t bytes) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(dev && host, INV_ERR_ARG, "null pointer");
        sync(c);
        INV_HIP(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost));
    });
}

int inv_state_get(inv_ctx *ctx, int slot, float *host) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid && host && slot >= 0 && slot < c.g.batch, INV_ERR_ARG, "bad slot");
        sync(c);
        const size_t b = c.N * sizeof(float);
        int f = 0;
        auto get = [&](const float *d) {
            INV_HIP(hipMemcpy(reinterpret_cast<char *>(host) + (size_t)f++ * b,
                              reinterpret_cast<const char *>(d) + (size_t)slot * b, b,
                              hipMemcpyDeviceToHost));
        };
        auto cur = [&](DevBuf &x, DevBuf &y) { return cur_field(c, x, y); };
        if (c.g.sh) {
            get(cur(c.vy, c.vy_b)); get(c.uy.as<float>());
            get(cur(c.sxy, c.sxy_b)); get(cur(c.szy, c.szy_b));
        }
        if (c.g.psv) {
            get(cur(c.vx, c.vx_b)); get(cur(c.vz, c.vz_b)); get(c.ux.as<float>());
            get(c.uz.as<float>()); get(cur(c.sxx, c.sxx_b)); get(cur(c.szz, c.szz_b));
            get(cur(c.sxz, c.sxz_b));
        }
    });
}

int inv_state_set(inv_ctx *ctx, int slot, const float *host) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid && host && slot >= 0 && slot < c.g.batch, INV_ERR_ARG, "bad slot");
        sync(c);
        const size_t b = c.N * sizeof(float);
        int f = 0;
        auto put = [&](float *d) {
            INV_HIP(hipMemcpy(reinterpret_cast<char *>(d) + (size_t)slot * b,
                              reinterpret_cast<const char *>(host) + (size_t)f++ * b, b,
                              hipMemcpyHostToDevice));
        };
        auto cur = [&](DevBuf &x, DevBuf &y) { return cur_field(c, x, y); };
        if (c.g.sh) {
            put(cur(c.vy, c.vy_b)); put(c.uy.as<float>());
            put(cur(c.sxy, c.sxy_b)); put(cur(c.szy, c.szy_b));
        }
        if (c.g.psv) {
            put(cur(c.vx, c.vx_b)); put(cur(c.vz, c.vz_b)); put(c.ux.as<float>());
            put(c.uz.as<float>()); put(cur(c.sxx, c.sxx_b)); put(cur(c.szz, c.szz_b));
            put(cur(c.sxz, c.sxz_b));
        }
    });
}

static void check_half_step(Ctx &c, int count) {
    require_ready(c);
    INV_REQUIRE(count >= 1 && count <= c.g.batch, INV_ERR_ARG, "count must be in [1, batch]");
    INV_REQUIRE(!c.slot_models, INV_ERR_UNSUPPORTED, "half-steps need the context's own model");
    INV_REQUIRE(c.run_kind == 0, INV_ERR_STATE, "an incremental run is in progress");
}

int inv_step_velocity(inv_ctx *ctx, int first, int count, int it) {
    return guard(ctx, [&](Ctx &c) {
        check_half_step(c, count);
        check_slots(c, first, count);
        INV_REQUIRE(it >= 0 && it < c.g.nt, INV_ERR_ARG, "time step out of range");
        INV_REQUIRE(!c.any_explosive(), INV_ERR_UNSUPPORTED,
                    "half-steps inject point forces only (no explosive sources)");
        fdm_half_step(c, 0, first, count, it);
    });
}

int inv_step_stress(inv_ctx *ctx, int count) {
    return guard(ctx, [&](Ctx &c) {
        check_half_step(c, count);
        fdm_half_step(c, 1, 0, count, 0);
    });
}

// IEEE binary16 to float (subnormals included)
static float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float v;
    if (e == 0) {
        v = std::ldexp((float)m, -24);
        return sign ? -v : v;
    }
    const uint32_t bits = e == 31 ? (sign | 0x7f800000u | (m << 13)) : (sign | ((e + 112u) << 23) | (m << 13));
    std::memcpy(&v, &bits, 4);
    return v;
}

int inv_snapshots_get(inv_ctx *ctx, int slot, float *host) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.snap.p && host && slot >= 0 && slot < c.g.batch, INV_ERR_ARG, "bad slot");
        sync(c);
        const size_t n = (size_t)c.nslots * c.nsnapf * c.N;
        if (c.snap_format == INV_SNAP_F32) {
            INV_HIP(hipMemcpy(host, c.snap.as<char>() + slot * n * sizeof(float), n * sizeof(float),
                              hipMemcpyDeviceToHost));
            return;
        }
        std::vector<uint16_t> q(n);
        const size_t tb = snap_tab_bytes(c.g.batch, c.nsnapf, c.g.nx, c.g.nz);
        std::vector<signed char> tab((size_t)c.nslots * tb);
        INV_HIP(hipMemcpy(q.data(), c.snap.as<char>() + slot * n * 2, n * 2, hipMemcpyDeviceToHost));
        INV_HIP(hipMemcpy(tab.data(), c.sexp.p, tab.size(), hipMemcpyDeviceToHost));
        const int nseg = snap_nseg(c.g.nz), nbb = snap_nbb(c.g.nz);
        const long sblk = snap_blk_off(c.g.batch, c.nsnapf, c.g.nx, c.g.nz);
        for (int l = 0; l < c.nslots; l++) {
            const signed char *t = tab.data() + (size_t)l * tb;
            int segw;
            std::memcpy(&segw, t, 4);
            segw = segw > 0 ? segw : 1;
            for (int f = 0; f < c.nsnapf; f++) {
                const size_t o = ((size_t)l * c.nsnapf + f) * c.N;
                const size_t r0 = ((size_t)slot * c.nsnapf + f) * c.g.nx;
                const signed char *rows = t + SNAPTAB_HDR + r0 * nseg;
                const unsigned char *blks = reinterpret_cast<const unsigned char *>(t) + sblk +
                                            r0 * nbb;
                for (long k = 0; k < c.N; k++) {
                    const int i = (int)(k / c.g.nz), j = (int)(k % c.g.nz);
                    const int e = snap_cell_exp(rows + (size_t)i * nseg, blks + (size_t)i * nbb,
                                                nseg, nbb, segw, j);
                    const uint16_t v = q[o + k];
                    float d;
                    if (c.snap_format == INV_SNAP_F16) {
                        d = half_to_float(v);
                    } else if (c.snap_format == INV_SNAP_I16) {
                        d = (float)(int16_t)v;
                    } else {
                        const uint32_t bits = (uint32_t)v << 16;
                        std::memcpy(&d, &bits, 4);
                    }
                    host[o + k] = d * pow2f(-e);
                }
            }
        }
    });
}

int inv_set_snapshots(inv_ctx *ctx, int format, int ckpt_interval) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(c.have_grid, INV_ERR_STATE, "grid not set");
        INV_REQUIRE(format >= INV_SNAP_F32 && format <= INV_SNAP_I16, INV_ERR_ARG,
                    "unknown snapshot format");
        INV_REQUIRE(ckpt_interval >= 0, INV_ERR_ARG, "checkpoint interval must be >= 0");
        if (ckpt_interval > 0)
            INV_REQUIRE(c.g.adj && c.g.sfe > 0 && ckpt_interval % c.g.sfe == 0 &&
                            c.g.nt % ckpt_interval == 0,
                        INV_ERR_ARG,
                        "checkpoint interval must be a multiple of sfe that divides nt (adj = 1)");
        sync(c);
        c.snap.release();
        c.sexp.release();
        c.ckbuf.release();
        c.rstate.release();
        c.snap_format = format;
        c.ckpt = ckpt_interval;
        c.nslots = ckpt_interval > 0 ? ckpt_interval / c.g.sfe : c.nsfe;
        invalidate(c);
    });
}

int inv_prof_enable(inv_ctx *ctx, int every) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(every >= 0, INV_ERR_ARG, "every must be >= 0");
        c.prof.every = every;
    });
}

int inv_prof_reset(inv_ctx *ctx) {
    return guard(ctx, [&](Ctx &c) {
        prof_resolve(c);
        for (int k = 0; k < PROF_NKIND; k++) {
            c.prof.count[k] = 0;
            c.prof.ms[k] = 0;
        }
    });
}

int inv_prof_read(inv_ctx *ctx, int kernel, long long *count, double *total_ms) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(kernel >= 0 && kernel < PROF_NKIND && count && total_ms, INV_ERR_ARG,
                    "bad kernel id");
        prof_resolve(c);
        *count = c.prof.count[kernel];
        *total_ms = c.prof.ms[kernel];
    });
}

int inv_set_variant(inv_ctx *ctx, int variant, int tile_x) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(variant == INV_VARIANT_FUSED || variant == INV_VARIANT_SPLIT ||
                        variant == INV_VARIANT_FUSED_LINEAR,
                    INV_ERR_ARG, "unknown variant");
        INV_REQUIRE(tile_x == 0 || tile_x == 8 || tile_x == 16 || tile_x == 32, INV_ERR_ARG,
                    "tile_x must be 0, 8, 16 or 32");
        sync(c);
        c.variant = variant;
        if (tile_x) c.tile_x = tile_x;
        c.col_v = 1;
        invalidate(c);
        fused_invalidate_lists(c);
    });
}

// tile heights the column kernel is built for, per columns per lane
static bool tile_ok(int tile_x, int v) {
    if (v == 1) return tile_x == 8 || tile_x == 16 || tile_x == 32;
    if (v == 2) return tile_x == 4 || tile_x == 8 || tile_x == 16;
    if (v == 4) return tile_x == 4 || tile_x == 8;
    return false;
}

int inv_set_tile(inv_ctx *ctx, int tile_x, int cols_per_lane) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(tile_ok(tile_x, cols_per_lane), INV_ERR_ARG,
                    "tile must be 8/16/32 rows x 1 column per lane, 4/8/16 x 2 or 4/8 x 4");
        INV_REQUIRE(!c.have_grid || c.g.nz % cols_per_lane == 0, INV_ERR_UNSUPPORTED,
                    "nz must be divisible by the columns per lane");
        sync(c);
        c.variant = INV_VARIANT_FUSED;
        c.tile_x = tile_x;
        c.col_v = cols_per_lane;
        invalidate(c);
        fused_invalidate_lists(c);
    });
}

static bool pair_tile_ok(int rows, int cols) {
    return rows == 0 ||
           (cols == 2 && (rows == 4 || rows == 8 || rows == 12 || rows == 16 || rows == 20)) ||
           (cols == 4 && (rows == 4 || rows == 8));
}

int inv_set_pair(inv_ctx *ctx, int tile_x, int cols_per_lane) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(pair_tile_ok(tile_x, cols_per_lane), INV_ERR_ARG,
                    "two-step tile must be 4/8/12/16/20 rows x 2 columns per lane or 4/8 x 4 (0 rows: off)");
        INV_REQUIRE(!c.have_grid || tile_x == 0 || c.g.nz % cols_per_lane == 0,
                    INV_ERR_UNSUPPORTED, "nz must be divisible by the columns per lane");
        sync(c);
        c.pair_tx = tile_x;
        c.pair_v = tile_x ? cols_per_lane : 2;
        invalidate(c);
        fused_invalidate_lists(c);
    });
}

int inv_set_persistent(inv_ctx *ctx, int mode) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(mode >= 0 && mode <= 2, INV_ERR_ARG, "persistent mode must be 0, 1 or 2");
        sync(c);
        c.persist = mode;
        invalidate(c);
    });
}

int inv_get_persistent(inv_ctx *ctx, int *mode) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(mode, INV_ERR_ARG, "null output");
        *mode = c.persist;
    });
}

int inv_get_pair(inv_ctx *ctx, int *tile_x, int *cols_per_lane) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(tile_x && cols_per_lane, INV_ERR_ARG, "null output");
        *tile_x = c.pair_tx;
        *cols_per_lane = c.pair_v;
    });
}

int inv_set_triple(inv_ctx *ctx, int tile_x) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(tile_x == 0 || tile_x == 8 || tile_x == 12 || tile_x == 16 || tile_x == 20,
                    INV_ERR_ARG, "three-step tile must be 8, 12, 16 or 20 rows (0: off)");
        INV_REQUIRE(!c.have_grid || tile_x == 0 || c.g.nz % 2 == 0, INV_ERR_UNSUPPORTED,
                    "the three-step kernel needs nz even");
        if (tile_x) {
            int dev = 0, optin = 0;
            INV_HIP(hipGetDevice(&dev));
            INV_HIP(hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev));
            INV_REQUIRE(trip_lds_bytes(c.have_grid ? c.g.psv != 0 : true, tile_x) <= optin,
                        INV_ERR_UNSUPPORTED,
                        "three-step tile needs more LDS per block than the device allows");
        }
        sync(c);
        c.trip_tx = tile_x;
        invalidate(c);
        fused_invalidate_lists(c);
    });
}

int inv_get_triple(inv_ctx *ctx, int *tile_x) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(tile_x, INV_ERR_ARG, "null output");
        *tile_x = c.trip_tx;
    });
}

int inv_set_walk(inv_ctx *ctx, int chunk_rows) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(chunk_rows == 0 || (chunk_rows >= 16 && chunk_rows <= 4096 &&
                                        chunk_rows % 8 == 0),
                    INV_ERR_ARG, "walk chunk must be a multiple of 8 rows in [16, 4096] (0: off)");
        INV_REQUIRE(!c.have_grid || chunk_rows == 0 || c.g.nz % 2 == 0, INV_ERR_UNSUPPORTED,
                    "the walking three-step kernel needs nz even");
        if (chunk_rows) {
            int dev = 0, optin = 0;
            INV_HIP(hipGetDevice(&dev));
            INV_HIP(hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev));
            INV_REQUIRE(walk_lds_bytes(true) <= optin, INV_ERR_UNSUPPORTED,
                        "the walking kernel needs more LDS per block than the device allows");
        }
        sync(c);
        c.walk_ch = chunk_rows;
        invalidate(c);
        fused_invalidate_lists(c);
    });
}

int inv_get_walk(inv_ctx *ctx, int *chunk_rows) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(chunk_rows, INV_ERR_ARG, "null output");
        *chunk_rows = c.walk_ch;
    });
}

// ---- autotuner -------------------------------------------------------------
// Candidate step kernels and tile shapes, timed on the context's own shape;
// the best is cached per (grid, slots, wave types, math mode) for the process.
namespace {
struct TuneChoice {
    int variant, tile, cols;
    int pair_tx = 0, pair_v = 2;  // two-step kernel tile (0 rows: off)
    int persist = 0;              // persistent two-step kernel
    int trip = 0;                 // three-step kernel tile rows (0: off)
    int walk = 0;                 // walking three-step kernel chunk rows (0: off)
};
std::mutex tune_mu;
std::map<std::string, TuneChoice> tune_cache;

void apply_choice(Ctx &c, const TuneChoice &t) {
    c.variant = t.variant;
    c.tile_x = t.tile;
    c.col_v = t.cols;
    c.pair_tx = t.pair_tx;
    c.pair_v = t.pair_v;
    c.persist = t.persist;
    c.trip_tx = t.trip;
    c.walk_ch = t.walk;
    invalidate(c);
    fused_invalidate_lists(c);
}
}  // namespace

int inv_autotune(inv_ctx *ctx, int first, int count, int steps, int *variant, int *tile_rows,
                 int *cols_per_lane) {
    return guard(ctx, [&](Ctx &c) {
        require_ready(c);
        check_slots(c, first, count);
        INV_REQUIRE(steps >= 1 && steps <= c.g.nt, INV_ERR_ARG, "steps must be in [1, nt]");
        // shots per launch scale the grid of tiles, not the tile's work: the
        // choice is kept per grid, wave types and math mode
        const std::string key = std::to_string(c.g.nx) + "x" + std::to_string(c.g.nz) + "/" +
                                std::to_string(c.g.sh) + std::to_string(c.g.psv) + "/" +
                                std::to_string(c.math) + (c.tune_trip ? "" : "/no-trip");
        TuneChoice best{-1, 0, 0};
        {
            std::lock_guard<std::mutex> lk(tune_mu);
            auto f = tune_cache.find(key);
            if (f != tune_cache.end()) best = f->second;
        }
        if (best.variant < 0) {
            std::vector<TuneChoice> cand;
            for (int v : {1, 2, 4})
                if (c.g.nz % v == 0)
                    for (int tx : {4, 8, 16, 32})
                        if (tile_ok(tx, v)) cand.push_back({INV_VARIANT_FUSED, tx, v});
            const int every = c.prof.every;
            c.prof.every = 0;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            double best_ms = 1e300;
            try {
                INV_HIP(hipEventCreate(&e0));
                INV_HIP(hipEventCreate(&e1));
                // what a gradient runs: forward steps storing snapshots, then
                // adjoint steps with their kernel accumulation (when those go
                // to per-slot accumulators, which the next adjoint clears;
                // a one-slot context adds into the kernels themselves)
                bool adj = c.g.adj && c.g.sfe > 0 && c.g.batch > 1;
                if (adj) {
                    // snapshot storage for the current format; without the
                    // memory for it, time forward steps only
                    try {
                        snap_storage(c);
                    } catch (const Error &e) {
                        if (e.code != INV_ERR_NOMEM) throw;
                        adj = false;
                    }
                }
                const int flags = adj ? INV_FWD_SNAPSHOTS : 0;
                const bool verbose = getenv("INV_AUTOTUNE_VERBOSE") != nullptr;
                auto pass = [&]() {
                    fused_forward(c, first, count, flags, 0, steps);
                    if (adj) fused_adjoint(c, first, count, 0, steps);
                };
                auto timed = [&](const TuneChoice &t, bool first_cand) {
                    apply_choice(c, t);
                    if (adj) fused_prepare(c, 2, first, count, 0);
                    // untimed passes (the first candidate's also bring the
                    // clocks up), then the best of three timed passes
                    for (int w = 0; w < (first_cand ? 4 : 1); w++) pass();
                    float best_t = 1e30f;
                    for (int rep = 0; rep < 3; rep++) {
                        INV_HIP(hipEventRecord(e0, c.stream));
                        pass();
                        INV_HIP(hipEventRecord(e1, c.stream));
                        INV_HIP(hipEventSynchronize(e1));
                        float ms = 0;
                        INV_HIP(hipEventElapsedTime(&ms, e0, e1));
                        best_t = std::min(best_t, ms);
                    }
                    if (verbose)
                        fprintf(stderr,
                                "inv_autotune %s: %d rows x %d cols per lane, two-step %d x %d%s, "
                                "three-step %d, walk %d: %.3f ms\n",
                                key.c_str(), t.tile, t.cols, t.pair_tx, t.pair_v,
                                t.persist ? " persistent" : "", t.trip, t.walk, best_t);
                    return best_t;
                };
                // one-step tiles, in order of preference (one column per
                // lane first): a later one must be 2 % faster, so that timing
                // noise does not pick between near-ties
                for (size_t k = 0; k < cand.size(); k++) {
                    const float ms = timed(cand[k], k == 0);
                    if (ms < 0.98 * best_ms) {
                        best_ms = ms;
                        best = cand[k];
                    }
                }
                // then two-step tiles over the chosen one-step tile (which
                // runs the steps that cannot be paired); a tile needing more
                // LDS than the device allows is skipped.  Tallest first: a
                // shorter tile recomputes more halo per owned row and must
                // be 2 % faster to replace it (20 and 16 rows are within a
                // few per cent on large P-SV grids, where the order decided)
                const TuneChoice single = best;
                for (int v : {2, 4})
                    for (int tx : {20, 16, 12, 8, 4}) {
                        if (c.g.nz % v || !pair_tile_ok(tx, v)) continue;
                        TuneChoice t = single;
                        t.pair_tx = tx;
                        t.pair_v = v;
                        float ms;
                        try {
                            ms = timed(t, false);
                        } catch (const Error &e) {
                            if (e.code != INV_ERR_UNSUPPORTED) throw;
                            continue;
                        }
                        if (ms < 0.98 * best_ms) {
                            best_ms = ms;
                            best = t;
                        }
                    }
                // then the persistent kernel (grids one tile wide, whose
                // tiles all fit the device at once), over the same one-step
                // tile
                const TuneChoice two = best;
                for (int tx : {8, 4}) {
                    if (c.g.nz % RES_V || c.g.nz > pair_strip(RES_V) || c.ckpt) break;
                    TuneChoice t = single;
                    t.pair_tx = tx;
                    t.pair_v = RES_V;
                    t.persist = 1;
                    try {
                        const float ms = timed(t, false);
                        sync(c);  // a wait that gave up is reported here
                        if (ms < 0.98 * best_ms) {
                            best_ms = ms;
                            best = t;
                        }
                    } catch (const Error &e) {
                        // not resident at once on this device, or a wait gave
                        // up: not a candidate
                        if (e.code != INV_ERR_UNSUPPORTED && e.code != INV_ERR_HIP) throw;
                    }
                }
                // then three steps per launch over the chosen two-step (or
                // one-step) tile, which keeps the steps that store
                // snapshots (P-SV tiles of 16 and 20 rows fit one per CU and
                // lose; SH ones fit two or three), against the best so far
                // (the persistent kernel included)
                if (c.tune_trip && c.g.nz % 2 == 0 && c.ckpt == 0) {
                    for (int tx : {12, 16, 20, 8}) {
                        TuneChoice t = two;
                        t.trip = tx;
                        float ms;
                        try {
                            ms = timed(t, false);
                        } catch (const Error &e) {
                            if (e.code != INV_ERR_UNSUPPORTED) throw;
                            continue;
                        }
                        if (ms < 0.98 * best_ms) {
                            best_ms = ms;
                            best = t;
                        }
                    }
                    // the walking kernel's chunks (each workgroup walks one
                    // chunk: long chunks leave CUs idle on small grids)
                    for (int ch : {128, 96, 160, 192, 64}) {
                        if (ch > c.g.nx) continue;
                        TuneChoice t = two;
                        t.walk = ch;
                        float ms;
                        try {
                            ms = timed(t, false);
                        } catch (const Error &e) {
                            if (e.code != INV_ERR_UNSUPPORTED) throw;
                            continue;
                        }
                        if (ms < 0.98 * best_ms) {
                            best_ms = ms;
                            best = t;
                        }
                    }
                }
            } catch (...) {
                c.prof.every = every;
                if (e0) (void)hipEventDestroy(e0);
                if (e1) (void)hipEventDestroy(e1);
                throw;
            }
            c.prof.every = every;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            std::lock_guard<std::mutex> lk(tune_mu);
            tune_cache[key] = best;
        }
        apply_choice(c, best);
        c.parity = 0;
        if (variant) *variant = best.variant;
        if (tile_rows) *tile_rows = best.tile;
        if (cols_per_lane) *cols_per_lane = best.cols;
    });
}

int inv_set_math(inv_ctx *ctx, int mode) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(mode == INV_MATH_EXACT || mode == INV_MATH_FAST, INV_ERR_ARG,
                    "unknown math mode");
        sync(c);
        c.math = mode;
        invalidate(c);
    });
}

int inv_set_graphs(inv_ctx *ctx, int on) {
    return guard(ctx, [&](Ctx &c) {
        sync(c);
        c.use_graphs = on != 0;
        invalidate(c);
    });
}

int inv_set_step_limit(inv_ctx *ctx, int nsteps) {
    return guard(ctx, [&](Ctx &c) {
        INV_REQUIRE(nsteps >= 0, INV_ERR_ARG, "nsteps must be >= 0");
        c.step_limit = nsteps;
    });
}

}  // extern "C"
