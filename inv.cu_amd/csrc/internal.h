// Internal declarations shared by the libinv_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/inv_hip.h"
#include "stencil.h"

namespace inv {

// Status helpers -------------------------------------------------------------
struct Error {
    int code;
    std::string msg;
};

#define INV_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw ::inv::Error{e_ == hipErrorOutOfMemory ? INV_ERR_NOMEM : INV_ERR_HIP, \
                               std::string(#call) + ": " + hipGetErrorString(e_)};     \
    } while (0)

#define INV_REQUIRE(cond, code, text)                 \
    do {                                              \
        if (!(cond)) throw ::inv::Error{(code), (text)}; \
    } while (0)

// Device buffer owned by a context.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void release();
    void ensure(size_t nbytes);  // (re)allocates when too small; contents undefined
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

// Sparse point set (sources or receivers) bucketed by propagation tile, so a
// fused stencil kernel finds the points of its own tile without scanning the
// whole list.  ptr has nsets*ntiles+1 entries; entries sorted by (set, tile,
// original order).
struct PointList {
    DevBuf ptr, cell, row;
    int nsets = 0, ntiles = 0, npoints = 0;
};

// Kernel launch geometry of the propagation tiles.
constexpr int TILE_Z = 64;  // one wavefront along z (the contiguous axis)
constexpr int TILE_X = 4;   // wavefronts per block along x

// Sampled kernel timing with HIP events on the context stream.
enum ProfKind {
    PROF_VELOCITY = 0, PROF_STRESS = 1, PROF_INTERACT = 2, PROF_STEP = 3,
    PROF_FWD_RUN = 4, PROF_ADJ_RUN = 5, PROF_PAIR = 6, PROF_RES = 7, PROF_TRIP = 8,
    PROF_NKIND = 9
};
struct Prof {
    int every = 0;  // sample every n-th time step (0 = off)
    struct Rec {
        int kind;
        hipEvent_t a, b;
        int units;  // time steps covered by the interval
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    long long count[PROF_NKIND] = {};
    double ms[PROF_NKIND] = {};
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // grid
    inv_grid_desc g{};
    bool have_grid = false, have_model = false, have_src = false, have_rec = false;
    long N = 0;
    int nsfe = 0;
    int ncomp = 0;  // trace components: y (sh) then x, z (psv)
    int nsnapf = 0;  // snapshot fields per slot
    int step_limit = 0;
    // forward snapshots for the adjoint: storage format, checkpoint interval
    // (0: every snapshot of the run is kept) and the slots held per batch slot
    // (nsfe, or ckpt / sfe with checkpoints)
    int snap_format = INV_SNAP_F32;
    int ckpt = 0;
    int nslots = 0;

    // model and absorbing profile (1-D factor tables: ab = ((ax*bz)*tz))
    DevBuf lam, mu, rho, abx, abzb, abzt;
    // kernels
    DevBuf k_lambda, k_mu, k_rho;
    DevBuf kslot;  // per-slot kernel accumulators [batch][3][N] (batched adjoint)

    // per-slot wavefields [batch][N].  The fused step kernel reads v and s at
    // step n from one buffer set and writes step n+1 to the other (a
    // neighbouring tile still reads the old values of its halo); `parity`
    // names the set holding the current state.  u is updated in place.
    DevBuf vx, vz, ux, uz, sxx, szz, sxz, vy, uy, sxy, szy;
    DevBuf vx_b, vz_b, sxx_b, szz_b, sxz_b, vy_b, sxy_b, szy_b;
    int parity = 0;
    int variant = 0;   // 0 fused step kernel, 1 velocity/stress kernel pair
    int tile_x = 8;    // rows per fused tile (4, 8, 16 or 32)
    int col_v = 1;     // columns per lane of the fused column kernel (1, 2 or 4)
    // two-step kernel (fused variant): tile rows (0 = off) and columns per lane
    int pair_tx = 0;
    int pair_v = 2;
    // three-step launches (fdm_trip.hip): tile rows 8, 12, 16 or 20, 0 off;
    // tune_trip: the autotuner times them too (the request server's lanes
    // turn it off)
    int trip_tx = 0;
    // rows per chunk of the walking three-step kernel (0: off); when set,
    // three-step launches walk chunks of this many rows instead of tiles
    int walk_ch = 0;
    bool tune_trip = true;
    // persistent two-step kernel (fdm_resident.hip): 0 off, 1 on, 2 on with
    // every wait given up at once (tests the error path); its halo exchange
    // buffer and control words {launch epoch, error}
    int persist = 0;
    DevBuf xbuf, res_ctl;
    int math = INV_MATH_FAST;
    DevBuf bdt;        // dt / rho, rebuilt at the start of each fast-mode run
    // per-slot models: [batch][3] device pointers {rho (exact) or dt/rho
    // (fast), lam, mu} owned by the caller (the request server's model
    // store); when set, lam/mu/rho/bdt above are not read by the fused
    // kernels.  inv_grid_set clears it.
    DevBuf slot_model;
    bool slot_models = false;
    // keep snapshot and per-slot kernel buffers across inv_grid_set (they only
    // grow), for contexts that run one batch after another of similar size
    bool retain = false;
    DevBuf snap;   // [batch][nslots][nsnapf][N] of FP32, FP16, BF16 or I16
    DevBuf sexp;   // 16-bit formats: exponent tables [nslots][snap_tab_bytes]
    DevBuf ckbuf;  // state checkpoints [nt/ckpt - 1][state fields][batch][N]
    DevBuf rstate; // wavefields of the checkpoint recomputation
    DevBuf out;    // [batch][ncomp][nrec][nt]
    DevBuf adstf;  // [batch][ncomp][nrec][nt]
    DevBuf obs;    // [nsrc][ncomp][nrec][nt]
    DevBuf chi;    // per-trace misfit [batch][nrec][ncomp] (double)
    DevBuf taper;  // [nt] waveform taper
    DevBuf red;    // reduction scratch
    DevBuf scal;   // a few device scalars (FP64 results)

    // acquisition
    int nsrc = 0, nrec = 0;
    std::vector<int> src_ix, src_iz, rec_ix, rec_iz;
    DevBuf stf;  // [3][nsrc][nt]  x, y, z
    // source mechanism per source: 0 point force (addSTF), 1 explosive (stf_y
    // added to sxx and szz after the stress update)
    std::vector<int> src_type;
    bool any_explosive() const {
        for (int t : src_type)
            if (t == 1) return true;
        return false;
    }
    PointList rec_list;  // receivers, one set shared by all slots
    // sources of a forward, one set per slot, kept per shot range (and tile
    // shape) so that a captured graph of one batch never reads another's
    std::map<long long, PointList> src_lists;
    // explosive sources of a forward, one set per slot, bucketed by the tile
    // that owns their cell (the fused kernels add them to the stress)
    std::map<long long, PointList> src_slists;
    // the same sets bucketed by fused tile: injection points by the tile's
    // velocity region (tile plus halo), recording points by the tile itself
    PointList f_rec_v, f_rec_t;
    int f_rec_tx = -1;
    // the same for the two-step kernel's tile geometry
    PointList p_rec_v, p_rec_t;
    int p_rec_key = -1;
    // and for the three-step kernel's
    PointList t_rec_v, t_rec_t;
    int t_rec_key = -1;

    // captured time loops: one executable graph per (run kind, shots, flags,
    // kernel configuration); `gen` changes whenever buffers or settings they
    // bake in change
    struct GraphEntry {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        long gen = -1;
        int parity = 0;  // state-set parity the captured loop ends on
    };
    std::map<std::string, GraphEntry> graphs;
    long gen = 0;
    bool use_graphs = true;

    // envelope misfit: FFT twiddles and radix plan for length nt, or (env_m >
    // 0) for the power-of-two length env_m of a Bluestein transform, whose
    // chirp and filter spectra live in env_blue; env_global: the transforms
    // run through per-trace buffers in env_gbuf instead of LDS (too long)
    DevBuf env_w, env_fac, env_scratch, env_blue, env_gbuf;
    int env_n = -1, env_nfac = 0, env_m = 0, env_global = 0;

    // filter
    DevBuf gauss_w, gauss_tmp, gauss_sx, gauss_sz;
    int gauss_sigma = -1, gauss_nx = -1, gauss_nz = -1, gauss_m = 0;

    std::vector<void *> user_allocs;
    Prof prof;

    // an incremental run (inv_*_begin / inv_run_advance): 0 none, 1 forward,
    // 2 adjoint
    int run_kind = 0, run_first = 0, run_count = 0, run_flags = 0, run_it = 0;

    int slot_offset_traces() const { return ncomp * nrec * g.nt; }
};

// capi.hip ------------------------------------------------------------------
// every device buffer a context owns (released by inv_ctx_destroy)
std::vector<DevBuf *> ctx_buffers(Ctx &c);
// device bytes a context holds now
size_t ctx_bytes_held(Ctx &c);
// upper bound of the device bytes a context holds after a forward (+ adjoint)
// run of desc with nsrc sources, nrec receivers and the given snapshot storage
size_t ctx_bytes_estimate(const inv_grid_desc &d, int nsrc, int nrec, int snap_format, int ckpt);

// fdm.hip -------------------------------------------------------------------
void build_point_list(Ctx &c, PointList &pl, int nsets, const std::vector<std::vector<int>> &cells,
                      const std::vector<std::vector<int>> &rows);
// propagate steps [it0, it1) of a run; it0 == 0 starts from rest
void fdm_forward(Ctx &c, int first, int count, int flags, int it0, int it1);
void fdm_adjoint(Ctx &c, int first, int count, int it0, int it1);
void fdm_zero_state(Ctx &c, int count);
// one split-kernel half-step on the current state set (inv_step_velocity:
// which 0, time step it of shots first..; inv_step_stress: which 1)
void fdm_half_step(Ctx &c, int which, int first, int count, int it);
void fdm_prepare_model(Ctx &c);
void fdm_kslot_reduce(Ctx &c, int count);
// dt / rho of the fast math mode into bdt[n] on a stream
void launch_bdt(float *bdt, const float *rho, float dt, long n, hipStream_t stream);
// out[3][N] = sum over slots s0..s0+n-1 (in slot order, from 0) of the
// per-slot kernel accumulators kslot[batch][3][N]
void kslot_sum(Ctx &c, float *out, int s0, int n);
Geo make_geo(const Ctx &c);
void fill_dev(Ctx &c, void *p, long n, float v = 0.0f);
// interaction kernels for adjoint step it (it % sfe == 0) with the adjoint
// velocities vx/vz/vy of the current state set, reading snapshot buffer slot
// `slot`
void fdm_interact(Ctx &c, int count, int it, int slot, const float *vx, const float *vz,
                  const float *vy);
// sizes the snapshot, max-abs, checkpoint and recomputation buffers for the
// current format and checkpoint interval (allocation happens before a run,
// never inside a captured one)
void snap_storage(Ctx &c);
void fused_forward(Ctx &c, int first, int count, int flags, int it0, int it1);
void fused_adjoint(Ctx &c, int first, int count, int it0, int it1);
void fused_invalidate_lists(Ctx &c);
// buffers, point lists and (for runs with snapshots) snapshot storage of a
// forward (kind 1, with its flags) or adjoint (kind 2)
void fused_prepare(Ctx &c, int kind, int first, int count, int flags);
void fdm_prepare(Ctx &c, int kind, int first, int count);
// the per-slot source list of shots first..first+count-1 for a tile shape
// (key 0: the kernel-pair tiles); built on first use
PointList &source_list(Ctx &c, int first, int count, int key);
// the explosive sources of the same shots, bucketed by owning fused tile
PointList &stress_source_list(Ctx &c, int first, int count, int key);
void release_list(PointList &pl);
// device pointer of the current buffer of a ping-pong field
inline float *cur_field(Ctx &c, DevBuf &a, DevBuf &b) {
    return (c.parity & 1) ? b.as<float>() : a.as<float>();
}
bool prof_begin(Ctx &c, int it, hipEvent_t *a);
void prof_end(Ctx &c, int kind, hipEvent_t a, int units = 1);
bool prof_begin_any(Ctx &c, hipEvent_t *a);
void prof_resolve(Ctx &c);
void prof_release(Ctx &c);

// misfit.hip ----------------------------------------------------------------
void misfit_waveform(Ctx &c, int count, int first);
// Waveform misfit of slots 0..count-1 whose traces are shorter than the run:
// slot b holds nt_b = slot_nt[b] samples at the start of each [nt] row of
// out and obs, its taper is row b of taper_tab [count][nt], and its adjoint
// source is written time-reversed into the last nt_b samples of each adstf
// row, the first nt - nt_b samples zeroed.  Device arrays.
void misfit_waveform_slots(Ctx &c, int count, const int *slot_nt, const float *taper_tab);
// waveform taper of misfit/waveform.h:9-27 for nt samples of step dt
void waveform_taper(int nt, float dt, float *tw);
// k_waveform over ntr traces [ntr][nt] on device pointers: adjoint sources
// into adstf, the misfit of trace r into chi[r] (device)
void launch_waveform_traces(const float *syn, const float *obs, float *adstf, const float *tw,
                            double *chi, int nt, int ntr, float dt, hipStream_t stream);
void misfit_envelope(Ctx &c, int count, int first);
void envelope_trace(Ctx &c, const float *syn, const float *obs, float *adstf, double *chi_dev);
double misfit_trace(Ctx &c, int type, const float *syn, const float *obs, float *adstf);

// noalloc.hip ---------------------------------------------------------------
// resolves, once per device (with that device current), the device address of
// the no-allocation entry point's scratch array
void noalloc_init(int device);

// linalg.hip ----------------------------------------------------------------
void gauss_filter(Ctx &c, float *field, int nx, int nz, int sigma);
void vec_axpby(Ctx &c, float *out, float ka, const float *a, float kb, const float *b, size_t n);
void vec_scale(Ctx &c, float *out, float ka, const float *a, size_t n);
void vec_mul(Ctx &c, float *out, const float *a, const float *b, size_t n);
void vec_copy(Ctx &c, float *out, const float *a, size_t n);
void vec_fill(Ctx &c, float *out, float v, size_t n);
double vec_dot(Ctx &c, const float *a, const float *b, size_t n);
float vec_amax(Ctx &c, const float *a, size_t n);
void sum_f64(Ctx &c, const double *x, size_t n, double *dev_out);

// mesh2grid on the device (model.hip): the nearest of npt host points for
// every node into c.lam / c.mu / c.rho (values as given, before vps2lm)
void model_nearest(Ctx &c, int npt, const float *px, const float *pz, const float *p1,
                   const float *p2, const float *prho, float dmax);

// Exponent tables of 16-bit snapshots (stencil.h).  A row segment is the
// owned columns of one tile of the kernel that writes the slot: 58 (k_col),
// 120 / 248 (k_colv, 2 / 4 columns per lane), 116 / 240 (k_pair, 2 / 4).
constexpr int SNAP_SEG_WIDTHS[5] = {58, 116, 120, 240, 248};
// segments per row (the narrowest segment sets the count)
inline int snap_nseg(int nz) { return (nz + 57) / 58 + 1; }
// block-offset bytes per row: enough for every segment width
inline int snap_nbb(int nz) {
    int m = 1;
    for (int w : SNAP_SEG_WIDTHS) m = std::max(m, (nz + w - 1) / w * snap_seg_bytes(w));
    return m;
}
// byte offset of the block-offset section in a slot's table
inline long snap_blk_off(int batch, int nsnapf, int nx, int nz) {
    return SNAPTAB_HDR + (long)batch * nsnapf * nx * snap_nseg(nz);
}
inline size_t snap_tab_bytes(int batch, int nsnapf, int nx, int nz) {
    const size_t b = (size_t)snap_blk_off(batch, nsnapf, nx, nz) +
                     (size_t)batch * nsnapf * nx * snap_nbb(nz);
    return (b + 255) / 256 * 256;
}

}  // namespace inv

// The opaque context of the C ABI is the internal one.
struct inv_ctx : public inv::Ctx {};
