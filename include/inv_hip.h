/*
 * inv_hip.h — C ABI of libinv_hip.so, the MI355X (gfx950) implementation of
 * inv.cu's forward+adjoint gradient path.
 *
 * The reference exposes this path as C++ plugin classes (Solver / Misfit /
 * Filter / Optimizer / Source, lib/index.h:32-65).  This header is the flat,
 * FFI-friendly boundary underneath them: plain pointers, sizes and status
 * codes, no C++ or torch types.  The C++ plugin classes in
 * inv.cu_amd/host/ (same names and ids as the reference) sit on top of it.
 *
 * Conventions
 *   - every function returns int: INV_OK (0) or a negative inv_status;
 *     inv_last_error(ctx) gives the message (HIP error text preserved).
 *   - a context owns one device, one HIP stream and all device memory it
 *     allocates.  Functions are thread-compatible per context, not
 *     thread-safe; different contexts may run concurrently.
 *   - fields are float32 [nx][nz], z fastest (k = i*nz + j), the reference's
 *     Dim layout (lib/device.h:7-65).  Traces are [nrec][nt] per
 *     component; components are ordered y (if sh) then x, z (if psv).
 *   - "host" pointers are caller-owned host memory copied in or out;
 *     "dev" pointers are device addresses (from inv_field_ptr / inv_alloc).
 *   - calls that return data to the host synchronise the context stream;
 *     the others are asynchronous on it.
 */
#ifndef INV_HIP_H
#define INV_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct inv_ctx inv_ctx;

enum inv_status {
    INV_OK = 0,
    INV_ERR_ARG = -1,         /* invalid argument or size */
    INV_ERR_HIP = -2,         /* HIP runtime error (text in inv_last_error) */
    INV_ERR_STATE = -3,       /* call out of order (e.g. forward before model) */
    INV_ERR_NOMEM = -4,       /* device allocation failed */
    INV_ERR_UNSUPPORTED = -5  /* feature/combination not supported */
};

/* Field ids for inv_field_ptr. */
enum inv_field {
    INV_LAMBDA = 0, INV_MU = 1, INV_RHO = 2,
    INV_K_LAMBDA = 3, INV_K_MU = 4, INV_K_RHO = 5,
    INV_OUT = 6,     /* receiver traces of batch slot 0: [ncomp][nrec][nt] */
    INV_ADSTF = 7,   /* adjoint sources of batch slot 0: [ncomp][nrec][nt] */
    INV_OBS = 8      /* resident observed data: [nsrc][ncomp][nrec][nt] */
};

/* Snapshot storage formats (arithmetic is always FP32). */
enum inv_snap_format {
    INV_SNAP_F32 = 0,
    INV_SNAP_F16 = 1,   /* FP16, power-of-two scale per row segment */
    INV_SNAP_BF16 = 2,  /* BF16, the same scaling */
    INV_SNAP_I16 = 3    /* 16-bit fixed point relative to the row segment's max-abs */
};

/* Forward flags. */
enum inv_forward_flags {
    INV_FWD_SNAPSHOTS = 1  /* store forward snapshots for a later adjoint */
};

/* Solver description: the FdmSolver parameters read in Solver::init
 * (solver/solver.h:158-224) and FdmSolver::init (solver/fdm.h:350-462). */
typedef struct inv_grid_desc {
    int nx, nz;            /* grid nodes */
    float dx, dz, dt;      /* spacing (m) and time step (s) */
    int nt;                /* time steps */
    int sh, psv;           /* wave types (config keys sh, psv) */
    int adj;               /* 1 in modes 0/2: record u, keep snapshots */
    int trace_type;        /* 0 velocity, 1 displacement traces */
    int sfe;               /* snapshot interval (key sfe); nt % sfe must be 0 when adj */
    int abs_width;         /* absorbing strip width in cells */
    float abs_alpha;       /* damping parameter */
    int abs_left, abs_right, abs_top, abs_bottom;
    int inv_lambda, inv_mu, inv_rho;  /* kernels accumulated by the adjoint */
    int batch;             /* shot slots per launch (1 = plain one-shot path) */
} inv_grid_desc;

/* ---- context ---------------------------------------------------------- */
int inv_device_count(int *count);
int inv_ctx_create(int device, inv_ctx **out);
int inv_ctx_destroy(inv_ctx *ctx);
const char *inv_last_error(const inv_ctx *ctx);
int inv_sync(inv_ctx *ctx);
/* free/total device memory in bytes (lib/query.h:3-22) */
int inv_mem_info(inv_ctx *ctx, size_t *free_bytes, size_t *total_bytes);

/* ---- host-side setup helpers (no device work) --------------------------- */
/* Grid derivation of FdmSolver::init (solver/fdm.h:356-389): nx from the
 * model point count and the aspect ratio, nz = npt / nx, dx = xmax/(nx-1). */
int inv_grid_dims(int npt, float xmax, float zmax, int *nx, int *nz, float *dx, float *dz);
/* calcIdx (solver/fdm.h:44-47): nearest grid index of a coordinate. */
int inv_calc_idx(float coord, float length, int n);
/* RickerSource::create + Source::init (source/ricker.h:5-21,
 * source/source.h:17-37): stf = -t exp(-(pi f0 t)^2) peak-normalised, then
 * stf_x = amp stf cos(angle), stf_y = amp stf, stf_z = amp stf sin(angle). */
int inv_source_ricker(int nt, float dt, float f0, float t0, float angle, float amp,
                      float *stf_x, float *stf_y, float *stf_z);

/* ---- grid, model, acquisition ------------------------------------------ */
/* FdmSolver::init (solver/fdm.h:350-462): allocates every field for
 * desc->batch slots and builds the absorbing-boundary profile
 * (initAbsbound, solver/fdm.h:48-74). */
int inv_grid_set(inv_ctx *ctx, const inv_grid_desc *desc);
/* Solver::importModel + mesh2grid + vps2lm (solver/solver.h:225-252,
 * solver/fdm.h:4-36,463-476) for on-grid input: host [N] arrays. */
int inv_model_upload_lmr(inv_ctx *ctx, const float *lambda, const float *mu, const float *rho);
int inv_model_download_lmr(inv_ctx *ctx, float *lambda, float *mu, float *rho);
/* mesh2grid (solver/fdm.h:4-24): nearest model point, strict <, first wins;
 * then vps2lm (fdm.h:25-36) when model_type == 0.  Host point arrays [npt]. */
int inv_model_from_points(inv_ctx *ctx, int npt, const float *px, const float *pz,
                          const float *p1, const float *p2, const float *prho, float xmax,
                          float zmax, int model_type);
/* lm2vps (solver/fdm.h:37-43) into host arrays (export in vp/vs form). */
int inv_model_to_vps(inv_ctx *ctx, float *vp, float *vs, float *rho);
/* Source positions (grid indices from calcIdx, fdm.h:44-47,454-457) and
 * wavelets stf_{x,y,z} [nsrc][nt] (Source::init, source/source.h:17-37). */
int inv_sources_set(inv_ctx *ctx, int nsrc, const int *ix, const int *iz, const float *stf_x,
                    const float *stf_y, const float *stf_z);
/* Source mechanism per source, after inv_sources_set (which resets all to 0):
 * 0 the reference's point force (addSTF, solver/fdm.h:163-180: stf_x and
 * stf_z added to the P-SV stress divergence, stf_y to the SH one); 1 an
 * explosive source with no counterpart in the reference (source id 1 of
 * sources.dat here): after the stress update of step it, stf_y[it] (= amp
 * times the wavelet) is added to sxx and to szz at the source cell; P-SV
 * only, fused column kernel only.  type == NULL resets to 0. */
int inv_sources_type(inv_ctx *ctx, const int *type);
int inv_receivers_set(inv_ctx *ctx, int nrec, const int *ix, const int *iz);

/* ---- propagation -------------------------------------------------------- */
/* FdmSolver::runForward (solver/fdm.h:515-570) for shots first..first+count-1
 * in batch slots 0..count-1 (count <= batch).  first = -1, count = 1 fires
 * every source in slot 0 (runForward(-1), modes 1 and 4 of inv.cu:46-54).
 * Traces land in the slots' OUT buffers. */
int inv_forward(inv_ctx *ctx, int first, int count, int flags);
/* FdmSolver::runAdjoint (solver/fdm.h:571-612) for the same slots, driven
 * by their ADSTF buffers; kernels accumulate into K_LAMBDA/K_MU/K_RHO. */
int inv_adjoint(inv_ctx *ctx, int first, int count);
/* Incremental runs: begin a forward (or adjoint) of the given slots, then
 * advance it nsteps at a time; *position returns the steps done.  Between
 * advances the wavefields can be read (inv_state_get), which is how the
 * wavefield exports every wfe / wae steps of exportSnapshot
 * (solver/fdm.h:312-347) are served. */
int inv_forward_begin(inv_ctx *ctx, int first, int count, int flags);
int inv_adjoint_begin(inv_ctx *ctx, int first, int count);
int inv_run_advance(inv_ctx *ctx, int nsteps, int *position);
/* initKernels (solver/fdm.h:495-499) */
int inv_kernels_zero(inv_ctx *ctx);
int inv_kernels_get(inv_ctx *ctx, float *k_lambda, float *k_mu, float *k_rho);
int inv_kernels_set(inv_ctx *ctx, const float *k_lambda, const float *k_mu, const float *k_rho);

/* Host copies of one slot's traces / adjoint sources, [ncomp][nrec][nt]. */
int inv_traces_get(inv_ctx *ctx, int slot, float *host);
int inv_adstf_set(inv_ctx *ctx, int slot, const float *host);
int inv_adstf_get(inv_ctx *ctx, int slot, float *host);
/* Observed data resident on the device (Misfit::init, misfit/misfit.h:91-151):
 * set from host [ncomp][nrec][nt] for one shot, or copy slots' traces in. */
int inv_obs_set(inv_ctx *ctx, int isrc, const float *host);
int inv_obs_get(inv_ctx *ctx, int isrc, float *host);
int inv_obs_from_traces(inv_ctx *ctx, int first, int count);

/* ---- misfit ------------------------------------------------------------- */
/* Misfit of slots 0..count-1 (shots first..) against resident obs:
 * type 0 waveform (misfit/waveform.h:4-50), 1 envelope (misfit/envelope.h;
 * nt <= 2^20, INV_ERR_UNSUPPORTED beyond).
 * Writes the slots' ADSTF and, if chi != NULL, per-trace misfits
 * chi[count][nrec][ncomp] (receiver-major as Misfit::run sums them,
 * misfit/misfit.h:31-40).  *total (FP64) = their sum. */
int inv_misfit(inv_ctx *ctx, int first, int count, int type, double *chi, double *total);

/* One trace on device pointers (the per-trace Misfit::run(syn, obs, adstf) of
 * misfit/misfit.h:11): writes adstf [nt], returns the trace misfit. */
int inv_misfit_trace(inv_ctx *ctx, int type, const float *syn, const float *obs, float *adstf,
                     double *chi);

/* The waveform misfit of ntr traces [ntr][nt] on device pointers (nt of the
 * grid), with no device allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Writes adstf [ntr][nt] (the
 * time-reversed adjoint sources of misfit/waveform.h:4-50, as
 * inv_misfit_trace) and, if chi != NULL, the per-trace misfits chi[ntr]
 * (host); *total (host) = their FP64 sum.  ntr <= 65536: the per-trace
 * misfits live in a 512 KiB device array of the library (one per GPU, held
 * until the process ends), shared by every call on that GPU, so calls that
 * overlap on one GPU race.  Needs inv_grid_set (the taper). */
int inv_misfit_traces_noalloc(inv_ctx *ctx, const float *syn, const float *obs, float *adstf,
                              int ntr, double *chi, double *total);

/* ---- whole gradient: Misfit::run(kernel) (misfit/misfit.h:15-51) ---------
 * Forward every shot in batches of `batch`, misfit against resident obs,
 * adjoint (if kernel), then the Gaussian filter (sigma > 0) on the three
 * kernels.  *misfit = summed misfit (FP64). */
int inv_gradient(inv_ctx *ctx, int misfit_type, int sigma, int kernel, double *misfit);

/* ---- filter ------------------------------------------------------------- */
/* GaussianFilter::apply (filter/gaussian.h:45-56), one-sided weights, in
 * place on a device field [nx][nz]. */
int inv_filter_gauss1s(inv_ctx *ctx, float *dev_field, int nx, int nz, int sigma);

/* ---- vector ops on device arrays (lib/device.h:157-192, the p_* helpers of
 * optimizer/optimizer.h:49-107) --------------------------------------- */
int inv_vec_axpby(inv_ctx *ctx, float *c, float ka, const float *a, float kb, const float *b,
                  size_t n);                                    /* c = ka*a + kb*b */
int inv_vec_scale(inv_ctx *ctx, float *c, float ka, const float *a, size_t n); /* c = ka*a */
int inv_vec_mul(inv_ctx *ctx, float *c, const float *a, const float *b, size_t n); /* c = a*b */
int inv_vec_copy(inv_ctx *ctx, float *dst, const float *src, size_t n);
int inv_vec_fill(inv_ctx *ctx, float *dst, float value, size_t n);
int inv_vec_dot_f64(inv_ctx *ctx, const float *a, const float *b, size_t n, double *out);
int inv_vec_amax(inv_ctx *ctx, const float *a, size_t n, float *out);  /* max |a| */

/* ---- device memory ------------------------------------------------------ */
int inv_field_ptr(inv_ctx *ctx, int field, float **out);
int inv_alloc(inv_ctx *ctx, size_t bytes, void **out);
int inv_free(inv_ctx *ctx, void *dev);
int inv_memcpy_h2d(inv_ctx *ctx, void *dev, const void *host, size_t bytes);
int inv_memcpy_d2h(inv_ctx *ctx, void *host, const void *dev, size_t bytes);

/* ---- test hooks ---------------------------------------------------------
 * Wavefields of one slot after the last forward/adjoint:
 * SH {vy, uy, sxy, szy} then P-SV {vx, vz, ux, uz, sxx, szz, sxz}, each [N].
 * Snapshots of one slot: [slots][nf][N], SH {uy, vy} then P-SV {ux, uz, vx, vz},
 * decoded to float; slots = nt/sfe (or ckpt/sfe with checkpoints, holding the
 * segment last used). */
int inv_state_get(inv_ctx *ctx, int slot, float *host);
int inv_snapshots_get(inv_ctx *ctx, int slot, float *host);
/* Writes one slot's wavefields (host layout of inv_state_get) into the state
 * that the next half-step or inv_state_get sees. */
int inv_state_set(inv_ctx *ctx, int slot, const float *host);
/* The reference's separable phases of a time step (divS and divV,
 * solver/fdm.h:273-294), on the current wavefields of slots 0..count-1, in
 * place:
 *   inv_step_velocity: stress divergence (divSY / divSXZ, fdm.h:76-102), the
 *     point forces of time step `it` of shots first.. (addSTF, fdm.h:163-180;
 *     first = -1 with count 1 fires every source) and the velocity and
 *     displacement update (updateVY / updateVXZ, fdm.h:150-161);
 *   inv_step_stress: velocity gradient and stress update (divVY / divVXZ,
 *     updateSY / updateSXZ, fdm.h:103-149).
 * One time step of runForward is inv_step_velocity(it) then inv_step_stress.
 * No receiver recording or snapshots; explosive sources are refused
 * (INV_ERR_UNSUPPORTED).  Test hooks: the runs above are faster. */
int inv_step_velocity(inv_ctx *ctx, int first, int count, int it);
int inv_step_stress(inv_ctx *ctx, int count);
/* ---- measurement -------------------------------------------------------
 * Sampled kernel timing: with every > 0, each propagation kernel launched
 * directly on a time step it with it % every == 0 is bracketed by HIP events
 * on the context stream, and so is every replay of a captured run (counted
 * in time steps); inv_prof_read returns the count and the summed duration.  No counterpart in the reference (which times only
 * the whole run, optimizer/optimizer.h:432,456). */
enum inv_prof_kernel {
    INV_PROF_VELOCITY = 0,  /* velocity half-step (kernel-pair variant) */
    INV_PROF_STRESS = 1,    /* stress half-step (kernel-pair variant) */
    INV_PROF_INTERACT = 2,  /* adjoint kernel accumulation */
    INV_PROF_STEP = 3,      /* whole time step (fused variant) */
    INV_PROF_FWD_RUN = 4,   /* replayed forward graphs (count = time steps) */
    INV_PROF_ADJ_RUN = 5,   /* replayed adjoint graphs (count = time steps) */
    INV_PROF_PAIR = 6,      /* two-step launches (count = launches) */
    INV_PROF_RESIDENT = 7,  /* persistent launches (count = time steps) */
    INV_PROF_TRIPLE = 8     /* three-step launches (count = launches) */
};
int inv_prof_enable(inv_ctx *ctx, int every);
int inv_prof_reset(inv_ctx *ctx);
int inv_prof_read(inv_ctx *ctx, int kernel, long long *count, double *total_ms);

/* Propagation kernel variant: INV_VARIANT_FUSED (default) runs one kernel per
 * time step on tiles of 58 (z) x tile_x (x) cells, one lane per column;
 * INV_VARIANT_FUSED_LINEAR the same step on 64 x tile_x tiles with a linear
 * cell index; tile_x in {8, 16, 32} (0 keeps the current value).
 * INV_VARIANT_SPLIT runs a velocity and a stress kernel per step.  Results
 * are identical between variants. */
enum inv_variant { INV_VARIANT_FUSED = 0, INV_VARIANT_SPLIT = 1, INV_VARIANT_FUSED_LINEAR = 2 };
int inv_set_variant(inv_ctx *ctx, int variant, int tile_x);
/* The fused column kernel with tiles of tile_x rows and cols_per_lane
 * adjacent columns per lane (float2 / float4 loads and stores): 1 column per
 * lane with 8, 16 or 32 rows (58 columns per tile), 2 with 4, 8 or 16 rows
 * (120 columns), 4 with 4 or 8 rows (248 columns); 2 and 4 need nz divisible
 * by them.  Results are identical for every tile.  Selects variant 0. */
int inv_set_tile(inv_ctx *ctx, int tile_x, int cols_per_lane);
/* Two time steps per launch (fused variant): the stress and velocity of the
 * tile stay in LDS between the steps, so each field crosses HBM once per two
 * steps.  tile_x rows x cols_per_lane columns per lane: 4, 8, 12, 16 or 20
 * rows x 2 (116 columns per tile) or 4 or 8 x 4 (240 columns); tile_x = 0 turns it off
 * (default).  Runs use it for every pair of steps that has no checkpoint
 * between them and no kernel accumulation after the first; the other steps
 * use the one-step kernel.  Results are identical with and without it.
 * Needs nz divisible by cols_per_lane.  No counterpart in the reference (one
 * kernel launch per stencil half-step, solver/fdm.h:534-551). */
int inv_set_pair(inv_ctx *ctx, int tile_x, int cols_per_lane);
/* The two-step tile in use (tile_x 0: off), as set by inv_set_pair or chosen
 * by inv_autotune. */
int inv_get_pair(inv_ctx *ctx, int *tile_x, int *cols_per_lane);
/* Three time steps per launch (tile_x 8, 12, 16 or 20 rows x 108 columns,
 * two per lane; 0 off, the default): forward runs use it for every three steps none
 * of which stores a snapshot, adjoint runs for every three steps with no
 * kernel accumulation after the first two, both only without checkpoints;
 * the other steps take the two-step or one-step launches they take without
 * it.  Results are identical with and without it.  Needs nz even.  No
 * counterpart in the reference (one launch per stencil half-step,
 * solver/fdm.h:534-551). */
int inv_set_triple(inv_ctx *ctx, int tile_x);
int inv_get_triple(inv_ctx *ctx, int *tile_x);
/* The same three-step launches by a walking kernel (chunk_rows a multiple of
 * 8 in [16, 4096]; 0 off, the default): a workgroup walks down chunk_rows
 * rows of a 108-column strip eight rows at a time, keeping every step's rows
 * in LDS, so a row is updated once per step instead of once per tile it is
 * a halo row of.  When set it replaces the tiles of inv_set_triple (and
 * turns three-step launches on by itself).  Results are identical.  Needs
 * nz even.  No counterpart in the reference (solver/fdm.h:534-551). */
int inv_set_walk(inv_ctx *ctx, int chunk_rows);
int inv_get_walk(inv_ctx *ctx, int *chunk_rows);
/* Persistent two-step kernel (mode 1; 0 off, the default): one launch runs
 * every pair of steps of a forward, or of an adjoint between two kernel
 * accumulations, on tiles of 4 or 8 rows spanning all nz columns, one
 * workgroup per tile.  A tile's wavefields stay in LDS and registers between
 * steps; only halo rows cross memory, exchanged with the tiles the halo
 * reaches through tagged words (no grid-wide barrier).  For launch-bound
 * grids: needs the two-step tile 8 x 4 or 4 x 4 (inv_set_pair(ctx, 8, 4)),
 * nz <= 240 and divisible by 4, no checkpoints, and every tile of a launch
 * resident on the device at once (shots x ceil(nx/rows) workgroups;
 * INV_ERR_UNSUPPORTED at run time otherwise).  Results are identical to those of inv_set_pair alone.  A wait
 * for a neighbour gives up after a bounded time instead of hanging; the run
 * then ends with INV_ERR_HIP at the next call that synchronises.  Mode 2 (a
 * test hook) gives up every wait at once.  No counterpart in the reference
 * (one kernel launch per stencil half-step, solver/fdm.h:534-551). */
int inv_set_persistent(inv_ctx *ctx, int mode);
int inv_get_persistent(inv_ctx *ctx, int *mode);
/* Autotuner: times `steps` forward steps of shots first..first+count-1 (with
 * snapshots and as many adjoint steps when the context keeps per-slot kernel
 * accumulators, batch > 1) with each tile of inv_set_tile that nz allows and
 * selects the fastest (a tile must beat the ones tried before it, in the
 * order above, by 2 %); then, over that tile, each two-step tile of
 * inv_set_pair that nz and the device's LDS allow, kept if 2 % faster than
 * the best so far (inv_get_pair reports it); then, for grids one tile wide,
 * the persistent kernel, and over the choice so far three-step tiles of 12,
 * 16, 20 and 8 rows (inv_set_triple; inv_get_triple reports the choice) and
 * walking chunks of 128, 96, 160, 192 and 64 rows (inv_set_walk,
 * inv_get_walk; a launch's workgroups are shots x chunks x strips, so the
 * best chunk depends on the shots per launch timed here),
 * each kept if 2 % faster.  The
 * choice is cached per (nx, nz, wave types, math mode) for the process, so
 * later contexts and batch sizes of the same grid take it without timing.
 * Overwrites the slots' wavefields, traces and snapshots.  No counterpart in the
 * reference (fixed 32x32 blocks, lib/device.h:7-65).  Outputs may be NULL. */
int inv_autotune(inv_ctx *ctx, int first, int count, int steps, int *variant, int *tile_rows,
                 int *cols_per_lane);
/* Stencil arithmetic: INV_MATH_EXACT evaluates the reference's expressions
 * (solver/fdm.h:76-161) with IEEE FP32 divisions, bitwise equal to them;
 * INV_MATH_FAST (default) multiplies by FP32-rounded coefficients 9/(8h),
 * 1/(24h) and dt/rho, the same FP32 accuracy against FP64 at a fraction of
 * the instructions. */
enum inv_math { INV_MATH_EXACT = 0, INV_MATH_FAST = 1 };
int inv_set_math(inv_ctx *ctx, int mode);

/* HIP graphs (default on): a whole forward or adjoint run is captured once
 * per (run, shot range, flags) and replayed; setting changes re-capture. */
int inv_set_graphs(inv_ctx *ctx, int on);

/* Run only the first nsteps steps of the next forward (0 = all). */
int inv_set_step_limit(inv_ctx *ctx, int nsteps);

/* Forward-snapshot storage for the adjoint.  No counterpart in the reference,
 * which keeps FP32 snapshots of every sfe-th step in host memory
 * (solver/fdm.h:403-425,521-560).
 *   format   INV_SNAP_F32, or a 16-bit format: each row segment of a field
 *            (the owned columns of one tile of the kernel that writes it) is
 *            scaled by the power of two 2^e that brings its max-abs into
 *            [2^14, 2^15), each block of 4 cells by a further 2^d (d in
 *            0..15, from the block's own max-abs), computed in the launch
 *            that writes the snapshot and kept in a per-slot exponent table;
 *            the scaled value is stored as INV_SNAP_I16 (nearest integer,
 *            |q| <= 32767), INV_SNAP_F16 or INV_SNAP_BF16 (nearest, ties to
 *            even) and decoded before use; all arithmetic stays FP32.
 *   ckpt     0: every snapshot of the run is kept (nt/sfe slots per shot);
 *            K > 0 (K % sfe == 0, nt % K == 0): the forward keeps the FP32
 *            state every K steps plus the snapshots of its last K steps, and
 *            the adjoint recomputes each earlier K-step segment from its
 *            checkpoint before using it (K/sfe slots per shot).
 * Needs the fused kernel (INV_VARIANT_FUSED) for anything but (F32, 0).
 * Call after inv_grid_set (which resets to F32, 0); buffers are sized at the
 * next run.  inv_snapshots_get then returns the slots held, decoded. */
int inv_set_snapshots(inv_ctx *ctx, int format, int ckpt);

/* ---- request server (continuous batching of shot solves) ----------------
 * No counterpart in the reference, which runs one inversion per process
 * (inv.cu:33-87).  A server owns resident models and a set of lanes, each a
 * context with its own HIP stream driven by its own host thread; a scheduler
 * thread admits queued requests into idle lanes as lanes and device memory
 * free up, so batches on different lanes overlap on the device.  Requests of
 * one batching class (same grid, spacing, dt, wave type, mode, sfe,
 * absorbing strips and receivers) share one launch, one shot per slot, each
 * slot reading its own request's model; shorter runs are padded to the
 * batch's longest, with results identical to the request run alone.
 * Admission keeps resident models plus every lane's buffers (wavefields,
 * snapshots, traces) within the HBM budget.
 *
 * Each request is one shot-solve job: forward of nshots shots on a resident
 * model (FdmSolver::runForward, solver/fdm.h:515-570) returning receiver
 * traces, and for adjoint = 1 the waveform misfit against the given
 * observed traces (misfit/waveform.h:4-50) and the adjoint
 * (solver/fdm.h:571-612) returning the lambda/mu/rho kernels summed over the
 * request's shots (Misfit::run(kernel=true) without the filter,
 * misfit/misfit.h:15-45).
 *
 * Host arrays of a request are read while it runs and its outputs written
 * when it completes: they must stay valid until inv_server_wait returns for
 * it.  Requests may be submitted from any thread; the server's functions are
 * thread-safe. */
typedef struct inv_server inv_server;

typedef struct inv_server_desc {
    int device;
    size_t hbm_budget;   /* device bytes for models + lanes (0: 90% of free memory) */
    int max_lanes;       /* batches in flight at once, one stream each (1..16) */
    int max_slots;       /* shots per batch */
    int max_bypass;      /* times the oldest request may be overtaken (0: strict FIFO) */
    int snap_format;     /* INV_SNAP_* of adjoint requests' forward snapshots */
    int math;            /* INV_MATH_* */
    int tile_x;          /* fused tile rows (8, 16, 32); 0: inv_autotune per grid */
    int plain;           /* 1: one request at a time, one shot per launch, in
                            arrival order (the plain baseline loop) */
} inv_server_desc;

typedef struct inv_request {
    int nx, nz;
    float dx, dz, dt;
    int nt;
    int sh, psv;         /* exactly one wave type */
    int adjoint;         /* 0 forward only; 1 forward + waveform misfit + adjoint */
    int trace_type;      /* forward-only traces: 0 velocity, 1 displacement */
    int sfe;             /* snapshot interval (adjoint: nt % sfe == 0) */
    int abs_width;
    float abs_alpha;
    int abs_left, abs_right, abs_top, abs_bottom;
    int model;           /* resident model id (inv_server_model_add) */
    int nshots;
    const int *src_ix, *src_iz;          /* [nshots] grid indices */
    const float *stf_x, *stf_y, *stf_z;  /* [nshots][nt] wavelets */
    int nrec;
    const int *rec_ix, *rec_iz;          /* [nrec] grid indices */
    const float *obs;    /* adjoint: observed traces [nshots][ncomp][nrec][nt] */
    float *traces;       /* out or NULL: [nshots][ncomp][nrec][nt]; displacement when adjoint */
    float *k_lambda, *k_mu, *k_rho;      /* out or NULL (adjoint): [nx*nz] */
} inv_request;

enum inv_request_status {
    INV_REQ_QUEUED = 0, INV_REQ_RUNNING = 1, INV_REQ_DONE = 2, INV_REQ_FAILED = 3
};

typedef struct inv_request_info {
    int status;          /* inv_request_status */
    int error;           /* inv_status of a failed request */
    double misfit;       /* adjoint: summed trace misfits (FP64) */
    double t_submit, t_start, t_done;  /* seconds on the server clock */
    int lane;            /* lane that ran it */
    int batch_shots;     /* shots in the launch it shared */
    int batch_requests;  /* requests in that launch */
} inv_request_info;

typedef struct inv_server_stats {
    long long submitted, completed, failed, batches, shots;
    double cell_updates;  /* completed work: sum over shots of nx*nz*nt*(1 or 2) */
    size_t budget, models_bytes;
    size_t bytes_held;    /* models + lane reservations now */
    size_t bytes_peak;    /* highest bytes_held so far */
    size_t bytes_actual;  /* device bytes the lanes' contexts hold now */
    int lanes_busy, queued;
    double now;           /* server clock, seconds since creation */
} inv_server_stats;

int inv_server_create(const inv_server_desc *desc, inv_server **out);
/* Waits for queued and running requests, then releases everything. */
int inv_server_destroy(inv_server *srv);
const char *inv_server_last_error(const inv_server *srv);
/* Uploads a model (lambda, mu, rho host arrays [nx*nz], z fastest) to stay
 * resident on the device; *id names it in requests. */
int inv_server_model_add(inv_server *srv, int nx, int nz, const float *lambda, const float *mu,
                         const float *rho, int *id);
/* Releases a model no queued or running request uses (INV_ERR_STATE otherwise). */
int inv_server_model_remove(inv_server *srv, int id);
int inv_server_submit(inv_server *srv, const inv_request *req, long long *id);
/* Waits up to timeout_s (< 0: forever) for a request; returns INV_OK once it
 * is done (its record is then released), its error code if it failed,
 * INV_ERR_STATE on timeout.  info may be NULL. */
int inv_server_wait(inv_server *srv, long long id, double timeout_s, inv_request_info *info);
/* Current state of a request without waiting or releasing it. */
int inv_server_poll(inv_server *srv, long long id, inv_request_info *info);
int inv_server_stats_get(inv_server *srv, inv_server_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
