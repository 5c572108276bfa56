/*
 * ORACLE — CPU restatement of libccy/inv.cu's forward+adjoint gradient path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (inv.cu_amd/) never links it.
 *
 * Every function restates one reference function; the file:line it follows is
 * given beside each declaration (paths relative to the reference checkout).
 * The arithmetic keeps the reference's expression order; build with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Parity pins (how each function is checked, see tests/test_oracle_*.py):
 *   - SH forward: reference golden traces example/checker/traces/vy_*.su
 *   - P-SV forward: two 1-D reductions (x-uniform z-forces equal the SH path
 *     with mu' = lambda + 2 mu, x-forces the SH path with mu, to FP64
 *     round-off), source-receiver reciprocity, energy without absorption
 *   - adjoint kernels: finite-difference gradient checks (SH within 1 %,
 *     the P-SV reductions within 1e-4, a 2-D P-SV model within 5e-3),
 *     snapshot pairing (sfe 10 against sfe 1)
 *   - waveform misfit: closed form; envelope: closed forms for cosine
 *     traces (both branches of the floor) and the numpy FFT Hilbert mask
 *   - Gaussian filter: brute-force one-sided weights
 */
#ifndef INV_ORACLE_H
#define INV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Simulation description.  Model fields are float [nx*nz], z fastest
 * (k = i*nz + j), as Dim lays them out (lib/device.h:7-65). */
typedef struct {
    int nx, nz;
    float dx, dz, dt;
    int nt;
    int sh, psv;
    int sfe;          /* snapshot interval; nsfe = nt / sfe (solver/fdm.h:391) */
    int adj;          /* mode 0/2: record u, store snapshots (solver/solver.h:166) */
    int trace_type;   /* 0 velocity, 1 displacement (solver/solver.h:168) */
    int inv_lambda, inv_mu, inv_rho;
    const float *lambda, *mu, *rho, *ab;
    int nsrc;
    const int *src_ix, *src_iz;
    const float *stf_x, *stf_y, *stf_z;   /* [nsrc][nt] */
    int nrec;
    const int *rec_ix, *rec_iz;
    /* [nsrc] or NULL (all 0).  0: the reference's point force (addSTF,
     * solver/fdm.h:163-180).  1: explosive source, no reference counterpart:
     * after the stress update of step it, stf_y[is][it] is added to sxx and
     * to szz at the source cell (P-SV only). */
    const int *src_type;
} orc_sim;

/* ---- setup (float, as the reference computes them) ---- */
/* grid derivation, solver/fdm.h:356-389 */
int orc_grid_dims(int npt, float xmax, float zmax, int *nx, int *nz, float *dx, float *dz);
/* initAbsbound, solver/fdm.h:48-74 */
void orc_absbound(int nx, int nz, int width, float alpha, int left, int right,
                  int bottom, int top, float *ab);
/* calcIdx, solver/fdm.h:44-47 */
int orc_calc_idx(float coord, float L, int n);
/* RickerSource::create, source/ricker.h:5-21 */
void orc_ricker(int nt, float dt, float f0, float t0, float *stf);
/* Source::init, source/source.h:17-37; cfg = {f0, t0, angle, amp} */
void orc_source(int nt, float dt, const float *cfg, float *stf_x, float *stf_y, float *stf_z);
/* mesh2grid, solver/fdm.h:4-24 (brute force, strict <, first wins) */
void orc_mesh2grid(int nx, int nz, float dx, float dz, float dmax, int npt,
                   const float *px, const float *pz, const float *pvp, const float *pvs,
                   const float *prho, float *lam, float *mu, float *rho);
/* vps2lm / lm2vps, solver/fdm.h:25-43 */
void orc_vps2lm(int n, float *lam, float *mu, const float *rho);
void orc_lm2vps(int n, float *vp, float *vs, const float *rho);

/* ---- propagation ----
 * orc_forward_*: runForward(isrc) of solver/fdm.h:515-570 for `nsteps`
 *   steps (nsteps <= nt; nsteps < 0 means nt).  isrc < 0 fires every source.
 *   out_{x,y,z}: [nrec][nt] receiver traces (u if adj or trace_type==1, else v).
 *   snap (nullable): [nsfe][nf][N] forward snapshots, nf fields in the order
 *   SH {uy, vy} then P-SV {ux, uz, vx, vz}; written only when adj.
 *   state (nullable): final wavefields, SH {vy, uy, sxy, szy} then
 *   P-SV {vx, vz, ux, uz, sxx, szz, sxz}, each [N].
 * orc_adjoint_*: runAdjoint of solver/fdm.h:571-612; adstf_{x,y,z} [nrec][nt];
 *   kernels are accumulated (k -= ...) into k_lambda/k_mu/k_rho [N]. */
int orc_forward_f32(const orc_sim *s, int isrc, int nsteps, float *out_x, float *out_y,
                    float *out_z, float *snap, float *state);
int orc_forward_f64(const orc_sim *s, int isrc, int nsteps, double *out_x, double *out_y,
                    double *out_z, double *snap, double *state);
int orc_adjoint_f32(const orc_sim *s, const float *adstf_x, const float *adstf_y,
                    const float *adstf_z, const float *snap, float *k_lambda, float *k_mu,
                    float *k_rho, float *state);
int orc_adjoint_f64(const orc_sim *s, const double *adstf_x, const double *adstf_y,
                    const double *adstf_z, const double *snap, double *k_lambda,
                    double *k_mu, double *k_rho, double *state);
/* number of snapshot fields per slot and number of state fields */
int orc_snap_fields(const orc_sim *s);
int orc_state_fields(const orc_sim *s);

/* ---- misfits (per trace) ----
 * waveform: misfit/waveform.h:4-50; returns chi = ||r||_2 * sqrt(dt)
 * envelope: misfit/envelope.h:4-106 (Hilbert mask with the n/2+1 quirk) */
double orc_waveform_f32(int nt, float dt, const float *syn, const float *obs, float *adstf);
double orc_waveform_f64(int nt, double dt, const double *syn, const double *obs, double *adstf);
double orc_envelope_f32(int nt, float dt, const float *syn, const float *obs, float *adstf);
double orc_envelope_f64(int nt, double dt, const double *syn, const double *obs, double *adstf);
/* analytic signal A = IFFT(mask * FFT(x)), unnormalised; re/im out [n] */
void orc_hilbert_f64(int n, const double *x, double *re, double *im);

/* ---- one-sided "Gaussian" smoother, filter/gaussian.h:4-56 (in place) ---- */
void orc_gauss_f32(int nx, int nz, int sigma, float *data);
void orc_gauss_f64(int nx, int nz, int sigma, double *data);

/* ---- gradient evaluation: Misfit::run(kernel=true), misfit/misfit.h:15-51 ----
 * obs: [nsrc][ncomp][nrec][nt], comps SH {y} then P-SV {x, z}.
 * misfit_type 0 waveform, 1 envelope; sigma 0 = no filter.
 * Kernels are zeroed first (initKernels).  Returns the summed misfit. */
double orc_gradient_f32(const orc_sim *s, const float *obs, int misfit_type, int sigma,
                        int kernel, float *k_lambda, float *k_mu, float *k_rho);
double orc_gradient_f64(const orc_sim *s, const double *obs, int misfit_type, int sigma,
                        int kernel, double *k_lambda, double *k_mu, double *k_rho);

/* OpenMP thread count used by the propagation loops (0 = library default). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
