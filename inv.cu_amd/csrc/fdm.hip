// Velocity-stress propagation on gfx950: the forward and adjoint time loops of
// FdmSolver (solver/fdm.h:515-612) as fused per-half-step kernels.
//
// Per time step the reference launches divS -> addSTF -> updateV -> divV ->
// updateS -> saveRec (six to eight kernels through scratch fields dsx.., dvxdx..).
// Here one kernel does the velocity half-step (stress divergence, source
// injection, velocity/displacement update, receiver recording, snapshot store)
// and one the stress half-step (velocity gradient and Hooke update); the
// divergence scratch never touches HBM.  Every shot slot of a batch is one
// grid z-layer of the same launch.  The arithmetic keeps the reference's
// expression order and is compiled with -ffp-contract=off, so the FP32 result
// of each cell is the same as the reference's formula evaluated in IEEE FP32.
#include <algorithm>

#include "internal.h"
#include "stencil.h"

namespace inv {

struct VelArgs {
    Geo g;
    const float *rho, *abx, *abzb, *abzt;
    float *vx, *vz, *ux, *uz, *vy, *uy;
    const float *sxx, *szz, *sxz, *sxy, *szy;
    // point-source injection (addSTF, fdm.h:163-180)
    const float *inj_x, *inj_y, *inj_z;
    long inj_slot_stride;
    const int *inj_ptr, *inj_cell, *inj_row;
    int inj_per_slot, ntiles, nt, it;
    // receiver recording (saveRec, fdm.h:181-196)
    float *out_x, *out_y, *out_z;
    long out_slot_stride;
    const int *rec_ptr, *rec_cell, *rec_row;
    int rec_disp;
    // snapshot store: u before the update, v after it (fdm.h:521-560)
    float *snap;  // slot-0 base of the current snapshot slot, or null
    long snap_slot_stride;
};

struct StrArgs {
    Geo g;
    const float *lam, *mu;
    const float *vx, *vz, *vy;
    float *sxx, *szz, *sxz, *sxy, *szy;
};

struct IntArgs {
    Geo g;
    const float *vx, *vz, *ux, *uz, *vy, *uy;
    const void *snap;  // slot-0 base of the snapshot slot being read
    long snap_slot_stride;  // elements per batch slot
    // 16-bit formats: the slot's exponent table (header: segment width, then
    // [batch][nsnapf][nx][nseg] exponents, then nbb block-offset bytes per
    // [batch][nsnapf][nx] from byte sblk; layout in stencil.h)
    const signed char *sexp;
    int nsnapf, nseg, nbb;
    long sblk;
    float *kl, *km, *kr;
    long k_slot_stride;  // 0: every slot accumulates into the same kernels
    float ndt;
    int inv_lambda, inv_mu, inv_rho;
};

template <bool SH, bool PSV, bool FAST>
__global__ __launch_bounds__(TILE_Z *TILE_X) void k_velocity(const VelArgs a) {
    const int j = blockIdx.x * TILE_Z + threadIdx.x;
    const int i = blockIdx.y * TILE_X + threadIdx.y;
    const int b = blockIdx.z;
    const Geo &g = a.g;
    if (i >= g.nx || j >= g.nz) return;
    const long nz = g.nz;
    const long k = (long)i * nz + j;
    const long o = (long)b * g.N + k;
    const bool xin = i >= 2 && i < g.nx - 2;
    const bool zin = j >= 2 && j < g.nz - 2;
    // initAbsbound product order: left, right (abx), bottom, top
    const float ab = (a.abx[i] * a.abzb[j]) * a.abzt[j];
    const float rho = a.rho[k];  // dt / rho in fast mode
    const int tile = blockIdx.y * gridDim.x + blockIdx.x;
    int p0 = 0, p1 = 0;
    if (a.inj_ptr) {
        const int set = a.inj_per_slot ? b : 0;
        p0 = a.inj_ptr[set * a.ntiles + tile];
        p1 = a.inj_ptr[set * a.ntiles + tile + 1];
    }
    int r0 = 0, r1 = 0;
    if (a.rec_ptr) {
        r0 = a.rec_ptr[tile];
        r1 = a.rec_ptr[tile + 1];
    }
    const long it = a.it;
    if (SH) {
        const float *sxy = a.sxy + (long)b * g.N, *szy = a.szy + (long)b * g.N;
        float d = xin ? d_minus<FAST>(sxy, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x) : 0.0f;
        if (zin) d += d_minus<FAST>(szy, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
        for (int p = p0; p < p1; p++)
            if (a.inj_cell[p] == k)
                d += a.inj_y[(long)b * a.inj_slot_stride + (long)a.inj_row[p] * a.nt + it];
        float v = a.vy[o], u = a.uy[o];
        if (a.snap) a.snap[(long)b * a.snap_slot_stride + k] = u;
        v = vel_update<FAST>(ab, v, g.dt, d, rho);
        u = u_step<FAST>(u, v, g.dt);
        a.vy[o] = v;
        a.uy[o] = u;
        if (a.snap) a.snap[(long)b * a.snap_slot_stride + g.N + k] = v;
        for (int p = r0; p < r1; p++)
            if (a.rec_cell[p] == k)
                a.out_y[(long)b * a.out_slot_stride + (long)a.rec_row[p] * a.nt + it] =
                    a.rec_disp ? u : v;
    }
    if (PSV) {
        const long so = (long)b * g.N;
        const float *sxx = a.sxx + so, *szz = a.szz + so, *sxz = a.sxz + so;
        float dx_, dz_;
        if (xin) {
            dx_ = d_minus<FAST>(sxx, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x);
            dz_ = d_minus<FAST>(sxz, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x);
        } else {
            dx_ = 0.0f;
            dz_ = 0.0f;
        }
        if (zin) {
            dx_ += d_minus<FAST>(sxz, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
            dz_ += d_minus<FAST>(szz, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
        }
        for (int p = p0; p < p1; p++)
            if (a.inj_cell[p] == k) {
                const long ks = (long)b * a.inj_slot_stride + (long)a.inj_row[p] * a.nt + it;
                dx_ += a.inj_x[ks];
                dz_ += a.inj_z[ks];
            }
        float vx = a.vx[o], vz = a.vz[o], ux = a.ux[o], uz = a.uz[o];
        const int f0 = SH ? 2 : 0;
        float *sp = a.snap ? a.snap + (long)b * a.snap_slot_stride + (long)f0 * g.N + k : nullptr;
        if (sp) {
            sp[0] = ux;
            sp[g.N] = uz;
        }
        vx = vel_update<FAST>(ab, vx, g.dt, dx_, rho);
        vz = vel_update<FAST>(ab, vz, g.dt, dz_, rho);
        ux = u_step<FAST>(ux, vx, g.dt);
        uz = u_step<FAST>(uz, vz, g.dt);
        a.vx[o] = vx;
        a.vz[o] = vz;
        a.ux[o] = ux;
        a.uz[o] = uz;
        if (sp) {
            sp[2 * g.N] = vx;
            sp[3 * g.N] = vz;
        }
        for (int p = r0; p < r1; p++)
            if (a.rec_cell[p] == k) {
                const long kr = (long)b * a.out_slot_stride + (long)a.rec_row[p] * a.nt + it;
                a.out_x[kr] = a.rec_disp ? ux : vx;
                a.out_z[kr] = a.rec_disp ? uz : vz;
            }
    }
}

template <bool SH, bool PSV, bool FAST>
__global__ __launch_bounds__(TILE_Z *TILE_X) void k_stress(const StrArgs a) {
    const int j = blockIdx.x * TILE_Z + threadIdx.x;
    const int i = blockIdx.y * TILE_X + threadIdx.y;
    const int b = blockIdx.z;
    const Geo &g = a.g;
    if (i >= g.nx || j >= g.nz) return;
    const long nz = g.nz;
    const long k = (long)i * nz + j;
    const long so = (long)b * g.N;
    const long o = so + k;
    const bool xin = i >= 1 && i < g.nx - 2;
    const bool zin = j >= 1 && j < g.nz - 2;
    const float mu = a.mu[k];
    if (SH) {
        const float *vy = a.vy + so;
        const float dx_ = xin ? d_plus<FAST>(vy, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x) : 0.0f;
        const float dz_ = zin ? d_plus<FAST>(vy, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z) : 0.0f;
        a.sxy[o] = s_sh<FAST>(a.sxy[o], g.dt, mu, dx_);
        a.szy[o] = s_sh<FAST>(a.szy[o], g.dt, mu, dz_);
    }
    if (PSV) {
        const float *vx = a.vx + so, *vz = a.vz + so;
        float xx, zx, xz, zz;
        if (xin) {
            xx = d_plus<FAST>(vx, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x);
            zx = d_plus<FAST>(vz, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x);
        } else {
            xx = 0.0f;
            zx = 0.0f;
        }
        if (zin) {
            xz = d_plus<FAST>(vx, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
            zz = d_plus<FAST>(vz, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
        } else {
            xz = 0.0f;
            zz = 0.0f;
        }
        const float lam = a.lam[k];
        float nxx, nzz;
        s_normal<FAST>(a.sxx[o], a.szz[o], g.dt, lam, mu, xx, zz, nxx, nzz);
        a.sxx[o] = nxx;
        a.szz[o] = nzz;
        a.sxz[o] = s_shear<FAST>(a.sxz[o], g.dt, mu, xz, zx);
    }
}

// interaction* kernels of solver/fdm.h:198-219 with the D+ derivatives of the
// adjoint and the forward displacement (fdm.h:586-605) computed in place.
// FP32 snapshot field read through the same accessor interface as Snap16
struct Snap32 {
    const float *p;
    __device__ __forceinline__ float operator[](long k) const { return p[k]; }
};

// the segment width of the slot's writer (at least 1; segments are clamped
// to the table in any case)
__device__ __forceinline__ int snap_segw(const IntArgs &a) {
    const int w = *reinterpret_cast<const int *>(a.sexp);
    return w > 0 ? w : 1;
}

template <int FMT>
struct SnapField {
    using type = Snap16<FMT>;
    __device__ static type at(const IntArgs &a, int b, int f) {
        const unsigned short *q =
            static_cast<const unsigned short *>(a.snap) + (long)b * a.snap_slot_stride + f * a.g.N;
        return type{q, a.sexp + snaptab_row(b, a.nsnapf, f, a.g.nx, 0, a.nseg),
                    reinterpret_cast<const unsigned char *>(a.sexp) +
                        snaptab_blk(a.sblk, b, a.nsnapf, f, a.g.nx, 0, a.nbb),
                    a.g.nz, a.nseg, a.nbb, snap_segw(a)};
    }
};
template <>
struct SnapField<INV_SNAP_F32> {
    using type = Snap32;
    __device__ static type at(const IntArgs &a, int b, int f) {
        return Snap32{static_cast<const float *>(a.snap) + (long)b * a.snap_slot_stride + f * a.g.N};
    }
};

template <bool SH, bool PSV, bool FAST, int FMT>
__global__ __launch_bounds__(TILE_Z *TILE_X) void k_interact(const IntArgs a) {
    const int j = blockIdx.x * TILE_Z + threadIdx.x;
    const int i = blockIdx.y * TILE_X + threadIdx.y;
    const int b = blockIdx.z;
    const Geo &g = a.g;
    if (i >= g.nx || j >= g.nz) return;
    const long nz = g.nz;
    const long k = (long)i * nz + j;
    const long so = (long)b * g.N;
    const bool xin = i >= 1 && i < g.nx - 2;
    const bool zin = j >= 1 && j < g.nz - 2;
    float *kl = a.kl + (long)b * a.k_slot_stride;
    float *km = a.km + (long)b * a.k_slot_stride;
    float *kr = a.kr + (long)b * a.k_slot_stride;
    const float ndt = a.ndt;
    if (SH) {
        const float *uy = a.uy + so;
        const auto uy_fw = SnapField<FMT>::at(a, b, 0), vy_fw = SnapField<FMT>::at(a, b, 1);
        const float dx_ = xin ? d_plus<FAST>(uy, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x) : 0.0f;
        const float dz_ = zin ? d_plus<FAST>(uy, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z) : 0.0f;
        const float dxf = xin ? d_plus_acc<FAST>(uy_fw, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x) : 0.0f;
        const float dzf = zin ? d_plus_acc<FAST>(uy_fw, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z) : 0.0f;
        if (a.inv_rho) kr[k] -= vy_fw[k] * a.vy[so + k] * ndt;
        if (a.inv_mu) km[k] -= (dx_ * dxf + dz_ * dzf) * ndt;
    }
    if (PSV) {
        const float *ux = a.ux + so, *uz = a.uz + so;
        const int f0 = SH ? 2 : 0;
        const auto ux_fw = SnapField<FMT>::at(a, b, f0), uz_fw = SnapField<FMT>::at(a, b, f0 + 1);
        const auto vx_fw = SnapField<FMT>::at(a, b, f0 + 2), vz_fw = SnapField<FMT>::at(a, b, f0 + 3);
        float xx, xz, zx, zz, xxf, xzf, zxf, zzf;
        if (xin) {
            xx = d_plus<FAST>(ux, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x);
            zx = d_plus<FAST>(uz, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x);
            xxf = d_plus_acc<FAST>(ux_fw, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x);
            zxf = d_plus_acc<FAST>(uz_fw, k, nz, g.c8dx, g.c24dx, g.a1x, g.a2x);
        } else {
            xx = zx = xxf = zxf = 0.0f;
        }
        if (zin) {
            xz = d_plus<FAST>(ux, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
            zz = d_plus<FAST>(uz, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
            xzf = d_plus_acc<FAST>(ux_fw, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
            zzf = d_plus_acc<FAST>(uz_fw, k, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
        } else {
            xz = zz = xzf = zzf = 0.0f;
        }
        if (a.inv_rho) kr[k] -= (vx_fw[k] * a.vx[so + k] + vz_fw[k] * a.vz[so + k]) * ndt;
        if (a.inv_mu)
            km[k] -= (2 * xx * xxf + 2 * zz * zzf + (xz + zx) * (zxf + xzf)) * ndt;
        if (a.inv_lambda) kl[k] -= ((xx + zz) * (xxf + zzf)) * ndt;
    }
}

// The same accumulation on column strips (nz even): a wavefront owns 124
// columns (2 adjacent columns per lane, read as float2, or as one 32-bit word
// holding two 16-bit snapshot values, with 2 halo columns each side) of RW
// rows.  The x-derivatives come from the rows i-1 .. i+2 held in registers,
// the z-derivatives from the neighbouring lanes (shuffles), so each field is
// read about once; k_interact re-reads every neighbour through the caches.
// Each cell's expressions and their order are k_interact's (SH before P-SV
// on the shared accumulators), so the kernels agree bit for bit.
constexpr int ICOLS = 124;  // columns a wavefront owns

// where a column's exponent lives in a row of a slot's table: segment seg,
// block-offset byte by, nibble shift sh (as snap_cell_exp)
struct SnapLoc {
    int seg, by, sh;
};
__device__ __forceinline__ SnapLoc snap_loc(const IntArgs &a, int j) {
    const int segw = snap_segw(a);
    int sg = j / segw;
    sg = sg < a.nseg - 1 ? sg : a.nseg - 1;
    const int lb = (j - sg * segw) >> 2;
    int by = sg * snap_seg_bytes(segw) + (lb >> 1);
    by = by < a.nbb - 1 ? by : a.nbb - 1;
    return SnapLoc{sg, by, (lb & 1) << 2};
}

// the two snapshot values of field f at cells k, k+1 (row i; their column's
// place in the table: sl)
template <int FMT>
__device__ __forceinline__ void snap_pair(const IntArgs &a, int b, int f, long k, int i,
                                          const SnapLoc &sl, float &x0, float &x1) {
    if constexpr (FMT == INV_SNAP_F32) {
        const float *q = static_cast<const float *>(a.snap) + (long)b * a.snap_slot_stride +
                         (long)f * a.g.N;
        const float2 t = *reinterpret_cast<const float2 *>(q + k);
        x0 = t.x;
        x1 = t.y;
    } else {
        const unsigned short *q = static_cast<const unsigned short *>(a.snap) +
                                  (long)b * a.snap_slot_stride + (long)f * a.g.N;
        const unsigned char *bt = reinterpret_cast<const unsigned char *>(a.sexp);
        const int e = a.sexp[snaptab_row(b, a.nsnapf, f, a.g.nx, i, a.nseg) + sl.seg] +
                      ((bt[snaptab_blk(a.sblk, b, a.nsnapf, f, a.g.nx, i, a.nbb) + sl.by] >> sl.sh) &
                       15);
        const unsigned w = *reinterpret_cast<const unsigned *>(q + k);
        x0 = snap_dec16((unsigned short)(w & 0xffffu), e, FMT);
        x1 = snap_dec16((unsigned short)(w >> 16), e, FMT);
    }
}

// D+ along z at the lane's two columns from their values e0, e1 (columns
// c, c+1) and the neighbouring lanes' (c-1 from the lane before, c+2, c+3
// from the lane after, by DPP; the edge lanes, which own no cell, get 0);
// every lane of the wavefront must call it
template <bool FAST>
__device__ __forceinline__ void dz_pair(float e0, float e1, const Geo &g, float &d0, float &d1) {
    const float m1 = lane_prev(e1);
    const float p2 = lane_next(e0);
    const float p3 = lane_next(e1);
    d0 = dp<FAST>(e1, e0, p2, m1, g.c8dz, g.c24dz, g.a1z, g.a2z);
    d1 = dp<FAST>(p2, e1, p3, e0, g.c8dz, g.c24dz, g.a1z, g.a2z);
}

template <bool SH, bool PSV, bool FAST, int FMT, int RW>
__global__ __launch_bounds__(256) void k_interact_col(const IntArgs a, int tiles_z) {
    constexpr int WR = RW + 3;  // rows i0-1 .. i0+RW+1 of the derivative fields
    const Geo &g = a.g;
    const int nx = g.nx, nz = g.nz;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tz = blockIdx.x % tiles_z, tx = blockIdx.x / tiles_z;
    const int b = blockIdx.y;
    const int i0 = (tx * 4 + wave) * RW;
    if (i0 >= nx) return;  // whole wavefront; no block barrier follows
    const int jb = tz * ICOLS - 2 + 2 * lane;
    const int jl = jb < 0 ? 0 : (jb > nz - 2 ? nz - 2 : jb);
    const bool own = lane >= 1 && lane <= 62 && jb < nz;
    const long so = (long)b * g.N;
    auto rowc = [&](int i) { return i < 0 ? 0 : (i > nx - 1 ? nx - 1 : i); };
    auto row = [&](int i) { return (long)rowc(i) * nz + jl; };
    // the lane's snapshot segment and block (columns jl, jl+1 share both:
    // segment widths are even and blocks start on even columns)
    SnapLoc sl{0, 0, 0};
    if constexpr (FMT != INV_SNAP_F32) sl = snap_loc(a, jl);
    float *kl = a.kl + (long)b * a.k_slot_stride;
    float *km = a.km + (long)b * a.k_slot_stride;
    float *kr = a.kr + (long)b * a.k_slot_stride;
    const float ndt = a.ndt;
    const int f0 = SH ? 2 : 0;
    // the derivative fields on rows i0-1 .. i0+RW+1
    float ay[WR][2], fy[WR][2], ax[WR][2], az[WR][2], fx[WR][2], fz[WR][2];
#pragma unroll
    for (int r = 0; r < WR; r++) {
        const long k = row(i0 - 1 + r);
        const int ic = rowc(i0 - 1 + r);
        if (SH) {
            const float2 t = *reinterpret_cast<const float2 *>(a.uy + so + k);
            ay[r][0] = t.x;
            ay[r][1] = t.y;
            snap_pair<FMT>(a, b, 0, k, ic, sl, fy[r][0], fy[r][1]);
        }
        if (PSV) {
            const float2 t = *reinterpret_cast<const float2 *>(a.ux + so + k);
            const float2 u = *reinterpret_cast<const float2 *>(a.uz + so + k);
            ax[r][0] = t.x;
            ax[r][1] = t.y;
            az[r][0] = u.x;
            az[r][1] = u.y;
            snap_pair<FMT>(a, b, f0, k, ic, sl, fx[r][0], fx[r][1]);
            snap_pair<FMT>(a, b, f0 + 1, k, ic, sl, fz[r][0], fz[r][1]);
        }
    }
#pragma unroll
    for (int t = 0; t < RW; t++) {
        const int i = i0 + t;
        if (i >= nx) break;
        const bool xin = i >= 1 && i < nx - 2;
        const long k = (long)i * nz + jl;
        float krv[2], kmv[2], klv[2];
        const bool z0 = jl >= 1 && jl < nz - 2, z1 = jl + 1 >= 1 && jl + 1 < nz - 2;
        if (own) {
            if (a.inv_rho) {
                const float2 q = *reinterpret_cast<const float2 *>(kr + k);
                krv[0] = q.x;
                krv[1] = q.y;
            }
            if (a.inv_mu) {
                const float2 q = *reinterpret_cast<const float2 *>(km + k);
                kmv[0] = q.x;
                kmv[1] = q.y;
            }
            if (PSV && a.inv_lambda) {
                const float2 q = *reinterpret_cast<const float2 *>(kl + k);
                klv[0] = q.x;
                klv[1] = q.y;
            }
        }
        if (SH) {
            float dzz[2], dzf[2];
            dz_pair<FAST>(ay[t + 1][0], ay[t + 1][1], g, dzz[0], dzz[1]);
            dz_pair<FAST>(fy[t + 1][0], fy[t + 1][1], g, dzf[0], dzf[1]);
            if (own) {
                float vf[2];
                snap_pair<FMT>(a, b, 1, k, i, sl, vf[0], vf[1]);
                const float2 vv = *reinterpret_cast<const float2 *>(a.vy + so + k);
                const float vy[2] = {vv.x, vv.y};
#pragma unroll
                for (int v = 0; v < 2; v++) {
                    const bool zin = v ? z1 : z0;
                    const float dx_ = xin ? dp<FAST>(ay[t + 2][v], ay[t + 1][v], ay[t + 3][v],
                                                     ay[t][v], g.c8dx, g.c24dx, g.a1x, g.a2x)
                                          : 0.0f;
                    const float dz_ = zin ? dzz[v] : 0.0f;
                    const float dxf = xin ? dp<FAST>(fy[t + 2][v], fy[t + 1][v], fy[t + 3][v],
                                                     fy[t][v], g.c8dx, g.c24dx, g.a1x, g.a2x)
                                          : 0.0f;
                    const float dzf_ = zin ? dzf[v] : 0.0f;
                    if (a.inv_rho) krv[v] -= vf[v] * vy[v] * ndt;
                    if (a.inv_mu) kmv[v] -= (dx_ * dxf + dz_ * dzf_) * ndt;
                }
            }
        }
        if (PSV) {
            float dxz[2], dzz[2], dxzf[2], dzzf[2];
            dz_pair<FAST>(ax[t + 1][0], ax[t + 1][1], g, dxz[0], dxz[1]);
            dz_pair<FAST>(az[t + 1][0], az[t + 1][1], g, dzz[0], dzz[1]);
            dz_pair<FAST>(fx[t + 1][0], fx[t + 1][1], g, dxzf[0], dxzf[1]);
            dz_pair<FAST>(fz[t + 1][0], fz[t + 1][1], g, dzzf[0], dzzf[1]);
            if (own) {
                float vfx[2], vfz[2];
                snap_pair<FMT>(a, b, f0 + 2, k, i, sl, vfx[0], vfx[1]);
                snap_pair<FMT>(a, b, f0 + 3, k, i, sl, vfz[0], vfz[1]);
                const float2 vxx = *reinterpret_cast<const float2 *>(a.vx + so + k);
                const float2 vzz = *reinterpret_cast<const float2 *>(a.vz + so + k);
                const float vx[2] = {vxx.x, vxx.y}, vz[2] = {vzz.x, vzz.y};
#pragma unroll
                for (int v = 0; v < 2; v++) {
                    const bool zin = v ? z1 : z0;
                    float xx, xz, zx, zz, xxf, xzf, zxf, zzf;
                    if (xin) {
                        xx = dp<FAST>(ax[t + 2][v], ax[t + 1][v], ax[t + 3][v], ax[t][v], g.c8dx,
                                      g.c24dx, g.a1x, g.a2x);
                        zx = dp<FAST>(az[t + 2][v], az[t + 1][v], az[t + 3][v], az[t][v], g.c8dx,
                                      g.c24dx, g.a1x, g.a2x);
                        xxf = dp<FAST>(fx[t + 2][v], fx[t + 1][v], fx[t + 3][v], fx[t][v], g.c8dx,
                                       g.c24dx, g.a1x, g.a2x);
                        zxf = dp<FAST>(fz[t + 2][v], fz[t + 1][v], fz[t + 3][v], fz[t][v], g.c8dx,
                                       g.c24dx, g.a1x, g.a2x);
                    } else {
                        xx = zx = xxf = zxf = 0.0f;
                    }
                    if (zin) {
                        xz = dxz[v];
                        zz = dzz[v];
                        xzf = dxzf[v];
                        zzf = dzzf[v];
                    } else {
                        xz = zz = xzf = zzf = 0.0f;
                    }
                    if (a.inv_rho) krv[v] -= (vfx[v] * vx[v] + vfz[v] * vz[v]) * ndt;
                    if (a.inv_mu)
                        kmv[v] -= (2 * xx * xxf + 2 * zz * zzf + (xz + zx) * (zxf + xzf)) * ndt;
                    if (a.inv_lambda) klv[v] -= ((xx + zz) * (xxf + zzf)) * ndt;
                }
            }
        }
        if (own) {
            if (a.inv_rho) *reinterpret_cast<float2 *>(kr + k) = make_float2(krv[0], krv[1]);
            if (a.inv_mu) *reinterpret_cast<float2 *>(km + k) = make_float2(kmv[0], kmv[1]);
            if (PSV && a.inv_lambda)
                *reinterpret_cast<float2 *>(kl + k) = make_float2(klv[0], klv[1]);
        }
    }
}

// k_total[k] += sum over slots (in slot order) of the per-slot accumulators.
__global__ void k_kslot_reduce(float *__restrict__ kl, float *__restrict__ km,
                               float *__restrict__ kr, const float *__restrict__ ks, long N,
                               int count) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    float a = kl[k], m = km[k], r = kr[k];
    for (int b = 0; b < count; b++) {
        const float *s = ks + (long)b * 3 * N;
        a += s[k];
        m += s[N + k];
        r += s[2 * N + k];
    }
    kl[k] = a;
    km[k] = m;
    kr[k] = r;
}

__global__ void k_bdt(float *__restrict__ bdt, const float *__restrict__ rho, float dt, long n) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) bdt[k] = dt / rho[k];
}

__global__ void k_fill(float *__restrict__ p, float v, long n) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
Geo make_geo(const Ctx &c) {
    Geo g;
    g.nx = c.g.nx;
    g.nz = c.g.nz;
    g.N = c.N;
    g.dt = c.g.dt;
    g.c8dx = 8 * c.g.dx;
    g.c24dx = 24 * c.g.dx;
    g.c8dz = 8 * c.g.dz;
    g.c24dz = 24 * c.g.dz;
    g.a1x = 9.0f / g.c8dx;
    g.a2x = 1.0f / g.c24dx;
    g.a1z = 9.0f / g.c8dz;
    g.a2z = 1.0f / g.c24dz;
    return g;
}

static dim3 tile_grid(const Ctx &c, int count) {
    return dim3((c.g.nz + TILE_Z - 1) / TILE_Z, (c.g.nx + TILE_X - 1) / TILE_X, count);
}

static int tile_of(const Ctx &c, long cell) {
    const int tz = (c.g.nz + TILE_Z - 1) / TILE_Z;
    const int i = (int)(cell / c.g.nz), j = (int)(cell % c.g.nz);
    return (i / TILE_X) * tz + (j / TILE_Z);
}

void build_point_list(Ctx &c, PointList &pl, int nsets, const std::vector<std::vector<int>> &cells,
                      const std::vector<std::vector<int>> &rows) {
    const int tz = (c.g.nz + TILE_Z - 1) / TILE_Z;
    const int tx = (c.g.nx + TILE_X - 1) / TILE_X;
    const int ntiles = tz * tx;
    std::vector<int> count((size_t)nsets * ntiles + 1, 0);
    int total = 0;
    for (int s = 0; s < nsets; s++)
        for (int cell : cells[s]) {
            count[(size_t)s * ntiles + tile_of(c, cell) + 1]++;
            total++;
        }
    for (size_t t = 1; t < count.size(); t++) count[t] += count[t - 1];
    std::vector<int> fill(count.begin(), count.end() - 1);
    std::vector<int> hc(std::max(total, 1)), hr(std::max(total, 1));
    for (int s = 0; s < nsets; s++)
        for (size_t p = 0; p < cells[s].size(); p++) {
            const int slot = fill[(size_t)s * ntiles + tile_of(c, cells[s][p])]++;
            hc[slot] = cells[s][p];
            hr[slot] = rows[s][p];
        }
    INV_HIP(hipStreamSynchronize(c.stream));
    pl.ptr.ensure(count.size() * sizeof(int));
    pl.cell.ensure(hc.size() * sizeof(int));
    pl.row.ensure(hr.size() * sizeof(int));
    INV_HIP(hipMemcpyAsync(pl.ptr.p, count.data(), count.size() * sizeof(int),
                           hipMemcpyHostToDevice, c.stream));
    INV_HIP(hipMemcpyAsync(pl.cell.p, hc.data(), hc.size() * sizeof(int), hipMemcpyHostToDevice,
                           c.stream));
    INV_HIP(hipMemcpyAsync(pl.row.p, hr.data(), hr.size() * sizeof(int), hipMemcpyHostToDevice,
                           c.stream));
    // the host vectors die at return: wait for the copies
    INV_HIP(hipStreamSynchronize(c.stream));
    pl.nsets = nsets;
    pl.ntiles = ntiles;
    pl.npoints = total;
}

void fill_dev(Ctx &c, void *p, long n, float v) {
    if (n <= 0 || !p) return;
    k_fill<<<(unsigned)((n + 255) / 256), 256, 0, c.stream>>>(static_cast<float *>(p), v, n);
    INV_HIP(hipGetLastError());
}

void launch_bdt(float *bdt, const float *rho, float dt, long n, hipStream_t stream) {
    k_bdt<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(bdt, rho, dt, n);
    INV_HIP(hipGetLastError());
}

void fdm_prepare_model(Ctx &c) {
    // per-slot models come with their own dt / rho
    if (c.math != INV_MATH_FAST || c.slot_models) return;
    // the model may have changed through its device pointers since the last
    // run (the optimizer updates it in place): rebuild dt / rho.  The buffer
    // is allocated by inv_grid_set with the model.
    INV_REQUIRE(c.bdt.p && c.bdt.bytes >= c.N * sizeof(float), INV_ERR_STATE,
                "dt/rho buffer missing");
    k_bdt<<<(unsigned)((c.N + 255) / 256), 256, 0, c.stream>>>(c.bdt.as<float>(),
                                                               c.rho.as<float>(), c.g.dt, c.N);
    INV_HIP(hipGetLastError());
}

void fdm_zero_state(Ctx &c, int count) {
    const long n = (long)count * c.N;
    if (c.g.sh) {
        fill_dev(c, c.vy.p, n); fill_dev(c, c.uy.p, n); fill_dev(c, c.sxy.p, n); fill_dev(c, c.szy.p, n);
    }
    if (c.g.psv) {
        fill_dev(c, c.vx.p, n); fill_dev(c, c.vz.p, n); fill_dev(c, c.ux.p, n); fill_dev(c, c.uz.p, n);
        fill_dev(c, c.sxx.p, n); fill_dev(c, c.szz.p, n); fill_dev(c, c.sxz.p, n);
    }
}

// ---- sampled event timing ----
static constexpr size_t PROF_MAX_PENDING = 1 << 15;

static hipEvent_t prof_event(Ctx &c) {
    if (!c.prof.pool.empty()) {
        hipEvent_t e = c.prof.pool.back();
        c.prof.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    INV_HIP(hipEventCreate(&e));
    return e;
}

bool prof_begin(Ctx &c, int it, hipEvent_t *a) {
    if (c.prof.every <= 0 || it % c.prof.every != 0) return false;
    if (c.prof.pending.size() >= PROF_MAX_PENDING) return false;
    *a = prof_event(c);
    INV_HIP(hipEventRecord(*a, c.stream));
    return true;
}

void prof_end(Ctx &c, int kind, hipEvent_t a, int units) {
    hipEvent_t b = prof_event(c);
    INV_HIP(hipEventRecord(b, c.stream));
    c.prof.pending.push_back({kind, a, b, units});
}

bool prof_begin_any(Ctx &c, hipEvent_t *a) {
    if (c.prof.every <= 0 || c.prof.pending.size() >= PROF_MAX_PENDING) return false;
    *a = prof_event(c);
    INV_HIP(hipEventRecord(*a, c.stream));
    return true;
}

void prof_resolve(Ctx &c) {
    if (c.prof.pending.empty()) return;
    INV_HIP(hipStreamSynchronize(c.stream));
    for (auto &r : c.prof.pending) {
        float ms = 0;
        INV_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        c.prof.count[r.kind] += r.units;
        c.prof.ms[r.kind] += ms;
        c.prof.pool.push_back(r.a);
        c.prof.pool.push_back(r.b);
    }
    c.prof.pending.clear();
}

void prof_release(Ctx &c) {
    (void)hipStreamSynchronize(c.stream);
    for (auto &r : c.prof.pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : c.prof.pool) (void)hipEventDestroy(e);
    c.prof.pending.clear();
    c.prof.pool.clear();
}

template <bool SH, bool PSV, bool FAST>
static void launch_step(Ctx &c, const VelArgs &va, const StrArgs &sa, int count, int it) {
    const dim3 grid = tile_grid(c, count), block(TILE_Z, TILE_X);
    hipEvent_t a;
    bool p = prof_begin(c, it, &a);
    k_velocity<SH, PSV, FAST><<<grid, block, 0, c.stream>>>(va);
    if (p) prof_end(c, PROF_VELOCITY, a);
    p = prof_begin(c, it, &a);
    k_stress<SH, PSV, FAST><<<grid, block, 0, c.stream>>>(sa);
    if (p) prof_end(c, PROF_STRESS, a);
}

template <bool FAST>
static void step_m(Ctx &c, const VelArgs &va, const StrArgs &sa, int count, int it) {
    if (c.g.sh && c.g.psv) launch_step<true, true, FAST>(c, va, sa, count, it);
    else if (c.g.sh) launch_step<true, false, FAST>(c, va, sa, count, it);
    else launch_step<false, true, FAST>(c, va, sa, count, it);
}

static void step(Ctx &c, const VelArgs &va, const StrArgs &sa, int count, int it) {
    if (c.math == INV_MATH_FAST) step_m<true>(c, va, sa, count, it);
    else step_m<false>(c, va, sa, count, it);
    INV_HIP(hipGetLastError());
}

static void fill_state_ptrs(Ctx &c, VelArgs &va, StrArgs &sa) {
    va.vx = c.vx.as<float>(); va.vz = c.vz.as<float>();
    va.ux = c.ux.as<float>(); va.uz = c.uz.as<float>();
    va.vy = c.vy.as<float>(); va.uy = c.uy.as<float>();
    va.sxx = c.sxx.as<float>(); va.szz = c.szz.as<float>(); va.sxz = c.sxz.as<float>();
    va.sxy = c.sxy.as<float>(); va.szy = c.szy.as<float>();
    sa.vx = va.vx; sa.vz = va.vz; sa.vy = va.vy;
    sa.sxx = c.sxx.as<float>(); sa.szz = c.szz.as<float>(); sa.sxz = c.sxz.as<float>();
    sa.sxy = c.sxy.as<float>(); sa.szy = c.szy.as<float>();
}

static void make_common(Ctx &c, VelArgs &va, StrArgs &sa) {
    va = VelArgs{};
    sa = StrArgs{};
    va.g = sa.g = make_geo(c);
    va.rho = c.math == INV_MATH_FAST ? c.bdt.as<float>() : c.rho.as<float>();
    va.abx = c.abx.as<float>();
    va.abzb = c.abzb.as<float>();
    va.abzt = c.abzt.as<float>();
    sa.lam = c.lam.as<float>();
    sa.mu = c.mu.as<float>();
    INV_REQUIRE(va.rho && va.abx && va.abzb && va.abzt && sa.lam && sa.mu, INV_ERR_STATE,
                "model or absorbing profile not allocated");
    fill_state_ptrs(c, va, sa);
    va.nt = c.g.nt;
    va.ntiles = c.rec_list.ntiles;
}

void release_list(PointList &pl) {
    pl.ptr.release();
    pl.cell.release();
    pl.row.release();
}

void fdm_prepare(Ctx &c, int kind, int first, int count) {
    if (kind == 1) source_list(c, first, count, 0);
}

void fdm_forward(Ctx &c, int first, int count, int flags, int it0, int it1) {
    // point lists of this batch's sources: set b = shot first+b (or every
    // source in set 0 for first == -1, runForward(-1) of fdm.h:171)
    PointList &sl = source_list(c, first, count, 0);
    VelArgs va;
    StrArgs sa;
    make_common(c, va, sa);
    const long nt = c.g.nt, nsrc = c.nsrc;
    va.inj_x = c.stf.as<float>();
    va.inj_y = c.stf.as<float>() + nsrc * nt;
    va.inj_z = c.stf.as<float>() + 2 * nsrc * nt;
    va.inj_slot_stride = 0;
    va.inj_ptr = sl.ptr.as<int>();
    va.inj_cell = sl.cell.as<int>();
    va.inj_row = sl.row.as<int>();
    va.inj_per_slot = 1;
    va.ntiles = sl.ntiles;
    const long tr = (long)c.nrec * nt;
    const int cy = 0, cx = c.g.sh ? 1 : 0, cz = cx + 1;
    va.out_y = c.out.as<float>() + cy * tr;
    va.out_x = c.out.as<float>() + cx * tr;
    va.out_z = c.out.as<float>() + cz * tr;
    va.out_slot_stride = (long)c.ncomp * tr;
    va.rec_ptr = c.nrec ? c.rec_list.ptr.as<int>() : nullptr;
    va.rec_cell = c.rec_list.cell.as<int>();
    va.rec_row = c.rec_list.row.as<int>();
    va.rec_disp = (c.g.adj || c.g.trace_type == 1) ? 1 : 0;
    va.snap_slot_stride = (long)c.nsfe * c.nsnapf * c.N;
    const bool store = (flags & INV_FWD_SNAPSHOTS) && c.g.adj && c.g.sfe > 0 && c.snap.p;

    if (it0 == 0) {
        c.parity = 0;
        fdm_zero_state(c, count);
        fdm_prepare_model(c);
    }
    for (int it = it0; it < it1; it++) {
        va.it = it;
        va.snap = nullptr;
        if (store && (it + 1) % c.g.sfe == 0) {
            const long isfe = c.nsfe - (it + 1) / c.g.sfe;
            va.snap = c.snap.as<float>() + isfe * c.nsnapf * c.N;
        }
        step(c, va, sa, count, it);
    }
}

// One half-step of the split kernels on the current state set of slots
// 0..count-1, in place (the velocity half-step reads only the stress and
// writes v and u; the stress half-step reads only v and writes the stress).
// which 0: divS + addSTF of time step `it` for shots first.. + updateV;
// which 1: divV + updateS (solver/fdm.h:273-294).  No receiver recording,
// snapshots or explosive sources.
template <bool SH, bool PSV, bool FAST>
static void launch_half(Ctx &c, int which, const VelArgs &va, const StrArgs &sa, int count) {
    const dim3 grid = tile_grid(c, count), block(TILE_Z, TILE_X);
    if (which == 0) k_velocity<SH, PSV, FAST><<<grid, block, 0, c.stream>>>(va);
    else k_stress<SH, PSV, FAST><<<grid, block, 0, c.stream>>>(sa);
}

template <bool FAST>
static void launch_half_m(Ctx &c, int which, const VelArgs &va, const StrArgs &sa, int count) {
    if (c.g.sh && c.g.psv) launch_half<true, true, FAST>(c, which, va, sa, count);
    else if (c.g.sh) launch_half<true, false, FAST>(c, which, va, sa, count);
    else launch_half<false, true, FAST>(c, which, va, sa, count);
}

void fdm_half_step(Ctx &c, int which, int first, int count, int it) {
    VelArgs va;
    StrArgs sa;
    make_common(c, va, sa);
    auto cur = [&](DevBuf &x, DevBuf &y) { return cur_field(c, x, y); };
    if (c.g.psv) {
        va.vx = cur(c.vx, c.vx_b);
        va.vz = cur(c.vz, c.vz_b);
        va.sxx = cur(c.sxx, c.sxx_b);
        va.szz = cur(c.szz, c.szz_b);
        va.sxz = cur(c.sxz, c.sxz_b);
    }
    if (c.g.sh) {
        va.vy = cur(c.vy, c.vy_b);
        va.sxy = cur(c.sxy, c.sxy_b);
        va.szy = cur(c.szy, c.szy_b);
    }
    sa.vx = va.vx; sa.vz = va.vz; sa.vy = va.vy;
    sa.sxx = const_cast<float *>(va.sxx);
    sa.szz = const_cast<float *>(va.szz);
    sa.sxz = const_cast<float *>(va.sxz);
    sa.sxy = const_cast<float *>(va.sxy);
    sa.szy = const_cast<float *>(va.szy);
    if (which == 0) {
        fdm_prepare_model(c);
        PointList &sl = source_list(c, first, count, 0);
        const long nt = c.g.nt, nsrc = c.nsrc;
        va.inj_x = c.stf.as<float>();
        va.inj_y = c.stf.as<float>() + nsrc * nt;
        va.inj_z = c.stf.as<float>() + 2 * nsrc * nt;
        va.inj_ptr = sl.ptr.as<int>();
        va.inj_cell = sl.cell.as<int>();
        va.inj_row = sl.row.as<int>();
        va.inj_per_slot = 1;
        va.ntiles = sl.ntiles;
        va.it = it;
    }
    if (c.math == INV_MATH_FAST) launch_half_m<true>(c, which, va, sa, count);
    else launch_half_m<false>(c, which, va, sa, count);
    INV_HIP(hipGetLastError());
}

template <bool SH, bool PSV, int FMT>
static void launch_interact_f(Ctx &c, const IntArgs &ia, int count) {
    if (c.math == INV_MATH_FAST)
        k_interact<SH, PSV, true, FMT><<<tile_grid(c, count), dim3(TILE_Z, TILE_X), 0, c.stream>>>(ia);
    else
        k_interact<SH, PSV, false, FMT><<<tile_grid(c, count), dim3(TILE_Z, TILE_X), 0, c.stream>>>(ia);
}

template <bool SH, bool PSV, int FMT, int RW>
static void launch_interact_col(Ctx &c, const IntArgs &ia, int count) {
    const int tiles_z = (c.g.nz + ICOLS - 1) / ICOLS;
    const int tiles_x = (c.g.nx + 4 * RW - 1) / (4 * RW);
    const dim3 grid((unsigned)(tiles_z * tiles_x), (unsigned)count);
    if (c.math == INV_MATH_FAST)
        k_interact_col<SH, PSV, true, FMT, RW><<<grid, 256, 0, c.stream>>>(ia, tiles_z);
    else
        k_interact_col<SH, PSV, false, FMT, RW><<<grid, 256, 0, c.stream>>>(ia, tiles_z);
}

template <bool SH, bool PSV, int FMT>
static void launch_interact_cf(Ctx &c, const IntArgs &ia, int count) {
    // 8 rows per wavefront where that still gives a few blocks per CU, else 4
    const long blocks8 = (long)((c.g.nz + ICOLS - 1) / ICOLS) * ((c.g.nx + 31) / 32) * count;
    if (blocks8 >= 2048) launch_interact_col<SH, PSV, FMT, 8>(c, ia, count);
    else launch_interact_col<SH, PSV, FMT, 4>(c, ia, count);
}

template <bool SH, bool PSV>
static void launch_interact(Ctx &c, const IntArgs &ia, int count) {
    if (c.g.nz % 2 == 0 && c.g.nz >= 4) {
        switch (c.snap_format) {
            case INV_SNAP_F16: launch_interact_cf<SH, PSV, INV_SNAP_F16>(c, ia, count); break;
            case INV_SNAP_BF16: launch_interact_cf<SH, PSV, INV_SNAP_BF16>(c, ia, count); break;
            case INV_SNAP_I16: launch_interact_cf<SH, PSV, INV_SNAP_I16>(c, ia, count); break;
            default: launch_interact_cf<SH, PSV, INV_SNAP_F32>(c, ia, count); break;
        }
        return;
    }
    switch (c.snap_format) {
        case INV_SNAP_F16: launch_interact_f<SH, PSV, INV_SNAP_F16>(c, ia, count); break;
        case INV_SNAP_BF16: launch_interact_f<SH, PSV, INV_SNAP_BF16>(c, ia, count); break;
        case INV_SNAP_I16: launch_interact_f<SH, PSV, INV_SNAP_I16>(c, ia, count); break;
        default: launch_interact_f<SH, PSV, INV_SNAP_F32>(c, ia, count); break;
    }
}

void fdm_interact(Ctx &c, int count, int it, int slot, const float *vx, const float *vz,
                  const float *vy) {
    IntArgs ia{};
    ia.g = make_geo(c);
    ia.vx = vx;
    ia.vz = vz;
    ia.vy = vy;
    ia.ux = c.ux.as<float>();
    ia.uz = c.uz.as<float>();
    ia.uy = c.uy.as<float>();
    const bool s16 = c.snap_format != INV_SNAP_F32;
    INV_REQUIRE(slot >= 0 && slot < c.nslots && c.snap.p && (!s16 || c.sexp.p), INV_ERR_STATE,
                "snapshot slot outside the snapshot buffer");
    const long esize = s16 ? 2 : 4;
    ia.snap_slot_stride = (long)c.nslots * c.nsnapf * c.N;
    ia.snap = c.snap.as<char>() + (long)slot * c.nsnapf * c.N * esize;
    ia.sexp = s16 ? c.sexp.as<signed char>() +
                        (long)slot * snap_tab_bytes(c.g.batch, c.nsnapf, c.g.nx, c.g.nz)
                  : nullptr;
    ia.nsnapf = c.nsnapf;
    ia.nseg = snap_nseg(c.g.nz);
    ia.nbb = snap_nbb(c.g.nz);
    ia.sblk = snap_blk_off(c.g.batch, c.nsnapf, c.g.nx, c.g.nz);
    ia.ndt = (float)c.g.sfe * c.g.dt;  // sfe * dt, fdm.h:585
    ia.inv_lambda = c.g.inv_lambda;
    ia.inv_mu = c.g.inv_mu;
    ia.inv_rho = c.g.inv_rho;
    if (c.g.batch == 1) {
        // one-shot path: accumulate straight into the kernels, shot after
        // shot, as Misfit::run does (misfit/misfit.h:29-44)
        ia.kl = c.k_lambda.as<float>();
        ia.km = c.k_mu.as<float>();
        ia.kr = c.k_rho.as<float>();
        ia.k_slot_stride = 0;
    } else {
        // batched path: one accumulator per slot, summed in slot order after
        // the adjoint (fdm_kslot_reduce), so a shot's contribution does not
        // depend on which shots share its launch
        ia.kl = c.kslot.as<float>();
        ia.km = c.kslot.as<float>() + c.N;
        ia.kr = c.kslot.as<float>() + 2 * c.N;
        ia.k_slot_stride = 3 * c.N;
    }
    hipEvent_t a;
    const bool p = prof_begin(c, it / c.g.sfe, &a);
    if (c.g.sh && c.g.psv) launch_interact<true, true>(c, ia, count);
    else if (c.g.sh) launch_interact<true, false>(c, ia, count);
    else launch_interact<false, true>(c, ia, count);
    INV_HIP(hipGetLastError());
    if (p) prof_end(c, PROF_INTERACT, a);
}

void fdm_adjoint(Ctx &c, int first, int count, int it0, int it1) {
    (void)first;
    VelArgs va;
    StrArgs sa;
    make_common(c, va, sa);
    const long nt = c.g.nt;
    const long tr = (long)c.nrec * nt;
    const int cy = 0, cx = c.g.sh ? 1 : 0, cz = cx + 1;
    va.inj_y = c.adstf.as<float>() + cy * tr;
    va.inj_x = c.adstf.as<float>() + cx * tr;
    va.inj_z = c.adstf.as<float>() + cz * tr;
    va.inj_slot_stride = (long)c.ncomp * tr;
    va.inj_ptr = c.nrec ? c.rec_list.ptr.as<int>() : nullptr;
    va.inj_cell = c.rec_list.cell.as<int>();
    va.inj_row = c.rec_list.row.as<int>();
    va.inj_per_slot = 0;
    va.ntiles = c.rec_list.ntiles;
    va.rec_ptr = nullptr;
    va.snap = nullptr;
    const bool direct = c.g.batch == 1;
    if (it0 == 0) {
        if (!direct) fill_dev(c, c.kslot.p, 3L * count * c.N);
        c.parity = 0;
        fdm_zero_state(c, count);
        fdm_prepare_model(c);
    }
    for (int it = it0; it < it1; it++) {
        va.it = it;
        step(c, va, sa, count, it);
        if (it % c.g.sfe == 0)
            fdm_interact(c, count, it, it / c.g.sfe, c.vx.as<float>(), c.vz.as<float>(),
                         c.vy.as<float>());
    }
    if (!direct && it1 == c.g.nt) fdm_kslot_reduce(c, count);
}

// out[3][N] = 0 + kslot[s0] + kslot[s0+1] + ... in slot order
__global__ void k_kslot_sum(float *__restrict__ out, const float *__restrict__ ks, long N, int s0,
                            int n) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    float a = 0.0f, m = 0.0f, r = 0.0f;
    for (int b = s0; b < s0 + n; b++) {
        const float *s = ks + (long)b * 3 * N;
        a += s[k];
        m += s[N + k];
        r += s[2 * N + k];
    }
    out[k] = a;
    out[N + k] = m;
    out[2 * N + k] = r;
}

void kslot_sum(Ctx &c, float *out, int s0, int n) {
    INV_REQUIRE(c.kslot.p && s0 >= 0 && n >= 1 && s0 + n <= c.g.batch, INV_ERR_STATE,
                "per-slot kernel accumulators missing or slot range out of bounds");
    k_kslot_sum<<<(unsigned)((c.N + 255) / 256), 256, 0, c.stream>>>(out, c.kslot.as<float>(),
                                                                    c.N, s0, n);
    INV_HIP(hipGetLastError());
}

void fdm_kslot_reduce(Ctx &c, int count) {
    k_kslot_reduce<<<(unsigned)((c.N + 255) / 256), 256, 0, c.stream>>>(
        c.k_lambda.as<float>(), c.k_mu.as<float>(), c.k_rho.as<float>(), c.kslot.as<float>(), c.N,
        count);
    INV_HIP(hipGetLastError());
}

}  // namespace inv
