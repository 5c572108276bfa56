// One kernel per time step: the whole velocity-stress step of FdmSolver
// (divS -> addSTF -> updateV -> divV -> updateS -> saveRec, solver/fdm.h:
// 534-551) for a tile of FZ x TX cells of one shot slot.
//
// Data flow of a tile (rows i0..i0+TX-1 along x, columns j0..j0+FZ-1 along z):
//   1. the stress of step n is staged in LDS on the tile plus a 3-cell halo;
//   2. the velocity of step n+1 is computed on the tile plus a halo of 1 cell
//      before and 2 after in each direction (what the forward-staggered
//      gradient of the tile needs), from that LDS stress, v^n, rho and the
//      absorbing profile; halo cells are recomputed, not exchanged, and get
//      the same bits as in their owner tile.  Tile cells also update u, record
//      receivers and store snapshots;
//   3. the stress of step n+1 is computed on the tile from LDS velocities.
// Every wavefield is read and written once per step from HBM (halo re-reads
// come from L2), 68 B per P-SV cell-update instead of the 88 B of a
// velocity/stress kernel pair.  v and s are double-buffered because
// neighbouring tiles still read step-n values of their halo.
//
// Tiles are ordered slot-major, then x, then z; the hardware deals workgroups
// to the 8 XCDs round-robin, so the block index is remapped to give each XCD
// one contiguous run of tiles and the halo rows of consecutive tiles hit the
// same L2.
#include <algorithm>

#include "internal.h"
#include "colv.h"

namespace inv {

constexpr int FZ = 64;        // tile width along z: one wavefront
constexpr int FBLOCK = 256;   // four wavefronts per tile

template <bool SH, bool PSV, int TX, bool FAST>
__global__ __launch_bounds__(FBLOCK) void k_step(const StepArgs a) {
    constexpr int SW = FZ + 6, SR = TX + 6;  // stress tile: halo 3 on each side
    constexpr int VW = FZ + 3, VR = TX + 3;  // velocity tile: halo 1 before, 2 after
    constexpr int NV = (VR * VW + FBLOCK - 1) / FBLOCK;  // velocity cells per thread
    constexpr int NT = TX * FZ / FBLOCK;                 // tile cells per thread
    __shared__ float l_sxx[PSV ? SR * SW : 1], l_szz[PSV ? SR * SW : 1], l_sxz[PSV ? SR * SW : 1];
    __shared__ float l_sxy[SH ? SR * SW : 1], l_szy[SH ? SR * SW : 1];
    __shared__ float l_vx[PSV ? VR * VW : 1], l_vz[PSV ? VR * VW : 1], l_vy[SH ? VR * VW : 1];

    // XCD-aware tile order
    const int nw = gridDim.x, w = blockIdx.x;
    const int q = nw >> 3, rmd = nw & 7, xcd = w & 7;
    const int L = xcd * q + (xcd < rmd ? xcd : rmd) + (w >> 3);
    const int per = a.tiles_x * a.tiles_z;
    const int b = L / per;
    const int t = L - b * per;
    const int tx = t / a.tiles_z, tz = t - tx * a.tiles_z;
    const int i0 = tx * TX, j0 = tz * FZ;
    const Geo &g = a.g;
    const int nx = g.nx, nz = g.nz;
    const long so = (long)b * g.N;
    const float *m_rho = a.rho, *m_lam = a.lam, *m_mu = a.mu;
    if (a.slot_model) {
        m_rho = a.slot_model[3 * b];
        m_lam = a.slot_model[3 * b + 1];
        m_mu = a.slot_model[3 * b + 2];
    }

    // 0. issue the global loads of phases 2 and 3 first, so that they are in
    //    flight together with the stress loads of phase 1 (the barriers below
    //    wait for LDS traffic only)
    float pvx[NV], pvz[NV], pvy[NV], prho[NV], pux[NV], puz[NV], puy[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const int e = threadIdx.x + k * FBLOCK;
        const int vr = e / VW, vc = e - vr * VW;
        const int i = i0 - 1 + vr, j = j0 - 1 + vc;
        const bool in = e < VR * VW && i >= 0 && i < nx && j >= 0 && j < nz;
        const bool own = in && vr >= 1 && vr <= TX && vc >= 1 && vc <= FZ;
        const long kc = (long)i * nz + j;
        prho[k] = in ? m_rho[kc] : 1.0f;
        if (PSV) {
            pvx[k] = in ? a.vx0[so + kc] : 0.0f;
            pvz[k] = in ? a.vz0[so + kc] : 0.0f;
            pux[k] = own ? a.ux[so + kc] : 0.0f;
            puz[k] = own ? a.uz[so + kc] : 0.0f;
        }
        if (SH) {
            pvy[k] = in ? a.vy0[so + kc] : 0.0f;
            puy[k] = own ? a.uy[so + kc] : 0.0f;
        }
    }
    float plam[NT], pmu[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) {
        const int e = threadIdx.x + k * FBLOCK;
        const int i = i0 + e / FZ, j = j0 + (e % FZ);
        const bool in = i < nx && j < nz;
        const long kc = (long)i * nz + j;
        pmu[k] = in ? m_mu[kc] : 0.0f;
        if (PSV) plam[k] = in ? m_lam[kc] : 0.0f;
    }

    // 1. stress of step n on the tile plus halo
    for (int e = threadIdx.x; e < SR * SW; e += FBLOCK) {
        const int rr = e / SW, cc = e - rr * SW;
        const int i = i0 - 3 + rr, j = j0 - 3 + cc;
        const bool in = i >= 0 && i < nx && j >= 0 && j < nz;
        const long k = so + (long)i * nz + j;
        if (PSV) {
            l_sxx[e] = in ? a.sxx0[k] : 0.0f;
            l_szz[e] = in ? a.szz0[k] : 0.0f;
            l_sxz[e] = in ? a.sxz0[k] : 0.0f;
        }
        if (SH) {
            l_sxy[e] = in ? a.sxy0[k] : 0.0f;
            l_szy[e] = in ? a.szy0[k] : 0.0f;
        }
    }
    __syncthreads();

    int p0 = 0, p1 = 0;
    if (a.inj_ptr) {
        const int set = a.inj_per_slot ? b : 0;
        p0 = a.inj_ptr[set * per + t];
        p1 = a.inj_ptr[set * per + t + 1];
    }
    int r0 = 0, r1 = 0;
    if (a.rec_ptr) {
        r0 = a.rec_ptr[t];
        r1 = a.rec_ptr[t + 1];
    }
    const long it = a.it;

    // 2. velocity of step n+1 on the tile plus halo (updateV*, fdm.h:150-161)
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const int e = threadIdx.x + k * FBLOCK;
        if (e >= VR * VW) break;
        const int vr = e / VW, vc = e - vr * VW;
        const int i = i0 - 1 + vr, j = j0 - 1 + vc;
        if (i < 0 || i >= nx || j < 0 || j >= nz) {
            if (PSV) {
                l_vx[e] = 0.0f;
                l_vz[e] = 0.0f;
            }
            if (SH) l_vy[e] = 0.0f;
            continue;
        }
        const long kc = (long)i * nz + j;
        const long o = so + kc;
        const bool xin = i >= 2 && i < nx - 2;
        const bool zin = j >= 2 && j < nz - 2;
        const bool own = vr >= 1 && vr <= TX && vc >= 1 && vc <= FZ;
        const int se = (vr + 2) * SW + (vc + 2);
        const float ab = (a.abx[i] * a.abzb[j]) * a.abzt[j];
        const float rho = prho[k];  // dt / rho in fast mode
        if (SH) {
            float d = xin ? d_minus<FAST>(l_sxy, se, SW, g.c8dx, g.c24dx, g.a1x, g.a2x) : 0.0f;
            if (zin) d += d_minus<FAST>(l_szy, se, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
            for (int p = p0; p < p1; p++)
                if (a.inj_cell[p] == kc)
                    d += a.inj_y[(long)b * a.inj_slot_stride + (long)a.inj_row[p] * a.nt + it];
            const float v = vel_update<FAST>(ab, pvy[k], g.dt, d, rho);
            l_vy[e] = v;
            if (own) {
                float u = puy[k];
                float *sp = static_cast<float *>(a.snap);
                if (sp) sp[(long)b * a.snap_slot_stride + kc] = u;
                u = u_step<FAST>(u, v, g.dt);
                a.vy1[o] = v;
                a.uy[o] = u;
                if (sp) sp[(long)b * a.snap_slot_stride + g.N + kc] = v;
                for (int p = r0; p < r1; p++)
                    if (a.rec_cell[p] == kc)
                        a.out_y[(long)b * a.out_slot_stride + (long)a.rec_row[p] * a.nt + it] =
                            a.rec_disp ? u : v;
            }
        }
        if (PSV) {
            float dx_, dz_;
            if (xin) {
                dx_ = d_minus<FAST>(l_sxx, se, SW, g.c8dx, g.c24dx, g.a1x, g.a2x);
                dz_ = d_minus<FAST>(l_sxz, se, SW, g.c8dx, g.c24dx, g.a1x, g.a2x);
            } else {
                dx_ = 0.0f;
                dz_ = 0.0f;
            }
            if (zin) {
                dx_ += d_minus<FAST>(l_sxz, se, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
                dz_ += d_minus<FAST>(l_szz, se, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
            }
            for (int p = p0; p < p1; p++)
                if (a.inj_cell[p] == kc) {
                    const long ks = (long)b * a.inj_slot_stride + (long)a.inj_row[p] * a.nt + it;
                    dx_ += a.inj_x[ks];
                    dz_ += a.inj_z[ks];
                }
            const float vx = vel_update<FAST>(ab, pvx[k], g.dt, dx_, rho);
            const float vz = vel_update<FAST>(ab, pvz[k], g.dt, dz_, rho);
            l_vx[e] = vx;
            l_vz[e] = vz;
            if (own) {
                float ux = pux[k], uz = puz[k];
                const int f0 = SH ? 2 : 0;
                float *sp = a.snap ? static_cast<float *>(a.snap) + (long)b * a.snap_slot_stride +
                                         (long)f0 * g.N + kc
                                   : nullptr;
                if (sp) {
                    sp[0] = ux;
                    sp[g.N] = uz;
                }
                ux = u_step<FAST>(ux, vx, g.dt);
                uz = u_step<FAST>(uz, vz, g.dt);
                a.vx1[o] = vx;
                a.vz1[o] = vz;
                a.ux[o] = ux;
                a.uz[o] = uz;
                if (sp) {
                    sp[2 * g.N] = vx;
                    sp[3 * g.N] = vz;
                }
                for (int p = r0; p < r1; p++)
                    if (a.rec_cell[p] == kc) {
                        const long kr =
                            (long)b * a.out_slot_stride + (long)a.rec_row[p] * a.nt + it;
                        a.out_x[kr] = a.rec_disp ? ux : vx;
                        a.out_z[kr] = a.rec_disp ? uz : vz;
                    }
            }
        }
    }
    __syncthreads();

    // 3. stress of step n+1 on the tile (divV* + updateS*, fdm.h:103-149)
#pragma unroll
    for (int k = 0; k < NT; k++) {
        const int e = threadIdx.x + k * FBLOCK;
        const int tr = e / FZ, tc = e - tr * FZ;
        const int i = i0 + tr, j = j0 + tc;
        if (i >= nx || j >= nz) continue;
        const long kc = (long)i * nz + j;
        const long o = so + kc;
        const bool xin = i >= 1 && i < nx - 2;
        const bool zin = j >= 1 && j < nz - 2;
        const int ve = (tr + 1) * VW + (tc + 1);
        const int se = (tr + 3) * SW + (tc + 3);
        const float mu = pmu[k];
        if (SH) {
            const float dx_ = xin ? d_plus<FAST>(l_vy, ve, VW, g.c8dx, g.c24dx, g.a1x, g.a2x) : 0.0f;
            const float dz_ = zin ? d_plus<FAST>(l_vy, ve, 1, g.c8dz, g.c24dz, g.a1z, g.a2z) : 0.0f;
            a.sxy1[o] = s_sh<FAST>(l_sxy[se], g.dt, mu, dx_);
            a.szy1[o] = s_sh<FAST>(l_szy[se], g.dt, mu, dz_);
        }
        if (PSV) {
            float xx, zx, xz, zz;
            if (xin) {
                xx = d_plus<FAST>(l_vx, ve, VW, g.c8dx, g.c24dx, g.a1x, g.a2x);
                zx = d_plus<FAST>(l_vz, ve, VW, g.c8dx, g.c24dx, g.a1x, g.a2x);
            } else {
                xx = 0.0f;
                zx = 0.0f;
            }
            if (zin) {
                xz = d_plus<FAST>(l_vx, ve, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
                zz = d_plus<FAST>(l_vz, ve, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
            } else {
                xz = 0.0f;
                zz = 0.0f;
            }
            const float lam = plam[k];
            float nxx, nzz;
            s_normal<FAST>(l_sxx[se], l_szz[se], g.dt, lam, mu, xx, zz, nxx, nzz);
            a.sxx1[o] = nxx;
            a.szz1[o] = nzz;
            a.sxz1[o] = s_shear<FAST>(l_sxz[se], g.dt, mu, xz, zx);
        }
    }
}

// ---------------------------------------------------------------------------
// Column kernel: the same step with one lane per column.  The stress tile is
// exactly one wavefront wide (CZ + 6 = 64 columns), the velocity tile uses
// lanes 0..CZ+2 and the output tile lanes 0..CZ-1; the four wavefronts stride
// over rows.  Addresses advance by nz per row, with no index division and no
// row wrap-around inside a wavefront.
// ---------------------------------------------------------------------------
constexpr int CZ = 58;
static_assert(CZ == SNAP_SEG_WIDTHS[0], "snapshot tables are sized for this segment width");

template <bool SH, bool PSV, int TX, bool FAST>
__global__ __launch_bounds__(FBLOCK) void k_col(const StepArgs a) {
    static_assert(SH != PSV, "one wave type per column-kernel launch");
    constexpr int W = 64;                 // LDS row = wavefront width
    constexpr int SR = TX + 6, VR = TX + 3;
    constexpr int NWAVE = FBLOCK / 64;
    constexpr int PS = (SR + NWAVE - 1) / NWAVE;  // stress rows per wave
    constexpr int PV = (VR + NWAVE - 1) / NWAVE;  // velocity rows per wave
    constexpr int PT = TX / NWAVE;                // output rows per wave
    __shared__ float l_sxx[PSV ? SR * W : 1], l_szz[PSV ? SR * W : 1], l_sxz[PSV ? SR * W : 1];
    __shared__ float l_sxy[SH ? SR * W : 1], l_szy[SH ? SR * W : 1];
    __shared__ float l_vx[PSV ? VR * W : 1], l_vz[PSV ? VR * W : 1], l_vy[SH ? VR * W : 1];

    const int nw = gridDim.x, w = blockIdx.x;
    const int q = nw >> 3, rmd = nw & 7, xcd = w & 7;
    const int L = xcd * q + (xcd < rmd ? xcd : rmd) + (w >> 3);
    const int per = a.tiles_x * a.tiles_z;
    const int b = L / per;
    const int t = L - b * per;
    const int tx = t / a.tiles_z, tz = t - tx * a.tiles_z;
    const int i0 = tx * TX, j0 = tz * CZ;
    const Geo &g = a.g;
    const int nx = g.nx, nz = g.nz;
    const long so = (long)b * g.N;
    const float *m_rho = a.rho, *m_lam = a.lam, *m_mu = a.mu;
    if (a.slot_model) {
        m_rho = a.slot_model[3 * b];
        m_lam = a.slot_model[3 * b + 1];
        m_mu = a.slot_model[3 * b + 2];
    }
    // the wavefront index is uniform across the wavefront: reading it through
    // readfirstlane lets every row quantity derived from it live in scalar
    // registers (scalar branches and address arithmetic instead of per-lane)
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };

    // All global loads of the three phases are issued first, at clamped (so
    // always valid) addresses; values outside the grid are replaced by 0
    // afterwards.  Branch-free loads let the whole batch be in flight at once.
    const int js = j0 - 3 + lane, jv = j0 - 1 + lane, jt = j0 + lane;
    const bool jsin = js >= 0 && js < nz;
    const bool jvin = lane < CZ + 3 && jv >= 0 && jv < nz;
    const bool jtin = lane < CZ && jt < nz;
    const int jsc = clampi(js, 0, nz - 1), jvc = clampi(jv, 0, nz - 1), jtc = clampi(jt, 0, nz - 1);

    float s1[PS], s2[PS], s3[PS];
#pragma unroll
    for (int p = 0; p < PS; p++) {
        const int r = wave + p * NWAVE;
        const long k = so + (long)clampi(i0 - 3 + r, 0, nx - 1) * nz + jsc;
        if (PSV) {
            s1[p] = a.sxx0[k];
            s2[p] = a.szz0[k];
            s3[p] = a.sxz0[k];
        } else {
            s1[p] = a.sxy0[k];
            s2[p] = a.szy0[k];
        }
    }
    float qv1[PV], qv2[PV], qr[PV], qu1[PV], qu2[PV];
#pragma unroll
    for (int p = 0; p < PV; p++) {
        const int vr = wave + p * NWAVE;
        const long kc = (long)clampi(i0 - 1 + vr, 0, nx - 1) * nz + jvc;
        qr[p] = m_rho[kc];
        if (PSV) {
            qv1[p] = a.vx0[so + kc];
            qv2[p] = a.vz0[so + kc];
            qu1[p] = a.ux[so + kc];
            qu2[p] = a.uz[so + kc];
        } else {
            qv1[p] = a.vy0[so + kc];
            qu1[p] = a.uy[so + kc];
        }
    }
    float qmu[PT], qlam[PT];
#pragma unroll
    for (int p = 0; p < PT; p++) {
        const int tr = wave + p * NWAVE;
        const long kc = (long)clampi(i0 + tr, 0, nx - 1) * nz + jtc;
        qmu[p] = m_mu[kc];
        if (PSV) qlam[p] = m_lam[kc];
    }
    const float abz = a.abzb[jvc], abt = a.abzt[jvc];

    // 1. stress of step n, rows i0-3 .. i0+TX+2, columns j0-3 .. j0+60
#pragma unroll
    for (int p = 0; p < PS; p++) {
        const int r = wave + p * NWAVE;
        if (r < SR) {
            const int i = i0 - 3 + r;
            const bool in = jsin && i >= 0 && i < nx;
            if (PSV) {
                l_sxx[r * W + lane] = in ? s1[p] : 0.0f;
                l_szz[r * W + lane] = in ? s2[p] : 0.0f;
                l_sxz[r * W + lane] = in ? s3[p] : 0.0f;
            } else {
                l_sxy[r * W + lane] = in ? s1[p] : 0.0f;
                l_szy[r * W + lane] = in ? s2[p] : 0.0f;
            }
        }
    }
    __syncthreads();

    int p0 = 0, p1 = 0;
    if (a.inj_ptr) {
        const int set = a.inj_per_slot ? b : 0;
        p0 = a.inj_ptr[set * per + t];
        p1 = a.inj_ptr[set * per + t + 1];
    }
    int r0 = 0, r1 = 0;
    if (a.rec_ptr) {
        r0 = a.rec_ptr[t];
        r1 = a.rec_ptr[t + 1];
    }
    const LanePts ip = lane_pts(a.inj_cell, a.inj_row, p0, p1, lane);
    const LanePts rp = lane_pts(a.rec_cell, a.rec_row, r0, r1, lane);
    const long it = a.it;

    // snapshot fields of this wave type: u at fu (and fu+1), v at fv (and fv+1)
    const int fmt = a.snap_fmt;
    const int fu = PSV ? a.psv_f0 : 0, fv = PSV ? a.psv_f0 + 2 : 1;
    const long sb = (long)b * a.snap_slot_stride;
    // 16-bit snapshots: one exponent per field and row over the tile's CZ
    // owned columns (segment tz of the slot's table, stored by lane 0), plus
    // the offset of the lane's block of 4 columns (lane l owns column l - 1
    // of the segment).  mag: |x| at an owned cell, else 0.  Every lane calls it.
    const bool seg16 = a.snap && fmt != INV_SNAP_F32;
    auto cell_exp = [&](int f, int i, float mag, bool own) {
        const int es = snap_exp(wave_max(mag));
        if (lane == 0) {
            a.sexp[snaptab_row(b, a.nsnapf, f, nx, i, a.nseg) + tz] = (signed char)es;
            *reinterpret_cast<int *>(a.sexp) = CZ;
        }
        unsigned char *blk = reinterpret_cast<unsigned char *>(a.sexp) +
                             snaptab_blk(a.sblk, b, a.nsnapf, f, nx, i, a.nbb) +
                             tz * snap_seg_bytes(CZ);
        return es + snap_blk_put<1>(mag, es, lane, lane - 1, own, CZ, blk);
    };

    // 2. velocity of step n+1, rows i0-1 .. i0+TX+1, columns j0-1 .. j0+CZ+1
    {
        const int j = jv;
        const bool zin = j >= 2 && j < nz - 2;
        const bool jown = lane >= 1 && lane <= CZ;
#pragma unroll
        for (int p = 0; p < PV; p++) {
            const int vr = wave + p * NWAVE;
            if (vr >= VR) break;
            const int i = i0 - 1 + vr;
            const int e = vr * W + lane;
            if (!(i >= 0 && i < nx)) {  // wave-uniform
                if (PSV) {
                    l_vx[e] = 0.0f;
                    l_vz[e] = 0.0f;
                } else {
                    l_vy[e] = 0.0f;
                }
                continue;
            }
            // every lane computes the row (the point-list scans need every
            // lane); lanes outside the grid store 0
            const bool cin = jvin;
            const long kc = (long)i * nz + j;
            const long o = so + kc;
            // cells of this row the wavefront covers: point lists are scanned
            // with one wave-uniform test per point before the per-lane one
            const int rlo = i * nz + (j0 - 1), rhi = rlo + 63;
            const bool xin = i >= 2 && i < nx - 2;
            const bool own = cin && jown && vr >= 1 && vr <= TX;
            const int se = (vr + 2) * W + (lane + 2);
            const float ab = (a.abx[i] * abz) * abt;
            const float rho = qr[p];  // dt / rho in fast mode
            if (!PSV) {
                float d = xin ? d_minus<FAST>(l_sxy, se, W, g.c8dx, g.c24dx, g.a1x, g.a2x) : 0.0f;
                if (zin) d += d_minus<FAST>(l_szy, se, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
                scan_pts(ip, a.inj_cell, a.inj_row, rlo, rhi, [&](int cp, int row) {
                    if (cin && cp == kc)
                        d += a.inj_y[(long)b * a.inj_slot_stride + (long)row * a.nt + it];
                });
                const float v = vel_update<FAST>(ab, qv1[p], g.dt, d, rho);
                l_vy[e] = cin ? v : 0.0f;
                float u = qu1[p];
                int eu = 0, ev = 0;
                if (seg16 && vr >= 1 && vr <= TX) {  // wave-uniform: every lane reduces
                    eu = cell_exp(0, i, own ? fabsf(u) : 0.0f, own);
                    ev = cell_exp(1, i, own ? fabsf(v) : 0.0f, own);
                }
                if (own) {
                    if (a.snap) snap_put(a.snap, fmt, sb + kc, u, eu);
                    u = u_step<FAST>(u, v, g.dt);
                    a.vy1[o] = v;
                    a.uy[o] = u;
                    if (a.snap) snap_put(a.snap, fmt, sb + g.N + kc, v, ev);
                }
                scan_pts(rp, a.rec_cell, a.rec_row, rlo, rhi, [&](int cp, int row) {
                    if (own && cp == kc)
                        a.out_y[(long)b * a.out_slot_stride + (long)row * a.nt + it] =
                            a.rec_disp ? u : v;
                });
            } else {
                float dx_, dz_;
                if (xin) {
                    dx_ = d_minus<FAST>(l_sxx, se, W, g.c8dx, g.c24dx, g.a1x, g.a2x);
                    dz_ = d_minus<FAST>(l_sxz, se, W, g.c8dx, g.c24dx, g.a1x, g.a2x);
                } else {
                    dx_ = 0.0f;
                    dz_ = 0.0f;
                }
                if (zin) {
                    dx_ += d_minus<FAST>(l_sxz, se, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
                    dz_ += d_minus<FAST>(l_szz, se, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
                }
                scan_pts(ip, a.inj_cell, a.inj_row, rlo, rhi, [&](int cp, int row) {
                    if (cin && cp == kc) {
                        const long ks = (long)b * a.inj_slot_stride + (long)row * a.nt + it;
                        dx_ += a.inj_x[ks];
                        dz_ += a.inj_z[ks];
                    }
                });
                const float vx = vel_update<FAST>(ab, qv1[p], g.dt, dx_, rho);
                const float vz = vel_update<FAST>(ab, qv2[p], g.dt, dz_, rho);
                l_vx[e] = cin ? vx : 0.0f;
                l_vz[e] = cin ? vz : 0.0f;
                float ux = qu1[p], uz = qu2[p];
                int ex = 0, ez = 0, evx = 0, evz = 0;
                if (seg16 && vr >= 1 && vr <= TX) {  // wave-uniform: every lane reduces
                    ex = cell_exp(fu, i, own ? fabsf(ux) : 0.0f, own);
                    ez = cell_exp(fu + 1, i, own ? fabsf(uz) : 0.0f, own);
                    evx = cell_exp(fv, i, own ? fabsf(vx) : 0.0f, own);
                    evz = cell_exp(fv + 1, i, own ? fabsf(vz) : 0.0f, own);
                }
                if (own) {
                    if (a.snap) {
                        const long su = sb + (long)fu * g.N + kc;
                        snap_put(a.snap, fmt, su, ux, ex);
                        snap_put(a.snap, fmt, su + g.N, uz, ez);
                    }
                    ux = u_step<FAST>(ux, vx, g.dt);
                    uz = u_step<FAST>(uz, vz, g.dt);
                    a.vx1[o] = vx;
                    a.vz1[o] = vz;
                    a.ux[o] = ux;
                    a.uz[o] = uz;
                    if (a.snap) {
                        const long sv = sb + (long)fv * g.N + kc;
                        snap_put(a.snap, fmt, sv, vx, evx);
                        snap_put(a.snap, fmt, sv + g.N, vz, evz);
                    }
                }
                scan_pts(rp, a.rec_cell, a.rec_row, rlo, rhi, [&](int cp, int row) {
                    if (own && cp == kc) {
                        const long kr = (long)b * a.out_slot_stride + (long)row * a.nt + it;
                        a.out_x[kr] = a.rec_disp ? ux : vx;
                        a.out_z[kr] = a.rec_disp ? uz : vz;
                    }
                });
            }
        }
    }
    __syncthreads();

    // 3. stress of step n+1, rows i0 .. i0+TX-1, columns j0 .. j0+CZ-1 (the
    // tile's explosions, one per lane, loaded while every lane is active)
    int q0 = 0, q1 = 0;
    if (PSV && a.sinj_ptr) {
        q0 = a.sinj_ptr[b * per + t];
        q1 = a.sinj_ptr[b * per + t + 1];
    }
    const LanePts xp = lane_pts(a.sinj_cell, a.sinj_row, q0, q1, lane);
    // every lane runs the rows (the explosion scan needs every lane); lanes
    // with a column of this tile in the grid store
    {
        const int j = jt;
        const bool zin = j >= 1 && j < nz - 2;
#pragma unroll
        for (int p = 0; p < PT; p++) {
            const int tr = wave + p * NWAVE;
            const int i = i0 + tr;
            if (i >= nx) break;
            const long kc = (long)i * nz + j;
            const long o = so + kc;
            const bool xin = i >= 1 && i < nx - 2;
            const int ve = (tr + 1) * W + (lane + 1);
            const int se = (tr + 3) * W + (lane + 3);
            const float mu = qmu[p];
            if (!PSV) {
                const float dx_ = xin ? d_plus<FAST>(l_vy, ve, W, g.c8dx, g.c24dx, g.a1x, g.a2x) : 0.0f;
                const float dz_ = zin ? d_plus<FAST>(l_vy, ve, 1, g.c8dz, g.c24dz, g.a1z, g.a2z) : 0.0f;
                if (jtin) {
                    a.sxy1[o] = s_sh<FAST>(l_sxy[se], g.dt, mu, dx_);
                    a.szy1[o] = s_sh<FAST>(l_szy[se], g.dt, mu, dz_);
                }
            } else {
                float xx, zx, xz, zz;
                if (xin) {
                    xx = d_plus<FAST>(l_vx, ve, W, g.c8dx, g.c24dx, g.a1x, g.a2x);
                    zx = d_plus<FAST>(l_vz, ve, W, g.c8dx, g.c24dx, g.a1x, g.a2x);
                } else {
                    xx = 0.0f;
                    zx = 0.0f;
                }
                if (zin) {
                    xz = d_plus<FAST>(l_vx, ve, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
                    zz = d_plus<FAST>(l_vz, ve, 1, g.c8dz, g.c24dz, g.a1z, g.a2z);
                } else {
                    xz = 0.0f;
                    zz = 0.0f;
                }
                const float lam = qlam[p];
                float nxx, nzz;
                s_normal<FAST>(l_sxx[se], l_szz[se], g.dt, lam, mu, xx, zz, nxx, nzz);
                // explosive sources this tile owns: stf_y into sxx and szz
                const int xlo = i * nz + j0;
                scan_pts(xp, a.sinj_cell, a.sinj_row, xlo, xlo + CZ - 1, [&](int cp, int row) {
                    if (jtin && cp == kc) {
                        const float f = a.inj_y[(long)b * a.inj_slot_stride + (long)row * a.nt + it];
                        nxx += f;
                        nzz += f;
                    }
                });
                if (jtin) {
                    a.sxx1[o] = nxx;
                    a.szz1[o] = nzz;
                    a.sxz1[o] = s_shear<FAST>(l_sxz[se], g.dt, mu, xz, zx);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// output columns per tile along z: 64 for the linear-index kernel, 58 for the
// column kernel (whose stress tile, 58 + 6 halo columns, is one wavefront wide)
// (with 2 or 4 columns per lane: 120 or 248)
static int fz_of(const Ctx &c) {
    if (c.variant == INV_VARIANT_FUSED_LINEAR) return FZ;
    return col_strip(c.col_v);
}
static int tile_rows(const Ctx &c) { return c.tile_x; }
static int tiles_x_of(const Ctx &c, int TX) { return (c.g.nx + TX - 1) / TX; }
static int tiles_z_of(const Ctx &c) { return (c.g.nz + fz_of(c) - 1) / fz_of(c); }
static int list_key(const Ctx &c) { return tile_rows(c) * 1000 + fz_of(c); }

// CSR of points by tile: a point (i, j) belongs to tile (tx, tz) when
// tx*TX - lo <= i < (tx+1)*TX + hi and tz*W - lo <= j < (tz+1)*W + hi, W the
// tile width along z.
// Point lists bucketed by tile (TX rows x FZ columns): a point goes to every
// tile whose rows and columns extended by lo before and hi after contain it.
static void build_tiled(Ctx &c, PointList &pl, int nsets, const std::vector<std::vector<int>> &cells,
                        const std::vector<std::vector<int>> &rows, int TX, int FZ, int lo, int hi) {
    const int ntx = tiles_x_of(c, TX), ntz = (c.g.nz + FZ - 1) / FZ;
    const int ntiles = ntx * ntz;
    auto tiles_of = [&](int cell, std::vector<int> &out) {
        out.clear();
        const int i = cell / c.g.nz, j = cell % c.g.nz;
        for (int tx = std::max(0, (i - hi) / TX - 1); tx <= std::min(ntx - 1, (i + lo) / TX); tx++) {
            if (!(tx * TX - lo <= i && i < (tx + 1) * TX + hi)) continue;
            for (int tz = std::max(0, (j - hi) / FZ - 1); tz <= std::min(ntz - 1, (j + lo) / FZ);
                 tz++) {
                if (!(tz * FZ - lo <= j && j < (tz + 1) * FZ + hi)) continue;
                out.push_back(tx * ntz + tz);
            }
        }
    };
    std::vector<int> count((size_t)nsets * ntiles + 1, 0), tl;
    long total = 0;
    for (int s = 0; s < nsets; s++)
        for (int cell : cells[s]) {
            tiles_of(cell, tl);
            for (int t : tl) count[(size_t)s * ntiles + t + 1]++;
            total += (long)tl.size();
        }
    for (size_t t = 1; t < count.size(); t++) count[t] += count[t - 1];
    std::vector<int> fillp(count.begin(), count.end() - 1);
    std::vector<int> hc(std::max(total, 1L)), hr(std::max(total, 1L));
    for (int s = 0; s < nsets; s++)
        for (size_t p = 0; p < cells[s].size(); p++) {
            tiles_of(cells[s][p], tl);
            for (int t : tl) {
                const int slot = fillp[(size_t)s * ntiles + t]++;
                hc[slot] = cells[s][p];
                hr[slot] = rows[s][p];
            }
        }
    // each bucket in cell order (stable: entries on one cell keep their
    // order, so every cell's sums are unchanged): the walking kernel finds a
    // row's entries by bisection
    {
        std::vector<int> idx, tc, tr;
        for (size_t t = 0; t + 1 < count.size(); t++) {
            const int b0 = count[t], b1 = count[t + 1];
            if (b1 - b0 < 2) continue;
            idx.resize(b1 - b0);
            for (int k = 0; k < b1 - b0; k++) idx[k] = b0 + k;
            std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return hc[x] < hc[y]; });
            tc.resize(idx.size());
            tr.resize(idx.size());
            for (size_t k = 0; k < idx.size(); k++) {
                tc[k] = hc[idx[k]];
                tr[k] = hr[idx[k]];
            }
            std::copy(tc.begin(), tc.end(), hc.begin() + b0);
            std::copy(tr.begin(), tr.end(), hr.begin() + b0);
        }
    }
    // queued launches may still read the old lists: wait before replacing them
    INV_HIP(hipStreamSynchronize(c.stream));
    pl.ptr.ensure(count.size() * sizeof(int));
    pl.cell.ensure(hc.size() * sizeof(int));
    pl.row.ensure(hr.size() * sizeof(int));
    INV_HIP(hipMemcpy(pl.ptr.p, count.data(), count.size() * sizeof(int), hipMemcpyHostToDevice));
    INV_HIP(hipMemcpy(pl.cell.p, hc.data(), hc.size() * sizeof(int), hipMemcpyHostToDevice));
    INV_HIP(hipMemcpy(pl.row.p, hr.data(), hr.size() * sizeof(int), hipMemcpyHostToDevice));
    pl.nsets = nsets;
    pl.ntiles = ntiles;
    pl.npoints = (int)total;
}

void fused_invalidate_lists(Ctx &c) {
    INV_HIP(hipStreamSynchronize(c.stream));
    for (auto &kv : c.src_lists) release_list(kv.second);
    c.src_lists.clear();
    for (auto &kv : c.src_slists) release_list(kv.second);
    c.src_slists.clear();
    c.f_rec_tx = -1;
    c.p_rec_key = -1;
    c.t_rec_key = -1;
}


static void ensure_rec_lists(Ctx &c) {
    if (c.f_rec_tx == list_key(c)) return;
    std::vector<std::vector<int>> cells(1), rows(1);
    for (int r = 0; r < c.nrec; r++) {
        cells[0].push_back(c.rec_ix[r] * c.g.nz + c.rec_iz[r]);
        rows[0].push_back(r);
    }
    build_tiled(c, c.f_rec_v, 1, cells, rows, tile_rows(c), fz_of(c), 1, 2);
    build_tiled(c, c.f_rec_t, 1, cells, rows, tile_rows(c), fz_of(c), 0, 0);
    c.f_rec_tx = list_key(c);
}

// Point forces (type 0) and explosions (type 1) of shots first .. first+count-1
// in batch slots 0 .. count-1 (first < 0: every source in slot 0).
static void source_points(const Ctx &c, int first, int count, int type,
                          std::vector<std::vector<int>> &cells, std::vector<std::vector<int>> &rows) {
    cells.assign(count, {});
    rows.assign(count, {});
    for (int b = 0; b < count; b++)
        for (int is = 0; is < c.nsrc; is++) {
            const int ty = is < (int)c.src_type.size() ? c.src_type[is] : 0;
            if ((first < 0 || is == first + b) && ty == type) {
                cells[b].push_back(c.src_ix[is] * c.g.nz + c.src_iz[is]);
                rows[b].push_back(is);
            }
        }
}

static long long list_id(int first, int count, int key) {
    return ((long long)(first + 2) << 40) | ((long long)count << 20) | key;
}

PointList &source_list(Ctx &c, int first, int count, int key) {
    const long long k = list_id(first, count, key);
    auto found = c.src_lists.find(k);
    if (found != c.src_lists.end()) return found->second;
    PointList &pl = c.src_lists[k];
    std::vector<std::vector<int>> cells, rows;
    source_points(c, first, count, 0, cells, rows);
    if (key == 0) build_point_list(c, pl, count, cells, rows);
    else build_tiled(c, pl, count, cells, rows, tile_rows(c), fz_of(c), 1, 2);
    return pl;
}

PointList &stress_source_list(Ctx &c, int first, int count, int key) {
    const long long k = list_id(first, count, key);
    auto found = c.src_slists.find(k);
    if (found != c.src_slists.end()) return found->second;
    PointList &pl = c.src_slists[k];
    std::vector<std::vector<int>> cells, rows;
    source_points(c, first, count, 1, cells, rows);
    build_tiled(c, pl, count, cells, rows, tile_rows(c), fz_of(c), 0, 0);
    return pl;
}

// Two-step kernel geometry: pair_tx rows x pair_strip(pair_v) columns.  Its
// lists cover the rows and columns each launch updates: forces and adjoint
// sources (velocity of the first step, rows and columns -4 .. +4 inclusive
// around the tile: lo 4, hi 5 past its end), explosions (stress of the first
// step, -3 .. +2: lo 3, hi 3), receivers (owned cells: 0, 0).
static bool pair_on(const Ctx &c) { return c.pair_tx > 0 && c.variant == INV_VARIANT_FUSED; }
static int pair_fz(const Ctx &c) { return pair_strip(c.pair_v); }
static int pair_key(const Ctx &c) { return 500000 + c.pair_tx * 1000 + pair_fz(c); }

static PointList &pair_source_list(Ctx &c, int first, int count) {
    const long long k = list_id(first, count, pair_key(c));
    auto found = c.src_lists.find(k);
    if (found != c.src_lists.end()) return found->second;
    PointList &pl = c.src_lists[k];
    std::vector<std::vector<int>> cells, rows;
    source_points(c, first, count, 0, cells, rows);
    build_tiled(c, pl, count, cells, rows, c.pair_tx, pair_fz(c), 4, 5);
    return pl;
}

static PointList &pair_stress_source_list(Ctx &c, int first, int count) {
    const long long k = list_id(first, count, pair_key(c));
    auto found = c.src_slists.find(k);
    if (found != c.src_slists.end()) return found->second;
    PointList &pl = c.src_slists[k];
    std::vector<std::vector<int>> cells, rows;
    source_points(c, first, count, 1, cells, rows);
    build_tiled(c, pl, count, cells, rows, c.pair_tx, pair_fz(c), 3, 3);
    return pl;
}

static void ensure_pair_rec_lists(Ctx &c) {
    if (c.p_rec_key == pair_key(c)) return;
    std::vector<std::vector<int>> cells(1), rows(1);
    for (int r = 0; r < c.nrec; r++) {
        cells[0].push_back(c.rec_ix[r] * c.g.nz + c.rec_iz[r]);
        rows[0].push_back(r);
    }
    build_tiled(c, c.p_rec_v, 1, cells, rows, c.pair_tx, pair_fz(c), 4, 5);
    build_tiled(c, c.p_rec_t, 1, cells, rows, c.pair_tx, pair_fz(c), 0, 0);
    c.p_rec_key = pair_key(c);
}

// Three-step kernel geometry: trip_tx rows x trip_strip() columns.  Its
// lists cover the rows and columns each launch updates: forces and adjoint
// sources (velocity of the first step, -7 .. +7 around the tile: lo 7, hi 8),
// explosions (stress of the first step, -6 .. +5: lo 6, hi 6), receivers
// (owned cells: 0, 0).
// The walking kernel's chunks take the lists of tiles of walk_ch rows.
static int trip_rows(const Ctx &c) { return c.walk_ch > 0 ? c.walk_ch : c.trip_tx; }
static bool trip_on(const Ctx &c) {
    return trip_rows(c) > 0 && c.variant == INV_VARIANT_FUSED && c.g.nz % 2 == 0;
}
static int trip_key(const Ctx &c) { return 600000 + trip_rows(c) * 1000 + trip_strip(); }

static PointList &trip_source_list(Ctx &c, int first, int count) {
    const long long k = list_id(first, count, trip_key(c));
    auto found = c.src_lists.find(k);
    if (found != c.src_lists.end()) return found->second;
    PointList &pl = c.src_lists[k];
    std::vector<std::vector<int>> cells, rows;
    source_points(c, first, count, 0, cells, rows);
    build_tiled(c, pl, count, cells, rows, trip_rows(c), trip_strip(), 7, 8);
    return pl;
}

static PointList &trip_stress_source_list(Ctx &c, int first, int count) {
    const long long k = list_id(first, count, trip_key(c));
    auto found = c.src_slists.find(k);
    if (found != c.src_slists.end()) return found->second;
    PointList &pl = c.src_slists[k];
    std::vector<std::vector<int>> cells, rows;
    source_points(c, first, count, 1, cells, rows);
    build_tiled(c, pl, count, cells, rows, trip_rows(c), trip_strip(), 6, 6);
    return pl;
}

static void ensure_trip_rec_lists(Ctx &c) {
    if (c.t_rec_key == trip_key(c)) return;
    std::vector<std::vector<int>> cells(1), rows(1);
    for (int r = 0; r < c.nrec; r++) {
        cells[0].push_back(c.rec_ix[r] * c.g.nz + c.rec_iz[r]);
        rows[0].push_back(r);
    }
    build_tiled(c, c.t_rec_v, 1, cells, rows, trip_rows(c), trip_strip(), 7, 8);
    build_tiled(c, c.t_rec_t, 1, cells, rows, trip_rows(c), trip_strip(), 0, 0);
    c.t_rec_key = trip_key(c);
}

static void ensure_second_set(Ctx &c) {
    const size_t b = (size_t)c.g.batch * c.N * sizeof(float);
    if (c.g.psv) {
        c.vx_b.ensure(b); c.vz_b.ensure(b);
        c.sxx_b.ensure(b); c.szz_b.ensure(b); c.sxz_b.ensure(b);
    }
    if (c.g.sh) {
        c.vy_b.ensure(b); c.sxy_b.ensure(b); c.szy_b.ensure(b);
    }
}

// ---------------------------------------------------------------------------
// Snapshot storage.  Forward snapshots for the adjoint hold FP32, FP16, BF16
// or I16 values (16-bit ones scaled per row segment by a power of two, with
// an exponent table per slot).  They cover
// either every snapshot step of the run (ckpt = 0) or, with checkpoints every
// ckpt steps, one ckpt-step segment: the forward keeps the FP32 state at each
// segment start plus the snapshots of its last segment, and the adjoint
// recomputes the forward steps of each earlier segment from its checkpoint
// just before it needs that segment's snapshots.
// ---------------------------------------------------------------------------
static size_t snap_esize(const Ctx &c) { return c.snap_format == INV_SNAP_F32 ? 4 : 2; }
static int nstate_fields(const Ctx &c) { return (c.g.sh ? 4 : 0) + (c.g.psv ? 7 : 0); }
static int nreplay_fields(const Ctx &c) { return (c.g.sh ? 7 : 0) + (c.g.psv ? 12 : 0); }

void snap_storage(Ctx &c) {
    if (!c.g.adj || c.nsfe <= 0) return;
    const size_t B = (size_t)c.g.batch, N = (size_t)c.N;
    c.nslots = c.ckpt > 0 ? c.ckpt / c.g.sfe : c.nsfe;
    c.snap.ensure(B * (size_t)c.nslots * c.nsnapf * N * snap_esize(c));
    if (c.snap_format != INV_SNAP_F32)
        c.sexp.ensure((size_t)c.nslots * snap_tab_bytes(c.g.batch, c.nsnapf, c.g.nx, c.g.nz));
    if (c.ckpt > 0) {
        const size_t nck = (size_t)(c.g.nt / c.ckpt - 1);
        if (nck > 0) c.ckbuf.ensure(nck * nstate_fields(c) * B * N * sizeof(float));
        c.rstate.ensure((size_t)nreplay_fields(c) * B * N * sizeof(float));
    }
}

// one wavefield set: v and s double-buffered, u in place
struct Fields {
    float *vx[2], *vz[2], *sxx[2], *szz[2], *sxz[2], *vy[2], *sxy[2], *szy[2];
    float *ux, *uz, *uy;
};

static Fields main_fields(Ctx &c) {
    Fields F{};
    auto two = [](DevBuf &a, DevBuf &b, float **out) {
        out[0] = a.as<float>();
        out[1] = b.as<float>();
    };
    two(c.vx, c.vx_b, F.vx); two(c.vz, c.vz_b, F.vz);
    two(c.sxx, c.sxx_b, F.sxx); two(c.szz, c.szz_b, F.szz); two(c.sxz, c.sxz_b, F.sxz);
    two(c.vy, c.vy_b, F.vy); two(c.sxy, c.sxy_b, F.sxy); two(c.szy, c.szy_b, F.szy);
    F.ux = c.ux.as<float>();
    F.uz = c.uz.as<float>();
    F.uy = c.uy.as<float>();
    return F;
}

// the recomputation's own wavefields, carved from one buffer
static Fields replay_fields(Ctx &c) {
    Fields F{};
    float *p = c.rstate.as<float>();
    const long n = (long)c.g.batch * c.N;
    auto take = [&]() {
        float *q = p;
        p += n;
        return q;
    };
    if (c.g.psv) {
        F.vx[0] = take(); F.vx[1] = take(); F.vz[0] = take(); F.vz[1] = take();
        F.sxx[0] = take(); F.sxx[1] = take(); F.szz[0] = take(); F.szz[1] = take();
        F.sxz[0] = take(); F.sxz[1] = take(); F.ux = take(); F.uz = take();
    }
    if (c.g.sh) {
        F.vy[0] = take(); F.vy[1] = take(); F.sxy[0] = take(); F.sxy[1] = take();
        F.szy[0] = take(); F.szy[1] = take(); F.uy = take();
    }
    return F;
}

// current fields of a set in checkpoint order: SH {vy, uy, sxy, szy}, then
// P-SV {vx, vz, ux, uz, sxx, szz, sxz}
static int state_list(const Ctx &c, const Fields &F, int par, float **out) {
    const int p = par & 1;
    int n = 0;
    if (c.g.sh) {
        out[n++] = F.vy[p]; out[n++] = F.uy; out[n++] = F.sxy[p]; out[n++] = F.szy[p];
    }
    if (c.g.psv) {
        out[n++] = F.vx[p]; out[n++] = F.vz[p]; out[n++] = F.ux; out[n++] = F.uz;
        out[n++] = F.sxx[p]; out[n++] = F.szz[p]; out[n++] = F.sxz[p];
    }
    return n;
}

// checkpoint ck >= 1 holds the state before step ck * ckpt
static float *ckpt_field(Ctx &c, int ck, int f) {
    return c.ckbuf.as<float>() + ((long)(ck - 1) * nstate_fields(c) + f) * (long)c.g.batch * c.N;
}

static void set_state_ptrs(StepArgs &a, const Fields &F, int par) {
    const int p0 = par & 1, p1 = p0 ^ 1;
    a.vx0 = F.vx[p0]; a.vx1 = F.vx[p1];
    a.vz0 = F.vz[p0]; a.vz1 = F.vz[p1];
    a.sxx0 = F.sxx[p0]; a.sxx1 = F.sxx[p1];
    a.szz0 = F.szz[p0]; a.szz1 = F.szz[p1];
    a.sxz0 = F.sxz[p0]; a.sxz1 = F.sxz[p1];
    a.vy0 = F.vy[p0]; a.vy1 = F.vy[p1];
    a.sxy0 = F.sxy[p0]; a.sxy1 = F.sxy[p1];
    a.szy0 = F.szy[p0]; a.szy1 = F.szy[p1];
    a.ux = F.ux;
    a.uz = F.uz;
    a.uy = F.uy;
}

// Where forward step it's snapshot goes: slot isfe = nsfe - (it+1)/sfe of the
// run (fdm.h:521-560), held at local index isfe - base when that lies in the
// buffer.
struct SnapPlan {
    bool on = false;
    int fmt = INV_SNAP_F32, sfe = 1, nsfe = 0, nslots = 0, base = 0;
    char *snap = nullptr;
    signed char *sexp = nullptr;
    long slot_bytes = 0;  // one snapshot slot of batch slot 0
    long tab_bytes = 0;   // one slot's exponent table
    int local(long it) const {
        if (!on || it < 0 || (it + 1) % sfe != 0) return -1;
        const long l = nsfe - (it + 1) / sfe - base;
        return (l >= 0 && l < nslots) ? (int)l : -1;
    }
    void *slot(int l) const { return snap + l * slot_bytes; }
    signed char *tab(int l) const { return sexp ? sexp + l * tab_bytes : nullptr; }
};

static SnapPlan snap_plan(Ctx &c, bool on, int base) {
    SnapPlan s;
    s.fmt = c.snap_format;
    s.on = on && c.snap.p && (s.fmt == INV_SNAP_F32 || c.sexp.p) && c.g.sfe > 0;
    s.sfe = c.g.sfe > 0 ? c.g.sfe : 1;
    s.nsfe = c.nsfe;
    s.nslots = c.nslots;
    s.base = base;
    s.snap = c.snap.as<char>();
    s.sexp = s.fmt == INV_SNAP_F32 ? nullptr : c.sexp.as<signed char>();
    s.slot_bytes = (long)c.nsnapf * c.N * (long)snap_esize(c);
    s.tab_bytes = (long)snap_tab_bytes(c.g.batch, c.nsnapf, c.g.nx, c.g.nz);
    return s;
}

// Snapshot role of forward step it: u before and v after the step go to the
// step's slot (and, for 16-bit formats, the exponents of their row segments
// to the slot's table) in the step itself.
static void snap_roles(const SnapPlan &s, StepArgs &a, long it) {
    a.snap = nullptr;
    a.sexp = nullptr;
    const int l = s.local(it);
    if (l >= 0) {
        a.snap = s.slot(l);
        a.sexp = s.tab(l);
    }
}

static bool res_on(const Ctx &c);
static void res_prepare(Ctx &c, int count);

void fused_prepare(Ctx &c, int kind, int first, int count, int flags) {
    ensure_second_set(c);
    if (res_on(c)) res_prepare(c, count);
    ensure_rec_lists(c);
    // snapshot storage only for runs that store or read snapshots
    if (kind == 2 || (flags & INV_FWD_SNAPSHOTS)) snap_storage(c);
    // the adjoint's recomputation injects the forward's sources
    if (kind == 1 || (c.ckpt > 0 && c.g.adj)) {
        source_list(c, first, count, list_key(c));
        if (c.any_explosive()) stress_source_list(c, first, count, list_key(c));
    }
    if (pair_on(c)) {
        ensure_pair_rec_lists(c);
        if (kind == 1 || (c.ckpt > 0 && c.g.adj)) {
            pair_source_list(c, first, count);
            if (c.any_explosive()) pair_stress_source_list(c, first, count);
        }
    }
    if (trip_on(c)) {
        ensure_trip_rec_lists(c);
        if (kind == 1) {
            trip_source_list(c, first, count);
            if (c.any_explosive()) trip_stress_source_list(c, first, count);
        }
    }
}

template <bool SH, bool PSV, bool FAST>
static void launch_lin(Ctx &c, const StepArgs &a, unsigned blocks) {
    switch (c.tile_x) {
        case 8: k_step<SH, PSV, 8, FAST><<<blocks, FBLOCK, 0, c.stream>>>(a); break;
        case 32: k_step<SH, PSV, 32, FAST><<<blocks, FBLOCK, 0, c.stream>>>(a); break;
        default: k_step<SH, PSV, 16, FAST><<<blocks, FBLOCK, 0, c.stream>>>(a); break;
    }
}

template <bool SH, bool PSV, bool FAST>
static void launch_col(Ctx &c, const StepArgs &a, unsigned blocks) {
    switch (c.tile_x) {
        case 8: k_col<SH, PSV, 8, FAST><<<blocks, FBLOCK, 0, c.stream>>>(a); break;
        case 32: k_col<SH, PSV, 32, FAST><<<blocks, FBLOCK, 0, c.stream>>>(a); break;
        default: k_col<SH, PSV, 16, FAST><<<blocks, FBLOCK, 0, c.stream>>>(a); break;
    }
}

template <bool FAST>
static void launch_m(Ctx &c, const StepArgs &a, unsigned blocks) {
    if (c.variant == INV_VARIANT_FUSED && c.col_v > 1) {
        if (c.g.sh) wide_launch<true, false, FAST>(c, a, blocks);
        if (c.g.psv) wide_launch<false, true, FAST>(c, a, blocks);
    } else if (c.variant == INV_VARIANT_FUSED_LINEAR) {
        if (c.g.sh && c.g.psv) launch_lin<true, true, FAST>(c, a, blocks);
        else if (c.g.sh) launch_lin<true, false, FAST>(c, a, blocks);
        else launch_lin<false, true, FAST>(c, a, blocks);
    } else {
        // SH and P-SV are independent systems: one launch each per step
        if (c.g.sh) launch_col<true, false, FAST>(c, a, blocks);
        if (c.g.psv) launch_col<false, true, FAST>(c, a, blocks);
    }
}

static void fused_step(Ctx &c, StepArgs &a, const Fields &F, int &par, int count, int it) {
    set_state_ptrs(a, F, par);
    a.it = it;
    const unsigned blocks = (unsigned)(count * a.tiles_x * a.tiles_z);
    hipEvent_t ev;
    const bool p = prof_begin(c, it, &ev);
    if (c.math == INV_MATH_FAST) launch_m<true>(c, a, blocks);
    else launch_m<false>(c, a, blocks);
    INV_HIP(hipGetLastError());
    if (p) prof_end(c, PROF_STEP, ev);
    par ^= 1;
}

template <bool FAST>
static void pair_m(Ctx &c, const StepArgs &a, const StepRoles &r2, unsigned blocks) {
    // SH and P-SV are independent systems: one launch each
    if (c.g.sh) pair_launch<true, false, FAST>(c, a, r2, blocks);
    if (c.g.psv) pair_launch<false, true, FAST>(c, a, r2, blocks);
}

template <bool FAST>
static void trip_m(Ctx &c, const StepArgs &a, unsigned blocks) {
    if (c.walk_ch > 0) {
        if (c.g.sh) walk_launch<true, false, FAST>(c, a, blocks);
        if (c.g.psv) walk_launch<false, true, FAST>(c, a, blocks);
        return;
    }
    if (c.g.sh) trip_launch<true, false, FAST>(c, a, blocks);
    if (c.g.psv) trip_launch<false, true, FAST>(c, a, blocks);
}

// Steps it, it+1 and it+2 in one launch of the three-step kernel (no
// snapshot roles).
static void trip_step(Ctx &c, StepArgs &a, const Fields &F, int &par, int count, int it) {
    set_state_ptrs(a, F, par);
    a.it = it;
    const unsigned blocks = (unsigned)(count * a.tiles_x * a.tiles_z);
    hipEvent_t ev;
    const bool p = prof_begin(c, it, &ev);
    if (c.math == INV_MATH_FAST) trip_m<true>(c, a, blocks);
    else trip_m<false>(c, a, blocks);
    INV_HIP(hipGetLastError());
    if (p) prof_end(c, PROF_TRIP, ev);
    par ^= 1;
}

// Steps it and it+1 in one launch of the two-step kernel (roles of step it
// in a, of step it+1 in r2).
static void pair_step(Ctx &c, StepArgs &a, const StepRoles &r2, const Fields &F, int &par,
                      int count, int it) {
    set_state_ptrs(a, F, par);
    a.it = it;
    const unsigned blocks = (unsigned)(count * a.tiles_x * a.tiles_z);
    hipEvent_t ev;
    const bool p = prof_begin(c, it, &ev);
    if (c.math == INV_MATH_FAST) pair_m<true>(c, a, r2, blocks);
    else pair_m<false>(c, a, r2, blocks);
    INV_HIP(hipGetLastError());
    if (p) prof_end(c, PROF_PAIR, ev);
    par ^= 1;
}

static StepRoles roles_of(const StepArgs &a) { return StepRoles{a.snap, a.sexp}; }

static void set_roles(StepArgs &a, const StepRoles &r) {
    a.snap = r.snap;
    a.sexp = r.sexp;
}

static void common_args(Ctx &c, StepArgs &a) {
    a = StepArgs{};
    a.g = make_geo(c);
    a.tiles_x = tiles_x_of(c, tile_rows(c));
    a.tiles_z = tiles_z_of(c);
    a.rho = c.math == INV_MATH_FAST ? c.bdt.as<float>() : c.rho.as<float>();
    a.lam = c.lam.as<float>();
    a.mu = c.mu.as<float>();
    a.slot_model = c.slot_models ? c.slot_model.as<const float *const>() : nullptr;
    a.abx = c.abx.as<float>();
    a.abzb = c.abzb.as<float>();
    a.abzt = c.abzt.as<float>();
    a.nt = c.g.nt;
    a.snap_fmt = c.snap_format;
    a.nsnapf = c.nsnapf;
    a.psv_f0 = c.g.sh ? 2 : 0;
    a.snap_slot_stride = (long)c.nslots * c.nsnapf * c.N;
    a.nseg = snap_nseg(c.g.nz);
    a.nbb = snap_nbb(c.g.nz);
    a.sblk = snap_blk_off(c.g.batch, c.nsnapf, c.g.nx, c.g.nz);
    INV_REQUIRE(a.rho && a.lam && a.mu && a.abx && a.abzb && a.abzt, INV_ERR_STATE,
                "model or absorbing profile not allocated");
    INV_REQUIRE((!c.g.psv || (c.ux.p && c.uz.p && c.vx.p && c.vx_b.p && c.sxx_b.p)) &&
                    (!c.g.sh || (c.uy.p && c.vy.p && c.vy_b.p && c.sxy_b.p)),
                INV_ERR_STATE, "wavefields not allocated");
}

// sources of shots first.. as the forward injects them; receivers recorded
// unless record is false (recomputation)
static void forward_args(Ctx &c, StepArgs &a, int first, int count, bool record) {
    PointList &sl = source_list(c, first, count, list_key(c));
    const long nt = c.g.nt, nsrc = c.nsrc;
    a.inj_x = c.stf.as<float>();
    a.inj_y = c.stf.as<float>() + nsrc * nt;
    a.inj_z = c.stf.as<float>() + 2 * nsrc * nt;
    a.inj_slot_stride = 0;
    a.inj_ptr = sl.ptr.as<int>();
    a.inj_cell = sl.cell.as<int>();
    a.inj_row = sl.row.as<int>();
    a.inj_per_slot = 1;
    a.sinj_ptr = a.sinj_cell = a.sinj_row = nullptr;
    if (c.any_explosive() && c.g.psv) {
        PointList &xl = stress_source_list(c, first, count, list_key(c));
        a.sinj_ptr = xl.ptr.as<int>();
        a.sinj_cell = xl.cell.as<int>();
        a.sinj_row = xl.row.as<int>();
    }
    const long tr = (long)c.nrec * nt;
    const int cy = 0, cx = c.g.sh ? 1 : 0, cz = cx + 1;
    a.out_y = c.out.as<float>() + cy * tr;
    a.out_x = c.out.as<float>() + cx * tr;
    a.out_z = c.out.as<float>() + cz * tr;
    a.out_slot_stride = (long)c.ncomp * tr;
    a.rec_ptr = (record && c.nrec) ? c.f_rec_t.ptr.as<int>() : nullptr;
    a.rec_cell = c.f_rec_t.cell.as<int>();
    a.rec_row = c.f_rec_t.row.as<int>();
    a.rec_disp = (c.g.adj || c.g.trace_type == 1) ? 1 : 0;
}

// Arguments of the two-step kernel: those of the one-step kernel with the
// pair geometry's tiles and point lists (built by fused_prepare).
static StepArgs pair_args(Ctx &c, const StepArgs &a, int first, int count, bool adjoint,
                          bool record) {
    StepArgs p = a;
    p.tiles_x = tiles_x_of(c, c.pair_tx);
    p.tiles_z = (c.g.nz + pair_fz(c) - 1) / pair_fz(c);
    ensure_pair_rec_lists(c);
    if (adjoint) {
        p.inj_ptr = c.nrec ? c.p_rec_v.ptr.as<int>() : nullptr;
        p.inj_cell = c.p_rec_v.cell.as<int>();
        p.inj_row = c.p_rec_v.row.as<int>();
        p.rec_ptr = nullptr;
        return p;
    }
    PointList &sl = pair_source_list(c, first, count);
    p.inj_ptr = sl.ptr.as<int>();
    p.inj_cell = sl.cell.as<int>();
    p.inj_row = sl.row.as<int>();
    if (a.sinj_ptr) {
        PointList &xl = pair_stress_source_list(c, first, count);
        p.sinj_ptr = xl.ptr.as<int>();
        p.sinj_cell = xl.cell.as<int>();
        p.sinj_row = xl.row.as<int>();
    }
    p.rec_ptr = (record && c.nrec) ? c.p_rec_t.ptr.as<int>() : nullptr;
    p.rec_cell = c.p_rec_t.cell.as<int>();
    p.rec_row = c.p_rec_t.row.as<int>();
    return p;
}

// Arguments of the three-step kernel: those of the one-step kernel with the
// triple geometry's tiles and point lists (built by fused_prepare).
static StepArgs trip_args(Ctx &c, const StepArgs &a, int first, int count, bool adjoint,
                          bool record) {
    StepArgs p = a;
    p.snap = nullptr;
    p.sexp = nullptr;
    p.tiles_x = tiles_x_of(c, trip_rows(c));
    p.tiles_z = (c.g.nz + trip_strip() - 1) / trip_strip();
    ensure_trip_rec_lists(c);
    if (adjoint) {
        p.inj_ptr = c.nrec ? c.t_rec_v.ptr.as<int>() : nullptr;
        p.inj_cell = c.t_rec_v.cell.as<int>();
        p.inj_row = c.t_rec_v.row.as<int>();
        p.rec_ptr = nullptr;
        return p;
    }
    PointList &sl = trip_source_list(c, first, count);
    p.inj_ptr = sl.ptr.as<int>();
    p.inj_cell = sl.cell.as<int>();
    p.inj_row = sl.row.as<int>();
    if (a.sinj_ptr) {
        PointList &xl = trip_stress_source_list(c, first, count);
        p.sinj_ptr = xl.ptr.as<int>();
        p.sinj_cell = xl.cell.as<int>();
        p.sinj_row = xl.row.as<int>();
    }
    p.rec_ptr = (record && c.nrec) ? c.t_rec_t.ptr.as<int>() : nullptr;
    p.rec_cell = c.t_rec_t.cell.as<int>();
    p.rec_row = c.t_rec_t.row.as<int>();
    return p;
}

// Persistent runs (fdm_resident.hip): every pair of steps of a forward, or
// of an adjoint between kernel accumulations, in one launch.
static bool res_on(const Ctx &c) { return c.persist != 0 && pair_on(c); }

static void res_require(const Ctx &c) {
    INV_REQUIRE((c.pair_tx == 4 || c.pair_tx == 8) && c.pair_v == RES_V, INV_ERR_UNSUPPORTED,
                "the persistent kernel runs the two-step tiles of 4 or 8 rows x 4 columns per "
                "lane");
    INV_REQUIRE(c.g.nz <= pair_strip(RES_V) && c.g.nz % RES_V == 0, INV_ERR_UNSUPPORTED,
                "the persistent kernel holds one tile across z: nz <= 240 and divisible by 4");
    INV_REQUIRE(c.ckpt == 0, INV_ERR_UNSUPPORTED,
                "the persistent kernel does not run checkpointed snapshots");
}

static void res_prepare(Ctx &c, int count) {
    res_require(c);
    if (!c.res_ctl.p) {
        c.res_ctl.ensure(2 * sizeof(unsigned));
        const unsigned z[2] = {0, 0};
        INV_HIP(hipMemcpy(c.res_ctl.p, z, sizeof z, hipMemcpyHostToDevice));
    }
    const size_t need = res_xbuf_bytes(count, c.g.nx, c.pair_tx);
    if (c.xbuf.bytes < need) {
        c.xbuf.ensure(need);
        // tag 0 is never expected: a fresh buffer holds no published flag
        fill_dev(c, c.xbuf.p, (long)(c.xbuf.bytes / sizeof(float)));
    }
}

template <bool FAST>
static void res_m(Ctx &c, const StepArgs &a, const ResArgs &ra, unsigned blocks) {
    // SH and P-SV are independent systems: one launch each
    if (c.g.sh) res_launch<true, false, FAST>(c, a, ra, blocks);
    if (c.g.psv) res_launch<false, true, FAST>(c, a, ra, blocks);
}

// iters iterations (steps it .. it + 2 iters - 1) from the current state set
// into the other one; snapshot roles from sp when it is on, or (adjoint, inl)
// the kernel accumulations of the steps that are multiples of sfe in the
// launch, reading the snapshots of plan sp
static void res_run(Ctx &c, StepArgs &a2, const SnapPlan &sp, const Fields &F, int count,
                    int it, int iters, bool inl = false) {
    set_state_ptrs(a2, F, c.parity);
    a2.it = it;
    set_roles(a2, StepRoles{nullptr, nullptr});
    ResArgs ra{};
    ra.xbuf = c.xbuf.as<float>();
    ra.xflag = reinterpret_cast<unsigned *>(ra.xbuf + res_xbuf_floats(count, c.g.nx, c.pair_tx));
    ra.ctl = c.res_ctl.as<unsigned>();
    ra.iters = iters;
    ra.spin_limit = c.persist == 2 ? -1 : RES_SPIN;
    if (inl) {
        // the accumulators of fdm_interact (fdm.hip): per-slot ones for a
        // batched adjoint, the kernels themselves for one slot
        ra.int_on = 1;
        ra.ndt = (float)c.g.sfe * c.g.dt;
        ra.inv_lambda = c.g.inv_lambda;
        ra.inv_mu = c.g.inv_mu;
        ra.inv_rho = c.g.inv_rho;
        if (c.g.batch == 1) {
            ra.kl = c.k_lambda.as<float>();
            ra.km = c.k_mu.as<float>();
            ra.kr = c.k_rho.as<float>();
            ra.k_slot_stride = 0;
        } else {
            ra.kl = c.kslot.as<float>();
            ra.km = c.kslot.as<float>() + c.N;
            ra.kr = c.kslot.as<float>() + 2 * c.N;
            ra.k_slot_stride = 3 * c.N;
        }
        INV_REQUIRE(sp.snap && (sp.fmt == INV_SNAP_F32 || sp.sexp) && c.nslots * c.g.sfe >= c.g.nt,
                    INV_ERR_STATE, "snapshots of the run not in the buffer");
    }
    if (sp.on || inl) {
        ra.snap = sp.snap;
        ra.sexp = sp.sexp;
        ra.slot_bytes = sp.slot_bytes;
        ra.tab_bytes = sp.tab_bytes;
        ra.sfe = sp.sfe;
        ra.nsfe = sp.nsfe;
        ra.nslots = sp.nslots;
        ra.sbase = sp.base;
    }
    INV_REQUIRE(c.xbuf.bytes >= res_xbuf_bytes(count, c.g.nx, c.pair_tx) && c.res_ctl.p, INV_ERR_STATE,
                "persistent kernel buffers not prepared");
    const unsigned blocks = (unsigned)(count * a2.tiles_x);
    hipEvent_t ev;
    const bool p = prof_begin_any(c, &ev);
    if (c.math == INV_MATH_FAST) res_m<true>(c, a2, ra, blocks);
    else res_m<false>(c, a2, ra, blocks);
    INV_HIP(hipGetLastError());
    if (p) prof_end(c, PROF_RES, ev, 2 * iters);
    c.parity ^= 1;
}

void fused_forward(Ctx &c, int first, int count, int flags, int it0, int it1) {
    fused_prepare(c, 1, first, count, flags);
    StepArgs a;
    common_args(c, a);
    forward_args(c, a, first, count, true);
    const bool store = (flags & INV_FWD_SNAPSHOTS) && c.g.adj && c.g.sfe > 0 && c.snap.p;
    const SnapPlan sp = snap_plan(c, store, 0);
    const Fields F = main_fields(c);
    if (it0 == 0) {
        c.parity = 0;
        fdm_zero_state(c, count);
        fdm_prepare_model(c);
    }
    // two steps per launch where the steps allow it: no checkpoint before the
    // second step
    const bool pr = pair_on(c);
    StepArgs a2;
    if (pr) a2 = pair_args(c, a, first, count, false, true);
    // three steps per launch where none of them stores a snapshot (runs with
    // checkpoints keep the two-step plan)
    const bool tr = trip_on(c) && !(sp.on && c.ckpt > 0);
    StepArgs a3;
    if (tr) a3 = trip_args(c, a, first, count, false, true);
    if (res_on(c)) {
        // every pair of steps in persistent launches, an odd last step alone
        int it = it0;
        while (it + 1 < it1) {
            const int K = std::min((it1 - it) / 2, RES_KMAX);
            res_run(c, a2, sp, F, count, it, K);
            it += 2 * K;
        }
        if (it < it1) {
            snap_roles(sp, a, it);
            fused_step(c, a, F, c.parity, count, it);
        }
        return;
    }
    for (int it = it0; it < it1;) {
        if (sp.on && c.ckpt > 0 && it > 0 && it % c.ckpt == 0) {
            float *cur[11];
            const int nf = state_list(c, F, c.parity, cur);
            for (int f = 0; f < nf; f++)
                vec_copy(c, ckpt_field(c, it / c.ckpt, f), cur[f], (size_t)count * c.N);
        }
        if (tr && it + 2 < it1 && sp.local(it) < 0 && sp.local(it + 1) < 0 &&
            sp.local(it + 2) < 0) {
            trip_step(c, a3, F, c.parity, count, it);
            it += 3;
            continue;
        }
        snap_roles(sp, a, it);
        if (pr && it + 1 < it1 && !(sp.on && c.ckpt > 0 && (it + 1) % c.ckpt == 0)) {
            StepArgs n = a;
            snap_roles(sp, n, it + 1);
            set_roles(a2, roles_of(a));
            pair_step(c, a2, roles_of(n), F, c.parity, count, it);
            it += 2;
        } else {
            fused_step(c, a, F, c.parity, count, it);
            it += 1;
        }
    }
}

// Recompute forward segment fseg (steps fseg*ckpt ..) from its checkpoint,
// storing its snapshots for adjoint segment aseg.
static void replay_segment(Ctx &c, int first, int count, int fseg, int aseg) {
    const int K = c.ckpt;
    const long f0 = (long)fseg * K;
    const SnapPlan sp = snap_plan(c, true, aseg * (K / c.g.sfe));
    const Fields R = replay_fields(c);
    int par = 0;
    float *dst[11];
    const int nf = state_list(c, R, 0, dst);
    const size_t n = (size_t)count * c.N;
    for (int f = 0; f < nf; f++) {
        if (fseg == 0) fill_dev(c, dst[f], (long)n);
        else vec_copy(c, dst[f], ckpt_field(c, fseg, f), n);
    }
    StepArgs a;
    common_args(c, a);
    forward_args(c, a, first, count, false);
    const bool pr = pair_on(c);
    StepArgs a2;
    if (pr) a2 = pair_args(c, a, first, count, false, false);
    for (long it = f0; it < f0 + K;) {
        snap_roles(sp, a, it);
        if (pr && it + 1 < f0 + K) {
            StepArgs n = a;
            snap_roles(sp, n, it + 1);
            set_roles(a2, roles_of(a));
            pair_step(c, a2, roles_of(n), R, par, count, (int)it);
            it += 2;
        } else {
            fused_step(c, a, R, par, count, (int)it);
            it += 1;
        }
    }
}

void fused_adjoint(Ctx &c, int first, int count, int it0, int it1) {
    fused_prepare(c, 2, first, count, 0);
    StepArgs a;
    common_args(c, a);
    const long nt = c.g.nt;
    const long tr = (long)c.nrec * nt;
    const int cy = 0, cx = c.g.sh ? 1 : 0, cz = cx + 1;
    a.inj_y = c.adstf.as<float>() + cy * tr;
    a.inj_x = c.adstf.as<float>() + cx * tr;
    a.inj_z = c.adstf.as<float>() + cz * tr;
    a.inj_slot_stride = (long)c.ncomp * tr;
    a.inj_ptr = c.nrec ? c.f_rec_v.ptr.as<int>() : nullptr;
    a.inj_cell = c.f_rec_v.cell.as<int>();
    a.inj_row = c.f_rec_v.row.as<int>();
    a.inj_per_slot = 0;
    a.rec_ptr = nullptr;
    const Fields F = main_fields(c);
    const bool direct = c.g.batch == 1;
    if (it0 == 0) {
        if (!direct) fill_dev(c, c.kslot.p, 3L * count * c.N);
        c.parity = 0;
        fdm_zero_state(c, count);
        fdm_prepare_model(c);
    }
    const int K = c.ckpt, S = K > 0 ? c.g.nt / K : 1, m = K > 0 ? K / c.g.sfe : 0;
    // two steps per launch where no kernel accumulation follows the first;
    // a segment recomputation due before the second step runs before both
    // (the interactions that read the old segment are done by then)
    const bool pr = pair_on(c) && c.g.sfe > 1;
    StepArgs a2;
    if (pr) a2 = pair_args(c, a, first, count, true, false);
    const bool res = res_on(c) && pr;
    // three steps per launch where no kernel accumulation follows the first
    // two (runs without checkpoints); r steps up to and including the next
    // accumulation split as triples and pairs, never leaving one step alone
    // (a one-step launch moves as many bytes as a two-step one): 10 = 3+3+2+2
    const bool three = trip_on(c) && K == 0 && c.g.sfe > 2 && !res;
    StepArgs a3;
    if (three) a3 = trip_args(c, a, first, count, true, false);
    for (int it = it0; it < it1;) {
        const int r = (c.g.sfe - it % c.g.sfe) % c.g.sfe + 1;
        if (three && it + 2 < it1 && r >= 3 && r != 4) {
            trip_step(c, a3, F, c.parity, count, it);
            const int last = it + 2;
            if (last % c.g.sfe == 0)
                fdm_interact(c, count, last, last / c.g.sfe, cur_field(c, c.vx, c.vx_b),
                             cur_field(c, c.vz, c.vz_b), cur_field(c, c.vy, c.vy_b));
            it = last + 1;
            continue;
        }
        const bool two = pr && it + 1 < it1 && it % c.g.sfe != 0;
        if (two && res) {
            // every pair up to the next step that cannot pair (one whose
            // kernel accumulation would fall between its two steps) in one
            // launch, with the accumulations of the pairs' second steps in it
            // (one wave type: the accumulation order of a cell is then that
            // of per-pair launches); with SH and P-SV together, up to the
            // next accumulation, which fdm_interact runs after the launch
            const bool inl = !(c.g.sh && c.g.psv);
            int iters = 0, nxt = it, last = it;
            while (nxt + 1 < it1 && nxt % c.g.sfe != 0 && iters < RES_KMAX) {
                iters++;
                last = nxt + 1;
                nxt += 2;
                if (!inl && last % c.g.sfe == 0) break;
            }
            res_run(c, a2, inl ? snap_plan(c, true, 0) : SnapPlan{}, F, count, it, iters, inl);
            if (!inl && last % c.g.sfe == 0)
                fdm_interact(c, count, last, last / c.g.sfe, cur_field(c, c.vx, c.vx_b),
                             cur_field(c, c.vz, c.vz_b), cur_field(c, c.vy, c.vy_b));
            it = last + 1;
            continue;
        }
        const int last = two ? it + 1 : it;
        if (K > 0 && last > 0 && last % K == 0)
            replay_segment(c, first, count, S - 1 - last / K, last / K);
        if (two) pair_step(c, a2, StepRoles{}, F, c.parity, count, it);
        else fused_step(c, a, F, c.parity, count, it);
        if (last % c.g.sfe == 0) {
            const int slot = last / c.g.sfe - (K > 0 ? (last / K) * m : 0);
            fdm_interact(c, count, last, slot, cur_field(c, c.vx, c.vx_b),
                         cur_field(c, c.vz, c.vz_b), cur_field(c, c.vy, c.vy_b));
        }
        it = last + 1;
    }
    if (!direct && it1 == c.g.nt) fdm_kslot_reduce(c, count);
}

}  // namespace inv
