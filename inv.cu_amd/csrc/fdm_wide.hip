// Wide column kernel: the fused time step of k_col (divS -> addSTF -> updateV
// -> divV -> updateS -> saveRec, solver/fdm.h:534-551) with V = 2 or 4
// adjacent columns per lane, so that every global access of a wavefront is a
// float2 / float4 vector covering 128 or 256 contiguous columns of a row.
// Grids that do not fit in the caches are bound by how well HBM serves this
// access pattern; for the step's 17 row streams, 2-4 columns per lane move
// about 20 % more bytes per second than one.
//
// Tile: TX rows x (64 V - 8) owned columns, with a halo of 4 columns on each
// side (3 are needed; 4 keeps the vectors aligned) and, as in k_col, 3 stress
// rows and 1 + 2 velocity rows above and below.  All global loads are issued
// first at clamped addresses; the stress of step n goes to LDS; the velocity
// of step n+1 is computed on the tile plus halo from LDS (owned cells also
// update u, snapshots and receivers) and goes to LDS; the stress of step n+1
// is computed on the owned cells.  Arithmetic is expression for expression
// that of k_col, so results are bitwise identical to it (16-bit snapshots
// are scaled per row segment of the owned columns, so their scales follow the
// tile width).  SNAP = false is the
// variant for steps without snapshot roles (with FP32 snapshots every sfe,
// nine forward steps in ten and every adjoint step): the roles are
// compile-time null there and their code and registers are gone; REC =
// false likewise drops receiver recording (adjoint, recomputation).  Needs
// nz % V == 0.
#include "colv.h"

namespace inv {

constexpr int WBLOCK = 256;  // four wavefronts per tile

template <bool SH, bool PSV, int TX, int V, bool FAST, bool SNAP, bool REC>
__global__ __launch_bounds__(WBLOCK) void k_colv(const StepArgs a) {
    static_assert(SH != PSV, "one wave type per launch");
    static_assert(V == 2 || V == 4, "2 or 4 columns per lane");
    constexpr int HL = col_halo(V), W = 64 * V, S = col_strip(V);
    constexpr int SR = TX + 6, VR = TX + 3, NW = WBLOCK / 64;
    constexpr int PS = (SR + NW - 1) / NW, PV = (VR + NW - 1) / NW, PT = (TX + NW - 1) / NW;
    __shared__ __align__(16) float l_s1[SR * W], l_s2[SR * W], l_s3[PSV ? SR * W : 4];
    __shared__ __align__(16) float l_v1[VR * W], l_v2[PSV ? VR * W : 4];

    // XCD-aware tile order (as k_col)
    const int nw = gridDim.x, w = blockIdx.x;
    const int q = nw >> 3, rmd = nw & 7, xcd = w & 7;
    const int L = xcd * q + (xcd < rmd ? xcd : rmd) + (w >> 3);
    const int per = a.tiles_x * a.tiles_z;
    const int b = L / per;
    const int t = L - b * per;
    const int tx = t / a.tiles_z, tz = t - tx * a.tiles_z;
    const int i0 = tx * TX, j0 = tz * S;
    const Geo &g = a.g;
    const int nx = g.nx, nz = g.nz;
    const long so = (long)b * g.N;
    // the wavefront index is uniform across the wavefront: reading it through
    // readfirstlane lets every row quantity derived from it live in scalar
    // registers (scalar branches and address arithmetic instead of per-lane)
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c0 = lane * V;          // the lane's first LDS column
    const int jb = j0 - HL + c0;      // its first grid column (a multiple of V)
    const int jl = jb < 0 ? 0 : (jb > nz - V ? nz - V : jb);
    const bool jin = jb >= 0 && jb < nz;  // all V columns in the grid, or none
    const bool cown = jin && c0 >= HL && c0 < HL + S;
    const ZCoef<V> zc = zcoef<V>(g, jb);  // fast mode's z coefficients
    auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
    const float *m_rho = a.rho, *m_lam = a.lam, *m_mu = a.mu;
    if (a.slot_model) {
        m_rho = a.slot_model[3 * b];
        m_lam = a.slot_model[3 * b + 1];
        m_mu = a.slot_model[3 * b + 2];
    }
    const float *S1 = PSV ? a.sxx0 : a.sxy0, *S2 = PSV ? a.szz0 : a.szy0, *S3 = a.sxz0;
    const float *V1 = PSV ? a.vx0 : a.vy0, *V2 = a.vz0;
    float *U1 = PSV ? a.ux : a.uy, *U2 = a.uz;
    float *V1o = PSV ? a.vx1 : a.vy1, *V2o = a.vz1;
    float *S1o = PSV ? a.sxx1 : a.sxy1, *S2o = PSV ? a.szz1 : a.szy1, *S3o = a.sxz1;

    // all global loads of the three phases first
    float s1[PS][V], s2[PS][V], s3[PS][V];
#pragma unroll
    for (int p = 0; p < PS; p++) {
        const int r = wave + p * NW;
        const long k = so + (long)clampi(i0 - 3 + r, 0, nx - 1) * nz + jl;
        vload<V>(S1 + k, s1[p]);
        vload<V>(S2 + k, s2[p]);
        if (PSV) vload<V>(S3 + k, s3[p]);
    }
    float qv1[PV][V], qv2[PV][V], qr[PV][V], qu1[PV][V], qu2[PV][V], qab[PV];
#pragma unroll
    for (int p = 0; p < PV; p++) {
        const int vr = wave + p * NW;
        const int ic = clampi(i0 - 1 + vr, 0, nx - 1);
        const long kc = (long)ic * nz + jl;
        vload_g<V>(m_rho + kc, qr[p]);
        qab[p] = vload_uniform(a.abx, ic);
        vload<V>(V1 + so + kc, qv1[p]);
        if (PSV) vload<V>(V2 + so + kc, qv2[p]);
        if (vr >= 1 && vr <= TX) {  // u only on owned rows (wave-uniform)
            vload<V>(U1 + so + kc, qu1[p]);
            if (PSV) vload<V>(U2 + so + kc, qu2[p]);
        }
    }
    float qmu[PT][V], qlam[PT][V];
#pragma unroll
    for (int p = 0; p < PT; p++) {
        const int tr = wave + p * NW;
        const long kc = (long)clampi(i0 + tr, 0, nx - 1) * nz + jl;
        vload_g<V>(m_mu + kc, qmu[p]);
        if (PSV) vload_g<V>(m_lam + kc, qlam[p]);
    }
    float abz[V], abt[V];
    vload<V>(a.abzb + jl, abz);
    vload<V>(a.abzt + jl, abt);

    // 1. stress of step n on rows i0-3 .. i0+TX+2
#pragma unroll
    for (int p = 0; p < PS; p++) {
        const int r = wave + p * NW;
        if (r < SR) {
            const int i = i0 - 3 + r;
            if (!(jin && i >= 0 && i < nx)) {
#pragma unroll
                for (int v = 0; v < V; v++) {
                    s1[p][v] = 0.0f;
                    s2[p][v] = 0.0f;
                    s3[p][v] = 0.0f;
                }
            }
            vstore<V>(&l_s1[r * W + c0], s1[p]);
            vstore<V>(&l_s2[r * W + c0], s2[p]);
            if (PSV) vstore<V>(&l_s3[r * W + c0], s3[p]);
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
    if (REC && a.rec_ptr) {
        r0 = a.rec_ptr[t];
        r1 = a.rec_ptr[t + 1];
    }
    const LanePts ip = lane_pts(a.inj_cell, a.inj_row, p0, p1, lane);
    const LanePts rp = lane_pts(a.rec_cell, a.rec_row, r0, r1, lane);
    const long it = a.it;
    const int fu = PSV ? a.psv_f0 : 0, fv = PSV ? a.psv_f0 + 2 : 1;
    const long sb = (long)b * a.snap_slot_stride;
    const StepRoles r1s = SNAP ? StepRoles{a.snap, a.sexp} : StepRoles{nullptr, nullptr};

    // 2. velocity of step n+1 on rows i0-1 .. i0+TX+1 (stress LDS rows vr .. vr+3)
#pragma unroll
    for (int p = 0; p < PV; p++) {
        const int vr = wave + p * NW;
        if (vr >= VR) break;
        const int i = i0 - 1 + vr;
        const bool rin = jin && i >= 0 && i < nx;
        // cells of this row the wavefront covers: point lists are scanned with
        // one wave-uniform test per point before the per-lane one
        const int rlo = i * nz + (j0 - HL), rhi = rlo + W - 1;
        float n1[V], n2[V];
        vel_row<PSV, V, W, FAST>(a, l_s1, l_s2, l_s3, vr + 2, vr + 2, c0, i, jb, rin, zc, qab[p],
                                 abz, abt, qv1[p], qv2[p], qr[p], ip, rlo, rhi, b, it, n1, n2);
        vstore<V>(&l_v1[vr * W + c0], n1);
        if (PSV) vstore<V>(&l_v2[vr * W + c0], n2);
        // owned cells: v, u, snapshots, receivers (the row test is
        // wave-uniform; owned_row takes the lane's ownership)
        if (vr >= 1 && vr <= TX && i < nx) {
            owned_row<PSV, V, FAST>(a, r1s, fu, fv, sb, i, jb, cown, it, b, tz, S, lane, rp, rlo,
                                    rhi, qu1[p], qu2[p], n1, n2);
            if (cown) {
                const long o = so + (long)i * nz + jb;
                vstore<V>(V1o + o, n1);
                vstore<V>(U1 + o, qu1[p]);
                if (PSV) {
                    vstore<V>(V2o + o, n2);
                    vstore<V>(U2 + o, qu2[p]);
                }
            }
        }
    }
    __syncthreads();

    // 3. stress of step n+1 on the owned cells (velocity LDS rows tr .. tr+3)
    // the tile's explosions, one per lane
    int q0 = 0, q1 = 0;
    if (PSV && a.sinj_ptr) {
        q0 = a.sinj_ptr[b * per + t];
        q1 = a.sinj_ptr[b * per + t + 1];
    }
    const LanePts xp = lane_pts(a.sinj_cell, a.sinj_row, q0, q1, lane);
#pragma unroll
    for (int p = 0; p < PT; p++) {
        const int tr = wave + p * NW;
        if (tr >= TX) break;
        const int i = i0 + tr;
        if (i >= nx) break;
        const long o = so + (long)i * nz + jb;
        float sa[V], sc[V], sxz[V], o1[V], o2[V], o3[V];
        vload<V>(&l_s1[(tr + 3) * W + c0], sa);
        vload<V>(&l_s2[(tr + 3) * W + c0], sc);
        if (PSV) vload<V>(&l_s3[(tr + 3) * W + c0], sxz);
        stress_row<PSV, V, W, FAST>(g, l_v1, l_v2, tr + 1, c0, i, jb, zc, sa, sc, sxz, qmu[p],
                                    qlam[p], o1, o2, o3);
        // explosive sources this tile owns: stf_y into sxx and szz (every
        // lane scans; owned lanes store)
        if (PSV) explode_row<V>(a, xp, i, jb, b, it, o1, o2);
        if (cown) {
            vstore<V>(S1o + o, o1);
            vstore<V>(S2o + o, o2);
            if (PSV) vstore<V>(S3o + o, o3);
        }
    }
}

template <bool SH, bool PSV, bool FAST>
void wide_launch(Ctx &c, const StepArgs &a, unsigned blocks) {
    const bool snap = a.snap != nullptr;
    const int var = (snap ? 2 : 0) + (a.rec_ptr ? 1 : 0);
#define INV_COLV(R, VV)                                                                          \
    (var == 3   ? k_colv<SH, PSV, R, VV, FAST, true, true><<<blocks, WBLOCK, 0, c.stream>>>(a)    \
     : var == 2 ? k_colv<SH, PSV, R, VV, FAST, true, false><<<blocks, WBLOCK, 0, c.stream>>>(a)   \
     : var == 1 ? k_colv<SH, PSV, R, VV, FAST, false, true><<<blocks, WBLOCK, 0, c.stream>>>(a)   \
                : k_colv<SH, PSV, R, VV, FAST, false, false><<<blocks, WBLOCK, 0, c.stream>>>(a))
    if (c.col_v == 2) {
        switch (c.tile_x) {
            case 4: INV_COLV(4, 2); break;
            case 8: INV_COLV(8, 2); break;
            default: INV_COLV(16, 2); break;
        }
    } else {
        switch (c.tile_x) {
            case 8: INV_COLV(8, 4); break;
            default: INV_COLV(4, 4); break;
        }
    }
#undef INV_COLV
}

template void wide_launch<true, false, true>(Ctx &, const StepArgs &, unsigned);
template void wide_launch<false, true, true>(Ctx &, const StepArgs &, unsigned);
template void wide_launch<true, false, false>(Ctx &, const StepArgs &, unsigned);
template void wide_launch<false, true, false>(Ctx &, const StepArgs &, unsigned);

}  // namespace inv
