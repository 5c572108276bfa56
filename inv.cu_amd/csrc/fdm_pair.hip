// Two time steps per launch.  The fused step of k_colv (divS -> addSTF ->
// updateV -> divV -> updateS -> saveRec, solver/fdm.h:534-551) is applied
// twice to a tile held in LDS, so every wavefield and model field crosses HBM
// once per two steps: 68 B per two P-SV cell-updates instead of per one (40 B
// for SH).  The price is recomputing a deeper halo: the stress of step n is
// read on rows i0-6 .. i0+TX+5 (szz / szy, which enter only through
// z-differences, on i0-4 .. i0+TX+4) and columns j0-6 .. j0+S+5, the velocity of
// step n+1 is computed on rows i0-4 .. i0+TX+4, the stress of step n+1 on
// rows i0-3 .. i0+TX+2, the velocity of step n+2 on rows i0-1 .. i0+TX+1 and
// the stress of step n+2 on the owned rows (and columns likewise; edge
// columns of a phase hold values no later owned cell depends on).
//
// Row x = i - i0 of the tile is handled by wavefront (x + 4) mod NW in every
// phase (NW = 4 wavefronts per tile, 8 for the 16- and 20-row tiles), so the
// registers of a row (velocity, u, rho, lambda, mu) stay with one wavefront
// from the loads to the last store.  The stress buffer in LDS holds
// step n and is overwritten in place by step n+1 (each cell's update reads
// its own old value and only velocity neighbours); the velocity buffer holds
// step n+1 and is overwritten by step n+2 likewise.  Four barriers.
//
// Every cell is evaluated by the row functions of colv.h, shared with
// k_colv, so two launches of this kernel give bitwise the results of four
// launches of k_colv (and of k_col and the FP32 oracle in exact math).
// Point lists are bucketed for this geometry: forces with a halo of 4 rows
// and columns, explosions 3, receivers 0.  Snapshot roles of the first step
// come in StepArgs, of the second in StepRoles; 16-bit snapshots are scaled
// per row segment of S columns (the tile's owned columns), segment tz.
// SNAP = false is the variant for launches with no snapshot role in either
// step (most of them: every adjoint launch and four of five forward ones with
// sfe = 10): the roles are compile-time null there, so the snapshot code and
// the scalar registers of its pointers are gone.  REC = false likewise drops
// receiver recording (adjoint and recomputation).
// Needs nz % V == 0 and dynamic LDS of pair_lds_bytes.
#include "colv.h"

namespace inv {

// The SH 16-row tile without snapshot roles asks for eight
// wavefronts per SIMD: its registers then fit 78 scalar registers (106
// otherwise, which allowed seven) and four of its 40 KB tiles share a CU.
template <bool SH, bool PSV, int TX, int V, int NW, bool FAST, bool SNAP, bool REC>
__global__ __launch_bounds__(64 * NW, SH && TX == 16 && !SNAP ? 8 : 1) void k_pair(
    const StepArgs a, const StepRoles r2in) {
    static_assert(SH != PSV, "one wave type per launch");
    static_assert(V == 2 || V == 4, "2 or 4 columns per lane");
    static_assert(TX % 4 == 0, "tile rows a multiple of the wavefronts");
    constexpr int HL = pair_halo(V), W = 64 * V, S = pair_strip(V);
    constexpr int SR = TX + 12;                 // stress rows x = -6 .. TX+5
    constexpr int VR = TX + 9;                  // velocity rows x = -4 .. TX+4
    // szz / szy enter the velocity only through z-differences within a row,
    // so their buffer holds the velocity rows x = -4 .. TX+4 (LDS row x + 4)
    constexpr int PS = (SR + NW - 1) / NW;      // stress rows loaded per wavefront
    constexpr int P = (VR + NW - 1) / NW;       // rows x = wave - 4 + NW p per wavefront
    extern __shared__ __align__(16) float smem[];
    float *const l_s1 = smem;
    float *const l_s2 = l_s1 + SR * W;
    float *const l_s3 = PSV ? l_s2 + VR * W : l_s2;  // unused for SH
    float *const l_v1 = PSV ? l_s3 + SR * W : l_s2 + VR * W;
    float *const l_v2 = PSV ? l_v1 + VR * W : l_v1;  // unused for SH

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

    // all global loads first, at clamped addresses
    float s1[PS][V], s2[PS][V], s3[PS][V];
#pragma unroll
    for (int p = 0; p < PS; p++) {
        const int r = wave + p * NW;
        const long k = so + (long)clampi(i0 - 6 + r, 0, nx - 1) * nz + jl;
        vload<V>(S1 + k, s1[p]);
        if (r >= 2 && r < SR - 1) vload<V>(S2 + k, s2[p]);  // wave-uniform
        if (PSV) vload<V>(S3 + k, s3[p]);
    }
    float qv1[P][V], qv2[P][V], qr[P][V], qu1[P][V], qu2[P][V], qmu[P][V], qlam[P][V], qab[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int x = wave - 4 + p * NW;
        const int ic = clampi(i0 + x, 0, nx - 1);
        const long kc = (long)ic * nz + jl;
        if (x < TX + 5) {
            vload_g<V>(m_rho + kc, qr[p]);
            vload<V>(V1 + so + kc, qv1[p]);
            if (PSV) vload<V>(V2 + so + kc, qv2[p]);
            qab[p] = vload_uniform(a.abx, ic);
        }
        if (x >= 0 && x < TX) {  // u only on owned rows (wave-uniform)
            vload<V>(U1 + so + kc, qu1[p]);
            if (PSV) vload<V>(U2 + so + kc, qu2[p]);
        }
        if (x >= -3 && x < TX + 3) {  // rows of the two stress updates
            vload_g<V>(m_mu + kc, qmu[p]);
            if (PSV) vload_g<V>(m_lam + kc, qlam[p]);
        }
    }
    float abz[V], abt[V];
    vload<V>(a.abzb + jl, abz);
    vload<V>(a.abzt + jl, abt);

    // stress of step n on rows x = -6 .. TX+5 (LDS row x + 6), 0 outside the grid
#pragma unroll
    for (int p = 0; p < PS; p++) {
        const int r = wave + p * NW;
        if (r < SR) {
            const int i = i0 - 6 + r;
            if (!(jin && i >= 0 && i < nx)) {
#pragma unroll
                for (int v = 0; v < V; v++) {
                    s1[p][v] = 0.0f;
                    s2[p][v] = 0.0f;
                    s3[p][v] = 0.0f;
                }
            }
            vstore<V>(&l_s1[r * W + c0], s1[p]);
            if (r >= 2 && r < SR - 1) vstore<V>(&l_s2[(r - 2) * W + c0], s2[p]);
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
    int q0 = 0, q1 = 0;
    if (PSV && a.sinj_ptr) {
        q0 = a.sinj_ptr[b * per + t];
        q1 = a.sinj_ptr[b * per + t + 1];
    }
    const LanePts ip = lane_pts(a.inj_cell, a.inj_row, p0, p1, lane);
    const LanePts rp = lane_pts(a.rec_cell, a.rec_row, r0, r1, lane);
    const LanePts xp = lane_pts(a.sinj_cell, a.sinj_row, q0, q1, lane);
    const long it = a.it;
    const int fu = PSV ? a.psv_f0 : 0, fv = PSV ? a.psv_f0 + 2 : 1;
    const long sb = (long)b * a.snap_slot_stride;
    const StepRoles none{nullptr, nullptr};
    const StepRoles r1s = SNAP ? StepRoles{a.snap, a.sexp} : none;
    const StepRoles r2 = SNAP ? r2in : none;

    // A. velocity of step n+1 on rows x = -4 .. TX+4 (LDS row x + 4)
    {
#pragma unroll
        for (int p = 0; p < P; p++) {
            const int x = wave - 4 + p * NW;
            if (x >= TX + 5) break;
            const int i = i0 + x;
            const bool rin = jin && i >= 0 && i < nx;
            const int rlo = i * nz + (j0 - HL), rhi = rlo + W - 1;
            float n1[V], n2[V];
            vel_row<PSV, V, W, FAST>(a, l_s1, l_s2, l_s3, x + 6, x + 4, c0, i, jb, rin, zc,
                                     qab[p], abz, abt, qv1[p], qv2[p], qr[p], ip, rlo, rhi, b, it,
                                     n1, n2);
            vstore<V>(&l_v1[(x + 4) * W + c0], n1);
            if (PSV) vstore<V>(&l_v2[(x + 4) * W + c0], n2);
            if (x >= 0 && x < TX && i < nx)  // wave-uniform
                owned_row<PSV, V, FAST>(a, r1s, fu, fv, sb, i, jb, cown, it, b, tz, S, lane, rp,
                                        rlo, rhi, qu1[p], qu2[p], n1, n2);
#pragma unroll
            for (int v = 0; v < V; v++) {
                qv1[p][v] = n1[v];
                if (PSV) qv2[p][v] = n2[v];
            }
        }
    }
    __syncthreads();

    // B. stress of step n+1 on rows x = -3 .. TX+2, in place (0 outside the grid)
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int x = wave - 4 + p * NW;
        if (x < -3) continue;
        if (x >= TX + 3) break;
        const int i = i0 + x;
        const bool rin = jin && i >= 0 && i < nx;
        const int r = x + 6;
        float sa[V], sc[V], sxz[V], o1[V], o2[V], o3[V];
        vload<V>(&l_s1[r * W + c0], sa);
        vload<V>(&l_s2[(x + 4) * W + c0], sc);
        if (PSV) vload<V>(&l_s3[r * W + c0], sxz);
        stress_row<PSV, V, W, FAST>(g, l_v1, l_v2, x + 4, c0, i, jb, zc, sa, sc, sxz, qmu[p],
                                    qlam[p], o1, o2, o3);
        if (PSV) explode_row<V>(a, xp, i, jb, b, it, o1, o2);
        if (!FAST && !rin) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                o1[v] = 0.0f;
                o2[v] = 0.0f;
                o3[v] = 0.0f;
            }
        }
        vstore<V>(&l_s1[r * W + c0], o1);
        vstore<V>(&l_s2[(x + 4) * W + c0], o2);
        if (PSV) vstore<V>(&l_s3[r * W + c0], o3);
    }
    __syncthreads();

    // C. velocity of step n+2 on rows x = -1 .. TX+1, in place; owned cells
    // store v and u
    {
#pragma unroll
        for (int p = 0; p < P; p++) {
            const int x = wave - 4 + p * NW;
            if (x < -1) continue;
            if (x >= TX + 2) break;
            const int i = i0 + x;
            const bool rin = jin && i >= 0 && i < nx;
            const int rlo = i * nz + (j0 - HL), rhi = rlo + W - 1;
            float n1[V], n2[V];
            vel_row<PSV, V, W, FAST>(a, l_s1, l_s2, l_s3, x + 6, x + 4, c0, i, jb, rin, zc,
                                     qab[p], abz, abt, qv1[p], qv2[p], qr[p], ip, rlo, rhi, b,
                                     it + 1, n1, n2);
            vstore<V>(&l_v1[(x + 4) * W + c0], n1);
            if (PSV) vstore<V>(&l_v2[(x + 4) * W + c0], n2);
            if (x >= 0 && x < TX && i < nx) {  // wave-uniform
                owned_row<PSV, V, FAST>(a, r2, fu, fv, sb, i, jb, cown, it + 1, b, tz, S, lane,
                                        rp, rlo, rhi, qu1[p], qu2[p], n1, n2);
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
    }
    __syncthreads();

    // D. stress of step n+2 on the owned cells (every lane runs the rows, for
    // the explosion scan; owned lanes store)
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int x = wave - 4 + p * NW;
        if (x < 0) continue;
        if (x >= TX) break;
        const int i = i0 + x;
        if (i >= nx) break;
        const int r = x + 6;
        const long o = so + (long)i * nz + jb;
        float sa[V], sc[V], sxz[V], o1[V], o2[V], o3[V];
        vload<V>(&l_s1[r * W + c0], sa);
        vload<V>(&l_s2[(x + 4) * W + c0], sc);
        if (PSV) vload<V>(&l_s3[r * W + c0], sxz);
        stress_row<PSV, V, W, FAST>(g, l_v1, l_v2, x + 4, c0, i, jb, zc, sa, sc, sxz, qmu[p],
                                    qlam[p], o1, o2, o3);
        if (PSV) explode_row<V>(a, xp, i, jb, b, it + 1, o1, o2);
        if (cown) {
            vstore<V>(S1o + o, o1);
            vstore<V>(S2o + o, o2);
            if (PSV) vstore<V>(S3o + o, o3);
        }
    }
}

template <class K>
static void pair_go(Ctx &c, K kern, int nw, int lds, const StepArgs &a, const StepRoles &r2,
                    unsigned blocks) {
    if (lds > 65536)
        INV_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    kern<<<blocks, 64 * nw, lds, c.stream>>>(a, r2);
}

template <bool SH, bool PSV, bool FAST>
void pair_launch(Ctx &c, const StepArgs &a, const StepRoles &r2, unsigned blocks) {
    const bool snap = a.snap || r2.snap;
    const int var = (snap ? 2 : 0) + (a.rec_ptr ? 1 : 0);
    const int rows = c.pair_tx, V = c.pair_v;
    const int lds = pair_lds_bytes(PSV, rows, V);
    if (lds > 65536) {
        // a tile larger than the default per-block LDS needs the opt-in limit
        int dev = 0, optin = 0;
        INV_HIP(hipGetDevice(&dev));
        INV_HIP(hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev));
        INV_REQUIRE(lds <= optin, INV_ERR_UNSUPPORTED,
                    "two-step tile needs more LDS per block than the device allows");
    }
#define INV_PAIR(R, VV, NWW)                                                                     \
    (var == 3   ? pair_go(c, k_pair<SH, PSV, R, VV, NWW, FAST, true, true>, NWW, lds, a, r2, blocks)  \
     : var == 2 ? pair_go(c, k_pair<SH, PSV, R, VV, NWW, FAST, true, false>, NWW, lds, a, r2, blocks) \
     : var == 1 ? pair_go(c, k_pair<SH, PSV, R, VV, NWW, FAST, false, true>, NWW, lds, a, r2, blocks) \
                : pair_go(c, k_pair<SH, PSV, R, VV, NWW, FAST, false, false>, NWW, lds, a, r2, blocks))
    if (V == 2) {
        switch (rows) {
            case 4: INV_PAIR(4, 2, 4); break;
            case 12: INV_PAIR(12, 2, 4); break;
            // 16 and 20 rows: eight wavefronts (20 rows keep the register arrays
            // of 16: four rows per wavefront in every phase)
            case 16: INV_PAIR(16, 2, 8); break;
            case 20: INV_PAIR(20, 2, 8); break;
            default: INV_PAIR(8, 2, 4); break;
        }
    } else {
        switch (rows) {
            case 8: INV_PAIR(8, 4, 4); break;
            default: INV_PAIR(4, 4, 4); break;
        }
    }
#undef INV_PAIR
}

template void pair_launch<true, false, true>(Ctx &, const StepArgs &, const StepRoles &, unsigned);
template void pair_launch<false, true, true>(Ctx &, const StepArgs &, const StepRoles &, unsigned);
template void pair_launch<true, false, false>(Ctx &, const StepArgs &, const StepRoles &, unsigned);
template void pair_launch<false, true, false>(Ctx &, const StepArgs &, const StepRoles &, unsigned);

}  // namespace inv
