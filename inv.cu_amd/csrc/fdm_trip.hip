// Three time steps per launch.  The two-step kernel's scheme (fdm_pair.hip)
// carried one step further: the fused step of k_colv (divS -> addSTF ->
// updateV -> divV -> updateS -> saveRec, solver/fdm.h:534-551) is applied
// three times to a tile held in LDS, so every wavefield and model field
// crosses HBM once per three steps: 68 B per three P-SV cell-updates (40 B
// for SH).  The halo is one and a half steps deeper: the stress of step n is
// read on rows i0-9 .. i0+TX+8 (szz / szy on i0-7 .. i0+TX+7) and columns
// j0-10 .. j0+S+9 (10 keeps the two-column vectors aligned), and the phases
// compute
//   A  velocity n+1 on rows -7 .. TX+7      B  stress n+1 on -6 .. TX+5
//   C  velocity n+2 on rows -4 .. TX+4      D  stress n+2 on -3 .. TX+2
//   E  velocity n+3 on rows -1 .. TX+1      F  stress n+3 on the owned rows
// each in place over the previous step in LDS (a cell's update reads its own
// old value and only neighbours of the other field).  Row x of the tile is
// handled by wavefront (x + 7) mod NW in every phase, so a row's velocity,
// u, rho, lambda and mu stay in that wavefront's registers.  Five barriers.
//
// Every cell is evaluated by the row functions of colv.h, so a launch gives
// bitwise the results of three steps of k_colv or k_pair (and of the FP32
// oracle in exact math).  Point lists are bucketed for this geometry: forces
// with a halo of 7 rows and columns (8 past the end), explosions 6,
// receivers 0.  No snapshot roles: the launch plans give a step that stores
// a snapshot to a two-step or one-step launch.  REC = false drops receiver
// recording (adjoint and recomputation).  Two columns per lane; needs nz even
// and dynamic LDS of trip_lds_bytes.
#include "colv.h"

namespace inv {

template <bool SH, bool PSV, int TX, int NW, bool FAST, bool REC>
__global__ __launch_bounds__(64 * NW, 1) void k_trip(const StepArgs a) {
    static_assert(SH != PSV, "one wave type per launch");
    static_assert(TX % 4 == 0, "tile rows a multiple of four");
    constexpr int V = 2, HL = trip_halo(), W = 64 * V, S = trip_strip();
    constexpr int SR = TX + 18;                 // stress rows x = -9 .. TX+8
    constexpr int VR = TX + 15;                 // velocity rows x = -7 .. TX+7
    constexpr int PS = (SR + NW - 1) / NW;      // stress rows loaded per wavefront
    constexpr int P = (VR + NW - 1) / NW;       // rows x = wave - 7 + NW p per wavefront
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
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c0 = lane * V;          // the lane's first LDS column
    const int jb = j0 - HL + c0;      // its first grid column (even)
    const int jl = jb < 0 ? 0 : (jb > nz - V ? nz - V : jb);
    const bool jin = jb >= 0 && jb < nz;  // both columns in the grid, or none
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
        const long k = so + (long)clampi(i0 - 9 + r, 0, nx - 1) * nz + jl;
        vload<V>(S1 + k, s1[p]);
        if (r >= 2 && r < SR - 1) vload<V>(S2 + k, s2[p]);  // wave-uniform
        if (PSV) vload<V>(S3 + k, s3[p]);
    }
    float qv1[P][V], qv2[P][V], qr[P][V], qu1[P][V], qu2[P][V], qmu[P][V], qlam[P][V], qab[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int x = wave - 7 + p * NW;
        const int ic = clampi(i0 + x, 0, nx - 1);
        const long kc = (long)ic * nz + jl;
        if (x < TX + 8) {
            vload_g<V>(m_rho + kc, qr[p]);
            vload<V>(V1 + so + kc, qv1[p]);
            if (PSV) vload<V>(V2 + so + kc, qv2[p]);
            qab[p] = vload_uniform(a.abx, ic);
        }
        if (x >= 0 && x < TX) {  // u only on owned rows (wave-uniform)
            vload<V>(U1 + so + kc, qu1[p]);
            if (PSV) vload<V>(U2 + so + kc, qu2[p]);
        }
        if (x >= -6 && x < TX + 6) {  // rows of the stress updates
            vload_g<V>(m_mu + kc, qmu[p]);
            if (PSV) vload_g<V>(m_lam + kc, qlam[p]);
        }
    }
    float abz[V], abt[V];
    vload<V>(a.abzb + jl, abz);
    vload<V>(a.abzt + jl, abt);

    // stress of step n on rows x = -9 .. TX+8 (LDS row x + 9; szz / szy on
    // the velocity rows, LDS row x + 7), 0 outside the grid
#pragma unroll
    for (int p = 0; p < PS; p++) {
        const int r = wave + p * NW;
        if (r < SR) {
            const int i = i0 - 9 + r;
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
    const StepRoles none{nullptr, nullptr};

    // velocity of step st + 1 on rows x = lo .. hi-1, in place (LDS row
    // x + 7); owned rows update u and record receivers; the last phase
    // stores v and u of the owned cells
    auto vel_phase = [&](int lo, int hi, long st, bool last) {
#pragma unroll
        for (int p = 0; p < P; p++) {
            const int x = wave - 7 + p * NW;
            if (x < lo) continue;
            if (x >= hi) break;
            const int i = i0 + x;
            const bool rin = jin && i >= 0 && i < nx;
            const int rlo = i * nz + (j0 - HL), rhi = rlo + W - 1;
            float n1[V], n2[V];
            vel_row<PSV, V, W, FAST>(a, l_s1, l_s2, l_s3, x + 9, x + 7, c0, i, jb, rin, zc,
                                     qab[p], abz, abt, qv1[p], qv2[p], qr[p], ip, rlo, rhi, b, st,
                                     n1, n2);
            vstore<V>(&l_v1[(x + 7) * W + c0], n1);
            if (PSV) vstore<V>(&l_v2[(x + 7) * W + c0], n2);
            if (x >= 0 && x < TX && i < nx) {  // wave-uniform
                owned_row<PSV, V, FAST>(a, none, fu, fv, 0, i, jb, cown, st, b, tz, S, lane, rp,
                                        rlo, rhi, qu1[p], qu2[p], n1, n2);
                if (last && cown) {
                    const long o = so + (long)i * nz + jb;
                    vstore<V>(V1o + o, n1);
                    vstore<V>(U1 + o, qu1[p]);
                    if (PSV) {
                        vstore<V>(V2o + o, n2);
                        vstore<V>(U2 + o, qu2[p]);
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < V; v++) {
                qv1[p][v] = n1[v];
                if (PSV) qv2[p][v] = n2[v];
            }
        }
    };
    // stress of step st + 1 on rows x = lo .. hi-1: in place in LDS (0
    // outside the grid in exact mode), or stored on the owned cells (last)
    auto str_phase = [&](int lo, int hi, long st, bool last) {
#pragma unroll
        for (int p = 0; p < P; p++) {
            const int x = wave - 7 + p * NW;
            if (x < lo) continue;
            if (x >= hi) break;
            const int i = i0 + x;
            if (last && i >= nx) break;
            const bool rin = jin && i >= 0 && i < nx;
            const int r = x + 9;
            float sa[V], sc[V], sxz[V], o1[V], o2[V], o3[V];
            vload<V>(&l_s1[r * W + c0], sa);
            vload<V>(&l_s2[(x + 7) * W + c0], sc);
            if (PSV) vload<V>(&l_s3[r * W + c0], sxz);
            stress_row<PSV, V, W, FAST>(g, l_v1, l_v2, x + 7, c0, i, jb, zc, sa, sc, sxz, qmu[p],
                                        qlam[p], o1, o2, o3);
            if (PSV) explode_row<V>(a, xp, i, jb, b, st, o1, o2);
            if (last) {
                if (cown) {
                    const long o = so + (long)i * nz + jb;
                    vstore<V>(S1o + o, o1);
                    vstore<V>(S2o + o, o2);
                    if (PSV) vstore<V>(S3o + o, o3);
                }
            } else {
                if (!FAST && !rin) {
#pragma unroll
                    for (int v = 0; v < V; v++) {
                        o1[v] = 0.0f;
                        o2[v] = 0.0f;
                        o3[v] = 0.0f;
                    }
                }
                vstore<V>(&l_s1[r * W + c0], o1);
                vstore<V>(&l_s2[(x + 7) * W + c0], o2);
                if (PSV) vstore<V>(&l_s3[r * W + c0], o3);
            }
        }
    };
    vel_phase(-7, TX + 8, it, false);      // A
    __syncthreads();
    str_phase(-6, TX + 6, it, false);      // B
    __syncthreads();
    vel_phase(-4, TX + 5, it + 1, false);  // C
    __syncthreads();
    str_phase(-3, TX + 3, it + 1, false);  // D
    __syncthreads();
    vel_phase(-1, TX + 2, it + 2, true);   // E
    __syncthreads();
    str_phase(0, TX, it + 2, true);        // F
}

template <class K>
static void trip_go(Ctx &c, K kern, int nw, int lds, const StepArgs &a, unsigned blocks) {
    if (lds > 65536)
        INV_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    kern<<<blocks, 64 * nw, lds, c.stream>>>(a);
}

template <bool SH, bool PSV, bool FAST>
void trip_launch(Ctx &c, const StepArgs &a, unsigned blocks) {
    INV_REQUIRE(!a.snap, INV_ERR_STATE, "three-step launches carry no snapshot role");
    const int rows = c.trip_tx;
    const int lds = trip_lds_bytes(PSV, rows);
    const bool rec = a.rec_ptr != nullptr;
#define INV_TRIP(R)                                                                     \
    (rec ? trip_go(c, k_trip<SH, PSV, R, 8, FAST, true>, 8, lds, a, blocks)             \
         : trip_go(c, k_trip<SH, PSV, R, 8, FAST, false>, 8, lds, a, blocks))
    switch (rows) {
        case 12: INV_TRIP(12); break;
        case 16: INV_TRIP(16); break;
        case 20: INV_TRIP(20); break;
        default: INV_TRIP(8); break;
    }
#undef INV_TRIP
}

template void trip_launch<true, false, true>(Ctx &, const StepArgs &, unsigned);
template void trip_launch<false, true, true>(Ctx &, const StepArgs &, unsigned);
template void trip_launch<true, false, false>(Ctx &, const StepArgs &, unsigned);
template void trip_launch<false, true, false>(Ctx &, const StepArgs &, unsigned);

}  // namespace inv
