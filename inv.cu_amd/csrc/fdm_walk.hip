// Three time steps per launch without recomputed halo rows.  The fused step
// of k_colv (divS -> addSTF -> updateV -> divV -> updateS -> saveRec,
// solver/fdm.h:534-551) applied three times, as in k_trip, but a workgroup
// owns a chunk of CH rows of one 108-column strip and walks down it eight
// rows at a time, keeping the rows of every intermediate step in LDS ring
// buffers: each phase of an iteration computes eight new rows, skewed so
// that their inputs were computed by the same phase of the iteration before
// or by the phase before it:
//   A  velocity n+1 on rows i0+8 .. i0+15    B  stress n+1 on i0+6 .. i0+13
//   C  velocity n+2 on rows i0+5 .. i0+12    D  stress n+2 on i0+3 .. i0+10
//   E  velocity n+3 on rows i0+2 .. i0+9     F  stress n+3 on i0 .. i0+7
// (velocity at row x reads stress rows x-2 .. x+1, stress reads velocity
// rows x-1 .. x+2).  A chunk's rows are computed once per phase, plus the
// halo of a three-step tile at its two ends (45 row updates per chunk
// against 7.5 per tile row in k_trip); every field still crosses HBM once
// per three steps (68 B per three P-SV cell-updates, 40 B for SH).
//
// Row x of any phase is handled by wavefront x mod 8, so a wavefront meets
// only two rows per iteration: o = i0 + w (its rows of F, and of B..E when w
// is at least the phase's skew) and n = i0 + 8 + w (A, and B..E otherwise).
// Their velocity, u, rho, lambda and mu stay in registers, the next
// iteration's row is loaded while this one computes, and the stress rows
// the next iteration needs are written to the ring while F runs.  Rings:
// 25 rows (the stress window i0 .. i0+16 and the eight rows being written)
// plus copies of slots 0-2 at 25-27, so the four rows a difference reads
// are always consecutive.  Six barriers per iteration.
//
// Every cell is evaluated by the row functions of colv.h, so a launch gives
// bitwise the results of k_trip, k_pair and k_colv (and of the FP32 oracle
// in exact math).  Point lists are those of a three-step tile of CH rows.
// Launches carry no snapshot roles; REC = false drops receiver recording.
// Two columns per lane; needs nz even and dynamic LDS of walk_lds_bytes.
#include "colv.h"

namespace inv {

template <bool SH, bool PSV, bool FAST, bool REC>
__global__ __launch_bounds__(512, 1) void k_walk(const StepArgs a, const int CH) {
    static_assert(SH != PSV, "one wave type per launch");
    constexpr int V = 2, HL = trip_halo(), W = 64 * V, S = trip_strip();
    constexpr int R = WALK_RING, RR = R + 3;
    extern __shared__ __align__(16) float smem[];
    float *const l_s1 = smem;
    float *const l_s2 = l_s1 + RR * W;
    float *const l_s3 = PSV ? l_s2 + RR * W : l_s2;  // unused for SH
    float *const l_v1 = PSV ? l_s3 + RR * W : l_s2 + RR * W;
    float *const l_v2 = PSV ? l_v1 + RR * W : l_v1;  // unused for SH

    // XCD-aware order (as k_trip): chunks play the part of tile rows
    const int nw = gridDim.x, wg = blockIdx.x;
    const int q = nw >> 3, rmd = nw & 7, xcd = wg & 7;
    const int L = xcd * q + (xcd < rmd ? xcd : rmd) + (wg >> 3);
    const int per = a.tiles_x * a.tiles_z;
    const int b = L / per;
    const int t = L - b * per;
    const int tx = t / a.tiles_z, tz = t - tx * a.tiles_z;
    const Geo &g = a.g;
    const int nx = g.nx, nz = g.nz;
    const int xa = tx * CH;
    const int xe = min(xa + CH, (nx + 7) & ~7);  // rows past it change no owned cell
    const int xo = min(xa + CH, nx);              // owned rows xa .. xo-1
    const int j0 = tz * S;
    const long so = (long)b * g.N;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c0 = lane * V;
    const int jb = j0 - HL + c0;
    const int jl = jb < 0 ? 0 : (jb > nz - V ? nz - V : jb);
    const bool jin = jb >= 0 && jb < nz;
    const bool cown = jin && c0 >= HL && c0 < HL + S;
    const ZCoef<V> zc = zcoef<V>(g, jb);
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
    float abz[V], abt[V];
    vload<V>(a.abzb + jl, abz);
    vload<V>(a.abzt + jl, abt);

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
    // the lists' buckets are in cell order: a row's entries are found by
    // bisection among the (up to 128) held in lanes
    const LanePts ip = lane_pts_sorted(a.inj_cell, a.inj_row, p0, p1, lane);
    const LanePts rp = lane_pts_sorted(a.rec_cell, a.rec_row, r0, r1, lane);
    const LanePts xp = lane_pts_sorted(a.sinj_cell, a.sinj_row, q0, q1, lane);
    // the grid rows a list's entries lie on (every row when the list is
    // longer than the 128 held in lanes): rows outside skip the scan
    auto row_span = [&](const LanePts &t_, int &lo, int &hi) {
        lo = 1 << 30;
        hi = -1;
        const int n = t_.p1 - t_.p0;
        if (n > 128) {
            lo = -(1 << 30);
            hi = 1 << 30;
            return;
        }
        if (n > 0) {  // cell order: the first and last entries bound the rows
            lo = __builtin_amdgcn_readlane(t_.cell, 0) / nz;
            hi = (n <= 64 ? __builtin_amdgcn_readlane(t_.cell, n - 1)
                          : __builtin_amdgcn_readlane(t_.cell2, n - 65)) / nz;
        }
    };
    int ilo, ihi, rlo_, rhi_, xlo, xhi;
    row_span(ip, ilo, ihi);
    row_span(rp, rlo_, rhi_);
    row_span(xp, xlo, xhi);
    const LanePts none_pts{0, 0, -1, 0};
    const long it = a.it;
    const int fu = PSV ? a.psv_f0 : 0, fv = PSV ? a.psv_f0 + 2 : 1;
    const StepRoles none{nullptr, nullptr};

    // ring slot of grid row x (x >= xa - 24)
    auto slot = [&](int x) { return (x - xa + 5 * R) % R; };
    auto ring_put = [&](float *ring, int x, const float (&o)[V]) {
        const int s = slot(x);
        vstore<V>(&ring[s * W + c0], o);
        if (s < 3) vstore<V>(&ring[(s + R) * W + c0], o);
    };

    // one row's registers
    struct Row {
        float v1[V], v2[V], rho[V], u1[V], u2[V], mu[V], lam[V], ab;
    };
    Row ro{}, rn{}, rp_{};  // old row i0+w, new row i0+8+w, the next new row
    float ps1[V] = {}, ps2[V] = {}, ps3[V] = {};  // stress of row i0+9+w, loaded ahead
    auto load_row = [&](Row &r, int x) {
        const int ic = clampi(x, 0, nx - 1);
        const long kc = (long)ic * nz + jl;
        if (x >= xa - 7) {  // velocity rows of A (wave-uniform)
            vload_g<V>(m_rho + kc, r.rho);
            vload<V>(V1 + so + kc, r.v1);
            if (PSV) vload<V>(V2 + so + kc, r.v2);
            r.ab = vload_uniform(a.abx, ic);
        }
        if (x >= xa && x < xo) {  // u on owned rows
            vload<V>(U1 + so + kc, r.u1);
            if (PSV) vload<V>(U2 + so + kc, r.u2);
        }
        if (x >= xa - 6) {  // stress rows of B
            vload_g<V>(m_mu + kc, r.mu);
            if (PSV) vload_g<V>(m_lam + kc, r.lam);
        }
    };
    auto load_stress = [&](int x) {
        const long k = so + (long)clampi(x, 0, nx - 1) * nz + jl;
        if (x >= xa - 9) {
            vload<V>(S1 + k, ps1);
            if (x >= xa - 7) vload<V>(S2 + k, ps2);  // szz / szy only on velocity rows
            if (PSV) vload<V>(S3 + k, ps3);
        }
    };
    auto put_stress = [&](int x) {  // stress of step n, 0 outside the grid
        if (x < xa - 9) return;
        if (!(jin && x >= 0 && x < nx)) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                ps1[v] = 0.0f;
                ps2[v] = 0.0f;
                ps3[v] = 0.0f;
            }
        }
        ring_put(l_s1, x, ps1);
        if (x >= xa - 7) ring_put(l_s2, x, ps2);
        if (PSV) ring_put(l_s3, x, ps3);
    };

    // velocity of step st + 1 at row x (in the ring, and in r), owned rows
    // update u and record receivers; last: store v and u of the owned cells
    auto vel = [&](Row &r, int x, long st, bool last) {
        const bool rin = jin && x >= 0 && x < nx;
        const int rlo = x * nz + (j0 - HL), rhi = rlo + W - 1;
        const bool own = x >= xa && x < xo;
        float n1[V], n2[V];
        const int sr = slot(x - 2) + 2;
        vel_row<PSV, V, W, FAST>(a, l_s1, l_s2, l_s3, sr, slot(x), c0, x, jb, rin, zc, r.ab, abz,
                                 abt, r.v1, r.v2, r.rho, x >= ilo && x <= ihi ? ip : none_pts,
                                 rlo, rhi, b, st, n1, n2);
        ring_put(l_v1, x, n1);
        if (PSV) ring_put(l_v2, x, n2);
        if (own) {
            owned_row<PSV, V, FAST>(a, none, fu, fv, 0, x, jb, cown, st, b, tz, S, lane,
                                    x >= rlo_ && x <= rhi_ ? rp : none_pts, rlo, rhi, r.u1,
                                    r.u2, n1, n2);
            if (last && cown) {
                const long o = so + (long)x * nz + jb;
                vstore<V>(V1o + o, n1);
                vstore<V>(U1 + o, r.u1);
                if (PSV) {
                    vstore<V>(V2o + o, n2);
                    vstore<V>(U2 + o, r.u2);
                }
            }
        }
#pragma unroll
        for (int v = 0; v < V; v++) {
            r.v1[v] = n1[v];
            if (PSV) r.v2[v] = n2[v];
        }
    };
    // stress of step st + 1 at row x: in place in the ring (0 outside the
    // grid in exact mode), or stored on the owned cells (last)
    auto str = [&](Row &r, int x, long st, bool last) {
        const bool rin = jin && x >= 0 && x < nx;
        const int s = slot(x);
        float sa[V], sc[V], sxz[V], o1[V], o2[V], o3[V];
        vload<V>(&l_s1[s * W + c0], sa);
        vload<V>(&l_s2[s * W + c0], sc);
        if (PSV) vload<V>(&l_s3[s * W + c0], sxz);
        stress_row<PSV, V, W, FAST>(g, l_v1, l_v2, slot(x - 1) + 1, c0, x, jb, zc, sa, sc, sxz,
                                    r.mu, r.lam, o1, o2, o3);
        if (PSV) explode_row<V>(a, x >= xlo && x <= xhi ? xp : none_pts, x, jb, b, st, o1, o2);
        if (last) {
            if (cown && x < xo) {
                const long o = so + (long)x * nz + jb;
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
            ring_put(l_s1, x, o1);
            ring_put(l_s2, x, o2);
            if (PSV) ring_put(l_s3, x, o3);
        }
    };

    // iterations i0 = xa - 24, xa - 16, ...: the first three build the
    // chunk's upper halo (rows below a phase's first row are skipped)
    const int K = (xe - xa) / 8 + 3;
    // iteration 0 computes nothing: its two stress rows (xa-9, xa-8) are
    // written here, and the loop starts at iteration 1
    load_stress(xa - 15 + w);
    put_stress(xa - 15 + w);
    load_row(rp_, xa - 8 + w);
    load_stress(xa - 7 + w);
    for (int k = 1; k < K; k++) {
        const int i0 = xa - 24 + 8 * k;
        const int xo_ = i0 + w, xn = i0 + 8 + w;
        put_stress(i0 + 9 + w);
        // rows move up: the new row becomes the old one, the row loaded
        // during the previous iteration the new one
        ro = rn;
        rn = rp_;
        if (k + 1 < K) {
            load_row(rp_, xn + 8);
            load_stress(i0 + 17 + w);
        }
        __syncthreads();
        if (xn >= xa - 7) vel(rn, xn, it, false);  // A
        __syncthreads();
        if (w >= 6) {  // B
            if (xo_ >= xa - 6) str(ro, xo_, it, false);
        } else if (xn >= xa - 6) {
            str(rn, xn, it, false);
        }
        __syncthreads();
        if (w >= 5) {  // C
            if (xo_ >= xa - 4) vel(ro, xo_, it + 1, false);
        } else if (xn >= xa - 4) {
            vel(rn, xn, it + 1, false);
        }
        __syncthreads();
        if (w >= 3) {  // D
            if (xo_ >= xa - 3) str(ro, xo_, it + 1, false);
        } else if (xn >= xa - 3) {
            str(rn, xn, it + 1, false);
        }
        __syncthreads();
        if (w >= 2) {  // E
            if (xo_ >= xa - 1) vel(ro, xo_, it + 2, true);
        } else if (xn >= xa - 1) {
            vel(rn, xn, it + 2, true);
        }
        __syncthreads();
        if (xo_ >= xa) str(ro, xo_, it + 2, true);  // F
    }
}

template <class K>
static void walk_go(Ctx &c, K kern, int lds, const StepArgs &a, int ch, unsigned blocks) {
    if (lds > 65536)
        INV_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    kern<<<blocks, 512, lds, c.stream>>>(a, ch);
}

template <bool SH, bool PSV, bool FAST>
void walk_launch(Ctx &c, const StepArgs &a, unsigned blocks) {
    INV_REQUIRE(!a.snap, INV_ERR_STATE, "three-step launches carry no snapshot role");
    INV_REQUIRE(c.walk_ch >= 8 && c.walk_ch % 8 == 0, INV_ERR_STATE, "walk chunk rows");
    const int lds = walk_lds_bytes(PSV);
    if (a.rec_ptr) walk_go(c, k_walk<SH, PSV, FAST, true>, lds, a, c.walk_ch, blocks);
    else walk_go(c, k_walk<SH, PSV, FAST, false>, lds, a, c.walk_ch, blocks);
}

template void walk_launch<true, false, true>(Ctx &, const StepArgs &, unsigned);
template void walk_launch<false, true, true>(Ctx &, const StepArgs &, unsigned);
template void walk_launch<true, false, false>(Ctx &, const StepArgs &, unsigned);
template void walk_launch<false, true, false>(Ctx &, const StepArgs &, unsigned);

}  // namespace inv
