// Persistent two-step kernel for grids whose shots fit the chip at one tile
// per workgroup.  One launch runs `iters` iterations of the two-step kernel's
// phases (k_pair, fdm_pair.hip: divS -> addSTF -> updateV -> divV -> updateS
// -> saveRec twice, solver/fdm.h:534-551) on tiles of TX rows spanning every
// column of the grid.  Between iterations a tile's wavefields stay on the
// chip: the stress in LDS, the velocity, displacement and model in
// registers.  Only the rows a neighbouring tile reads as its halo cross
// memory: a tile writes its owned rows of stress and velocity to an exchange
// buffer during an iteration; at its end every wavefront waits for its
// stores, the workgroup meets at a barrier and one thread publishes the
// iteration's tag in the tile's flag with a device-scope release.  At the
// start of the next iteration one wavefront polls the flags of the tiles
// its halo rows (six each side) come from (the neighbour on each side for
// 8-row tiles, two on each side for 4-row tiles, and the tile itself, whose
// rows the clamped edge rows read), then acquires, and after a barrier the
// workgroup reads the rows.  A tile waits only for the tiles its halo
// reaches, never for the whole grid.
//
// Two buffers alternate by iteration parity.  The tiles a halo reaches are
// the tiles that read it (the relation is symmetric), and a tile cannot
// finish iteration k+1 before each of them has published iteration k, which
// they do after reading its iteration k-1 rows: the buffer it overwrites is
// never still being read.  Tags are epoch * 8192 + k + 1 with the epoch
// counted up on the device before every launch (graph replays included), so
// a flag left over from another launch never matches.
//
// Every tile waits only for tiles that are running: the host launches the
// kernel only when the whole grid fits the device at once (occupancy), and
// every wait gives up after a bounded number of polls, setting an error word
// the host turns into an error, so that a grid which is not co-resident ends
// instead of hanging.
//
// Results are bitwise those of the same number of k_pair launches: every cell
// is evaluated by the row functions of colv.h, and a halo row holds what
// k_pair would load from memory: the owning tile's value inside the grid,
// 0 for stress outside it, the value at the clamped row and column for
// velocity outside it.  The last iteration stores the state as k_pair does.
#include "colv.h"

namespace inv {

// exchanged fields of a tile row: s1, s2, s3 (stress) and v1, v2 (velocity)
constexpr int RES_NF = 5;

// snapshot role of forward step s (SnapPlan::local in fdm_fused.hip)
__device__ __forceinline__ StepRoles res_role(const ResArgs &ra, long s) {
    if (!ra.snap || (s + 1) % ra.sfe != 0) return StepRoles{nullptr, nullptr};
    const long l = ra.nsfe - (s + 1) / ra.sfe - ra.sbase;
    if (l < 0 || l >= ra.nslots) return StepRoles{nullptr, nullptr};
    return StepRoles{ra.snap + l * ra.slot_bytes, ra.sexp ? ra.sexp + l * ra.tab_bytes : nullptr};
}

__global__ void k_res_epoch(unsigned *ctl) { ctl[0] += 1u; }

// the lane's V values of snapshot field f of shot slot b, row i, columns
// jl .. jl+V-1 of snapshot slot `slot`, decoded to FP32 (the reads of
// k_interact_col, fdm.hip)
template <int V>
__device__ __forceinline__ void res_snap_row(const StepArgs &a, const ResArgs &ra, long slot,
                                             int b, int f, int i, int jl, float (&o)[V]) {
    const long k = (long)b * a.snap_slot_stride + (long)f * a.g.N + (long)i * a.g.nz + jl;
    const char *base = ra.snap + slot * ra.slot_bytes;
    if (a.snap_fmt == INV_SNAP_F32) {
        vload<V>(reinterpret_cast<const float *>(base) + k, o);
        return;
    }
    const unsigned short *q = reinterpret_cast<const unsigned short *>(base) + k;
    const signed char *tab = ra.sexp + slot * ra.tab_bytes;
    const int w = *reinterpret_cast<const int *>(tab);
    const int segw = w > 0 ? w : 1;
    const signed char *erow = tab + snaptab_row(b, a.nsnapf, f, a.g.nx, i, a.nseg);
    const unsigned char *brow = reinterpret_cast<const unsigned char *>(tab) +
                                snaptab_blk(a.sblk, b, a.nsnapf, f, a.g.nx, i, a.nbb);
#pragma unroll
    for (int v = 0; v < V; v++)
        o[v] = snap_dec16(q[v], snap_cell_exp(erow, brow, a.nseg, a.nbb, segw, jl + v),
                          a.snap_fmt);
}

// Kernel accumulation (interaction*, solver/fdm.h:198-219, with the D+
// derivatives of fdm.h:586-605) on the owned cells of grid row i after step
// `last`: the expressions and their order of k_interact_col, so the kernels
// agree bit for bit.  lu1 / lu2: the adjoint displacement in LDS rows of W
// columns, row i at LDS row q (rows q-1 .. q+2 read); v1 / v2: the adjoint
// velocity at the lane's cells.  Every lane of the wavefront calls it.
template <bool PSV, int V, int W, bool FAST>
__device__ __forceinline__ void res_interact_row(const StepArgs &a, const ResArgs &ra, int b,
                                                 int i, int q, int c0, int jb, int jl, bool own,
                                                 long last, const float *lu1, const float *lu2,
                                                 const float (&v1)[V], const float (&v2)[V]) {
    const Geo &g = a.g;
    const int nx = g.nx, nz = g.nz;
    const long slot = last / ra.sfe;
    const int fu = PSV ? a.psv_f0 : 0, fv = PSV ? a.psv_f0 + 2 : 1;
    const bool xin = i >= 1 && i < nx - 2;
    auto rowc = [&](int r) { return r < 0 ? 0 : (r > nx - 1 ? nx - 1 : r); };
    // D+ along x (rows i-1 .. i+2) and z (columns j-1 .. j+2, neighbouring
    // lanes) of one field at the lane's cells, one field at a time (every
    // lane of the wavefront takes part in the z window)
    auto derivs = [&](const float (&r0)[V], const float (&r1)[V], const float (&r2)[V],
                      const float (&r3)[V], float (&dx)[V], float (&dz)[V]) {
        float e[3 * V];
        window_reg<V>(r1, e);
#pragma unroll
        for (int v = 0; v < V; v++) {
            dx[v] = dp<FAST>(r2[v], r1[v], r3[v], r0[v], g.c8dx, g.c24dx, g.a1x, g.a2x);
            dz[v] = dp<FAST>(e[V + v + 1], e[V + v], e[V + v + 2], e[V + v - 1], g.c8dz, g.c24dz,
                             g.a1z, g.a2z);
        }
    };
    auto lds_derivs = [&](const float *lu, float (&dx)[V], float (&dz)[V]) {
        float r0[V], r1[V], r2[V], r3[V];
        vload<V>(&lu[(q - 1) * W + c0], r0);
        vload<V>(&lu[q * W + c0], r1);
        vload<V>(&lu[(q + 1) * W + c0], r2);
        vload<V>(&lu[(q + 2) * W + c0], r3);
        derivs(r0, r1, r2, r3, dx, dz);
    };
    auto snap_derivs = [&](int f, float (&dx)[V], float (&dz)[V]) {
        float r0[V], r1[V], r2[V], r3[V];
        res_snap_row<V>(a, ra, slot, b, f, rowc(i - 1), jl, r0);
        res_snap_row<V>(a, ra, slot, b, f, i, jl, r1);
        res_snap_row<V>(a, ra, slot, b, f, rowc(i + 1), jl, r2);
        res_snap_row<V>(a, ra, slot, b, f, rowc(i + 2), jl, r3);
        derivs(r0, r1, r2, r3, dx, dz);
    };
    // adjoint and forward displacement derivatives: x and z of u1 (ux or
    // uy), of u2 (uz), and of their snapshots
    float d1x[V], d1z[V], f1x[V], f1z[V], d2x[V], d2z[V], f2x[V], f2z[V];
    lds_derivs(lu1, d1x, d1z);
    snap_derivs(fu, f1x, f1z);
    if (PSV) {
        lds_derivs(lu2, d2x, d2z);
        snap_derivs(fu + 1, f2x, f2z);
    }
    if (!own) return;
    float vf1[V], vf2[V];
    res_snap_row<V>(a, ra, slot, b, fv, i, jl, vf1);
    if (PSV) res_snap_row<V>(a, ra, slot, b, fv + 1, i, jl, vf2);
    const long k = (long)b * ra.k_slot_stride + (long)i * nz + jb;
    float kr[V], km[V], kl[V];
    if (ra.inv_rho) vload<V>(ra.kr + k, kr);
    if (ra.inv_mu) vload<V>(ra.km + k, km);
    if (PSV && ra.inv_lambda) vload<V>(ra.kl + k, kl);
    const float ndt = ra.ndt;
#pragma unroll
    for (int v = 0; v < V; v++) {
        const int j = jb + v;
        const bool zin = j >= 1 && j < nz - 2;
        if (!PSV) {
            const float dx_ = xin ? d1x[v] : 0.0f;
            const float dz_ = zin ? d1z[v] : 0.0f;
            const float dxf = xin ? f1x[v] : 0.0f;
            const float dzf_ = zin ? f1z[v] : 0.0f;
            if (ra.inv_rho) kr[v] -= vf1[v] * v1[v] * ndt;
            if (ra.inv_mu) km[v] -= (dx_ * dxf + dz_ * dzf_) * ndt;
        } else {
            const float xx = xin ? d1x[v] : 0.0f, zx = xin ? d2x[v] : 0.0f;
            const float xxf = xin ? f1x[v] : 0.0f, zxf = xin ? f2x[v] : 0.0f;
            const float xz = zin ? d1z[v] : 0.0f, zz = zin ? d2z[v] : 0.0f;
            const float xzf = zin ? f1z[v] : 0.0f, zzf = zin ? f2z[v] : 0.0f;
            if (ra.inv_rho) kr[v] -= (vf1[v] * v1[v] + vf2[v] * v2[v]) * ndt;
            if (ra.inv_mu) km[v] -= (2 * xx * xxf + 2 * zz * zzf + (xz + zx) * (zxf + xzf)) * ndt;
            if (ra.inv_lambda) kl[v] -= ((xx + zz) * (xxf + zzf)) * ndt;
        }
    }
    if (ra.inv_rho) vstore<V>(ra.kr + k, kr);
    if (ra.inv_mu) vstore<V>(ra.km + k, km);
    if (PSV && ra.inv_lambda) vstore<V>(ra.kl + k, kl);
}

// INT: the adjoint variant that accumulates the kernels in the launch (one
// wave type per context); it tracks the displacement on the halo rows -1, TX
// and TX+1 (u += v dt with the velocity it computes there anyway, the
// owner's values bit for bit), which the D+ x-derivatives read.
template <bool SH, bool PSV, int TX, int V, int NW, bool FAST, bool SNAP, bool REC, bool INT>
__global__ __launch_bounds__(64 * NW, 1) void k_res(const StepArgs a, const ResArgs ra) {
    static_assert(SH != PSV, "one wave type per launch");
    static_assert(TX % 4 == 0, "tile rows a multiple of four");
    constexpr int HL = pair_halo(V), W = 64 * V, S = pair_strip(V);
    constexpr int SR = TX + 12;  // stress rows x = -6 .. TX+5
    constexpr int VR = TX + 9;   // velocity rows x = -4 .. TX+4
    constexpr int PS = (SR + NW - 1) / NW;
    constexpr int P = (VR + NW - 1) / NW;
    extern __shared__ __align__(16) float smem[];
    __shared__ int l_abort;
    float *const l_s1 = smem;
    float *const l_s2 = l_s1 + SR * W;
    float *const l_s3 = PSV ? l_s2 + VR * W : l_s2;
    float *const l_v1 = PSV ? l_s3 + SR * W : l_s2 + VR * W;
    float *const l_v2 = PSV ? l_v1 + VR * W : l_v1;

    // one tile across z: tile tx of shot slot b (XCD-aware order, as k_pair)
    const int nw = gridDim.x, w = blockIdx.x;
    const int q = nw >> 3, rmd = nw & 7, xcd = w & 7;
    const int L = xcd * q + (xcd < rmd ? xcd : rmd) + (w >> 3);
    const int per = a.tiles_x;
    const int b = L / per;
    const int tx = L - b * per;
    const int i0 = tx * TX;
    const Geo &g = a.g;
    const int nx = g.nx, nz = g.nz;
    const long so = (long)b * g.N;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c0 = lane * V;
    const int jb = -HL + c0;
    const int jl = jb < 0 ? 0 : (jb > nz - V ? nz - V : jb);
    const bool jin = jb >= 0 && jb < nz;
    const bool cown = jin && c0 >= HL && c0 < HL + S;
    const int csrc = jl + HL;  // LDS column of grid column jl (the lane's own when jin)
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

    // the state at the launch's first step, as k_pair loads it
    float s1[PS][V], s2[PS][V], s3[PS][V];
#pragma unroll
    for (int p = 0; p < PS; p++) {
        const int r = wave + p * NW;
        const long k = so + (long)clampi(i0 - 6 + r, 0, nx - 1) * nz + jl;
        vload<V>(S1 + k, s1[p]);
        if (r >= 2 && r < SR - 1) vload<V>(S2 + k, s2[p]);
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
        if (x >= (INT ? -1 : 0) && x < TX + (INT ? 2 : 0)) {
            vload<V>(U1 + so + kc, qu1[p]);
            if (PSV) vload<V>(U2 + so + kc, qu2[p]);
        }
        if (x >= -3 && x < TX + 3) {
            vload_g<V>(m_mu + kc, qmu[p]);
            if (PSV) vload_g<V>(m_lam + kc, qlam[p]);
        }
    }
    float abz[V], abt[V];
    vload<V>(a.abzb + jl, abz);
    vload<V>(a.abzt + jl, abt);
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
    if (threadIdx.x == 0) l_abort = 0;
    __syncthreads();

    int p0 = 0, p1 = 0;
    if (a.inj_ptr) {
        const int set = a.inj_per_slot ? b : 0;
        p0 = a.inj_ptr[set * per + tx];
        p1 = a.inj_ptr[set * per + tx + 1];
    }
    int r0 = 0, r1 = 0;
    if (REC && a.rec_ptr) {
        r0 = a.rec_ptr[tx];
        r1 = a.rec_ptr[tx + 1];
    }
    int q0 = 0, q1 = 0;
    if (PSV && a.sinj_ptr) {
        q0 = a.sinj_ptr[b * per + tx];
        q1 = a.sinj_ptr[b * per + tx + 1];
    }
    const LanePts ip = lane_pts(a.inj_cell, a.inj_row, p0, p1, lane);
    const LanePts rp = lane_pts(a.rec_cell, a.rec_row, r0, r1, lane);
    const LanePts xp = lane_pts(a.sinj_cell, a.sinj_row, q0, q1, lane);
    const int fu = PSV ? a.psv_f0 : 0, fv = PSV ? a.psv_f0 + 2 : 1;
    const long sb = (long)b * a.snap_slot_stride;
    const StepRoles none{nullptr, nullptr};
    const unsigned ep = __builtin_amdgcn_readfirstlane(ra.ctl[0]);
    // exchange rows of this slot: [tile][parity][field][row][64 V columns],
    // and a flag per tile and parity holding the tag of the iteration whose
    // rows it last published
    float *const xb = ra.xbuf + (long)b * per * (2L * RES_NF * TX * W);
    auto xrow = [&](int tile, int par, int f, int r) {
        return xb + ((((long)tile * 2 + par) * RES_NF + f) * TX + r) * W;
    };
    unsigned *const xf = ra.xflag + (long)b * per * 2;
    constexpr int REACH = (6 + TX - 1) / TX;  // tiles a halo reaches on each side

    for (int k = 0; k < ra.iters; k++) {
        const long it = a.it + 2L * k;
        const int par = k & 1;
        const unsigned tag = ep * 8192u + (unsigned)k + 1u;
        const bool last = k == ra.iters - 1;
        const StepRoles r1s = SNAP ? res_role(ra, it) : none;
        const StepRoles r2 = SNAP ? res_role(ra, it + 1) : none;

        if (k > 0) {
            // wait until the tiles this tile's halo reaches (and the tile
            // itself, whose rows the clamped edge rows read) have published
            // the previous iteration (parity par ^ 1): one lane polls each
            // flag, then an acquire makes their rows visible to the
            // workgroup (after the barrier)
            const unsigned want = tag - 1u;
            const int pp = par ^ 1;
            if (wave == 0) {
                const int t = tx - REACH + lane;
                const bool need = lane <= 2 * REACH && t >= 0 && t < per;
                unsigned *const fl = xf + (need ? t : tx) * 2 + pp;
                bool timed_out = ra.spin_limit < 0;
                for (int spins = 0; !timed_out; spins++) {
                    const bool ok =
                        !need || __hip_atomic_load(fl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == want;
                    if (__all(ok)) break;
                    if (spins >= ra.spin_limit) timed_out = true;
                    else __builtin_amdgcn_s_sleep(1);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                if (timed_out && lane == 0) {
                    l_abort = 1;
                    __hip_atomic_store(ra.ctl + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            __syncthreads();
            if (l_abort) return;  // the whole workgroup, after the barrier
            // halo rows: stress 0 outside the grid, velocity at the clamped
            // row and column
#pragma unroll
            for (int p = 0; p < PS; p++) {
                const int r = wave + p * NW, x = r - 6, i = i0 + x;
                if (r >= SR || (x >= 0 && x < TX && i < nx)) continue;  // wave-uniform
                const bool has2 = r >= 2 && r < SR - 1;
                float t1[V], t2[V], t3[V];
#pragma unroll
                for (int v = 0; v < V; v++) t1[v] = t2[v] = t3[v] = 0.0f;
                if (i >= 0 && i < nx) {  // wave-uniform
                    const int tt = i / TX;
                    const float *const q = xrow(tt, pp, 0, i - tt * TX) + c0;
                    if (jin) {
                        vload<V>(q, t1);
                        if (has2) vload<V>(q + TX * W, t2);
                        if (PSV) vload<V>(q + 2 * TX * W, t3);
                    }
                }
                vstore<V>(&l_s1[r * W + c0], t1);
                if (has2) vstore<V>(&l_s2[(r - 2) * W + c0], t2);
                if (PSV) vstore<V>(&l_s3[r * W + c0], t3);
            }
#pragma unroll
            for (int p = 0; p < P; p++) {
                const int x = wave - 4 + p * NW, i = i0 + x;
                if (x >= TX + 5) break;
                if (x >= 0 && x < TX && i < nx) continue;  // owned (phase D): wave-uniform
                const int ic = clampi(i, 0, nx - 1);
                const int tt = ic / TX;
                const float *const q = xrow(tt, pp, 3, ic - tt * TX) + csrc;
                vload<V>(q, qv1[p]);
                if (PSV) vload<V>(q + TX * W, qv2[p]);
            }
            __syncthreads();
        }

        // A. velocity of step n+1 on rows x = -4 .. TX+4 (LDS row x + 4)
#pragma unroll
        for (int p = 0; p < P; p++) {
            const int x = wave - 4 + p * NW;
            if (x >= TX + 5) break;
            const int i = i0 + x;
            const bool rin = jin && i >= 0 && i < nx;
            const int rlo = i * nz - HL, rhi = rlo + W - 1;
            float n1[V], n2[V];
            vel_row<PSV, V, W, FAST>(a, l_s1, l_s2, l_s3, x + 6, x + 4, c0, i, jb, rin, zc,
                                     qab[p], abz, abt, qv1[p], qv2[p], qr[p], ip, rlo, rhi, b, it,
                                     n1, n2);
            vstore<V>(&l_v1[(x + 4) * W + c0], n1);
            if (PSV) vstore<V>(&l_v2[(x + 4) * W + c0], n2);
            if (x >= 0 && x < TX && i < nx)  // wave-uniform
                owned_row<PSV, V, FAST>(a, r1s, fu, fv, sb, i, jb, cown, it, b, 0, S, lane, rp,
                                        rlo, rhi, qu1[p], qu2[p], n1, n2);
            if (INT && (x == -1 || x == TX || x == TX + 1)) {
#pragma unroll
                for (int v = 0; v < V; v++) {
                    qu1[p][v] = u_step<FAST>(qu1[p][v], n1[v], g.dt);
                    if (PSV) qu2[p][v] = u_step<FAST>(qu2[p][v], n2[v], g.dt);
                }
            }
#pragma unroll
            for (int v = 0; v < V; v++) {
                qv1[p][v] = n1[v];
                if (PSV) qv2[p][v] = n2[v];
            }
        }
        __syncthreads();

        // B. stress of step n+1 on rows x = -3 .. TX+2, in place
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

        // C. velocity of step n+2 on rows x = -1 .. TX+1, in place; owned rows
        // keep it (and publish it, or store it with u after the last iteration)
#pragma unroll
        for (int p = 0; p < P; p++) {
            const int x = wave - 4 + p * NW;
            if (x < -1) continue;
            if (x >= TX + 2) break;
            const int i = i0 + x;
            const bool rin = jin && i >= 0 && i < nx;
            const int rlo = i * nz - HL, rhi = rlo + W - 1;
            float n1[V], n2[V];
            vel_row<PSV, V, W, FAST>(a, l_s1, l_s2, l_s3, x + 6, x + 4, c0, i, jb, rin, zc,
                                     qab[p], abz, abt, qv1[p], qv2[p], qr[p], ip, rlo, rhi, b,
                                     it + 1, n1, n2);
            vstore<V>(&l_v1[(x + 4) * W + c0], n1);
            if (PSV) vstore<V>(&l_v2[(x + 4) * W + c0], n2);
            if (INT && (x == -1 || x == TX || x == TX + 1)) {
#pragma unroll
                for (int v = 0; v < V; v++) {
                    qu1[p][v] = u_step<FAST>(qu1[p][v], n1[v], g.dt);
                    if (PSV) qu2[p][v] = u_step<FAST>(qu2[p][v], n2[v], g.dt);
                }
            }
            if (x >= 0 && x < TX && i < nx) {  // wave-uniform
                owned_row<PSV, V, FAST>(a, r2, fu, fv, sb, i, jb, cown, it + 1, b, 0, S, lane, rp,
                                        rlo, rhi, qu1[p], qu2[p], n1, n2);
                if (cown) {
                    if (last) {
                        const long o = so + (long)i * nz + jb;
                        vstore<V>(V1o + o, n1);
                        vstore<V>(U1 + o, qu1[p]);
                        if (PSV) {
                            vstore<V>(V2o + o, n2);
                            vstore<V>(U2 + o, qu2[p]);
                        }
                    } else {
                        vstore<V>(xrow(tx, par, 3, x) + c0, n1);
                        if (PSV) vstore<V>(xrow(tx, par, 4, x) + c0, n2);
                    }
                }
#pragma unroll
                for (int v = 0; v < V; v++) {
                    qv1[p][v] = n1[v];
                    if (PSV) qv2[p][v] = n2[v];
                }
            }
        }
        __syncthreads();

        // D. stress of step n+2 on the owned rows: kept in LDS in place (0 in
        // lanes outside the grid, as k_pair loads them) and published, or
        // stored after the last iteration
#pragma unroll
        for (int p = 0; p < P; p++) {
            const int x = wave - 4 + p * NW;
            if (x < 0) continue;
            if (x >= TX) break;
            const int i = i0 + x;
            if (i >= nx) break;
            const int r = x + 6;
            float sa[V], sc[V], sxz[V], o1[V], o2[V], o3[V];
            vload<V>(&l_s1[r * W + c0], sa);
            vload<V>(&l_s2[(x + 4) * W + c0], sc);
            if (PSV) vload<V>(&l_s3[r * W + c0], sxz);
            stress_row<PSV, V, W, FAST>(g, l_v1, l_v2, x + 4, c0, i, jb, zc, sa, sc, sxz, qmu[p],
                                        qlam[p], o1, o2, o3);
            if (PSV) explode_row<V>(a, xp, i, jb, b, it + 1, o1, o2);
            if (last) {
                if (cown) {
                    const long o = so + (long)i * nz + jb;
                    vstore<V>(S1o + o, o1);
                    vstore<V>(S2o + o, o2);
                    if (PSV) vstore<V>(S3o + o, o3);
                }
                continue;
            }
            // the velocity of lanes outside the grid becomes the clamped
            // column's (phase C's LDS row), as k_pair would load it
            if (!jin) {
                vload<V>(&l_v1[(x + 4) * W + csrc], qv1[p]);
                if (PSV) vload<V>(&l_v2[(x + 4) * W + csrc], qv2[p]);
            }
            if (cown) {
                vstore<V>(xrow(tx, par, 0, x) + c0, o1);
                vstore<V>(xrow(tx, par, 1, x) + c0, o2);
                if (PSV) vstore<V>(xrow(tx, par, 2, x) + c0, o3);
            } else {
#pragma unroll
                for (int v = 0; v < V; v++) o1[v] = o2[v] = o3[v] = 0.0f;
            }
            vstore<V>(&l_s1[r * W + c0], o1);
            vstore<V>(&l_s2[(x + 4) * W + c0], o2);
            if (PSV) vstore<V>(&l_s3[r * W + c0], o3);
        }
        if (!last) {
            // publish: every wavefront's stores of the iteration have
            // completed before the barrier, then one release (write-back to
            // device scope) and the flag
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (threadIdx.x == 0)
                __hip_atomic_store(xf + tx * 2 + par, tag, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (INT && (it + 1) % ra.sfe == 0) {
            // kernel accumulation after step it + 1: the displacement of rows
            // -1 .. TX+1 staged in the velocity rows of LDS (free until the
            // next phase A, behind the next iteration's barriers)
            __syncthreads();
#pragma unroll
            for (int p = 0; p < P; p++) {
                const int x = wave - 4 + p * NW;
                if (x < -1) continue;
                if (x > TX + 1) break;
                vstore<V>(&l_v1[(x + 4) * W + c0], qu1[p]);
                if (PSV) vstore<V>(&l_v2[(x + 4) * W + c0], qu2[p]);
            }
            __syncthreads();
#pragma unroll
            for (int p = 0; p < P; p++) {
                const int x = wave - 4 + p * NW;
                if (x < 0) continue;
                if (x >= TX) break;
                const int i = i0 + x;
                if (i >= nx) break;
                res_interact_row<PSV, V, W, FAST>(a, ra, b, i, x + 4, c0, jb, jl, cown, it + 1,
                                                  l_v1, l_v2, qv1[p], qv2[p]);
                // one row at a time: interleaving the rows' loads would need
                // more registers than the kernel has left
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
}

template <class K>
static void res_go(Ctx &c, K kern, int nw, int lds, const StepArgs &a, const ResArgs &ra,
                   unsigned blocks) {
    // every tile waits for its neighbours: the whole grid must be resident
    static int fit = -1;  // blocks per CU of this instantiation
    if (fit < 0) {
        if (lds > 65536)
            INV_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        int nb = 0;
        INV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(kern), 64 * nw, lds));
        fit = nb;
    }
    int dev = 0, cus = 0;
    INV_HIP(hipGetDevice(&dev));
    INV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    INV_REQUIRE((long)fit * cus >= (long)blocks, INV_ERR_UNSUPPORTED,
                "persistent kernel: the tiles of this launch do not fit the device at once "
                "(fewer shots per launch or a smaller grid)");
    k_res_epoch<<<1, 1, 0, c.stream>>>(ra.ctl);
    kern<<<blocks, 64 * nw, lds, c.stream>>>(a, ra);
}

template <bool SH, bool PSV, int TX, bool FAST>
static void res_launch_tx(Ctx &c, const StepArgs &a, const ResArgs &ra, unsigned blocks) {
    constexpr int V = RES_V, NW = RES_NW;
    const int lds = pair_lds_bytes(PSV, TX, V);
    if (ra.int_on)
        res_go(c, k_res<SH, PSV, TX, V, NW, FAST, false, false, true>, NW, lds, a, ra, blocks);
    else if (ra.snap)
        res_go(c, k_res<SH, PSV, TX, V, NW, FAST, true, true, false>, NW, lds, a, ra, blocks);
    else if (a.rec_ptr)
        res_go(c, k_res<SH, PSV, TX, V, NW, FAST, false, true, false>, NW, lds, a, ra, blocks);
    else res_go(c, k_res<SH, PSV, TX, V, NW, FAST, false, false, false>, NW, lds, a, ra, blocks);
}

template <bool SH, bool PSV, bool FAST>
void res_launch(Ctx &c, const StepArgs &a, const ResArgs &ra, unsigned blocks) {
    if (c.pair_tx == 4) res_launch_tx<SH, PSV, 4, FAST>(c, a, ra, blocks);
    else res_launch_tx<SH, PSV, 8, FAST>(c, a, ra, blocks);
}

template void res_launch<true, false, true>(Ctx &, const StepArgs &, const ResArgs &, unsigned);
template void res_launch<false, true, true>(Ctx &, const StepArgs &, const ResArgs &, unsigned);
template void res_launch<true, false, false>(Ctx &, const StepArgs &, const ResArgs &, unsigned);
template void res_launch<false, true, false>(Ctx &, const StepArgs &, const ResArgs &, unsigned);

}  // namespace inv
