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
// that of k_col, so results are bitwise identical to it.  Needs nz % V == 0.
#include "internal.h"
#include "step.h"

namespace inv {

constexpr int WBLOCK = 256;  // four wavefronts per tile

template <int V>
__device__ __forceinline__ void vload(const float *__restrict__ p, float (&o)[V]) {
    if constexpr (V == 2) {
        const float2 t = *reinterpret_cast<const float2 *>(p);
        o[0] = t.x;
        o[1] = t.y;
    } else {
        const float4 t = *reinterpret_cast<const float4 *>(p);
        o[0] = t.x;
        o[1] = t.y;
        o[2] = t.z;
        o[3] = t.w;
    }
}

template <int V>
__device__ __forceinline__ void vstore(float *__restrict__ p, const float (&o)[V]) {
    if constexpr (V == 2) *reinterpret_cast<float2 *>(p) = make_float2(o[0], o[1]);
    else *reinterpret_cast<float4 *>(p) = make_float4(o[0], o[1], o[2], o[3]);
}

// e[V + k] = row[c0 + k] for k in [-V, 2V): the lane's vector and its two
// neighbours from an LDS row of W columns (indices clamped into the row; the
// clamped values only reach columns no owned cell depends on)
template <int V, int W>
__device__ __forceinline__ void window(const float *row, int c0, float (&e)[3 * V]) {
    const int cl = c0 - V < 0 ? 0 : c0 - V;
    const int cr = c0 + V > W - V ? W - V : c0 + V;
    float t[V];
    vload<V>(row + cl, t);
#pragma unroll
    for (int v = 0; v < V; v++) e[v] = t[v];
    vload<V>(row + c0, t);
#pragma unroll
    for (int v = 0; v < V; v++) e[V + v] = t[v];
    vload<V>(row + cr, t);
#pragma unroll
    for (int v = 0; v < V; v++) e[2 * V + v] = t[v];
}

template <bool SH, bool PSV, int TX, int V, bool FAST>
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
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = lane * V;          // the lane's first LDS column
    const int jb = j0 - HL + c0;      // its first grid column (a multiple of V)
    const int jl = jb < 0 ? 0 : (jb > nz - V ? nz - V : jb);
    const bool jin = jb >= 0 && jb < nz;  // all V columns in the grid, or none
    const bool cown = jin && c0 >= HL && c0 < HL + S;
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
    const float *I1 = PSV ? a.inj_x : a.inj_y, *I2 = a.inj_z;
    float *O1 = PSV ? a.out_x : a.out_y, *O2 = a.out_z;

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
    float qv1[PV][V], qv2[PV][V], qr[PV][V], qu1[PV][V], qu2[PV][V];
#pragma unroll
    for (int p = 0; p < PV; p++) {
        const int vr = wave + p * NW;
        const long kc = (long)clampi(i0 - 1 + vr, 0, nx - 1) * nz + jl;
        vload<V>(m_rho + kc, qr[p]);
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
        vload<V>(m_mu + kc, qmu[p]);
        if (PSV) vload<V>(m_lam + kc, qlam[p]);
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
    if (a.rec_ptr) {
        r0 = a.rec_ptr[t];
        r1 = a.rec_ptr[t + 1];
    }
    const long it = a.it;
    const int fmt = a.snap_fmt;
    const int fu = PSV ? a.psv_f0 : 0, fv = PSV ? a.psv_f0 + 2 : 1;
    const long sb = (long)b * a.snap_slot_stride;
    float su1 = 1.0f, su2 = 1.0f, sv1 = 1.0f, sv2 = 1.0f;
    if (fmt != INV_SNAP_F32) {
        const int mb = b * a.nsnapf;
        if (a.snap_u) {
            su1 = snap_enc_scale(a.mx_u[mb + fu]);
            if (PSV) su2 = snap_enc_scale(a.mx_u[mb + fu + 1]);
        }
        if (a.snap_v) {
            sv1 = snap_enc_scale(a.mx_v[mb + fv]);
            if (PSV) sv2 = snap_enc_scale(a.mx_v[mb + fv + 1]);
        }
    }
    const bool track = a.acc_u || a.acc_v;
    float mu1 = 0.0f, mu2 = 0.0f, mv1 = 0.0f, mv2 = 0.0f;

    // 2. velocity of step n+1 on rows i0-1 .. i0+TX+1 (stress LDS rows vr .. vr+3)
#pragma unroll
    for (int p = 0; p < PV; p++) {
        const int vr = wave + p * NW;
        if (vr >= VR) break;
        const int i = i0 - 1 + vr;
        const bool rin = jin && i >= 0 && i < nx;
        const bool xin = i >= 2 && i < nx - 2;
        const float abx = a.abx[clampi(i, 0, nx - 1)];
        // cells of this row the wavefront covers: point lists are scanned with
        // one wave-uniform test per point before the per-lane one
        const long rlo = (long)i * nz + (j0 - HL), rhi = rlo + W - 1;
        float n1[V], n2[V];
        if (!PSV) {
            float x0[V], x1[V], x2[V], x3[V], ez[3 * V];
            vload<V>(&l_s1[vr * W + c0], x0);
            vload<V>(&l_s1[(vr + 1) * W + c0], x1);
            vload<V>(&l_s1[(vr + 2) * W + c0], x2);
            vload<V>(&l_s1[(vr + 3) * W + c0], x3);
            window<V, W>(&l_s2[(vr + 2) * W], c0, ez);
#pragma unroll
            for (int v = 0; v < V; v++) {
                const int j = jb + v;
                const long kc = (long)i * nz + j;
                float d = xin ? dm<FAST>(x2[v], x1[v], x3[v], x0[v], g.c8dx, g.c24dx, g.a1x, g.a2x)
                              : 0.0f;
                if (j >= 2 && j < nz - 2)
                    d += dm<FAST>(ez[V + v], ez[V + v - 1], ez[V + v + 1], ez[V + v - 2], g.c8dz,
                                  g.c24dz, g.a1z, g.a2z);
                for (int pp = p0; pp < p1; pp++) {
                    const int cp = a.inj_cell[pp];
                    if (cp < rlo || cp > rhi) continue;
                    if (cp == kc)
                        d += I1[(long)b * a.inj_slot_stride + (long)a.inj_row[pp] * a.nt + it];
                }
                const float ab = (abx * abz[v]) * abt[v];
                n1[v] = rin ? vel_update<FAST>(ab, qv1[p][v], g.dt, d, qr[p][v]) : 0.0f;
            }
        } else {
            float x0[V], x1[V], x2[V], x3[V], z0[V], z1[V], z2[V], z3[V], e3[3 * V], e2[3 * V];
            vload<V>(&l_s1[vr * W + c0], x0);
            vload<V>(&l_s1[(vr + 1) * W + c0], x1);
            vload<V>(&l_s1[(vr + 2) * W + c0], x2);
            vload<V>(&l_s1[(vr + 3) * W + c0], x3);
            vload<V>(&l_s3[vr * W + c0], z0);
            vload<V>(&l_s3[(vr + 1) * W + c0], z1);
            vload<V>(&l_s3[(vr + 2) * W + c0], z2);
            vload<V>(&l_s3[(vr + 3) * W + c0], z3);
            window<V, W>(&l_s3[(vr + 2) * W], c0, e3);
            window<V, W>(&l_s2[(vr + 2) * W], c0, e2);
#pragma unroll
            for (int v = 0; v < V; v++) {
                const int j = jb + v;
                const long kc = (long)i * nz + j;
                float dx_, dz_;
                if (xin) {
                    dx_ = dm<FAST>(x2[v], x1[v], x3[v], x0[v], g.c8dx, g.c24dx, g.a1x, g.a2x);
                    dz_ = dm<FAST>(z2[v], z1[v], z3[v], z0[v], g.c8dx, g.c24dx, g.a1x, g.a2x);
                } else {
                    dx_ = 0.0f;
                    dz_ = 0.0f;
                }
                if (j >= 2 && j < nz - 2) {
                    dx_ += dm<FAST>(e3[V + v], e3[V + v - 1], e3[V + v + 1], e3[V + v - 2],
                                    g.c8dz, g.c24dz, g.a1z, g.a2z);
                    dz_ += dm<FAST>(e2[V + v], e2[V + v - 1], e2[V + v + 1], e2[V + v - 2],
                                    g.c8dz, g.c24dz, g.a1z, g.a2z);
                }
                for (int pp = p0; pp < p1; pp++) {
                    const int cp = a.inj_cell[pp];
                    if (cp < rlo || cp > rhi) continue;
                    if (cp == kc) {
                        const long ks =
                            (long)b * a.inj_slot_stride + (long)a.inj_row[pp] * a.nt + it;
                        dx_ += I1[ks];
                        dz_ += I2[ks];
                    }
                }
                const float ab = (abx * abz[v]) * abt[v];
                n1[v] = rin ? vel_update<FAST>(ab, qv1[p][v], g.dt, dx_, qr[p][v]) : 0.0f;
                n2[v] = rin ? vel_update<FAST>(ab, qv2[p][v], g.dt, dz_, qr[p][v]) : 0.0f;
            }
        }
        vstore<V>(&l_v1[vr * W + c0], n1);
        if (PSV) vstore<V>(&l_v2[vr * W + c0], n2);
        // owned cells: v, u, snapshots, receivers
        if (vr >= 1 && vr <= TX && cown && i < nx) {
            const long o = so + (long)i * nz + jb;
            float nu1[V], nu2[V];
#pragma unroll
            for (int v = 0; v < V; v++) {
                const long kc = (long)i * nz + jb + v;
                float u = qu1[p][v];
                float uu = PSV ? qu2[p][v] : 0.0f;
                if (a.snap_u) {
                    snap_put(a.snap_u, fmt, sb + (long)fu * g.N + kc, u, su1);
                    if (PSV) snap_put(a.snap_u, fmt, sb + (long)(fu + 1) * g.N + kc, uu, su2);
                }
                u += n1[v] * g.dt;
                if (PSV) uu += n2[v] * g.dt;
                nu1[v] = u;
                nu2[v] = uu;
                if (a.snap_v) {
                    snap_put(a.snap_v, fmt, sb + (long)fv * g.N + kc,
                             a.v_after ? n1[v] : qv1[p][v], sv1);
                    if (PSV)
                        snap_put(a.snap_v, fmt, sb + (long)(fv + 1) * g.N + kc,
                                 a.v_after ? n2[v] : qv2[p][v], sv2);
                }
                if (track) {
                    mu1 = fmaxf(mu1, fabsf(u));
                    mv1 = fmaxf(mv1, fabsf(n1[v]));
                    if (PSV) {
                        mu2 = fmaxf(mu2, fabsf(uu));
                        mv2 = fmaxf(mv2, fabsf(n2[v]));
                    }
                }
                for (int pp = r0; pp < r1; pp++) {
                    const int cp = a.rec_cell[pp];
                    if (cp < rlo || cp > rhi) continue;
                    if (cp == kc) {
                        const long kr = (long)b * a.out_slot_stride + (long)a.rec_row[pp] * a.nt + it;
                        O1[kr] = a.rec_disp ? u : n1[v];
                        if (PSV) O2[kr] = a.rec_disp ? uu : n2[v];
                    }
                }
            }
            vstore<V>(V1o + o, n1);
            vstore<V>(U1 + o, nu1);
            if (PSV) {
                vstore<V>(V2o + o, n2);
                vstore<V>(U2 + o, nu2);
            }
        }
    }
    // max-abs of this tile's u and v after the step, one atomic per wavefront
    if (track) {
        mu1 = wave_max(mu1);
        mv1 = wave_max(mv1);
        if (PSV) {
            mu2 = wave_max(mu2);
            mv2 = wave_max(mv2);
        }
        if (lane == 0) {
            const int mb = b * a.nsnapf;
            if (a.acc_u) {
                atomicMax(a.acc_u + mb + fu, __float_as_uint(mu1));
                if (PSV) atomicMax(a.acc_u + mb + fu + 1, __float_as_uint(mu2));
            }
            if (a.acc_v) {
                atomicMax(a.acc_v + mb + fv, __float_as_uint(mv1));
                if (PSV) atomicMax(a.acc_v + mb + fv + 1, __float_as_uint(mv2));
            }
        }
    }
    __syncthreads();

    // 3. stress of step n+1 on the owned cells (velocity LDS rows tr .. tr+3)
    if (!cown) return;
    int q0 = 0, q1 = 0;
    if (PSV && a.sinj_ptr) {
        q0 = a.sinj_ptr[b * per + t];
        q1 = a.sinj_ptr[b * per + t + 1];
    }
#pragma unroll
    for (int p = 0; p < PT; p++) {
        const int tr = wave + p * NW;
        if (tr >= TX) break;
        const int i = i0 + tr;
        if (i >= nx) break;
        const bool xin = i >= 1 && i < nx - 2;
        const long o = so + (long)i * nz + jb;
        float o1[V], o2[V], o3[V];
        float y0[V], y1[V], y2[V], y3[V], e1[3 * V];
        vload<V>(&l_v1[tr * W + c0], y0);
        vload<V>(&l_v1[(tr + 1) * W + c0], y1);
        vload<V>(&l_v1[(tr + 2) * W + c0], y2);
        vload<V>(&l_v1[(tr + 3) * W + c0], y3);
        window<V, W>(&l_v1[(tr + 1) * W], c0, e1);
        float sa[V], sc[V];
        vload<V>(&l_s1[(tr + 3) * W + c0], sa);
        vload<V>(&l_s2[(tr + 3) * W + c0], sc);
        if (!PSV) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                const int j = jb + v;
                const float mu = qmu[p][v];
                const float dx_ =
                    xin ? dp<FAST>(y2[v], y1[v], y3[v], y0[v], g.c8dx, g.c24dx, g.a1x, g.a2x) : 0.0f;
                const float dz_ = (j >= 1 && j < nz - 2)
                                      ? dp<FAST>(e1[V + v + 1], e1[V + v], e1[V + v + 2],
                                                 e1[V + v - 1], g.c8dz, g.c24dz, g.a1z, g.a2z)
                                      : 0.0f;
                o1[v] = sa[v] + g.dt * mu * dx_;
                o2[v] = sc[v] + g.dt * mu * dz_;
            }
            vstore<V>(S1o + o, o1);
            vstore<V>(S2o + o, o2);
        } else {
            float w0[V], w1[V], w2[V], w3[V], e2[3 * V], sxz[V];
            vload<V>(&l_v2[tr * W + c0], w0);
            vload<V>(&l_v2[(tr + 1) * W + c0], w1);
            vload<V>(&l_v2[(tr + 2) * W + c0], w2);
            vload<V>(&l_v2[(tr + 3) * W + c0], w3);
            window<V, W>(&l_v2[(tr + 1) * W], c0, e2);
            vload<V>(&l_s3[(tr + 3) * W + c0], sxz);
#pragma unroll
            for (int v = 0; v < V; v++) {
                const int j = jb + v;
                const bool zin = j >= 1 && j < nz - 2;
                float xx, zx, xz, zz;
                if (xin) {
                    xx = dp<FAST>(y2[v], y1[v], y3[v], y0[v], g.c8dx, g.c24dx, g.a1x, g.a2x);
                    zx = dp<FAST>(w2[v], w1[v], w3[v], w0[v], g.c8dx, g.c24dx, g.a1x, g.a2x);
                } else {
                    xx = 0.0f;
                    zx = 0.0f;
                }
                if (zin) {
                    xz = dp<FAST>(e1[V + v + 1], e1[V + v], e1[V + v + 2], e1[V + v - 1], g.c8dz,
                                  g.c24dz, g.a1z, g.a2z);
                    zz = dp<FAST>(e2[V + v + 1], e2[V + v], e2[V + v + 2], e2[V + v - 1], g.c8dz,
                                  g.c24dz, g.a1z, g.a2z);
                } else {
                    xz = 0.0f;
                    zz = 0.0f;
                }
                const float mu = qmu[p][v], lam = qlam[p][v];
                o1[v] = sa[v] + g.dt * ((lam + 2 * mu) * xx + lam * zz);
                o2[v] = sc[v] + g.dt * ((lam + 2 * mu) * zz + lam * xx);
                o3[v] = sxz[v] + g.dt * (mu * (xz + zx));
                // explosive sources this tile owns: stf_y into sxx and szz
                const long kc = (long)i * nz + j;
                for (int pp = q0; pp < q1; pp++)
                    if (a.sinj_cell[pp] == kc) {
                        const float f =
                            a.inj_y[(long)b * a.inj_slot_stride + (long)a.sinj_row[pp] * a.nt + it];
                        o1[v] += f;
                        o2[v] += f;
                    }
            }
            vstore<V>(S1o + o, o1);
            vstore<V>(S2o + o, o2);
            vstore<V>(S3o + o, o3);
        }
    }
}

template <bool SH, bool PSV, bool FAST>
void wide_launch(Ctx &c, const StepArgs &a, unsigned blocks) {
    if (c.col_v == 2) {
        switch (c.tile_x) {
            case 4: k_colv<SH, PSV, 4, 2, FAST><<<blocks, WBLOCK, 0, c.stream>>>(a); break;
            case 8: k_colv<SH, PSV, 8, 2, FAST><<<blocks, WBLOCK, 0, c.stream>>>(a); break;
            default: k_colv<SH, PSV, 16, 2, FAST><<<blocks, WBLOCK, 0, c.stream>>>(a); break;
        }
    } else {
        switch (c.tile_x) {
            case 8: k_colv<SH, PSV, 8, 4, FAST><<<blocks, WBLOCK, 0, c.stream>>>(a); break;
            default: k_colv<SH, PSV, 4, 4, FAST><<<blocks, WBLOCK, 0, c.stream>>>(a); break;
        }
    }
}

template void wide_launch<true, false, true>(Ctx &, const StepArgs &, unsigned);
template void wide_launch<false, true, true>(Ctx &, const StepArgs &, unsigned);
template void wide_launch<true, false, false>(Ctx &, const StepArgs &, unsigned);
template void wide_launch<false, true, false>(Ctx &, const StepArgs &, unsigned);

}  // namespace inv
