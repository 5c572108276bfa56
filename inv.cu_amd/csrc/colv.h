// Row updates of the column kernels with V adjacent columns per lane (the
// wide one-step kernel k_colv, fdm_wide.hip, and the two-step kernel k_pair,
// fdm_pair.hip).  One function per half step, so both kernels evaluate every
// cell with the same expressions in the same order (those of k_col and of the
// reference's divS / updateV / divV / updateS, solver/fdm.h:534-551) and give
// bitwise the same wavefields.
#pragma once

#include "step.h"

namespace inv {

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

// Loads through a pointer the compiler cannot place (one read from memory,
// such as a per-slot model pointer) are flat loads; flat loads count on the
// same counter as LDS, so every wait for an LDS read would wait for them too.
// Global memory is what every such pointer addresses.
// (Native vector types: a copy of HIP's float2 class would pass through a
// generic reference and lose the address space.)
using gfloat = __attribute__((address_space(1))) const float;
template <int V>
__device__ __forceinline__ void vload_g(const float *p, float (&o)[V]) {
    typedef float vec __attribute__((ext_vector_type(V)));
    const vec t = *(const __attribute__((address_space(1))) vec *)p;
#pragma unroll
    for (int v = 0; v < V; v++) o[v] = t[v];
}

// A wave-uniform value read with a vector load, into a vector register: a
// scalar load would be waited for together with the LDS reads that follow
// (both count on one counter) in every row that uses it.
__device__ __forceinline__ float vload_uniform(const float *p, int k) {
    int kv;
    asm volatile("v_mov_b32 %0, %1" : "=v"(kv) : "s"(k));
    return *(const gfloat *)(p + kv);
}

template <int V>
__device__ __forceinline__ void vstore(float *__restrict__ p, const float (&o)[V]) {
    if constexpr (V == 2) *reinterpret_cast<float2 *>(p) = make_float2(o[0], o[1]);
    else *reinterpret_cast<float4 *>(p) = make_float4(o[0], o[1], o[2], o[3]);
}

// e[V + k] = row[c0 + k] for k in [-V, 2V): the lane's vector of an LDS row
// and its two neighbours, from the lane's own vector c = row[c0 .. c0+V-1]
// and the neighbouring lanes' (DPP): no LDS reads beyond c (reading the
// neighbours from LDS took narrowed single-word reads at a two-word stride,
// two-way bank conflicts).  The edge lanes get 0 for the missing neighbour; they are halo
// lanes, whose cells no owned cell depends on.  Every lane must be active.
template <int V>
__device__ __forceinline__ void window_reg(const float (&c)[V], float (&e)[3 * V]) {
#pragma unroll
    for (int v = 0; v < V; v++) {
        e[v] = lane_prev(c[v]);
        e[V + v] = c[v];
        e[2 * V + v] = lane_next(c[v]);
    }
}

// A tile's entries p0 .. p1-1 of a point list, the first 64 held one per
// lane (read once per block, with the other loads).  Scanning them per row
// is then a lane read per entry instead of a scalar load and its round trip
// (tiles holding many receivers spent most of their time in those).  Entries
// past the first 64 are read from memory.  A lane read of an inactive lane's
// register is undefined, so lane_pts and scan_pts are only called where
// every lane of the wavefront is active (wave-uniform control flow); per-lane
// conditions go inside the scan's callback.
struct LanePts {
    int p0, p1, cell, row;
    bool sorted = false;        // entries in cell order (scan_pts may bisect)
    int cell2 = -1, row2 = 0;   // entries 64 .. 127 (lane_pts_sorted)
};

__device__ __forceinline__ LanePts lane_pts(const int *cell, const int *row, int p0, int p1,
                                            int lane) {
    LanePts t{p0, p1, -1, 0};
    if (lane < p1 - p0) {
        t.cell = cell[p0 + lane];
        t.row = row[p0 + lane];
    }
    return t;
}

// The same for a list whose buckets are in cell order, holding up to 128
// entries in lanes (two per lane) for scan_pts' bisection.
__device__ __forceinline__ LanePts lane_pts_sorted(const int *cell, const int *row, int p0,
                                                   int p1, int lane) {
    LanePts t = lane_pts(cell, row, p0, p1, lane);
    t.sorted = true;
    if (lane + 64 < p1 - p0) {
        t.cell2 = cell[p0 + 64 + lane];
        t.row2 = row[p0 + 64 + lane];
    }
    return t;
}

// f(cell, row) for the entries whose cell lies in rlo .. rhi, in list order.
// Cell indices are those of one shot's grid (int); callers compute what
// only a matching entry needs inside f, so a tile without entries (most of
// them) pays for nothing but the empty loop.
template <class F>
__device__ __forceinline__ void scan_pts(const LanePts &t, const int *cell, const int *row,
                                         int rlo, int rhi, F &&f) {
    const int n = t.p1 - t.p0, m = n < 64 ? n : 64;
    if (t.sorted && n <= 128) {
        // the first entry at or past rlo by bisection, then the run up to rhi
        auto cell_at = [&](int q) {
            return q < 64 ? __builtin_amdgcn_readlane(t.cell, q)
                          : __builtin_amdgcn_readlane(t.cell2, q - 64);
        };
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (cell_at(mid) < rlo) lo = mid + 1;
            else hi = mid;
        }
        for (int q = lo; q < n; q++) {
            const int cp = cell_at(q);
            if (cp > rhi) break;
            f(cp, q < 64 ? __builtin_amdgcn_readlane(t.row, q)
                         : __builtin_amdgcn_readlane(t.row2, q - 64));
        }
        return;
    }
    for (int q = 0; q < m; q++) {
        const int cp = __builtin_amdgcn_readlane(t.cell, q);
        if (cp < rlo || cp > rhi) continue;
        f(cp, __builtin_amdgcn_readlane(t.row, q));
    }
    for (int pp = t.p0 + 64; pp < t.p1; pp++) {
        const int cp = cell[pp];
        if (cp < rlo || cp > rhi) continue;
        f(cp, row[pp]);
    }
}

// Fast mode drops a difference outside its range through its coefficients
// instead of a select (dm / dp give 0 there; every operand is finite):
// those of x per row (scalars), those of z per column of the lane, set once
// per kernel here.  Cells outside the grid are then left as computed, not
// zeroed: no cell inside depends on them except through a dropped
// difference.  Exact mode keeps the reference's conditionals.
template <int V>
struct ZCoef {
    float v1[V], v2[V];  // velocity (D-, 2 <= j < nz - 2)
    float s1[V], s2[V];  // stress (D+, 1 <= j < nz - 2)
};
template <int V>
__device__ __forceinline__ ZCoef<V> zcoef(const Geo &g, int jb) {
    ZCoef<V> z;
#pragma unroll
    for (int v = 0; v < V; v++) {
        const int j = jb + v;
        const bool zv = j >= 2 && j < g.nz - 2, zs = j >= 1 && j < g.nz - 2;
        z.v1[v] = zv ? g.a1z : 0.0f;
        z.v2[v] = zv ? g.a2z : 0.0f;
        z.s1[v] = zs ? g.a1z : 0.0f;
        z.s2[v] = zs ? g.a2z : 0.0f;
    }
    return z;
}

// Velocity of step it+1 (divS, addSTF, updateV) on the lane's V columns
// jb .. jb+V-1 of grid row i.  ls1/ls2/ls3: stress of step it in LDS rows of
// W columns (sxx, szz, sxz for P-SV; sxy, szy for SH), grid row i at LDS row
// r of ls1 and ls3 (rows r-2 .. r+1 read) and at row r2 of ls2 (only that row
// is read: szz and szy enter through z-differences alone).  v1/v2: the lane's velocity of step it, rho: rho
// or dt/rho (fast math).  Point forces of step it from the list entries
// p0 .. p1 whose cell lies in grid cells rlo .. rhi (this row's span of the
// tile, held per lane in ip) are added in list order after the derivatives.
// Cells outside the grid (rin false) give 0 in exact mode; zc: fast mode's z
// coefficients.
template <bool PSV, int V, int W, bool FAST>
__device__ __forceinline__ void vel_row(const StepArgs &a, const float *ls1, const float *ls2,
                                        const float *ls3, int r, int r2, int c0, int i, int jb,
                                        bool rin, const ZCoef<V> &zc,
                                        float abx, const float (&abz)[V], const float (&abt)[V],
                                        const float (&v1)[V], const float (&v2)[V],
                                        const float (&rho)[V], const LanePts &ip, int rlo,
                                        int rhi, int b, long it, float (&n1)[V],
                                        float (&n2)[V]) {
    const Geo &g = a.g;
    const int nx = g.nx, nz = g.nz;
    const bool xin = i >= 2 && i < nx - 2;
    const float cx1 = !FAST || xin ? g.a1x : 0.0f, cx2 = !FAST || xin ? g.a2x : 0.0f;
    const float *I1 = PSV ? a.inj_x : a.inj_y, *I2 = a.inj_z;
    if (!PSV) {
        float x0[V], x1[V], x2[V], x3[V], ez[3 * V];
        vload<V>(&ls1[(r - 2) * W + c0], x0);
        vload<V>(&ls1[(r - 1) * W + c0], x1);
        vload<V>(&ls1[r * W + c0], x2);
        vload<V>(&ls1[(r + 1) * W + c0], x3);
        float zc0[V];
        vload<V>(&ls2[r2 * W + c0], zc0);
        window_reg<V>(zc0, ez);
        // every difference is evaluated and the out-of-range ones dropped
        // by selects: branches here make the compiler sink each column's
        // LDS reads into its own branch, splitting the vector reads into
        // single words (twice the LDS instructions, two-way bank conflicts)
        float d[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
            const int j = jb + v;
            const float tx = dm<FAST>(x2[v], x1[v], x3[v], x0[v], g.c8dx, g.c24dx, cx1, cx2);
            const float tz = dm<FAST>(ez[V + v], ez[V + v - 1], ez[V + v + 1], ez[V + v - 2],
                                      g.c8dz, g.c24dz, FAST ? zc.v1[v] : g.a1z,
                                      FAST ? zc.v2[v] : g.a2z);
            if (FAST) {
                d[v] = tx + tz;
            } else {
                d[v] = xin ? tx : 0.0f;
                d[v] = (j >= 2 && j < nz - 2) ? d[v] + tz : d[v];
            }
        }
        scan_pts(ip, a.inj_cell, a.inj_row, rlo, rhi, [&](int cp, int row) {
            const int e = cp - (i * nz + jb);
            if (e < 0 || e >= V) return;
            const float f = I1[(long)b * a.inj_slot_stride + (long)row * a.nt + it];
            // a select per column keeps the array in registers
#pragma unroll
            for (int v = 0; v < V; v++) d[v] = e == v ? d[v] + f : d[v];
        });
#pragma unroll
        for (int v = 0; v < V; v++) {
            const float ab = (abx * abz[v]) * abt[v];
            const float n = vel_update<FAST>(ab, v1[v], g.dt, d[v], rho[v]);
            n1[v] = FAST || rin ? n : 0.0f;
        }
    } else {
        float x0[V], x1[V], x2[V], x3[V], z0[V], z1[V], z2[V], z3[V], e3[3 * V], e2[3 * V];
        vload<V>(&ls1[(r - 2) * W + c0], x0);
        vload<V>(&ls1[(r - 1) * W + c0], x1);
        vload<V>(&ls1[r * W + c0], x2);
        vload<V>(&ls1[(r + 1) * W + c0], x3);
        vload<V>(&ls3[(r - 2) * W + c0], z0);
        vload<V>(&ls3[(r - 1) * W + c0], z1);
        vload<V>(&ls3[r * W + c0], z2);
        vload<V>(&ls3[(r + 1) * W + c0], z3);
        float zc0[V];
        vload<V>(&ls2[r2 * W + c0], zc0);
        window_reg<V>(z2, e3);
        window_reg<V>(zc0, e2);
        float dx_[V], dz_[V];
#pragma unroll
        for (int v = 0; v < V; v++) {
            const int j = jb + v;
            const bool zin = j >= 2 && j < nz - 2;
            const float cz1 = FAST ? zc.v1[v] : g.a1z, cz2 = FAST ? zc.v2[v] : g.a2z;
            const float xx = dm<FAST>(x2[v], x1[v], x3[v], x0[v], g.c8dx, g.c24dx, cx1, cx2);
            const float zx = dm<FAST>(z2[v], z1[v], z3[v], z0[v], g.c8dx, g.c24dx, cx1, cx2);
            const float xz = dm<FAST>(e3[V + v], e3[V + v - 1], e3[V + v + 1], e3[V + v - 2],
                                      g.c8dz, g.c24dz, cz1, cz2);
            const float zz = dm<FAST>(e2[V + v], e2[V + v - 1], e2[V + v + 1], e2[V + v - 2],
                                      g.c8dz, g.c24dz, cz1, cz2);
            if (FAST) {
                dx_[v] = xx + xz;
                dz_[v] = zx + zz;
            } else {
                dx_[v] = xin ? xx : 0.0f;
                dz_[v] = xin ? zx : 0.0f;
                dx_[v] = zin ? dx_[v] + xz : dx_[v];
                dz_[v] = zin ? dz_[v] + zz : dz_[v];
            }
        }
        scan_pts(ip, a.inj_cell, a.inj_row, rlo, rhi, [&](int cp, int row) {
            const int e = cp - (i * nz + jb);
            if (e < 0 || e >= V) return;
            const long ks = (long)b * a.inj_slot_stride + (long)row * a.nt + it;
            const float fx = I1[ks], fz = I2[ks];
#pragma unroll
            for (int v = 0; v < V; v++) {
                dx_[v] = e == v ? dx_[v] + fx : dx_[v];
                dz_[v] = e == v ? dz_[v] + fz : dz_[v];
            }
        });
#pragma unroll
        for (int v = 0; v < V; v++) {
            const float ab = (abx * abz[v]) * abt[v];
            const float m1 = vel_update<FAST>(ab, v1[v], g.dt, dx_[v], rho[v]);
            const float m2 = vel_update<FAST>(ab, v2[v], g.dt, dz_[v], rho[v]);
            n1[v] = FAST || rin ? m1 : 0.0f;
            n2[v] = FAST || rin ? m2 : 0.0f;
        }
    }
}

// Stress of step it+1 (divV, updateS) on the lane's V columns jb .. jb+V-1
// of grid row i, from the velocity of step it+1 in LDS rows of W columns
// (vx, vz for P-SV; vy for SH), grid row i at LDS row q, rows q-1 .. q+2
// read, and the stress of step it at the lane's cells (sa, sc, sxz: sxx,
// szz, sxz or sxy, szy).  Explosive sources, and zeroing cells outside the
// grid in exact mode, are the caller's; zc: fast mode's z coefficients.
template <bool PSV, int V, int W, bool FAST>
__device__ __forceinline__ void stress_row(const Geo &g, const float *lv1, const float *lv2, int q,
                                           int c0, int i, int jb, const ZCoef<V> &zc,
                                           const float (&sa)[V],
                                           const float (&sc)[V], const float (&sxz)[V],
                                           const float (&mu)[V], const float (&lam)[V],
                                           float (&o1)[V], float (&o2)[V], float (&o3)[V]) {
    const int nx = g.nx, nz = g.nz;
    const bool xin = i >= 1 && i < nx - 2;
    const float cx1 = !FAST || xin ? g.a1x : 0.0f, cx2 = !FAST || xin ? g.a2x : 0.0f;
    float y0[V], y1[V], y2[V], y3[V], e1[3 * V];
    vload<V>(&lv1[(q - 1) * W + c0], y0);
    vload<V>(&lv1[q * W + c0], y1);
    vload<V>(&lv1[(q + 1) * W + c0], y2);
    vload<V>(&lv1[(q + 2) * W + c0], y3);
    window_reg<V>(y1, e1);
    if (!PSV) {
#pragma unroll
        for (int v = 0; v < V; v++) {
            const int j = jb + v;
            // differences evaluated unconditionally, as in vel_row
            const float tx = dp<FAST>(y2[v], y1[v], y3[v], y0[v], g.c8dx, g.c24dx, cx1, cx2);
            const float tz = dp<FAST>(e1[V + v + 1], e1[V + v], e1[V + v + 2], e1[V + v - 1],
                                      g.c8dz, g.c24dz, FAST ? zc.s1[v] : g.a1z,
                                      FAST ? zc.s2[v] : g.a2z);
            const float dx_ = FAST || xin ? tx : 0.0f;
            const float dz_ = FAST || (j >= 1 && j < nz - 2) ? tz : 0.0f;
            o1[v] = s_sh<FAST>(sa[v], g.dt, mu[v], dx_);
            o2[v] = s_sh<FAST>(sc[v], g.dt, mu[v], dz_);
        }
    } else {
        float w0[V], w1[V], w2[V], w3[V], e2[3 * V];
        vload<V>(&lv2[(q - 1) * W + c0], w0);
        vload<V>(&lv2[q * W + c0], w1);
        vload<V>(&lv2[(q + 1) * W + c0], w2);
        vload<V>(&lv2[(q + 2) * W + c0], w3);
        window_reg<V>(w1, e2);
#pragma unroll
        for (int v = 0; v < V; v++) {
            const int j = jb + v;
            const bool zin = j >= 1 && j < nz - 2;
            const float cz1 = FAST ? zc.s1[v] : g.a1z, cz2 = FAST ? zc.s2[v] : g.a2z;
            const float txx = dp<FAST>(y2[v], y1[v], y3[v], y0[v], g.c8dx, g.c24dx, cx1, cx2);
            const float tzx = dp<FAST>(w2[v], w1[v], w3[v], w0[v], g.c8dx, g.c24dx, cx1, cx2);
            const float txz = dp<FAST>(e1[V + v + 1], e1[V + v], e1[V + v + 2], e1[V + v - 1],
                                       g.c8dz, g.c24dz, cz1, cz2);
            const float tzz = dp<FAST>(e2[V + v + 1], e2[V + v], e2[V + v + 2], e2[V + v - 1],
                                       g.c8dz, g.c24dz, cz1, cz2);
            const float xx = FAST || xin ? txx : 0.0f, zx = FAST || xin ? tzx : 0.0f;
            const float xz = FAST || zin ? txz : 0.0f, zz = FAST || zin ? tzz : 0.0f;
            const float m = mu[v], l = lam[v];
            s_normal<FAST>(sa[v], sc[v], g.dt, l, m, xx, zz, o1[v], o2[v]);
            o3[v] = s_shear<FAST>(sxz[v], g.dt, m, xz, zx);
        }
    }
}

// Exponents of one row for the 16-bit formats: the max-abs of the fields
// over the lanes that own cells of the row (wave reductions; every lane of
// the wavefront calls it) gives the segment exponents (snap_exp), stored by
// lane 0 into the slot's table at segment seg, with the segment width segw in
// the table header; the block offsets of the lane's block (snap_blk_put, oc:
// the lane's first column in the segment) are stored alongside.  e returns
// the lane's exponents, segment plus block offset: e[0], e[1] for u (two
// fields for P-SV), e[2], e[3] for v.
template <bool PSV, int V>
__device__ __forceinline__ void row_exponents(const StepArgs &a, const StepRoles &ro, int fu,
                                              int fv, int i, int b, int seg, int segw, int lane,
                                              int oc, bool act, const float (&u1)[V],
                                              const float (&u2)[V], const float (&n1)[V],
                                              const float (&n2)[V], int (&e)[4]) {
    float m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (act) {
#pragma unroll
        for (int v = 0; v < V; v++) {
            m[0] = fmaxf(m[0], fabsf(u1[v]));
            m[2] = fmaxf(m[2], fabsf(n1[v]));
            if (PSV) {
                m[1] = fmaxf(m[1], fabsf(u2[v]));
                m[3] = fmaxf(m[3], fabsf(n2[v]));
            }
        }
    }
    const int nx = a.g.nx;
    const int fld[4] = {fu, fu + 1, fv, fv + 1};
    int es[4] = {0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (!PSV && q % 2) continue;
        es[q] = snap_exp(wave_max(m[q]));
        unsigned char *blk = reinterpret_cast<unsigned char *>(ro.sexp) +
                             snaptab_blk(a.sblk, b, a.nsnapf, fld[q], nx, i, a.nbb) +
                             seg * snap_seg_bytes(segw);
        e[q] = es[q] + snap_blk_put<V>(m[q], es[q], lane, oc, act, segw, blk);
    }
    if (lane == 0) {
        signed char *t = ro.sexp;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (PSV || q % 2 == 0) t[snaptab_row(b, a.nsnapf, fld[q], nx, i, a.nseg) + seg] = (signed char)es[q];
        *reinterpret_cast<int *>(t) = segw;
    }
}

// the lane's V values of one snapshot field from cell k (a multiple of V):
// FP32 as a vector, 16-bit formats scaled by 2^e and packed two to a 32-bit
// word, low half first
template <int V>
__device__ __forceinline__ void snap_store(void *base, int fmt, long k, const float (&x)[V],
                                           int e) {
    if (fmt == INV_SNAP_F32) {
        vstore<V>(static_cast<float *>(base) + k, x);
        return;
    }
    unsigned w[V / 2];
#pragma unroll
    for (int h = 0; h < V / 2; h++)
        w[h] = (unsigned)snap_enc16(x[2 * h], e, fmt) |
               ((unsigned)snap_enc16(x[2 * h + 1], e, fmt) << 16);
    unsigned short *q = static_cast<unsigned short *>(base) + k;
    if constexpr (V == 2) *reinterpret_cast<unsigned *>(q) = w[0];
    else *reinterpret_cast<uint2 *>(q) = make_uint2(w[0], w[1]);
}

// one snapshot value: FP32 as is, 16-bit formats scaled by 2^e
__device__ __forceinline__ void snap_put(void *base, int fmt, long idx, float x, int e) {
    if (fmt == INV_SNAP_F32) static_cast<float *>(base)[idx] = x;
    else static_cast<unsigned short *>(base)[idx] = snap_enc16(x, e, fmt);
}

// The owned cells of grid row i after the velocity update of step it, in
// the order of k_colv: u(it) into its snapshot, u(it+1) = u(it) + v(it+1) dt,
// v(it+1) into its snapshot, receivers of step it (list entries held per lane
// in rp).  u1/u2 hold u(it) on entry and u(it+1) on return; n1/n2: v(it+1).
// 16-bit snapshots are scaled per segment: this wavefront's owned cells of
// the row, segment `seg` of width segw of the slot's exponent table.
// Every lane of the wavefront calls it (wave-uniform row); act: the lane
// owns its cells of the row.
template <bool PSV, int V, bool FAST>
__device__ __forceinline__ void owned_row(const StepArgs &a, const StepRoles &ro, int fu, int fv,
                                          long sb, int i, int jb, bool act, long it, int b,
                                          int seg, int segw, int lane, const LanePts &rp,
                                          int rlo, int rhi, float (&u1)[V], float (&u2)[V],
                                          const float (&n1)[V], const float (&n2)[V]) {
    const Geo &g = a.g;
    const int fmt = a.snap_fmt;
    float *O1 = PSV ? a.out_x : a.out_y, *O2 = a.out_z;
    int e[4] = {0, 0, 0, 0};
    if (ro.snap && fmt != INV_SNAP_F32)  // oc: the lane's first column in the segment
        row_exponents<PSV, V>(a, ro, fu, fv, i, b, seg, segw, lane, jb - seg * segw, act, u1, u2,
                              n1, n2, e);
    if (act) {
        const long kc = (long)i * g.nz + jb;
        if (ro.snap) {
            snap_store<V>(ro.snap, fmt, sb + (long)fu * g.N + kc, u1, e[0]);
            if (PSV) snap_store<V>(ro.snap, fmt, sb + (long)(fu + 1) * g.N + kc, u2, e[1]);
        }
#pragma unroll
        for (int v = 0; v < V; v++) {
            u1[v] = u_step<FAST>(u1[v], n1[v], g.dt);
            if (PSV) u2[v] = u_step<FAST>(u2[v], n2[v], g.dt);
        }
        if (ro.snap) {
            snap_store<V>(ro.snap, fmt, sb + (long)fv * g.N + kc, n1, e[2]);
            if (PSV) snap_store<V>(ro.snap, fmt, sb + (long)(fv + 1) * g.N + kc, n2, e[3]);
        }
    }
    scan_pts(rp, a.rec_cell, a.rec_row, rlo, rhi, [&](int cp, int row) {
        const int e_ = cp - (i * g.nz + jb);
        if (!act || e_ < 0 || e_ >= V) return;
        float x = 0.0f, y = 0.0f;
#pragma unroll
        for (int v = 0; v < V; v++) {
            x = e_ == v ? (a.rec_disp ? u1[v] : n1[v]) : x;
            if (PSV) y = e_ == v ? (a.rec_disp ? u2[v] : n2[v]) : y;
        }
        const long kr = (long)b * a.out_slot_stride + (long)row * a.nt + it;
        O1[kr] = x;
        if (PSV) O2[kr] = y;
    });
}

// Explosive sources of step it (source type 1) among the tile's list
// entries (held per lane in xp) at the lane's cells of grid row i: their
// wavelet is added to sxx and szz.  Every lane of the wavefront calls it.
template <int V>
__device__ __forceinline__ void explode_row(const StepArgs &a, const LanePts &xp, int i, int jb,
                                            int b, long it, float (&o1)[V], float (&o2)[V]) {
    const int r0 = i * a.g.nz;
    scan_pts(xp, a.sinj_cell, a.sinj_row, r0, r0 + a.g.nz - 1, [&](int cp, int row) {
        const int e = cp - (r0 + jb);
        if (e < 0 || e >= V) return;
        const float f = a.inj_y[(long)b * a.inj_slot_stride + (long)row * a.nt + it];
#pragma unroll
        for (int v = 0; v < V; v++) {
            o1[v] = e == v ? o1[v] + f : o1[v];
            o2[v] = e == v ? o2[v] + f : o2[v];
        }
    });
}

}  // namespace inv
