// Arguments of the fused time-step kernels (fdm_fused.hip, fdm_wide.hip).
#pragma once

#include "internal.h"

namespace inv {

struct StepArgs {
    Geo g;
    int tiles_x, tiles_z;
    const float *rho, *lam, *mu, *abx, *abzb, *abzt;
    // Per-slot models ([batch][3] pointers {rho or dt/rho, lam, mu}), or null
    // when every slot shares rho/lam/mu above.  Batches of shots from
    // different models (the request server's lanes) set it.
    const float *const *slot_model;
    const float *vx0, *vz0, *sxx0, *szz0, *sxz0, *vy0, *sxy0, *szy0;  // step n
    float *vx1, *vz1, *sxx1, *szz1, *sxz1, *vy1, *sxy1, *szy1;        // step n+1
    float *ux, *uz, *uy;                                              // in place
    // point injection (addSTF, fdm.h:163-180), bucketed by velocity region
    const float *inj_x, *inj_y, *inj_z;
    long inj_slot_stride;
    const int *inj_ptr, *inj_cell, *inj_row;
    int inj_per_slot;
    // explosive sources (per slot, bucketed by owning tile): inj_y of the
    // source's row is added to sxx and szz after the stress update; null when
    // there are none
    const int *sinj_ptr, *sinj_cell, *sinj_row;
    int nt, it;
    // receiver recording (saveRec, fdm.h:181-196), bucketed by tile
    float *out_x, *out_y, *out_z;
    long out_slot_stride;
    const int *rec_ptr, *rec_cell, *rec_row;
    int rec_disp;
    // Snapshots (fdm.h:521-560: a slot holds u before and v after a step).
    // snap: slot base (batch slot 0) that receives this step's input u and
    // its result v, or null.  16-bit formats scale each row segment of a
    // field (the owned columns of one tile row) by a power of two chosen
    // from the segment's max-abs (snap_exp), refined per block of 4 cells,
    // and store the exponents in sexp, the slot's exponent table (layout in
    // stencil.h); nseg: segments per row, sblk: byte offset of the block
    // offsets in the table, nbb: their bytes per row.  The linear kernel
    // supports FP32 only.
    void *snap;
    signed char *sexp;
    int snap_fmt, nsnapf, psv_f0, nseg, nbb;
    long sblk;
    long snap_slot_stride;  // elements per batch slot
};

// The maximum of x >= 0 over the wavefront, wave-uniform: DPP within each
// row of 16 lanes (quad swaps, half-row and row mirrors), then the four
// rows' results read into scalars and compared as integers (the order of
// non-negative floats).  Every lane must be active.
__device__ __forceinline__ float dpp_max(float x, int ctrl) {
    const int y = ctrl == 0xB1 ? __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xf, 0xf, false)
                : ctrl == 0x4E ? __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xf, 0xf, false)
                : ctrl == 0x141 ? __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x141, 0xf, 0xf, false)
                                : __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x140, 0xf, 0xf, false);
    return fmaxf(x, __int_as_float(y));
}
__device__ __forceinline__ float wave_max(float x) {
    x = dpp_max(x, 0xB1);   // quad_perm [1, 0, 3, 2]
    x = dpp_max(x, 0x4E);   // quad_perm [2, 3, 0, 1]
    x = dpp_max(x, 0x141);  // row_half_mirror
    x = dpp_max(x, 0x140);  // row_mirror
    const unsigned b = (unsigned)__float_as_int(x);
    unsigned m = __builtin_amdgcn_readlane(b, 0);
    const unsigned m1 = __builtin_amdgcn_readlane(b, 16), m2 = __builtin_amdgcn_readlane(b, 32),
                   m3 = __builtin_amdgcn_readlane(b, 48);
    m = m > m1 ? m : m1;
    m = m > m2 ? m : m2;
    m = m > m3 ? m : m3;
    return __int_as_float((int)m);
}

// Block offsets of one field of a row segment (stencil.h): the lane holds V
// consecutive columns from owned column oc of the segment (oc a multiple of
// V; a block of 4 columns spans 4 / V lanes when V < 4); m is the max-abs
// of its cells (0 for a lane that owns none), e the segment's exponent.
// Returns the offset d of the lane's block; the first lane of each
// even-numbered block stores the byte of that block and the next one into
// seg_blk, the segment's block-offset bytes (own: the lane owns its cells).
// Every lane of the wavefront calls it.
template <int V>
__device__ __forceinline__ int snap_blk_put(float m, int e, int lane, int oc, bool own, int segw,
                                            unsigned char *seg_blk) {
    constexpr int LPB = V >= 4 ? 1 : 4 / V;  // lanes per block
    float bm = m;
    int d, dn;
    if constexpr (V == 4) {
        d = snap_blk_d(bm, e);
        dn = __builtin_amdgcn_update_dpp(0, d, 0x130, 0xf, 0xf, true);  // lane + 1
    } else if constexpr (V == 2) {
        // the block's other lane is the next one for its first lane, else the
        // previous one (neighbour exchange by DPP wavefront shifts)
        const int mb = __float_as_int(m);
        const float nx = __int_as_float(__builtin_amdgcn_update_dpp(0, mb, 0x130, 0xf, 0xf, true));
        const float pv = __int_as_float(__builtin_amdgcn_update_dpp(0, mb, 0x138, 0xf, 0xf, true));
        bm = fmaxf(m, (oc & 3) == 0 ? nx : pv);
        d = snap_blk_d(bm, e);
        const int d1 = __builtin_amdgcn_update_dpp(0, d, 0x130, 0xf, 0xf, true);
        dn = __builtin_amdgcn_update_dpp(0, d1, 0x130, 0xf, 0xf, true);  // lane + 2
    } else {
        const int lead = lane - ((oc & 3) / V);
#pragma unroll
        for (int k = 0; k < LPB; k++) bm = fmaxf(bm, __shfl(m, lead + k));
        d = snap_blk_d(bm, e);
        dn = __shfl(d, lane + LPB);  // the next block's offset, from its first lane
    }
    if (own && (oc & 7) == 0) {
        const int lb = oc >> 2;
        const int hi = (lb + 1) * 4 < segw ? dn : 0;
        seg_blk[lb >> 1] = (unsigned char)(d | (hi << 4));
    }
    return d;
}

// Launch the wide column kernel (fdm_wide.hip): tile_x rows, c.col_v = 2 or
// 4 columns per lane; `blocks` tiles (shot slots x tile rows x tile columns).
template <bool SH, bool PSV, bool FAST>
void wide_launch(Ctx &c, const StepArgs &a, unsigned blocks);

// Columns a column-kernel tile owns with V columns per lane: 64 V minus a
// halo of 3 columns each side (V = 1) or 4 (V = 2, 4: keeps vectors aligned).
constexpr int col_halo(int V) { return V == 1 ? 3 : 4; }
constexpr int col_strip(int V) { return 64 * V - 2 * col_halo(V); }

// The snapshot fields of StepArgs for one time step (see StepArgs).
struct StepRoles {
    void *snap;
    signed char *sexp;
};

// Two-step kernel (fdm_pair.hip): two time steps per launch on tiles of
// tile rows x pair_strip(V) columns; the stress of step n is read with a
// halo of 6 rows and 6 columns (8 with V = 4, for alignment) each side.
constexpr int pair_halo(int V) { return V == 2 ? 6 : 8; }
constexpr int pair_strip(int V) { return 64 * V - 2 * pair_halo(V); }
static_assert(col_strip(2) == SNAP_SEG_WIDTHS[2] && col_strip(4) == SNAP_SEG_WIDTHS[4] &&
                  pair_strip(2) == SNAP_SEG_WIDTHS[1] && pair_strip(4) == SNAP_SEG_WIDTHS[3],
              "snapshot tables are sized for these segment widths");
// LDS bytes of a two-step tile: stress of steps n / n+1 on rows+12 rows
// (szz or szy on rows+9: they enter only through z-differences) and
// velocity of steps n+1 / n+2 on rows+9 rows, 64 V columns each.
constexpr int pair_lds_bytes(bool psv, int rows, int V) {
    return 4 * 64 * V * ((psv ? 2 : 1) * (rows + 12) + (rows + 9) + (psv ? 2 : 1) * (rows + 9));
}

// Launch the two-step kernel for steps a.it and a.it + 1 (roles of the
// second step in r2): tile rows c.pair_tx, c.pair_v columns per lane.
template <bool SH, bool PSV, bool FAST>
void pair_launch(Ctx &c, const StepArgs &a, const StepRoles &r2, unsigned blocks);

// Three-step kernel (fdm_trip.hip): three time steps per launch on tiles of
// c.trip_tx = 8, 12, 16 or 20 rows x trip_strip() columns (2 per lane); the stress
// of step n is read with a halo of 9 rows and trip_halo() columns each side.
// Launches without snapshot roles only.
constexpr int trip_halo() { return 10; }
constexpr int trip_strip() { return 128 - 2 * trip_halo(); }
// LDS bytes of a three-step tile: stress on rows+18 rows (szz or szy on
// rows+15) and velocity on rows+15 rows, 128 columns each.
constexpr int trip_lds_bytes(bool psv, int rows) {
    return 4 * 128 * ((psv ? 2 : 1) * (rows + 18) + (rows + 15) + (psv ? 2 : 1) * (rows + 15));
}
template <bool SH, bool PSV, bool FAST>
void trip_launch(Ctx &c, const StepArgs &a, unsigned blocks);

// Walking three-step kernel (fdm_walk.hip): a workgroup walks down c.walk_ch
// rows of one trip_strip() column strip; its point lists are those of
// three-step tiles of walk_ch rows.  LDS: ring buffers of WALK_RING rows
// (and three copied rows) of the stress fields and the velocity fields.
// Launches without snapshot roles only.
constexpr int WALK_RING = 25;
constexpr int walk_lds_bytes(bool psv) { return 4 * 128 * (WALK_RING + 3) * (psv ? 5 : 3); }
template <bool SH, bool PSV, bool FAST>
void walk_launch(Ctx &c, const StepArgs &a, unsigned blocks);

// Persistent two-step kernel (fdm_resident.hip): `iters` iterations of two
// steps per launch on tiles of c.pair_tx = 4 or 8 rows x every column (nz <=
// 240 with RES_V = 4 columns per lane), one workgroup of RES_NW wavefronts
// per tile, halo rows exchanged through xbuf ([batch][tiles_x][2][5][rows]
// [64 RES_V] floats) with a flag per tile and parity in xflag
// ([batch][tiles_x][2]: the tag of the iteration last published).  ctl: [0]
// the launch epoch (counted up before each launch), [1] the error word (1: a
// tile gave up waiting for a neighbour).  Snapshot
// roles of each step follow from the plan below (snap null: none).  The
// state is read from the `0` set of the StepArgs and written to the `1` set.
constexpr int RES_V = 4, RES_NW = 8;
constexpr int RES_KMAX = 8191;     // iterations per launch (tag space)
constexpr int RES_SPIN = 1 << 20;  // polls of one wait before it gives up
struct ResArgs {
    float *xbuf;
    unsigned *xflag;
    unsigned *ctl;
    int iters;
    int spin_limit;  // < 0: give up at the first wait (tests the error path)
    char *snap;
    signed char *sexp;
    long slot_bytes, tab_bytes;
    int sfe, nsfe, nslots, sbase;
    // adjoint (int_on): the kernel accumulation after each iteration whose
    // second step `last` is a multiple of sfe, in the launch, reading
    // snapshot slot last / sfe of the buffer above (snap, slot_bytes apart;
    // exponent tables from sexp, tab_bytes apart) and adding into kl, km, kr
    // (+ shot slot x k_slot_stride) times ndt
    int int_on;
    float ndt;
    float *kl, *km, *kr;
    long k_slot_stride;
    int inv_lambda, inv_mu, inv_rho;
};
template <bool SH, bool PSV, bool FAST>
void res_launch(Ctx &c, const StepArgs &a, const ResArgs &ra, unsigned blocks);
// floats of the exchange rows for count shot slots of a grid of nx rows in
// tiles of tx rows, and bytes of the rows and flags together
inline size_t res_xbuf_floats(int count, int nx, int tx) {
    const long tiles = (nx + tx - 1) / tx;
    return (size_t)count * tiles * 2 * 5 * tx * 64 * RES_V;
}
inline size_t res_xbuf_bytes(int count, int nx, int tx) {
    const long tiles = (nx + tx - 1) / tx;
    return res_xbuf_floats(count, nx, tx) * sizeof(float) + (size_t)count * tiles * 2 * sizeof(unsigned);
}

}  // namespace inv
