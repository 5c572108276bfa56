// Model points to grid nodes on the device (mesh2grid, solver/fdm.h:4-24):
// each grid node takes the values of its nearest model point, distances
// compared as the reference computes them (FP32 (x - px)^2 + (z - pz)^2,
// strict <, so the first point at the least distance wins; nodes with no
// point closer than dmax keep 0).
//
// The reference scans every point for every node, O(N npt): hours at
// 4096 x 2048.  Here the points are bucketed on a uniform grid (CSR, each
// bucket in point order) and a node searches square rings of buckets around
// its own until no unsearched bucket can hold a point as close as the best
// one found.  Among points at equal distance the lowest index wins, which is
// what the sequential strict-< scan picks, so the result is the reference's
// value for value.
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace inv {

struct BucketGrid {
    float x0, z0, bs;  // origin and bucket side
    int nbx, nbz;
};

// the bucket of a point, clamped into the bucket grid
__device__ __host__ __forceinline__ void bucket_of(const BucketGrid &g, float x, float z, int &bx,
                                                   int &bz) {
    float fx = (x - g.x0) / g.bs, fz = (z - g.z0) / g.bs;
    fx = fx < 0.0f ? 0.0f : (fx > (float)(g.nbx - 1) ? (float)(g.nbx - 1) : fx);
    fz = fz < 0.0f ? 0.0f : (fz > (float)(g.nbz - 1) ? (float)(g.nbz - 1) : fz);
    bx = (int)fx;
    bz = (int)fz;
}

// px/pz/v1/v2/vr/idx: the points in bucket order; start[b] .. start[b+1]-1
// are bucket b's (b = bx * nbz + bz)
__global__ __launch_bounds__(256) void k_mesh2grid(
    const BucketGrid g, const int *__restrict__ start, const float *__restrict__ px,
    const float *__restrict__ pz, const float *__restrict__ v1, const float *__restrict__ v2,
    const float *__restrict__ vr, const int *__restrict__ idx, int nx, int nz, float dx, float dz,
    float dmax, float *__restrict__ lam, float *__restrict__ mu, float *__restrict__ rho) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)nx * nz) return;
    const int i = (int)(k / nz), j = (int)(k - (long)i * nz);
    const float x = i * dx, z = j * dz;
    int cbx, cbz;
    bucket_of(g, x, z, cbx, cbz);
    float best = dmax;
    int bi = -1;  // original index of the best point
    int bp = -1;  // its position in bucket order
    const int rmax = max(g.nbx, g.nbz);
    for (int r = 0; r <= rmax; r++) {
        // every bucket at Chebyshev distance r from the node's bucket
        for (int bx = cbx - r; bx <= cbx + r; bx++) {
            if (bx < 0 || bx >= g.nbx) continue;
            const bool edge = bx == cbx - r || bx == cbx + r;
            const int step = edge ? 1 : 2 * r;
            for (int bz = cbz - r; bz <= cbz + r; bz += (step > 0 ? step : 1)) {
                if (bz < 0 || bz >= g.nbz) continue;
                const int b = bx * g.nbz + bz;
                for (int q = start[b]; q < start[b + 1]; q++) {
                    const float ex = x - px[q], ez = z - pz[q];
                    const float d = ex * ex + ez * ez;
                    const int id = idx[q];
                    if (d < best || (d == best && bi >= 0 && id < bi)) {
                        best = d;
                        bi = id;
                        bp = q;
                    }
                }
            }
        }
        // unsearched points are at least r bucket sides away (the node may
        // sit anywhere in its bucket); a relative margin covers the rounding
        // of the FP32 distances
        const float lb = (float)r * g.bs;
        if (bi >= 0 && lb * lb * 0.999f > best) break;
    }
    if (bp >= 0) {
        lam[k] = v1[bp];
        mu[k] = v2[bp];
        rho[k] = vr[bp];
    } else {
        lam[k] = 0.0f;
        mu[k] = 0.0f;
        rho[k] = 0.0f;
    }
}

void model_nearest(Ctx &c, int npt, const float *px, const float *pz, const float *p1,
                   const float *p2, const float *prho, float dmax) {
    // bucket grid over the points' bounding box, about 4 points per bucket
    float xmin = px[0], xmaxp = px[0], zmin = pz[0], zmaxp = pz[0];
    for (int p = 1; p < npt; p++) {
        xmin = std::min(xmin, px[p]);
        xmaxp = std::max(xmaxp, px[p]);
        zmin = std::min(zmin, pz[p]);
        zmaxp = std::max(zmaxp, pz[p]);
    }
    const double w = std::max((double)xmaxp - xmin, 1e-30);
    const double h = std::max((double)zmaxp - zmin, 1e-30);
    double bs = std::sqrt(w * h * 4.0 / npt);
    if (!(bs > 0)) bs = std::max(w, h);
    // (degenerate point sets, e.g. all on one line, get larger buckets)
    while (std::ceil(w / bs + 1) * std::ceil(h / bs + 1) > std::max(4.0 * npt, 1024.0)) bs *= 2;
    BucketGrid g;
    g.x0 = xmin;
    g.z0 = zmin;
    g.bs = (float)bs;
    g.nbx = (int)(std::ceil(w / bs) + 1);
    g.nbz = (int)(std::ceil(h / bs) + 1);
    const long nb = (long)g.nbx * g.nbz;
    std::vector<int> count(nb + 1, 0), bucket(npt);
    for (int p = 0; p < npt; p++) {
        int bx, bz;
        bucket_of(g, px[p], pz[p], bx, bz);
        bucket[p] = bx * g.nbz + bz;
        count[bucket[p] + 1]++;
    }
    for (long b = 0; b < nb; b++) count[b + 1] += count[b];
    std::vector<int> pos(count.begin(), count.end() - 1), idx(npt);
    std::vector<float> sx(npt), sz(npt), s1(npt), s2(npt), sr(npt);
    for (int p = 0; p < npt; p++) {  // stable: a bucket keeps point order
        const int q = pos[bucket[p]]++;
        idx[q] = p;
        sx[q] = px[p];
        sz[q] = pz[p];
        s1[q] = p1[p];
        s2[q] = p2[p];
        sr[q] = prho[p];
    }
    DevBuf d_start, d_pts, d_idx;
    d_start.ensure(count.size() * sizeof(int));
    d_pts.ensure((size_t)5 * npt * sizeof(float));
    d_idx.ensure((size_t)npt * sizeof(int));
    float *f = d_pts.as<float>();
    INV_HIP(hipMemcpy(d_start.p, count.data(), count.size() * sizeof(int),
                      hipMemcpyHostToDevice));
    INV_HIP(hipMemcpy(d_idx.p, idx.data(), npt * sizeof(int), hipMemcpyHostToDevice));
    const float *src[5] = {sx.data(), sz.data(), s1.data(), s2.data(), sr.data()};
    for (int a = 0; a < 5; a++)
        INV_HIP(hipMemcpy(f + (size_t)a * npt, src[a], npt * sizeof(float),
                          hipMemcpyHostToDevice));
    const long N = c.N;
    k_mesh2grid<<<(unsigned)((N + 255) / 256), 256, 0, c.stream>>>(
        g, d_start.as<int>(), f, f + npt, f + 2 * (size_t)npt, f + 3 * (size_t)npt,
        f + 4 * (size_t)npt, d_idx.as<int>(), c.g.nx, c.g.nz, c.g.dx, c.g.dz, dmax,
        c.lam.as<float>(), c.mu.as<float>(), c.rho.as<float>());
    INV_HIP(hipGetLastError());
    INV_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace inv
