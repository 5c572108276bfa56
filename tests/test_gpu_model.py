"""mesh2grid on the device (inv_model_from_points with points off the grid
nodes, solver/fdm.h:4-24): the bucketed nearest-point search must choose the
point the reference's full scan chooses (strict <, first point at the least
distance), checked value for value against the oracle's scan, including
duplicated points, points at equal distance from a node, points outside the
grid and nodes no point is close to."""
import numpy as np
import pytest

from conftest import inv, ROOT

pytestmark = pytest.mark.gpu


def _oracle(nx, nz, dx, dz, xmax, zmax, px, pz, a, b, r, model_type):
    import sys
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    lam, mu, rho = pyoracle.mesh2grid(nx, nz, dx, dz, xmax, zmax, px, pz, a, b, r)
    if model_type == 0:
        lam, mu = pyoracle.vps2lm(lam, mu, rho)
    return lam, mu, rho


def _ctx(nx, nz, dx, dz):
    P = inv.problems.tiny(nx=nx, nz=nz, dx=dx, sh=1, psv=1, nsrc=1, nrec=2, nt=20)
    ctx = P.context(batch=1)
    return ctx


@pytest.mark.parametrize("seed,npt", [(1, 900), (2, 4000), (3, 37)])
def test_scattered_points_match_full_scan(gpu, seed, npt):
    nx, nz, dx = 48, 40, 10.0
    rng = np.random.default_rng(seed)
    xmax, zmax = (nx - 1) * dx, (nz - 1) * dx
    px = rng.uniform(-30, xmax + 30, npt).astype(np.float32)
    pz = rng.uniform(-30, zmax + 30, npt).astype(np.float32)
    # duplicates (the first must win) and points on a node's circle (ties)
    px[npt // 2:npt // 2 + 5] = px[3]
    pz[npt // 2:npt // 2 + 5] = pz[3]
    px[7], pz[7] = np.float32(100.0 + 3.0), np.float32(50.0 + 4.0)
    px[9], pz[9] = np.float32(100.0 - 3.0), np.float32(50.0 - 4.0)
    px[8], pz[8] = np.float32(100.0 + 4.0), np.float32(50.0 - 3.0)
    vp = rng.uniform(2000, 4000, npt).astype(np.float32)
    vs = (vp / rng.uniform(1.5, 2.0, npt)).astype(np.float32)
    rho = rng.uniform(1800, 2600, npt).astype(np.float32)
    for model_type in (0, 1):
        ctx = _ctx(nx, nz, dx, dx)
        ctx.model_from_points(px, pz, vp, vs, rho, xmax, zmax, model_type)
        got = ctx.model_download()
        ctx.close()
        ref = _oracle(nx, nz, dx, dx, xmax, zmax, px, pz, vp, vs, rho, model_type)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)


def test_sparse_points_leave_far_nodes(gpu):
    """Few points in one corner and a small dmax (xmax, zmax of the call):
    nodes farther than dmax from every point keep 0, as in the scan."""
    nx, nz, dx = 40, 36, 10.0
    px = np.array([5.0, 12.0, 12.0, 33.0], np.float32)
    pz = np.array([5.0, 9.0, 9.0, 2.0], np.float32)
    v = np.array([3000, 3100, 3200, 3300], np.float32)
    s = np.array([1700, 1800, 1900, 2000], np.float32)
    r = np.array([2000, 2100, 2200, 2300], np.float32)
    ctx = _ctx(nx, nz, dx, dx)
    ctx.model_from_points(px, pz, v, s, r, 60.0, 40.0, 1)
    got = ctx.model_download()
    ctx.close()
    ref = _oracle(nx, nz, dx, dx, 60.0, 40.0, px, pz, v, s, r, 1)
    for g, e in zip(got, ref):
        assert np.array_equal(g, e)
    assert np.count_nonzero(ref[2] == 0) > 0 and np.count_nonzero(ref[2]) > 0
