"""Block-element assembly on the device across its row-length variants:
the stress-coupled Cahn-Hilliard elements stage a node's bonds in LDS when
its CSR block row has at most 64 (2D) or 128 node blocks and fall back to a
chunked kernel above that; larger horizons than the golden fixtures' move
the same elements onto the other variants.  Checked against the oracle on
the device's own g-weights and bond geometry: K and the residual's c and u
rows bit for bit, the mu rows (device logarithm) within 1e-14 of their
scale.  The oracle is pinned bit for bit to the reference's assembly by the
golden fixture tests."""
import numpy as np
import pytest

import perix_amd
from perix_amd import pxg, problems

pytestmark = pytest.mark.gpu


def _case(dim, nx, ny, nz, factor, seed):
    # isotropic spacing dx = 0.1 (2D: 0.05)
    g = perix_amd.StructuredGrid(dim, nx, ny, nz, hi=_hi(dim, nx, ny, nz))
    cloud = perix_amd.build_structured_cloud(g)
    delta = perix_amd.horizon_radius(g, factor)
    ctx = pxg.Context(0)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, dim)
    ctx.set_cracks(segs=[(0.5, 0.0, 0.5, 0.3)])
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    nd = dim + 2
    ctx.build_csr(nd)
    rs = np.random.RandomState(seed)
    n = cloud.n
    u = np.zeros(n * nd)
    u[0::nd] = 0.2 + 0.6 * rs.rand(n)
    u[np.arange(0, n, 7) * nd] = -0.01           # the log barrier / clamp ranges
    u[np.arange(3, n, 11) * nd] = 1.01
    u[1::nd] = 0.1 * rs.standard_normal(n)
    for a in range(dim):
        u[2 + a::nd] = 1e-3 * rs.standard_normal(n)
    uold = u + 1e-3 * rs.standard_normal(n * nd)
    return ctx, cloud, g, u, uold


def _hi(dim, nx, ny, nz):
    h = 0.05 if dim == 2 else 0.1
    return (nx * h, ny * h, nz * h if dim == 3 else 1.0)


def _compare(k, r, ok, orr, nd):
    assert np.array_equal(k.view(np.uint64), ok.view(np.uint64))
    mu = np.zeros(r.size, bool)
    mu[1::nd] = True
    assert np.array_equal(r[~mu].view(np.uint64), orr[~mu].view(np.uint64))
    assert np.abs(r[mu] - orr[mu]).max() <= 1e-14 * np.abs(orr[mu]).max()


# most node blocks per row: 81 (2D, 128-block staging), 123 (3D, 128-block
# staging), 344 (3D, chunked fallback)
VARIANTS = [(2, 20, 16, 1, 5.0), (3, 8, 7, 6, 3.015), (3, 7, 6, 5, 4.5)]


@pytest.mark.parametrize("dim,nx,ny,nz,factor", VARIANTS)
def test_stress_ch_row_variants(dim, nx, ny, nz, factor, oracle):
    print("stress CH dim %d grid %dx%dx%d factor %.3f" % (dim, nx, ny, nz, factor), flush=True)
    ctx, cloud, g, u, uold = _case(dim, nx, ny, nz, factor, 5)
    nd = dim + 2
    p = problems.SILICON
    ctx.stressch_setup(pxg.StressChParams(p["E"], p["nu"], p["D"], p["Omega"], p["cref"], p["chi"], p["kappa"], 1))
    dt = 1e-3
    k, r = ctx.assemble_stressch(dt, u, uold)
    m = oracle.Mesh(oracle.make_grid(dim, nx, ny, nz, hi=_hi(dim, nx, ny, nz)), factor=factor)
    off, nbr = ctx.families()
    assert np.array_equal(off, m.off) and np.array_equal(nbr, m.nbr)
    ops, _ = ctx.operators()
    rp, cols = ctx.csr()
    blocks = np.diff(rp)[::nd] // nd
    print("most blocks in a row: %d" % blocks.max())
    lap, rev, isg = ctx.ch_geometry()
    ok, orr = oracle.assemble_stressch(m, ops, lap, rev, isg, ctx.stressch_constants(), dt, rp, cols, u, uold)
    _compare(k, r, ok, orr, nd)
    ctx.close()


@pytest.mark.parametrize("dim,nx,ny,nz,factor", VARIANTS)
def test_fracture_stress_ch_row_variants(dim, nx, ny, nz, factor, oracle):
    print("fracture stress CH dim %d grid %dx%dx%d factor %.3f" % (dim, nx, ny, nz, factor), flush=True)
    ctx, cloud, g, u, uold = _case(dim, nx, ny, nz, factor, 8)
    nd = dim + 2
    p = problems.SILICON
    base = pxg.StressChParams(p["E"], p["nu"], p["D"], p["Omega"], p["cref"], p["chi"], p["kappa"], 1)
    ctx.fsch_setup(pxg.FschParams(base, 2.0, 2e-3, 1e-3, 0, 1), uold)
    ctx.fsch_commit(u, 1e-3)               # velocity and a few broken bonds
    h, v = ctx.fsch_state()
    assert (h == 0).any()
    dt = 5e-4
    k, r = ctx.assemble_fsch(dt, u, uold)
    m = oracle.Mesh(oracle.make_grid(dim, nx, ny, nz, hi=_hi(dim, nx, ny, nz)), factor=factor)
    ops, _ = ctx.operators()
    rp, cols = ctx.csr()
    _, rev, isg = ctx.ch_geometry()
    ok, orr = oracle.assemble_fsch(m, ops, oracle.fsch_lapc(m, ops), rev, isg, ctx.fsch_constants()[:13], dt,
                                   v, h.astype(float), rp, cols, u, uold)
    _compare(k, r, ok, orr, nd)
    ctx.close()
