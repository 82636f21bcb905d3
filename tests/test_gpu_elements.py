"""Block-element assembly on the device across its row-length variants:
the stress-coupled Cahn-Hilliard, Cahn-Hilliard and dynamic fracture
elements stage a node's bonds in LDS when its CSR block row has at most 32,
64 or 128 node blocks (16, 32 or 64 lanes per node) and fall back to a
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


@pytest.mark.parametrize("dim,nx,ny,nz,factor", VARIANTS + [(2, 20, 16, 1, 3.015)])
def test_dynamic_fracture_row_variants(dim, nx, ny, nz, factor, oracle):
    print("dynamic fracture dim %d grid %dx%dx%d factor %.3f" % (dim, nx, ny, nz, factor), flush=True)
    g = perix_amd.StructuredGrid(dim, nx, ny, nz, hi=_hi(dim, nx, ny, nz))
    cloud = perix_amd.build_structured_cloud(g)
    delta = perix_amd.horizon_radius(g, factor)
    ctx = pxg.Context(0)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, dim)
    ctx.set_cracks(segs=[(0.5, 0.0, 0.5, 0.3)])
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(dim)
    E, nu, rho = problems.E_STEEL, problems.NU, problems.RHO
    c11, c12, c66 = problems.elastic_constants(dim, False, E, nu)
    rs = np.random.RandomState(3)
    nd = cloud.n * dim
    uold = 1e-6 * rs.standard_normal(nd)
    ctx.dynfrac_setup(pxg.FracParams(dim, c11, c12, c66, rho, 1e5, -2e5, 3e4, 2e-3, 1, 1), uold,
                      v0=[0.3, -0.1, 0.2][:dim])
    uold2 = uold + 2e-4 * rs.standard_normal(nd)
    ctx.dynfrac_commit(uold2, 1e-7)
    h, v = ctx.dynfrac_state()
    assert (h == 0).any()
    u = uold2 + 1e-7 * rs.standard_normal(nd)
    dt = 1e-7
    k, r = ctx.assemble_dynfrac(dt, u, uold2)
    m = oracle.Mesh(oracle.make_grid(dim, nx, ny, nz, hi=_hi(dim, nx, ny, nz)), factor=factor)
    ops, _ = ctx.operators()
    rp, cols = ctx.csr()
    print("most blocks in a row: %d" % (np.diff(rp)[::dim] // dim).max())
    ok, orr = oracle.assemble_dynfrac(m, ops, rp, cols, c11, c12, c66, rho, [1e5, -2e5, 3e4][:dim], dt, v,
                                      h.astype(float), u, uold2)
    assert np.array_equal(k.view(np.uint64), ok.view(np.uint64))
    assert np.array_equal(r.view(np.uint64), orr.view(np.uint64))
    ctx.close()


@pytest.mark.parametrize("dim,nx,ny,nz,factor", VARIANTS + [(2, 20, 16, 1, 3.015)])
def test_cahn_hilliard_row_variants(dim, nx, ny, nz, factor, oracle):
    print("Cahn-Hilliard dim %d grid %dx%dx%d factor %.3f" % (dim, nx, ny, nz, factor), flush=True)
    g = perix_amd.StructuredGrid(dim, nx, ny, nz, hi=_hi(dim, nx, ny, nz))
    cloud = perix_amd.build_structured_cloud(g)
    delta = perix_amd.horizon_radius(g, factor)
    ctx = pxg.Context(0)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, dim)
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(2)
    ctx.ch_setup()
    rs = np.random.RandomState(9)
    u = np.zeros(2 * cloud.n)
    u[0::2] = 0.5 + 0.02 * rs.standard_normal(cloud.n)
    u[0::2][::13] = -0.02                        # the clamp range
    u[1::2] = 0.1 * rs.standard_normal(cloud.n)
    uold = u + 1e-3 * rs.standard_normal(2 * cloud.n)
    kappa, chi, M0, dt = 2e-4, 3.0, 1.0, 1e-3
    k, r = ctx.assemble_ch(kappa, chi, M0, dt, u, uold)
    m = oracle.Mesh(oracle.make_grid(dim, nx, ny, nz, hi=_hi(dim, nx, ny, nz)), factor=factor)
    ops, _ = ctx.operators()
    rp, cols = ctx.csr()
    print("most blocks in a row: %d" % (np.diff(rp)[::2] // 2).max())
    lap, rev, isg = ctx.ch_geometry()
    mn = oracle.ch_mobility(m, u, M0)
    ok, orr = oracle.assemble_ch(m, lap, rev, isg, mn, kappa, chi, dt, rp, cols, u, uold)
    _compare(k, r, ok, orr, 2)
    ctx.close()
