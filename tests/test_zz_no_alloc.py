"""pxg_assemble_poisson_noalloc (the allocation-free Poisson assembly) against
pxg_assemble_poisson_dev through the C ABI: bit-identical K and residual on
two calls in a row, the second on a smaller problem (the scratch array keeps
the first call's Laplacian, so stale entries would show), and a problem
beyond the scratch limit refused with PXG_ERR_ARG before any launch."""
import math

import numpy as np
import pytest

import perix_amd
from perix_amd import pxg

pytestmark = pytest.mark.gpu


def _poisson(nx, ny):
    g = perix_amd.StructuredGrid(2, nx, ny)
    cloud = perix_amd.build_structured_cloud(g)
    delta = perix_amd.horizon_radius(g, 3.015)
    ctx = pxg.Context(0)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, 2)
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(1)
    return ctx, cloud


def _both(ctx, cloud, seed):
    import torch
    n, nnz = cloud.n, ctx.nnz
    rs = np.random.RandomState(seed)
    u = torch.from_numpy(rs.standard_normal(n)).cuda()
    x, y = cloud.coords[0::3], cloud.coords[1::3]
    f = torch.from_numpy(np.ascontiguousarray(2 * math.pi ** 2 * np.sin(math.pi * x) * np.sin(math.pi * y))).cuda()
    k1 = torch.full((nnz,), float("nan"), dtype=torch.float64, device="cuda")
    r1 = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    k2, r2 = k1.clone(), r1.clone()
    torch.cuda.synchronize()
    ctx.assemble_poisson_noalloc(1.3, u.data_ptr(), k1.data_ptr(), r1.data_ptr(), f_ptr=f.data_ptr())
    ctx.sync()
    ctx.assemble_poisson_dev(1.3, u.data_ptr(), k2.data_ptr(), r2.data_ptr(), f_ptr=f.data_ptr())
    ctx.sync()
    k1, r1, k2, r2 = (t.cpu().numpy() for t in (k1, r1, k2, r2))
    assert np.isfinite(k1).all() and np.isfinite(r1).all()
    assert np.array_equal(k1.view(np.uint64), k2.view(np.uint64))
    assert np.array_equal(r1.view(np.uint64), r2.view(np.uint64))


def test_noalloc_poisson_matches_dev_on_two_calls():
    big, cb = _poisson(120, 100)
    small, cs = _poisson(40, 30)
    assert small.nb < big.nb <= 2097152
    _both(big, cb, 1)
    _both(small, cs, 2)          # smaller input: scratch holds the larger call's data
    _both(big, cb, 3)
    big.close()
    small.close()


def test_noalloc_poisson_refuses_beyond_scratch():
    ctx, cloud = _poisson(300, 300)
    assert ctx.nb > 2097152
    with pytest.raises(pxg.PxgError) as e:
        ctx.assemble_poisson_noalloc(1.0, 1, 1, 1)
    assert e.value.code == -1 and "scratch limit" in str(e.value)
    ctx.close()
