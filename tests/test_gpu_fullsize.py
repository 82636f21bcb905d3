"""Size-independent properties of the device's families and g-weights at the
full 160 x 80 x 40 block (556,800 nodes, 64.7 M directed bonds), where the
oracle is too slow to run beside it: no duplicate family members, the horizon
relation is symmetric (j in H_i <=> i in H_j), an interior family holds the
closed-form 122 members at 3.015 dx, and the g-weights of sampled
non-degenerate nodes reproduce polynomials up to order 2:
sum_j g^q(xi_j) xi_j^p = p! delta_pq to 1e-12 of the term scale."""
import math

import numpy as np
import pytest

import perix_amd
from perix_amd import pxg

pytestmark = pytest.mark.gpu

MULTI = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]


def family_checks(n, off, nbr):
    """(no duplicates, symmetric) of a CSR family list with 1-based ids."""
    i = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    j = nbr.astype(np.int64) - 1
    fwd = np.sort(i * n + j)
    rev = np.sort(j * n + i)
    return bool(np.all(np.diff(fwd) > 0)), bool(np.array_equal(fwd, rev))


def reproduction_error(X, off, nbr, ops, nodes):
    """Largest |sum_j g^q xi^p - p! delta_pq| / term scale over `nodes`."""
    worst = 0.0
    for i in nodes:
        t0, t1 = off[i], off[i + 1]
        xi = X[nbr[t0:t1] - 1] - X[i]
        for a, p in enumerate(MULTI):
            mono = xi[:, 0] ** p[0] * xi[:, 1] ** p[1] * xi[:, 2] ** p[2]
            terms = ops[t0:t1, :] * mono[:, None]
            want = np.zeros(len(MULTI))
            want[a] = math.factorial(p[0]) * math.factorial(p[1]) * math.factorial(p[2])
            err = np.abs(terms.sum(axis=0) - want) / np.maximum(1.0, np.abs(terms).sum(axis=0))
            worst = max(worst, float(err.max()))
    return worst


def test_block_families_and_weights_full_size():
    g = perix_amd.StructuredGrid(3, 160, 80, 40, hi=(0.16, 0.08, 0.04))
    cloud = perix_amd.build_structured_cloud(g)
    delta = perix_amd.horizon_radius(g, 3.015)
    print("block: %d nodes" % cloud.n, flush=True)
    ctx = pxg.Context(0)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, 3)
    nb = ctx.build_families(delta)
    nd = ctx.build_operators(delta, float(max(cloud.dxyz)))
    off, nbr = ctx.families()
    print("bonds %d, degenerate %d" % (nb, nd), flush=True)
    assert nb == 64718816 and nd == 0
    unique, symmetric = family_checks(cloud.n, off, nbr)
    assert unique and symmetric
    # interior bulk node (cell 80, 40, 20): the closed-form family size
    centre = (20 * 80 + 40) * 160 + 80
    assert off[centre + 1] - off[centre] == 122
    ops, flags = ctx.operators()
    assert not flags.any()
    X = cloud.coords.reshape(-1, 3)
    nodes = np.arange(0, cloud.n, 1009)
    err = reproduction_error(X, off, nbr, ops, nodes)
    print("polynomial reproduction over %d nodes: max rel %.2e" % (len(nodes), err))
    assert err <= 1e-12
    ctx.close()


def test_block_rigid_and_linear_fields_full_size():
    """Full-size step arithmetic (pxg_frac_accel, FP64): a rigid translation
    gives exactly zero acceleration and a linear field zero to round-off
    (the second-derivative operators reproduce linear fields), measured
    against the acceleration of a random field of the same amplitude."""
    from perix_amd import problems
    s = problems.explicit_plate(3, 160, 80, 40, hi=(0.16, 0.08, 0.04), cracks=[],
                                box=(0.04, 0.06, 1.0, -1.0, 1.0, -1.0), nsteps=10)
    ctx, cl = s.ctx, s.cloud
    print("block: %d nodes, %d bonds" % (cl.n, ctx.nb), flush=True)
    X = cl.coords.reshape(-1, 3)
    eps = 1.0e-4
    rng = np.random.default_rng(5)
    A = rng.standard_normal((3, 3))
    h = np.ones(ctx.nb)
    a, broke = ctx.frac_accel(np.tile([1e-4, -2e-4, 3e-4], cl.n), h)
    assert broke == 0 and not np.any(a)
    a_lin, broke = ctx.frac_accel((eps * X @ A.T).reshape(-1), h)
    assert broke == 0
    a_rnd, _ = ctx.frac_accel(rng.standard_normal(cl.n * 3) * eps * s.delta, h)
    assert np.all(h == 1.0)
    ratio = np.abs(a_lin).max() / np.abs(a_rnd).max()
    print("linear field: max |a| %.3e, random field %.3e, ratio %.2e" %
          (np.abs(a_lin).max(), np.abs(a_rnd).max(), ratio))
    assert ratio <= 1e-9
    ctx.close()
