"""GPU parity of the explicit fracture loop against the oracle.  FP64 must
reproduce the oracle trajectory bit for bit (displacements, u_old, broken
bond set and break steps) when both start from the same operators; the
FP32C form reports its error against FP64."""
import numpy as np
import pytest

from perix_amd import pxg, problems

pytestmark = pytest.mark.gpu


def _oracle_run(oracle, setup, nsteps_total, ops=None, factor=3.015):
    """Oracle trajectory; ops=None uses the oracle's own g-weights, otherwise
    the given ones (e.g. the device's, to compare the step arithmetic alone)."""
    s = setup
    cl = s.cloud
    g = cl.grid
    m = oracle.Mesh(oracle.make_grid(g.dim, g.nx, g.ny, g.nz, hi=g.hi), factor=factor)
    if ops is None:
        ops, _ = oracle.operators(m)
    opc = oracle.frac_opcache(m, ops)
    p = oracle.frac_params(g.dim, problems.E_STEEL, problems.NU, problems.RHO, s.params.sc)
    h = oracle.seed_health(m, segs=s.cracks)
    bcs = oracle.make_bcs(s.bcs)
    u = s.u0.copy()
    uo = np.zeros_like(u)
    steps, bs = oracle.explicit_run(m, p, opc, bcs, s.dt, s.total, u, uo, h)
    return steps, u, uo, h, bs


@pytest.mark.parametrize("dim", [2, 3])
def test_explicit_fp64_bit_exact(oracle, dim):
    if dim == 2:
        s = problems.explicit_plate(2, 80, 40, hi=(0.02, 0.01, 1.0), sc=0.003,
                                    cracks=[(0.005, 0.0, 0.005, 0.005), (0.015, 0.0, 0.015, 0.005)],
                                    box=(0.0075, 0.0125, 1.0, -1.0, 1.0, -1.0), nsteps=300)
    else:
        s = problems.explicit_plate(3, 24, 12, 6, hi=(0.024, 0.012, 0.006), sc=0.003,
                                    cracks=[(0.012, 0.0, 0.012, 0.0012)],
                                    box=(0.006, 0.009, 1.0, -1.0, 1.0, -1.0), nsteps=200)
    print("dim", dim, "N", s.cloud.n, flush=True)
    ctx = s.ctx
    ctx.frac_init(s.u0, s.dt)
    done = ctx.frac_run(s.dt, s.total)
    u, uo, h, bs = ctx.frac_state()
    dev_ops, _ = ctx.operators()
    # same operators: the step arithmetic must agree bit for bit
    steps, ru, ruo, rh, rbs = _oracle_run(oracle, s, s.nsteps, ops=dev_ops)
    assert done == steps
    assert (rh == 0).sum() > 0, "case must break bonds"
    assert np.array_equal(h.astype(float), rh)
    assert np.array_equal(bs, rbs)
    assert np.array_equal(u.view(np.uint64), ru.view(np.uint64))
    assert np.array_equal(uo.view(np.uint64), ruo.view(np.uint64))
    # end to end (each side with its own g-weights): same broken-bond set,
    # displacements within FP64 round-off of the operators
    steps2, ru2, _, rh2, rbs2 = _oracle_run(oracle, s, s.nsteps)
    rel = np.linalg.norm(u - ru2) / np.linalg.norm(ru2)
    print("end-to-end rel-L2 %.3e, broken-bond mismatches %d" % (rel, int((h.astype(float) != rh2).sum())))
    assert rel <= 1e-12
    assert np.array_equal(h.astype(float), rh2)
    ctx.close()


def test_explicit_fp32c_error(oracle):
    s64 = problems.explicit_plate(2, 80, 40, hi=(0.02, 0.01, 1.0), sc=0.003,
                                  cracks=[(0.005, 0.0, 0.005, 0.005), (0.015, 0.0, 0.015, 0.005)],
                                  box=(0.0075, 0.0125, 1.0, -1.0, 1.0, -1.0), nsteps=300)
    s32 = problems.explicit_plate(2, 80, 40, hi=(0.02, 0.01, 1.0), sc=0.003,
                                  cracks=[(0.005, 0.0, 0.005, 0.005), (0.015, 0.0, 0.015, 0.005)],
                                  box=(0.0075, 0.0125, 1.0, -1.0, 1.0, -1.0), nsteps=300,
                                  precision=pxg.FP32C)
    for s in (s64, s32):
        s.ctx.frac_init(s.u0, s.dt)
        s.ctx.frac_run(s.dt, s.total)
    u64, _, h64, _ = s64.ctx.frac_state()
    u32, _, h32, _ = s32.ctx.frac_state()
    rel = np.linalg.norm(u32 - u64) / np.linalg.norm(u64)
    mism = int((h32 != h64).sum())
    print("FP32C rel-L2 %.3e, max abs %.3e, broken-bond mismatches %d of %d broken"
          % (rel, np.abs(u32 - u64).max(), mism, int((h64 == 0).sum())))
    assert rel <= 1e-5
    s64.ctx.close(); s32.ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_fp32c_is_fp64_on_fp32_operators(dim):
    """FP32C stores the operators in FP32 and forms and sums the bond forces in
    FP64, so its trajectory is that of the FP64 step run on the FP32-rounded
    operators (PXG_FRAC_FP32_OPERATORS): the two differ only by the order of
    the family sum (lane partial sums vs family order), far below the last
    bit of u, and break the same bonds at the same steps."""
    if dim == 2:
        kw = dict(hi=(0.02, 0.01, 1.0), sc=0.003, cracks=[(0.005, 0.0, 0.005, 0.005), (0.015, 0.0, 0.015, 0.005)],
                  box=(0.0075, 0.0125, 1.0, -1.0, 1.0, -1.0), nsteps=300)
        args = (2, 80, 40)
    else:
        kw = dict(hi=(0.024, 0.012, 0.006), sc=0.003, cracks=[(0.012, 0.0, 0.012, 0.0012)],
                  box=(0.006, 0.009, 1.0, -1.0, 1.0, -1.0), nsteps=200)
        args = (3, 24, 12, 6)
    runs = []
    for prec, opts in ((pxg.FP32C, pxg.Context.FRAC_NO_SHARED_RUNS), (pxg.FP64, pxg.Context.FRAC_FP32_OPERATORS)):
        ctx = pxg.Context(0)
        ctx.frac_options(opts)
        s = problems.explicit_plate(*args, precision=prec, ctx=ctx, **kw)
        s.ctx.frac_init(s.u0, s.dt)
        s.ctx.frac_run(s.dt, s.total)
        runs.append(s.ctx.frac_state())
        s.ctx.close()
    (u32, _, h32, b32), (u64, _, h64, b64) = runs
    rel = np.linalg.norm(u32 - u64) / np.linalg.norm(u64)
    print("FP32C vs FP64 on FP32 operators: rel-L2 %.2e, bit-equal fraction %.4f, broken %d"
          % (rel, np.mean(u32 == u64), int((h64 == 0).sum())))
    assert np.array_equal(h32, h64) and np.array_equal(b32, b64)
    assert rel <= 1e-13


def test_drop_in_accel(oracle):
    """pxg_frac_accel = one preprocessIteration (damage + force + accel)."""
    s = problems.explicit_plate(2, 40, 20, hi=(0.02, 0.01, 1.0), sc=0.003,
                                cracks=[(0.005, 0.0, 0.005, 0.005)],
                                box=(0.0075, 0.0125, 1.0, -1.0, 1.0, -1.0), nsteps=10)
    cl = s.cloud
    m = oracle.Mesh(oracle.make_grid(2, 40, 20, hi=(0.02, 0.01, 1.0)))
    ops, _ = s.ctx.operators()
    opc = oracle.frac_opcache(m, ops)
    p = oracle.frac_params(2, problems.E_STEEL, problems.NU, problems.RHO, s.params.sc)
    h_ref = oracle.seed_health(m, segs=s.cracks)
    u = np.random.RandomState(1).standard_normal(cl.n * 2) * 2e-5
    a_ref = oracle.frac_accel(m, p, opc, h_ref, u)
    h = oracle.seed_health(m, segs=s.cracks)
    a, broke = s.ctx.frac_accel(u, h)
    assert np.array_equal(a.view(np.uint64), a_ref.view(np.uint64))
    assert np.array_equal(h, h_ref)
    assert broke == int((h_ref == 0).sum() > (oracle.seed_health(m, segs=s.cracks) == 0).sum())
    s.ctx.close()


def test_blowup_guard(oracle):
    """dt far above the CFL limit trips the guard at the same step as the oracle."""
    s = problems.explicit_plate(2, 40, 20, hi=(0.02, 0.01, 1.0), sc=0.0, cracks=[],
                                box=(0.0075, 0.0125, 1.0, -1.0, 1.0, -1.0), nsteps=400, cfl=3.0)
    s.ctx.frac_init(s.u0, s.dt)
    with pytest.raises(pxg.PxgError) as ei:
        s.ctx.frac_run(s.dt, s.total)
    steps, ru, ruo, rh, rbs = _oracle_run(oracle, s, s.nsteps)
    assert steps < 0
    assert ei.value.code == -6
    assert ei.value.steps == -steps - 1
    s.ctx.close()


def test_explicit_fp32c_3d_shared_runs():
    """3D FP32C on the tiled stream layout: most interior nodes read a shared
    operator run, and the result stays within 1e-5 relative-L2 of FP64 with
    the same broken-bond set."""
    kw = dict(hi=(0.032, 0.016, 0.012), sc=0.003, cracks=[(0.016, 0.0, 0.016, 0.004)],
              box=(0.008, 0.012, 1.0, -1.0, 1.0, -1.0), nsteps=200)
    s64 = problems.explicit_plate(3, 32, 16, 12, **kw)
    s32 = problems.explicit_plate(3, 32, 16, 12, precision=pxg.FP32C, **kw)
    pst, tiled, nclasses, nshared = s32.ctx.frac_layout()
    print("layout: %d slots/component, %d shared runs, %d of %d nodes on them"
          % (pst, nclasses, nshared, s32.cloud.n))
    assert tiled == 1 and nshared > 0.2 * s32.cloud.n
    for s in (s64, s32):
        s.ctx.frac_init(s.u0, s.dt)
        s.ctx.frac_run(s.dt, s.total)
    u64, _, h64, _ = s64.ctx.frac_state()
    u32, _, h32, _ = s32.ctx.frac_state()
    rel = np.linalg.norm(u32 - u64) / np.linalg.norm(u64)
    print("3D FP32C rel-L2 %.3e, broken-bond mismatches %d of %d broken"
          % (rel, int((h32 != h64).sum()), int((h64 == 0).sum())))
    assert (h64 == 0).sum() > 0
    assert rel <= 1e-5 and np.array_equal(h32, h64)
    s64.ctx.close(); s32.ctx.close()


def test_explicit_fp32c_2d_exact_shared_runs():
    """2D FP32C shares only operator runs whose FP32 bits equal the node's
    own, so the run is bit for bit the run with every node on its own
    operators (PXG_FRAC_NO_SHARED_RUNS): u^n, u^{n-1}, health and break
    steps, with most nodes on a shared run."""
    kw = dict(hi=(0.02, 0.01, 1.0), sc=0.003, cracks=[(0.005, 0.0, 0.005, 0.005), (0.015, 0.0, 0.015, 0.005)],
              box=(0.0075, 0.0125, 1.0, -1.0, 1.0, -1.0), nsteps=300)
    runs = []
    for opts in (0, pxg.Context.FRAC_NO_SHARED_RUNS):
        ctx = pxg.Context(0)
        ctx.frac_options(opts)
        s = problems.explicit_plate(2, 80, 40, precision=pxg.FP32C, ctx=ctx, **kw)
        pst, tiled, nclasses, nshared = s.ctx.frac_layout()
        print("opts %d: %d shared runs, %d of %d nodes on them" % (opts, nclasses, nshared, s.cloud.n))
        if opts == 0:
            assert nshared > 0.4 * s.cloud.n and nclasses < 0.25 * s.cloud.n
        else:
            assert nshared == 0
        s.ctx.frac_init(s.u0, s.dt)
        assert s.ctx.frac_run(s.dt, s.total) == 300
        runs.append(s.ctx.frac_state())
        s.ctx.close()
    (u, uo, h, b), (u2, uo2, h2, b2) = runs
    assert (h == 0).sum() > 0
    assert np.array_equal(u.view(np.uint64), u2.view(np.uint64))
    assert np.array_equal(uo.view(np.uint64), uo2.view(np.uint64))
    assert np.array_equal(h, h2) and np.array_equal(b, b2)


def test_explicit_large_horizon_global_gathers(oracle):
    """A 4.5 dx horizon (about 380 bonds per family): every interior step
    block's neighbourhood exceeds the LDS row budget, so those blocks gather
    their neighbours from global memory.  FP64 stays bit for bit with the
    oracle on the device's g-weights; FP32C stays within 1e-5 of FP64."""
    kw = dict(hi=(0.016, 0.012, 0.008), sc=0.003, cracks=[(0.008, 0.0, 0.008, 0.003)],
              box=(0.004, 0.008, 1.0, -1.0, 1.0, -1.0), nsteps=60, factor=4.5)
    s = problems.explicit_plate(3, 16, 12, 8, **kw)
    s32 = problems.explicit_plate(3, 16, 12, 8, precision=pxg.FP32C, **kw)
    print("N %d, bonds %d" % (s.cloud.n, s.ctx.nb), flush=True)
    for q in (s, s32):
        q.ctx.frac_init(q.u0, q.dt)
        q.ctx.frac_run(q.dt, q.total)
    u, uo, h, bs = s.ctx.frac_state()
    dev_ops, _ = s.ctx.operators()
    steps, ru, ruo, rh, rbs = _oracle_run(oracle, s, s.nsteps, ops=dev_ops, factor=4.5)
    assert (rh == 0).sum() > 0, "case must break bonds"
    assert np.array_equal(h.astype(float), rh) and np.array_equal(bs, rbs)
    assert np.array_equal(u.view(np.uint64), ru.view(np.uint64))
    u32, _, h32, _ = s32.ctx.frac_state()
    rel = np.linalg.norm(u32 - u) / np.linalg.norm(u)
    print("FP32C rel-L2 %.3e, mismatches %d" % (rel, int((h32 != h).sum())))
    assert rel <= 1e-5
    s.ctx.close(); s32.ctx.close()


def test_step_timing_spans_every_launch():
    """pxg_frac_kernel_time counts every step launch of the runs made while
    profiling (event pairs around each run of launches between guard polls)
    and its average cannot exceed the host clock's time per step."""
    import time
    s = problems.explicit_plate(2, 80, 40, hi=(0.02, 0.01, 1.0), nsteps=600)
    ctx = s.ctx
    ctx.frac_init(s.u0, s.dt)
    ctx.sync()
    ctx.frac_profile(True)
    t0 = time.perf_counter()
    done = ctx.frac_run(s.dt, s.total)
    ctx.sync()
    wall = time.perf_counter() - t0
    ms, launches = ctx.frac_kernel_time()
    ctx.frac_profile(False)
    print("launches %d, event ms %.3f, host ms %.3f" % (launches, ms, wall * 1e3))
    assert done == 600 and launches == 600
    assert 0.0 < ms <= wall * 1e3
    ctx.close()
