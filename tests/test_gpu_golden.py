"""GPU results against the golden fixtures written from the reference's own
build (tests/golden/make_golden.py): the device path is pinned to the
reference directly, not only through the oracle.

Bit-exact: families (members and order), CSR pattern, g-weight degeneracy
flags.  Within 1e-12 relative: g-weights (device exp/sqrt differ from glibc
by an ulp), Poisson and Cahn-Hilliard assembly (normalised by the scale of
the terms they sum).  FP64 explicit trajectory end to end: the same
broken-bond set and displacements within 1e-9 relative L2 (the operators'
last-bit differences are amplified by the dynamics).
"""
import os

import numpy as np
import pytest

import perix_amd
from perix_amd import pxg, problems

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def device_mesh(z, factor=3.015):
    nx, ny, nz = (int(v) for v in z["grid"])
    dim = int(z["dim"])
    g = perix_amd.StructuredGrid(dim, nx, ny, nz, hi=tuple(float(v) for v in z["hi"]))
    cloud = perix_amd.build_structured_cloud(g)
    delta = perix_amd.horizon_radius(g, factor)
    ctx = pxg.Context(0)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, dim)
    ctx.build_families(delta)
    return ctx, cloud, delta


@pytest.mark.parametrize("name", ["mesh_fam_2d.npz", "mesh_fam_3d.npz"])
def test_families_csr_ops_poisson_vs_reference(name):
    print("fixture", name, flush=True)
    z = load(name)
    ctx, cloud, delta = device_mesh(z)
    assert delta == float(z["delta"])
    off, nbr = ctx.families()
    assert np.array_equal(off, z["off"]) and np.array_equal(nbr, z["nbr"])
    ctx.build_csr(1)
    rp, cols = ctx.csr()
    assert np.array_equal(rp, z["rowptr"]) and np.array_equal(cols, z["cols"])
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ops, flags = ctx.operators()
    ref = z["ops"]
    assert np.array_equal(flags & 1, z["flags"])
    assert (np.abs(ops - ref) / np.abs(ref).max(axis=0)).max() <= 1e-12
    k, r = ctx.assemble_poisson(float(z["sigma"]), z["u"], f_const=float(z["f"]))
    kscale = np.abs(z["K"]).max()
    assert np.abs(k - z["K"]).max() <= 1e-12 * kscale
    # residual terms: f and coefficient * (u_j - u_i) with |u| < 1
    assert np.abs(r - z["R"]).max() <= 1e-12 * (abs(float(z["f"])) + kscale)
    ctx.close()


@pytest.mark.parametrize("name", ["explicit_2d.npz", "explicit_3d.npz"])
def test_explicit_fp64_vs_reference(name, oracle):
    print("fixture", name, flush=True)
    z = load(name)
    dim = int(z["dim"])
    nx, ny, nz = (int(v) for v in z["grid"])
    s = problems.explicit_plate(dim, nx, ny, nz, hi=tuple(float(v) for v in z["hi"]),
                                sc=0.003, cracks=[tuple(c) for c in z["cracks"]],
                                box=tuple(float(v) for v in z["box"]), nsteps=int(z["nsteps"]))
    assert s.params.sc == float(z["sc"]), "critical stretch from the same G0"
    dt, total = float(z["dt"]), float(z["total"])
    s.ctx.frac_init(s.u0, dt)
    done = s.ctx.frac_run(dt, total)
    assert done == int(z["steps"])
    u, uo, h, _ = s.ctx.frac_state()
    mism = int((h != z["health"]).sum())
    rel = np.linalg.norm(u - z["u"]) / np.linalg.norm(z["u"])
    print("broken %d, mismatches %d, rel-L2 %.3e" % (int((h == 0).sum()), mism, rel))
    assert mism == 0
    assert rel <= 1e-9
    # output projections of the final state: bit for bit against the oracle
    # on the device's own state and g-weights, and within the trajectory's
    # tolerance against the reference's projections
    m = oracle.Mesh(oracle.make_grid(dim, nx, ny, nz, hi=tuple(float(v) for v in z["hi"])))
    ops, _ = s.ctx.operators()
    opc = oracle.frac_opcache(m, ops)
    p = oracle.frac_params(dim, float(z["E"]), float(z["nu"]), float(z["rho"]), s.params.sc)
    for nm in oracle.PROJECTIONS:
        d = s.ctx.frac_project(nm)
        o = oracle.frac_project(m, p, opc, h.astype(float), u, nm)
        assert np.array_equal(d.view(np.uint64), o.view(np.uint64)), nm
        ref = z["proj_" + nm]
        err = np.abs(d - ref).max() / max(np.abs(ref).max(), 1e-300)
        print("projection %s: max rel to the reference %.2e" % (nm, err))
        assert err <= 1e-7 if nm != "damage" else np.array_equal(d, ref)
    s.ctx.close()


def test_cahn_hilliard_vs_reference():
    z = load("ch_2d.npz")
    nx = int(z["nx"])
    g = perix_amd.StructuredGrid(2, nx, nx)
    cloud = perix_amd.build_structured_cloud(g)
    delta = perix_amd.horizon_radius(g, 3.015)
    ctx = pxg.Context(0)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, 2)
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(2)
    ctx.ch_setup()
    u0 = np.zeros(2 * cloud.n)
    pxg.random_ic(int(z["seed"]), 0.49, 0.51, np.arange(1, cloud.n + 1), 2, [1], u0)
    assert np.array_equal(u0[0::2][: cloud.nbulk], z["u0"][0::2][: cloud.nbulk]), "RandomIC"
    k, r = ctx.assemble_ch(float(z["kappa"]), float(z["chi"]), float(z["M0"]), float(z["dt"]),
                           z["u1"], z["u0"])
    kscale = np.abs(z["K"]).max()
    assert np.abs(k - z["K"]).max() <= 1e-12 * kscale
    assert np.abs(r - z["R"]).max() <= 1e-12 * max(kscale, np.abs(z["R"]).max())
    ctx.close()


@pytest.mark.parametrize("name", ["diffusion_2d.npz", "diffusion_3d.npz"])
def test_diffusion_vs_reference(name, oracle):
    """DiffusionElement assembly: bit for bit against the oracle on the
    device's g-weights, within 1e-12 of the term scale against the reference."""
    z = load(name)
    ctx, cloud, delta = device_mesh(z)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(1)
    D, f, dt = float(z["D"]), float(z["f"]), float(z["dt"])
    k, r = ctx.assemble_diffusion(D, dt, z["u"], z["uold"], f_const=f)
    g = cloud.grid
    m = oracle.Mesh(oracle.make_grid(g.dim, g.nx, g.ny, g.nz, hi=g.hi))
    ops, _ = ctx.operators()
    rp, cols = ctx.csr()
    ok, orr = oracle.assemble_diffusion(m, ops, rp, cols, D, f, dt, z["u"], z["uold"])
    assert np.array_equal(k.view(np.uint64), ok.view(np.uint64))
    assert np.array_equal(r.view(np.uint64), orr.view(np.uint64))
    kscale = np.abs(z["K"]).max()
    assert np.abs(k - z["K"]).max() <= 1e-12 * kscale
    assert np.abs(r - z["R"]).max() <= 1e-12 * (kscale + np.abs(z["R"]).max())
    # device pointers give the same bits
    import torch
    ud, uod = torch.from_numpy(np.ascontiguousarray(z["u"])).cuda(), torch.from_numpy(np.ascontiguousarray(z["uold"])).cuda()
    kd = torch.empty(ctx.nnz, dtype=torch.float64, device="cuda")
    rd = torch.empty(cloud.n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.assemble_diffusion_dev(D, dt, ud.data_ptr(), uod.data_ptr(), kd.data_ptr(), rd.data_ptr(), f_const=f)
    ctx.sync()
    assert np.array_equal(kd.cpu().numpy().view(np.uint64), k.view(np.uint64))
    assert np.array_equal(rd.cpu().numpy().view(np.uint64), r.view(np.uint64))
    ctx.close()


@pytest.mark.parametrize("name", ["dynfrac_2d.npz", "dynfrac_3d.npz"])
def test_dynamic_fracture_vs_reference(name, oracle):
    """PDDODynamicFracElement on the device over the initialising and a
    committing step: health and velocity bit for bit against the reference,
    K and the residual bit for bit against the oracle on the device's
    g-weights and within 1e-12 of the term scale against the reference."""
    z = load(name)
    dim = int(z["dim"])
    nx, ny, nz = (int(v) for v in z["grid"])
    g = perix_amd.StructuredGrid(dim, nx, ny, nz, hi=tuple(float(v) for v in z["hi"]))
    cloud = perix_amd.build_structured_cloud(g)
    delta = perix_amd.horizon_radius(g, 3.015)
    ctx = pxg.Context(0)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, dim)
    ctx.set_cracks(segs=[tuple(float(v) for v in z["crack"])])
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(dim)
    E, nu, rho, dt = float(z["E"]), float(z["nu"]), float(z["rho"]), float(z["dt"])
    c11, c12, c66 = problems.elastic_constants(dim, False, E, nu)
    sc = problems.critical_stretch(float(z["G0"]), E, delta, dim)
    b = list(z["b"]) + [0.0] * (3 - dim)
    ctx.dynfrac_setup(pxg.FracParams(dim, c11, c12, c66, rho, b[0], b[1], b[2], sc, 1, 1), z["uold1"],
                      v0=list(z["v0"]))
    h, v = ctx.dynfrac_state()
    assert np.array_equal(h.astype(float), z["health1"]) and np.array_equal(v, z["vel1"])
    m = oracle.Mesh(oracle.make_grid(dim, nx, ny, nz, hi=tuple(float(v) for v in z["hi"])))
    ops, _ = ctx.operators()
    rp, cols = ctx.csr()
    for step in (1, 2):
        if step == 2:
            ctx.dynfrac_commit(z["uold2"], dt)
            h, v = ctx.dynfrac_state()
            assert np.array_equal(v.view(np.uint64), z["vel2"].view(np.uint64))
            assert np.array_equal(h.astype(float), z["health2"])
            print("%s: broken bonds %d -> %d" % (name, int((z["health1"] == 0).sum()), int((h == 0).sum())))
        u, uold = z["u%d" % step], z["uold%d" % step]
        k, r = ctx.assemble_dynfrac(dt, u, uold)
        ok, orr = oracle.assemble_dynfrac(m, ops, rp, cols, c11, c12, c66, rho, z["b"], dt, v,
                                          h.astype(float), u, uold)
        assert np.array_equal(k.view(np.uint64), ok.view(np.uint64)), "K, step %d" % step
        assert np.array_equal(r.view(np.uint64), orr.view(np.uint64)), "R, step %d" % step
        K, R = z["K%d" % step], z["R%d" % step]
        kscale = np.abs(K).max()
        assert np.abs(k - K).max() <= 1e-12 * kscale
        assert np.abs(r - R).max() <= 1e-12 * (np.abs(R).max() + kscale * np.abs(u).max())
    ctx.close()


@pytest.mark.parametrize("name", ["stressch_2d.npz", "stressch_2d_ps.npz", "stressch_3d.npz"])
def test_stress_cahn_hilliard_vs_reference(name, oracle):
    """StressCahnHilliardElement on the device: the element constants bit for
    bit against the reference; K and the c and u rows of the residual bit for
    bit against the oracle on the device's g-weights and bond geometry (the mu
    rows carry the device logarithm: within 1e-14 of their scale); every
    dof's rows within 1e-11 of that dof's scale against the reference."""
    z = load(name)
    print("fixture", name, flush=True)
    dim = int(z["dim"])
    nd = dim + 2
    ctx, cloud, delta = device_mesh(z)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(nd)
    ref = z["consts"]
    p = pxg.StressChParams(70e9, 0.3, float(ref[6]), float(ref[9]), float(ref[10]), float(ref[8]),
                           float(ref[7]), int(str(z["state"]) == "plane_stress"))
    ctx.stressch_setup(p)
    kc = ctx.stressch_constants()
    assert np.array_equal(kc.view(np.uint64), ref.view(np.uint64))
    dt = float(z["dt"])
    k, r = ctx.assemble_stressch(dt, z["u"], z["uold"])
    g = cloud.grid
    m = oracle.Mesh(oracle.make_grid(g.dim, g.nx, g.ny, g.nz, hi=g.hi))
    ops, _ = ctx.operators()
    rp, cols = ctx.csr()
    lap, rev, isg = ctx.ch_geometry()
    ok, orr = oracle.assemble_stressch(m, ops, lap, rev, isg, kc, dt, rp, cols, z["u"], z["uold"])
    assert np.array_equal(k.view(np.uint64), ok.view(np.uint64))
    mu_rows = np.zeros(r.size, bool)
    mu_rows[1::nd] = True
    assert np.array_equal(r[~mu_rows].view(np.uint64), orr[~mu_rows].view(np.uint64))
    dmu = np.abs(r[mu_rows] - orr[mu_rows]).max()
    print("mu rows: max |device - oracle| %.3e of scale %.3e" % (dmu, np.abs(orr[mu_rows]).max()))
    assert dmu <= 1e-14 * np.abs(orr[mu_rows]).max()
    row = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    for a in range(nd):
        sel = (row % nd) == a
        ks = np.abs(z["K"][sel]).max()
        rs = np.abs(z["R"][a::nd]).max()
        ek = np.abs(k[sel] - z["K"][sel]).max() / ks
        er = np.abs(r[a::nd] - z["R"][a::nd]).max() / rs
        print("dof %d: K rel %.2e, R rel %.2e" % (a, ek, er))
        assert ek <= 1e-11 and er <= 1e-11
    # device pointers give the same bits
    import torch
    ud = torch.from_numpy(np.ascontiguousarray(z["u"])).cuda()
    uod = torch.from_numpy(np.ascontiguousarray(z["uold"])).cuda()
    kd = torch.empty(ctx.nnz, dtype=torch.float64, device="cuda")
    rd = torch.empty(nd * cloud.n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.assemble_stressch_dev(dt, ud.data_ptr(), uod.data_ptr(), kd.data_ptr(), rd.data_ptr())
    ctx.sync()
    assert np.array_equal(kd.cpu().numpy().view(np.uint64), k.view(np.uint64))
    assert np.array_equal(rd.cpu().numpy().view(np.uint64), r.view(np.uint64))
    ctx.close()


@pytest.mark.parametrize("name", ["fsch_2d.npz", "fsch_3d.npz"])
def test_fracture_stress_cahn_hilliard_vs_reference(name, oracle):
    """FracStressCahnHilliardElement on the device over the initialising and
    a committing step: constants, health and velocity bit for bit against
    the reference; K and the c and u rows of the residual bit for bit against
    the oracle on the device's g-weights (mu rows within 1e-14 of their
    scale: device logarithm); every dof's rows within 1e-11 of that dof's
    scale against the reference."""
    z = load(name)
    print("fixture", name, flush=True)
    dim = int(z["dim"])
    nd = dim + 2
    nx, ny, nz = (int(v) for v in z["grid"])
    g = perix_amd.StructuredGrid(dim, nx, ny, nz, hi=tuple(float(v) for v in z["hi"]))
    cloud = perix_amd.build_structured_cloud(g)
    delta = perix_amd.horizon_radius(g, 3.015)
    ctx = pxg.Context(0)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, dim)
    ctx.set_cracks(segs=[tuple(float(v) for v in z["crack"])])
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(nd)
    kc = z["consts"]
    ps = dim == 2 and str(z["state"]) == "plane_stress"
    E, nu = float(z["E"]), float(z["nu"])
    sc = problems.critical_stretch(float(z["G0"]), E, delta, dim, ps)
    base = pxg.StressChParams(E, nu, float(kc[6]), float(kc[9]), float(kc[10]), float(kc[8]), float(kc[7]), int(ps))
    ctx.fsch_setup(pxg.FschParams(base, float(kc[11]), sc, float(kc[12]), int(z["tension_only"]), 1), z["uold1"])
    assert np.array_equal(ctx.fsch_constants().view(np.uint64), kc.view(np.uint64))
    h, v = ctx.fsch_state()
    assert np.array_equal(h.astype(float), z["health1"]) and np.array_equal(v, z["vel1"])
    m = oracle.Mesh(oracle.make_grid(dim, nx, ny, nz, hi=tuple(float(v) for v in z["hi"])))
    ops, _ = ctx.operators()
    rp, cols = ctx.csr()
    lapc = oracle.fsch_lapc(m, ops)
    _, rev, isg = ctx.ch_geometry()
    row = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    dts = {1: float(z["dt1"]), 2: float(z["dt2"])}
    for step in (1, 2):
        if step == 2:
            ctx.fsch_commit(z["uold2"], dts[1])
            h, v = ctx.fsch_state()
            assert np.array_equal(v.view(np.uint64), z["vel2"].view(np.uint64))
            assert np.array_equal(h.astype(float), z["health2"])
            print("%s: broken bonds %d -> %d" % (name, int((z["health1"] == 0).sum()), int((h == 0).sum())))
        u, uold = z["u%d" % step], z["uold%d" % step]
        k, r = ctx.assemble_fsch(dts[step], u, uold)
        ok, orr = oracle.assemble_fsch(m, ops, lapc, rev, isg, kc[:13], dts[step], v, h.astype(float), rp, cols,
                                       u, uold)
        assert np.array_equal(k.view(np.uint64), ok.view(np.uint64)), "K, step %d" % step
        mu_rows = np.zeros(r.size, bool)
        mu_rows[1::nd] = True
        assert np.array_equal(r[~mu_rows].view(np.uint64), orr[~mu_rows].view(np.uint64)), "R, step %d" % step
        assert np.abs(r[mu_rows] - orr[mu_rows]).max() <= 1e-14 * np.abs(orr[mu_rows]).max()
        K, R = z["K%d" % step], z["R%d" % step]
        for a in range(nd):
            sel = (row % nd) == a
            ek = np.abs(k[sel] - K[sel]).max() / np.abs(K[sel]).max()
            er = np.abs(r[a::nd] - R[a::nd]).max() / np.abs(R[a::nd]).max()
            print("step %d dof %d: K rel %.2e, R rel %.2e" % (step, a, ek, er))
            assert ek <= 1e-11 and er <= 1e-11
    ctx.close()


def _sch_device(z, frac):
    """Device context set up as the fixture's element: (ctx, oracle mesh, ops,
    device health or None)."""
    dim = int(z["dim"])
    nx, ny, nz = (int(v) for v in z["grid"])
    hi = tuple(float(v) for v in z["hi"])
    g = perix_amd.StructuredGrid(dim, nx, ny, nz, hi=hi)
    cloud = perix_amd.build_structured_cloud(g)
    delta = perix_amd.horizon_radius(g, 3.015)
    ctx = pxg.Context(0)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, dim)
    if frac:
        ctx.set_cracks(segs=[tuple(float(v) for v in z["crack"])])
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(dim + 2)
    kc = z["consts"]
    ps = dim == 2 and str(z["state"]) == "plane_stress"
    E = float(z["E"]) if frac else 70e9
    nu = float(z["nu"]) if frac else 0.3
    base = pxg.StressChParams(E, nu, float(kc[6]), float(kc[9]), float(kc[10]), float(kc[8]), float(kc[7]), int(ps))
    h = None
    if frac:
        sc = problems.critical_stretch(float(z["G0"]), E, delta, dim, ps)
        ctx.fsch_setup(pxg.FschParams(base, float(kc[11]), sc, float(kc[12]), int(z["tension_only"]), 1), z["uold1"])
        ctx.fsch_commit(z["uold2"], float(z["dt1"]))
        h, _ = ctx.fsch_state()
    else:
        ctx.stressch_setup(base)
    m = _oracle_mesh(z)
    ops, _ = ctx.operators()
    return ctx, m, ops, h


def _oracle_mesh(z):
    from oracle import pxo
    nx, ny, nz = (int(v) for v in z["grid"])
    return pxo.Mesh(pxo.make_grid(int(z["dim"]), nx, ny, nz, hi=tuple(float(v) for v in z["hi"])))


@pytest.mark.parametrize("name", ["stressch_2d.npz", "stressch_2d_ps.npz", "stressch_3d.npz",
                                  "fsch_2d.npz", "fsch_3d.npz"])
def test_stress_cahn_hilliard_projections_vs_reference(name, oracle):
    """The chemo-mechanical elements' projections on the device: bit for bit
    against the oracle on the device's g-weights and state (f' within 1e-14:
    device logarithm), within 1e-9 of each field's scale against the
    reference's computeNodalProjection (the g-weights differ by 1e-12)."""
    z = load(name)
    print("fixture", name, flush=True)
    frac = name.startswith("fsch")
    dim = int(z["dim"])
    ctx, m, ops, h = _sch_device(z, frac)
    k = z["consts"]
    consts = [k[0], k[1], k[2], k[3], k[10], k[8]]
    plane_strain = dim == 3 or str(z["state"]) != "plane_stress"
    u = z["u2"] if frac else z["u"]
    names = [n[len("proj_"):] for n in z.files if n.startswith("proj_")]
    for nm in names:
        d = ctx.fsch_project(nm, u) if frac else ctx.stressch_project(nm, u)
        o = oracle.sch_project(m, ops, consts, frac, plane_strain, u, nm,
                               health=None if h is None else h.astype(float),
                               seed=z["health1"] if frac else None)
        if nm == "fPrime":
            assert np.abs(d - o).max() <= 1e-14 * np.abs(o).max(), nm
        else:
            assert np.array_equal(d.view(np.uint64), o.view(np.uint64)), nm
        ref = z["proj_" + nm]
        err = np.abs(d - ref).max() / max(np.abs(ref).max(), 1e-300)
        print("%s: max rel to the reference %.2e" % (nm, err))
        assert err <= 1e-9, nm
    import torch
    ud = torch.from_numpy(np.ascontiguousarray(u)).cuda()
    od = torch.empty(m.n * 9, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    (ctx.fsch_project_dev if frac else ctx.stressch_project_dev)(pxg.SCH_PROJECTIONS["stress"], ud.data_ptr(),
                                                                od.data_ptr())
    ctx.sync()
    host = ctx.fsch_project("stress", u) if frac else ctx.stressch_project("stress", u)
    assert np.array_equal(od.cpu().numpy().reshape(-1, 9), host)
    ctx.close()
