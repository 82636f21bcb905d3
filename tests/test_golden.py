"""The C oracle against the golden fixtures written from the reference's own
build (tests/golden/make_golden.py).  CPU only.

Bit-exact: point cloud, horizon, families (members and order), CSR pattern,
g-weights, Poisson and Cahn-Hilliard assembly, the FP64 explicit trajectory
and its broken-bond set, the RandomIC field.
"""
import os

import numpy as np
import pytest

from oracle import pxo

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def mesh_of(z):
    nx, ny, nz = (int(v) for v in z["grid"])
    dim = int(z["dim"])
    return pxo.Mesh(pxo.make_grid(dim, nx, ny, nz, hi=tuple(z["hi"])), factor=float(z["factor"]))


def same(a, b):
    a = np.asarray(a); b = np.asarray(b)
    if a.dtype.kind == "f":
        return a.shape == b.shape and np.array_equal(a.view(np.uint64), np.asarray(b, np.float64).view(np.uint64))
    return a.shape == b.shape and np.array_equal(a, b)


@pytest.mark.parametrize("name", ["mesh_fam_2d.npz", "mesh_fam_3d.npz"])
def test_mesh_families_ops_csr_poisson(name):
    z = load(name)
    m = mesh_of(z)
    assert m.n == int(z["n"]) and m.nbulk == int(z["nbulk"])
    assert same(m.coords, z["coords"]), "node coordinates"
    assert same(m.vols, z["vols"]), "node volumes"
    assert np.array_equal(m.mirror, z["mirror"]), "ghost mirrors"
    assert m.delta == float(z["delta"]), "horizon radius"
    assert same(m.dxyz, z["dxyz"])
    assert np.array_equal(m.off, z["off"]) and np.array_equal(m.nbr, z["nbr"]), "families (members and order)"
    ops, flags = pxo.operators(m)
    assert same(ops, z["ops"]), "g-weights (Eigen FullPivLU restated bit for bit)"
    assert np.array_equal(flags & 1, z["flags"])
    rp, cols = pxo.csr_pattern(m, 1)
    assert np.array_equal(rp, z["rowptr"]) and np.array_equal(cols, z["cols"]), "CSR pattern"
    k, r = pxo.assemble_poisson(m, ops, rp, cols, float(z["sigma"]), z["u"], f_const=float(z["f"]))
    assert same(k, z["K"]) and same(r, z["R"]), "Poisson Jacobian / residual"


@pytest.mark.parametrize("name", ["explicit_2d.npz", "explicit_3d.npz"])
def test_explicit_trajectory(name):
    z = load(name)
    dim = int(z["dim"])
    m = mesh_of(dict(z, factor=3.015))
    E, nu, rho = float(z["E"]), float(z["nu"]), float(z["rho"])
    sc = pxo.critical_stretch(float(z["G0"]), E, m.delta, dim, False)
    assert sc == float(z["sc"]), "critical stretch from G0 (FractureCriterion.h:30-43)"
    ops, _ = pxo.operators(m)
    opc = pxo.frac_opcache(m, ops)
    p = pxo.frac_params(dim, E, nu, rho, sc)
    h = pxo.seed_health(m, segs=[tuple(c) for c in z["cracks"]])
    walls = ["left", "right", "bottom", "top"] + (["back", "front"] if dim == 3 else [])
    comps = list(range(dim))
    specs = [{"type": "neumann", "nodes": m.group(pxo.WALLS[w]), "comps": comps} for w in walls]
    specs.append({"type": "dirichlet", "nodes": m.group(2), "comps": comps, "value": [0.0] * dim,
                  "velocity": [0.0, 16.5] + ([0.0] if dim == 3 else []), "direct": True,
                  "box": list(z["box"])})
    bcs = pxo.make_bcs(specs)
    u = np.zeros(m.n * dim); uo = np.zeros(m.n * dim)
    pxo.bc_preset(m, dim, bcs, u, 0.0)
    steps, _ = pxo.explicit_run(m, p, opc, bcs, float(z["dt"]), float(z["total"]), u, uo, h)
    assert steps == int(z["steps"])
    assert same(u, z["u"]) and same(uo, z["uold"]), "FP64 displacement history"
    assert np.array_equal((h > 0).astype(np.uint8), z["health"]), "broken-bond set"
    # the element's output projections of the final state
    for nm in pxo.PROJECTIONS:
        assert same(pxo.frac_project(m, p, opc, h, u, nm), z["proj_" + nm]), nm


def test_cahn_hilliard_assembly_and_random_ic():
    z = load("ch_2d.npz")
    nx = int(z["nx"])
    m = pxo.Mesh(pxo.make_grid(2, nx, nx))
    u = np.zeros(2 * m.n)
    pxo.random_ic(int(z["seed"]), 0.49, 0.51, np.arange(1, m.n + 1), 2, [1], u)
    specs = [{"type": "neumann", "nodes": m.group(w), "comps": [0, 1]} for w in range(4)]
    pxo.bc_preset(m, 2, pxo.make_bcs(specs), u, 0.0)
    assert same(u, z["u0"]), "RandomIC (std::mt19937_64) + Neumann presets"
    ops, _ = pxo.operators(m)
    lap, rev, isg = pxo.ch_geometry(m, ops)
    assert same(lap, z["lap"]) and np.array_equal(rev, z["rev"])
    mn = pxo.ch_mobility(m, z["u1"], float(z["M0"]))
    assert same(mn, z["mnode"])
    rp, cols = pxo.csr_pattern(m, 2)
    k, r = pxo.assemble_ch(m, lap, rev, isg, mn, float(z["kappa"]), float(z["chi"]), float(z["dt"]),
                           rp, cols, z["u1"], z["u0"])
    assert same(k, z["K"]) and same(r, z["R"]), "Cahn-Hilliard Jacobian / residual"


@pytest.mark.parametrize("name", ["bcrows_poisson_2d.npz", "bcrows_ch_2d.npz"])
def test_bc_row_elimination(name):
    """BCSystem::applyBCs: the oracle's sequential restatement reproduces the
    reference's K and RHS bit for bit, including rows hit by several entries
    and rows whose scale falls back to the row maximum or to 1."""
    from conftest import bcrow_specs
    z = load(name)

    class M:
        n = int(z["n"]); coords = z["coords"]; mirror = z["mirror"]
    bcs = pxo.make_bcs(bcrow_specs(z))
    rc, k, r = pxo.bc_rows(M, int(z["ndof"]), bcs, float(z["t"]), z["u"], z["rowptr"], z["cols"],
                           z["K"], z["R"])
    assert rc == 0
    assert same(k, z["K_bc"]) and same(r, z["R_bc"])
    # the fixture exercises every scale rule: the stored diagonal, the row
    # maximum (zeroed diagonal) and 1 (zeroed row)
    rp, cols, K = z["rowptr"], z["cols"], z["K"]
    kinds = set()
    for g in range(len(rp) - 1):
        s, e = rp[g], rp[g + 1]
        d = s + int(np.searchsorted(cols[s:e], g))
        if np.array_equal(z["K_bc"][s:e], K[s:e]):
            continue
        scale = z["K_bc"][d]
        if scale == K[d] and K[d] != 0.0:
            kinds.add("diag")
        elif scale == np.abs(K[s:e]).max() and scale != 0.0:
            kinds.add("rowmax")
        elif scale == 1.0 and not K[s:e].any():
            kinds.add("unit")
    assert kinds == {"diag", "rowmax", "unit"}


@pytest.mark.parametrize("name", ["diffusion_2d.npz", "diffusion_3d.npz"])
def test_diffusion_assembly(name):
    """DiffusionElement through the assembler: K and the residual bit for bit."""
    z = load(name)
    m = mesh_of(dict(z, factor=3.015))
    ops, _ = pxo.operators(m)
    rp, cols = pxo.csr_pattern(m, 1)
    k, r = pxo.assemble_diffusion(m, ops, rp, cols, float(z["D"]), float(z["f"]), float(z["dt"]),
                                  z["u"], z["uold"])
    assert same(k, z["K"]) and same(r, z["R"])


@pytest.mark.parametrize("name", ["dynfrac_2d.npz", "dynfrac_3d.npz"])
def test_dynamic_fracture_element(name):
    """PDDODynamicFracElement through the assembler over two steps: the
    initialising step (seeded health, initial velocity) and a committing step
    (velocity from u_old, bond failure on u_old); K, the residual, the
    committed velocity and the bond health bit for bit."""
    z = load(name)
    dim = int(z["dim"])
    m = mesh_of(dict(z, factor=3.015))
    ops, _ = pxo.operators(m)
    rp, cols = pxo.csr_pattern(m, dim)
    E, nu, rho, dt = float(z["E"]), float(z["nu"]), float(z["rho"]), float(z["dt"])
    C11, C12, C66 = pxo.elastic_constants(dim, False, E, nu)
    sc = pxo.critical_stretch(float(z["G0"]), E, m.delta, dim, False)
    health = pxo.seed_health(m, segs=[tuple(z["crack"])])
    assert same(health, z["health1"]), "force-only seeded health"
    vel = np.tile(z["v0"], m.n)
    uprev = z["uold1"].copy()
    assert same(vel, z["vel1"])
    k, r = pxo.assemble_dynfrac(m, ops, rp, cols, C11, C12, C66, rho, z["b"], dt, vel, health,
                                z["u1"], z["uold1"])
    assert same(k, z["K1"]) and same(r, z["R1"]), "step 1"
    pxo.dynfrac_commit(m, sc, True, True, dt, z["uold2"], uprev, vel, health)
    assert same(vel, z["vel2"]) and same(health, z["health2"]), "commit"
    assert (health == 0).sum() > (z["health1"] == 0).sum()
    k, r = pxo.assemble_dynfrac(m, ops, rp, cols, C11, C12, C66, rho, z["b"], dt, vel, health,
                                z["u2"], z["uold2"])
    assert same(k, z["K2"]) and same(r, z["R2"]), "step 2"


@pytest.mark.parametrize("name", ["stressch_2d.npz", "stressch_2d_ps.npz", "stressch_3d.npz"])
def test_stress_cahn_hilliard_assembly(name):
    """StressCahnHilliardElement through the assembler (plane strain, plane
    stress, 3D; concentrations also below and above the log barrier and the
    mobility clamp): the element constants, K and the residual bit for bit."""
    z = load(name)
    dim = int(z["dim"])
    m = mesh_of(dict(z, factor=3.015))
    ops, _ = pxo.operators(m)
    nd = dim + 2
    rp, cols = pxo.csr_pattern(m, nd)
    ref = z["consts"]
    E, nu, Omega = 70e9, 0.3, float(ref[9])
    kc = pxo.stressch_consts(dim, str(z["state"]) == "plane_stress", E, nu, Omega)
    assert same(kc, ref[:6])
    lap, rev, isg = pxo.ch_geometry(m, ops)
    k, r = pxo.assemble_stressch(m, ops, lap, rev, isg, ref, float(z["dt"]), rp, cols, z["u"], z["uold"])
    assert same(k, z["K"]) and same(r, z["R"])
    # the barrier: outside [1e-3, 1 - 1e-3] the residual's nodal part is the
    # tangent line of ln(c/(1-c)), so the mu rows stay finite for c <= 0, c >= 1
    c = z["u"][0::nd]
    assert (c < 0).any() and (c > 1).any() and np.isfinite(r).all()


@pytest.mark.parametrize("name", ["fsch_2d.npz", "fsch_3d.npz"])
def test_fracture_stress_cahn_hilliard_element(name):
    """FracStressCahnHilliardElement through the assembler over an
    initialising step and a committing step (2D plane stress with inertia
    and a residual stiffness of 1e-3; 3D quasi-static with tension and
    compression failure): constants, species Laplacian, committed velocity,
    bond health, K and the residual bit for bit."""
    z = load(name)
    dim = int(z["dim"])
    nd = dim + 2
    m = mesh_of(dict(z, factor=3.015))
    ops, _ = pxo.operators(m)
    rp, cols = pxo.csr_pattern(m, nd)
    kc = z["consts"]
    ps = dim == 2 and str(z["state"]) == "plane_stress"
    E, nu = float(z["E"]), float(z["nu"])
    assert same(pxo.stressch_consts(dim, ps, E, nu, float(kc[9])), kc[:6])
    sc = pxo.critical_stretch(float(z["G0"]), E, m.delta, dim, ps)
    assert sc == kc[13]
    lapc = pxo.fsch_lapc(m, ops)
    assert same(lapc, z["lapc"])
    _, rev, isg = pxo.ch_geometry(m, ops)
    rho = float(kc[11])
    health = pxo.seed_health(m, segs=[tuple(z["crack"])])
    assert same(health, z["health1"]), "force-only seeded health"
    vel = np.zeros(m.n * dim)
    k, r = pxo.assemble_fsch(m, ops, lapc, rev, isg, kc[:13], float(z["dt1"]), vel, health, rp, cols,
                             z["u1"], z["uold1"])
    assert same(k, z["K1"]) and same(r, z["R1"]), "step 1"
    uprev = np.ascontiguousarray(z["uold1"].reshape(-1, nd)[:, 2:].ravel())
    pxo.fsch_commit(m, sc, bool(z["tension_only"]), True, float(kc[9]), float(kc[10]), rho, float(z["dt1"]),
                    z["uold2"], uprev, vel, health)
    if rho > 0:
        assert same(vel, z["vel2"])
    assert same(health, z["health2"]), "commit"
    assert (health == 0).sum() > (z["health1"] == 0).sum()
    k, r = pxo.assemble_fsch(m, ops, lapc, rev, isg, kc[:13], float(z["dt2"]), vel, health, rp, cols,
                             z["u2"], z["uold2"])
    assert same(k, z["K2"]) and same(r, z["R2"]), "step 2"


@pytest.mark.parametrize("name", ["stressch_2d.npz", "stressch_2d_ps.npz", "stressch_3d.npz",
                                  "fsch_2d.npz", "fsch_3d.npz"])
def test_stress_cahn_hilliard_projections(name):
    """The chemo-mechanical elements' computeNodalProjection (strain, stress,
    von Mises, sigma_h, f', f'', grad c; damage and DamageI after the commit
    of the fracture element) bit for bit."""
    z = load(name)
    dim = int(z["dim"])
    m = mesh_of(dict(z, factor=3.015))
    ops, _ = pxo.operators(m)
    k = z["consts"]
    consts = [k[0], k[1], k[2], k[3], k[10], k[8]]
    frac = name.startswith("fsch")
    plane_strain = dim == 3 or str(z["state"]) != "plane_stress"
    u = z["u2"] if frac else z["u"]
    names = [n[len("proj_"):] for n in z.files if n.startswith("proj_")]
    assert len(names) == 7
    for nm in names:
        o = pxo.sch_project(m, ops, consts, frac, plane_strain, u, nm,
                            health=z["health2"] if frac else None, seed=z["health1"] if frac else None)
        assert same(o, z["proj_" + nm]), nm
        if nm in ("damage", "DamageI"):
            assert o.max() > 0, nm
