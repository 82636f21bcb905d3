"""Writes the golden fixtures of tests/golden/ from the reference's own build.

Run in the development container, where oracle/ref/Makefile has compiled the
reference's C++ sources into oracle/_ref/libperix_ref.so:

    make -C oracle/ref && python tests/golden/make_golden.py

Every fixture holds inputs (the JSON deck parameters) and the reference's
outputs, nothing else.  The tests compare the C oracle and the HIP library
against these arrays, so neither needs the reference at test time.

Fixtures (all small, .npz):
  mesh_fam_2d.npz   20x12 grid on [0,1]x[0,0.6]: cloud, horizon, families,
                    g-weights, CSR, Poisson K/residual for a seeded field
  mesh_fam_3d.npz   6x5x4 grid on [0,0.6]x[0,0.5]x[0,0.4]: same, 10-term basis
  explicit_2d.npz   40x20 plate, two force-only notches, 16.5 m/s impactor,
                    400 FP64 steps: final u, u_prev, bond health and the
                    element's damage / von Mises / stress / strain projections
  explicit_3d.npz   16x8x4 block, one notch, impactor, 200 FP64 steps
  ch_2d.npz         16x16 Cahn-Hilliard: RandomIC + Neumann presets, CSR K and
                    residual for a perturbed state (Picard-frozen mobility)
  diffusion_*.npz   DiffusionElement assembly (D = 0.7, f = 0.3, dt = 2e-3) of
                    the mesh_fam grids for a seeded (u, u_old)
  dynfrac_*.npz     PDDODynamicFracElement: two backward-Euler steps (the
                    initialising step and a committing step with bond failure)
  stressch_*.npz    StressCahnHilliardElement assembly (c, mu, u): plane
                    strain, plane stress and 3D, concentrations also outside
                    the log barrier and the mobility clamp, with the
                    element's output projections of that state
  fsch_*.npz        FracStressCahnHilliardElement over an initialising and a
                    committing step (2D plane stress with inertia and a
                    residual stiffness; 3D quasi-static, tension and
                    compression failure), projections after the commit
  vtu_stressch_2d.npz  the reference's VTU output (savePDResults) of the
                    stressch_2d state with every projection, parsed
  bcrows_*_2d.npz   BC row elimination (BCSystem::applyBCs) of an assembled
                    Poisson (20x12) / Cahn-Hilliard (16x16) system under an
                    overlapping Dirichlet/Neumann list, with zeroed diagonals
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import refpy  # noqa: E402

E_STEEL, NU, RHO = 191e9, 0.25, 8000.0


def deck_poisson(dim, nx, ny, nz, hi):
    return {"Mesh": refpy.mesh_block(dim, nx, ny, nz, hi=hi),
            "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
            "DOFs": ["u"],
            "ElmtSystem": {"p": {"type": "poisson", "params": {"sigma": 1.3, "f": 0.7}}},
            "BCSystem": {"l": {"type": "dirichlet", "phygroup": "leftnodes_ghost", "dofs": ["u"],
                               "value": 0.0}},
            "JobSystem": {"type": "static", "assemble": "serial"}}


def mesh_fixture(name, dim, nx, ny, nz, hi):
    c = refpy.Case(deck_poisson(dim, nx, ny, nz, hi))
    coords, vols, mirror = c.mesh()
    off, nbr = c.families()
    nop = 9 if dim == 3 else 5
    ops, flags = c.ops(nop)
    rowptr, cols = c.csr()
    u = np.random.RandomState(7).rand(c.n)
    k, r = c.assemble(u, u, 0.0)
    np.savez_compressed(os.path.join(HERE, name), dim=dim, grid=np.array([nx, ny, nz]),
                        hi=np.array(hi, float), factor=3.015, sigma=1.3, f=0.7,
                        n=c.n, nbulk=c.nbulk, delta=c.delta, dxyz=c.dxyz(), coords=coords,
                        vols=vols, mirror=mirror, off=off, nbr=nbr, ops=ops, flags=flags,
                        rowptr=rowptr, cols=cols, u=u, K=k, R=r)
    c.close()


def deck_frac(dim, nx, ny, nz, hi, G0, dt, total, cracks, box):
    d = {"Mesh": refpy.mesh_block(dim, nx, ny, nz, hi=hi),
         "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
         "MeshModify": {"type": "pre_existing_cracks", "treatment": "force_only",
                        "Cracks": [{"x1": q[0], "y1": q[1], "x2": q[2], "y2": q[3]} for q in cracks]},
         "DOFs": ["ux", "uy"] + (["uz"] if dim == 3 else []),
         "ElmtSystem": {"f": {"type": "explicit_pddo_frac",
                              "params": {"E": E_STEEL, "nu": NU, "rho": RHO, "G0": G0,
                                         "tension_only": True, "damage": True}}},
         "BCSystem": {}, "JobSystem": {"type": "transient", "assemble": "openmp"},
         "TimeStepping": {"dt": dt, "total_time": total}}
    if dim == 2:
        d["ElmtSystem"]["f"]["params"]["state"] = "plane_strain"
    walls = ["left", "right", "bottom", "top"] + (["back", "front"] if dim == 3 else [])
    dofs = d["DOFs"]
    for w in walls:
        d["BCSystem"]["n_" + w] = {"type": "neumann", "phygroup": w + "nodes_ghost", "dofs": dofs,
                                   "value": [0.0] * len(dofs)}
    d["BCSystem"]["imp"] = {"type": "dirichlet", "phygroup": "bottomnodes_ghost", "dofs": dofs,
                            "value": [0.0] * len(dofs),
                            "velocity": [0.0, 16.5] + ([0.0] if dim == 3 else []), "direct": True,
                            "box": {"xmin": box[0], "xmax": box[1]}}
    return d


def explicit_fixture(name, dim, nx, ny, nz, hi, cracks, box, nsteps, sc_target=0.003):
    # G0 from the target stretch (FractureCriterion.h:30-43 inverted); the
    # reference recomputes sc from G0, and the fixture records both
    c0 = refpy.Case(deck_poisson(dim, nx, ny, nz, hi))
    delta, dxyz = c0.delta, c0.dxyz()
    c0.close()
    G0 = sc_target ** 2 * (12 * E_STEEL * delta / (5 * np.pi) if dim == 2 else 6 * E_STEEL * delta / 5)
    c11 = E_STEEL * (1 - NU) / ((1 + NU) * (1 - 2 * NU))
    cfl = float(max(dxyz)) / np.sqrt(c11 / RHO)
    dt = 0.5 * cfl
    total = nsteps * dt
    c = refpy.Case(deck_frac(dim, nx, ny, nz, hi, G0, dt, total, cracks, box))
    steps, u, uold, health, opc, sc = c.explicit()
    proj = {"proj_" + nm: c.projection(nm) for nm in ("damage", "vonMises", "stress", "strain")}
    np.savez_compressed(os.path.join(HERE, name), dim=dim, grid=np.array([nx, ny, nz]),
                        hi=np.array(hi, float), cracks=np.array(cracks, float),
                        box=np.array(box, float), E=E_STEEL, nu=NU, rho=RHO, G0=G0, sc=sc, dt=dt,
                        total=total, nsteps=nsteps, steps=steps, u=u, uold=uold,
                        health=health.astype(np.uint8), **proj)
    c.close()


def ch_fixture(name, nx):
    kappa = 1e-2 * (1.0 / nx) ** 2
    d = {"Mesh": refpy.mesh_block(2, nx, nx),
         "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2}, "DOFs": ["c", "mu"],
         "ElmtSystem": {"ch": {"type": "cahnhilliard", "params": {"chi": 3.0, "kappa": kappa, "M": 1.0}}},
         "BCSystem": {w: {"type": "neumann", "phygroup": w + "nodes_ghost", "dofs": ["c", "mu"],
                          "value": [0.0, 0.0]} for w in ["left", "right", "bottom", "top"]},
         "ICSystem": {"noise": {"type": "random", "dof": "c", "min": 0.49, "max": 0.51, "seed": 12345},
                      "mu0": {"type": "constant", "dof": "mu", "value": 0.0}},
         "JobSystem": {"type": "transient", "assemble": "serial"},
         "TimeStepping": {"dt": 1e-3, "total_time": 2e-2}}
    c = refpy.Case(d)
    u0, _ = c.initial_state()
    u1 = u0.copy()
    u1[1::2] = np.random.RandomState(3).rand(c.n) * 0.1
    u1[0::2] += 1e-3
    k, r = c.assemble(u1, u0, 1e-3)
    lap, rev, mnode = c.ch_geometry()
    np.savez_compressed(os.path.join(HERE, name), nx=nx, kappa=kappa, chi=3.0, M0=1.0, dt=1e-3,
                        seed=12345, u0=u0, u1=u1, K=k, R=r, lap=lap, rev=rev, mnode=mnode)
    c.close()


def diffusion_fixture(name, dim, nx, ny, nz, hi):
    """DiffusionElement assembly (backward Euler) for a seeded (u, uold)."""
    d = {"Mesh": refpy.mesh_block(dim, nx, ny, nz, hi=hi),
         "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
         "DOFs": ["c"],
         "ElmtSystem": {"d": {"type": "diffusion", "params": {"D": 0.7, "f": 0.3}}},
         "BCSystem": {"l": {"type": "dirichlet", "phygroup": "leftnodes_ghost", "dofs": ["c"],
                            "value": 0.0}},
         "JobSystem": {"type": "transient", "assemble": "serial"},
         "TimeStepping": {"dt": 2e-3, "total_time": 2e-2}}
    c = refpy.Case(d)
    rs = np.random.RandomState(9)
    u = rs.rand(c.n); uold = u + 0.01 * rs.standard_normal(c.n)
    k, r = c.assemble(u, uold, 2e-3)
    np.savez_compressed(os.path.join(HERE, name), dim=dim, grid=np.array([nx, ny, nz]),
                        hi=np.array(hi, float), D=0.7, f=0.3, dt=2e-3, u=u, uold=uold, K=k, R=r)
    c.close()


def dynfrac_fixture(name, dim, nx, ny, nz, hi, crack, amp):
    """PDDODynamicFracElement: two backward-Euler steps through the assembler.
    Step 1 initialises the element (velocity v0, force-only seeded health);
    step 2 commits (velocity from u_old, bond failure on u_old) and assembles."""
    dofs = ["ux", "uy"] + (["uz"] if dim == 3 else [])
    G0 = 50.0
    d = {"Mesh": refpy.mesh_block(dim, nx, ny, nz, hi=hi),
         "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
         "MeshModify": {"type": "pre_existing_cracks", "treatment": "force_only",
                        "Cracks": [{"x1": crack[0], "y1": crack[1], "x2": crack[2], "y2": crack[3]}]},
         "DOFs": dofs,
         "ElmtSystem": {"f": {"type": "pddo_dynamic_frac",
                              "params": {"E": E_STEEL, "nu": NU, "rho": RHO, "G0": G0,
                                         "tension_only": True, "damage": True}}},
         "BCSystem": {"l": {"type": "dirichlet", "phygroup": "leftnodes_ghost", "dofs": dofs,
                            "value": [0.0] * dim}},
         "JobSystem": {"type": "transient", "assemble": "serial"},
         "TimeStepping": {"dt": 1e-7, "total_time": 1e-6}}
    if dim == 2:
        d["ElmtSystem"]["f"]["params"]["state"] = "plane_strain"
    c = refpy.Case(d)
    b, v0, dt = [1e5, -2e5, 3e4], [0.3, -0.1, 0.2], 1e-7
    c.dynfrac_set(b, v0)
    rs = np.random.RandomState(1)
    nd = c.n * dim
    uold1 = 1e-6 * rs.standard_normal(nd); u1 = uold1 + 1e-7 * rs.standard_normal(nd)
    k1, r1 = c.assemble_step(u1, uold1, dt, 1)
    h1, v1 = c.dynfrac_state()
    uold2 = uold1 + amp * rs.standard_normal(nd); u2 = uold2 + 1e-7 * rs.standard_normal(nd)
    k2, r2 = c.assemble_step(u2, uold2, dt, 2)
    h2, v2 = c.dynfrac_state()
    np.savez_compressed(os.path.join(HERE, name), dim=dim, grid=np.array([nx, ny, nz]),
                        hi=np.array(hi, float), crack=np.array(crack, float), E=E_STEEL, nu=NU, rho=RHO,
                        G0=G0, b=np.array(b[:dim]), v0=np.array(v0[:dim]), dt=dt,
                        uold1=uold1, u1=u1, K1=k1, R1=r1, health1=h1, vel1=v1,
                        uold2=uold2, u2=u2, K2=k2, R2=r2, health2=h2, vel2=v2)
    c.close()
    return int((h1 == 0).sum()), int((h2 == 0).sum()), len(h2)


# the chemo-mechanical elements' projections and their component counts
# (getAvailableProjections, StressCahnHilliardElement.cpp:469-479,
# FracStressCahnHilliardElement.cpp:563-573)
SCH_PROJECTIONS = [("strain", 9), ("stress", 9), ("vonMises", 1), ("sigmaH", 1), ("fPrime", 1),
                   ("fDoublePrime", 1), ("gradC", 3)]
FSCH_PROJECTIONS = [("strain", 9), ("stress", 9), ("vonMises", 1), ("sigmaH", 1), ("gradC", 3),
                    ("damage", 1), ("DamageI", 1)]


def stressch_fixture(name, dim, nx, ny, nz, hi, state="plane_strain"):
    """StressCahnHilliardElement assembly (c, mu, u) for a seeded state whose
    concentrations also visit the regularised-log and mobility-clamp ranges."""
    dofs = ["c", "mu", "ux", "uy"] + (["uz"] if dim == 3 else [])
    params = {"E": 70e9, "nu": 0.3, "D": 1e-3, "Omega": 3.5e-6, "cref": 0.5, "chi": 2.5, "kappa": 2e-4}
    if dim == 2:
        params["state"] = state
    d = {"Mesh": refpy.mesh_block(dim, nx, ny, nz, hi=hi),
         "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
         "DOFs": dofs,
         "ElmtSystem": {"s": {"type": "stress_cahnhilliard", "params": params}},
         "BCSystem": {"l": {"type": "neumann", "phygroup": "leftnodes_ghost", "dofs": ["c", "mu"],
                            "value": [0.0, 0.0]}},
         "JobSystem": {"type": "transient", "assemble": "serial"},
         "TimeStepping": {"dt": 1e-3, "total_time": 1e-2}}
    c = refpy.Case(d)
    nd = dim + 2
    rs = np.random.RandomState(6)
    u = np.zeros(c.n * nd)
    u[0::nd] = 0.2 + 0.6 * rs.rand(c.n)
    pick = rs.choice(c.n, 12, replace=False)
    u[pick[:3] * nd] = [-0.01, 5e-4, 1e-9]             # below the barrier / clamp
    u[pick[3:6] * nd] = [1.02, 1 - 5e-4, 1 - 1e-9]     # above
    u[1::nd] = 0.1 * rs.standard_normal(c.n)
    for a in range(dim):
        u[2 + a::nd] = 1e-6 * rs.standard_normal(c.n)
    uold = u.copy()
    uold[0::nd] -= 0.01 * rs.rand(c.n)
    k, r = c.assemble(u, uold, 1e-3)
    consts = c.stressch_consts()
    proj = {"proj_" + nm: c.projection_u(nm, u, nc) for nm, nc in SCH_PROJECTIONS}
    np.savez_compressed(os.path.join(HERE, name), dim=dim, grid=np.array([nx, ny, nz]),
                        hi=np.array(hi, float), state=state, consts=consts, dt=1e-3, u=u, uold=uold,
                        K=k, R=r, **proj)
    c.close()


def fsch_fixture(name, dim, nx, ny, nz, hi, crack, rho, tension_only, resid=None, state="plane_stress"):
    """FracStressCahnHilliardElement: two steps through the assembler (step
    sizes 1e-3 then 5e-4).  Step 1 initialises the element (force-only seeded
    health, zero velocity, u_prev = u_old); step 2 commits on u_old (velocity
    when rho > 0, bond failure on the swelling-corrected stretch) and
    assembles.  Concentrations also visit the mobility / log clamp range."""
    from perix_amd import problems
    nd = dim + 2
    dofs = ["c", "mu", "ux", "uy"] + (["uz"] if dim == 3 else [])
    E, nu, sc = 87.9336, 0.29, 2e-3
    dx = hi[0] / nx
    delta = 3.015 * dx
    G0 = problems.g0_for_stretch(sc, E, delta, dim, plane_stress=(dim == 2 and state == "plane_stress"))
    params = {"E": E, "nu": nu, "D": 1e-2, "Omega": 0.01, "cref": 0.5, "chi": 2.6, "kappa": 2e-4,
              "rho": rho, "Gc": G0, "tension_only": tension_only, "damage_on": True}
    if resid is not None:
        params["residual_stiffness"] = resid
    if dim == 2:
        params["state"] = state
    d = {"Mesh": refpy.mesh_block(dim, nx, ny, nz, hi=hi),
         "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
         "MeshModify": {"type": "pre_existing_cracks", "treatment": "force_only",
                        "Cracks": [{"x1": crack[0], "y1": crack[1], "x2": crack[2], "y2": crack[3]}]},
         "DOFs": dofs,
         "ElmtSystem": {"s": {"type": "frac_stress_cahnhilliard", "params": params}},
         "BCSystem": {"l": {"type": "neumann", "phygroup": "leftnodes_ghost", "dofs": ["c", "mu"],
                            "value": [0.0, 0.0]}},
         "JobSystem": {"type": "transient", "assemble": "serial"},
         "TimeStepping": {"dt": 1e-3, "total_time": 1e-2}}
    c = refpy.Case(d)
    rs = np.random.RandomState(11)
    n = c.n

    def state_pair(amp_u):
        u = np.zeros(n * nd)
        u[0::nd] = 0.2 + 0.6 * rs.rand(n)
        pick = rs.choice(n, 6, replace=False)
        u[pick * nd] = [-0.02, 1e-9, 0.0, 1.0, 1 - 1e-9, 1.03]
        u[1::nd] = 0.1 * rs.standard_normal(n)
        for a in range(dim):
            u[2 + a::nd] = amp_u * rs.standard_normal(n)
        return u

    uold1 = state_pair(1e-5 * dx)
    u1 = uold1 + 1e-6 * rs.standard_normal(n * nd)
    dt1, dt2 = 1e-3, 5e-4
    k1, r1 = c.assemble_step(u1, uold1, dt1, 1)
    h1, v1, lapc, kc = c.fsch_state(dim)
    uold2 = state_pair(0.6 * sc * dx)
    u2 = uold2 + 1e-6 * rs.standard_normal(n * nd)
    k2, r2 = c.assemble_step(u2, uold2, dt2, 2)
    h2, v2, _, _ = c.fsch_state(dim)
    proj = {"proj_" + nm: c.projection_u(nm, u2, nc) for nm, nc in FSCH_PROJECTIONS}
    np.savez_compressed(os.path.join(HERE, name), **proj, dim=dim, grid=np.array([nx, ny, nz]),
                        hi=np.array(hi, float), crack=np.array(crack, float), state=state,
                        E=E, nu=nu, G0=G0, consts=kc, tension_only=int(tension_only),
                        dt1=dt1, dt2=dt2, lapc=lapc,
                        uold1=uold1, u1=u1, K1=k1, R1=r1, health1=h1, vel1=v1,
                        uold2=uold2, u2=u2, K2=k2, R2=r2, health2=h2, vel2=v2)
    c.close()
    return int((h1 == 0).sum()), int((h2 == 0).sum()), len(h2)


def vtu_fixture(name):
    """PDProblem::savePDResults (the reference's ASCII VTU) of the plane-strain
    stress-coupled Cahn-Hilliard state of stressch_2d.npz with every projection
    the element lists; stored as the parsed arrays in file order."""
    import tempfile
    sys.path.insert(0, os.path.dirname(HERE))
    from vtu_io import read_vtu
    z = np.load(os.path.join(HERE, "stressch_2d.npz"))
    params = {"E": 70e9, "nu": 0.3, "D": 1e-3, "Omega": 3.5e-6, "cref": 0.5, "chi": 2.5, "kappa": 2e-4,
              "state": "plane_strain"}
    d = {"Mesh": refpy.mesh_block(2, 14, 10, 1, hi=(1.4, 1.0, 1.0)),
         "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
         "DOFs": ["c", "mu", "ux", "uy"],
         "ElmtSystem": {"s": {"type": "stress_cahnhilliard", "params": params}},
         "BCSystem": {"l": {"type": "neumann", "phygroup": "leftnodes_ghost", "dofs": ["c", "mu"],
                            "value": [0.0, 0.0]}},
         "JobSystem": {"type": "transient", "assemble": "serial"},
         "TimeStepping": {"dt": 1e-3, "total_time": 1e-2},
         "OutputSystem": {"interval": 1, "Fields": [nm for nm, _ in SCH_PROJECTIONS]}}
    c = refpy.Case(d)
    stem = os.path.join(tempfile.mkdtemp(prefix="pxvtu_"), "case")
    path = c.save_vtu(z["u"], stem, 3)
    v = read_vtu(path)
    c.close()
    names = [a[0] for a in v["point_data"]]
    comps = [a[1] for a in v["point_data"]]
    arrays = {"pd_%d" % k: a[2] for k, a in enumerate(v["point_data"])}
    np.savez_compressed(os.path.join(HERE, name), names=np.array(names), comps=np.array(comps),
                        vtu_name=os.path.basename(path), points=v["points"],
                        connectivity=v["cells"][0][2], offsets=v["cells"][1][2], types=v["cells"][2][2],
                        u=z["u"], **arrays)


# BC lists of the row-elimination fixtures: (name, type, group, dofs, value,
# velocity, direct, box) in deck order; overlapping entries on one row
# exercise the list order
BCROW_SPECS = {
    "poisson": (["u"], [
        ("l", "dirichlet", "left", ["u"], [0.3], [0.5], False, None),
        ("r", "dirichlet", "right", ["u"], [-0.4], None, True, {"ymin": 0.1, "ymax": 0.45}),
        ("b", "neumann", "bottom", ["u"], [0.0], None, False, None),
        ("t", "dirichlet", "top", ["u"], [0.25], None, False, {"xmin": 0.2, "xmax": 0.7}),
        ("t2", "neumann", "top", ["u"], [0.0], None, False, None),
        ("t3", "dirichlet", "top", ["u"], [0.6], [-1.5], False, {"xmin": 0.5, "xmax": 0.9}),
    ]),
    "ch": (["c", "mu"], [
        # the reference rejects Dirichlet rows on Cahn-Hilliard ghosts
        # (BCSystem.cpp:61-112): Neumann rows only, per component
        ("lm", "neumann", "left", ["mu"], [0.0], None, False, None),
        ("lc", "neumann", "left", ["c"], [0.0], None, False, None),
        ("b", "neumann", "bottom", ["c", "mu"], [0.0, 0.0], None, False, None),
        ("rc", "neumann", "right", ["c"], [0.0], None, False, None),
        ("t", "neumann", "top", ["mu"], [0.0], None, False, None),
        ("t2", "neumann", "top", ["c", "mu"], [0.0, 0.0], None, False, None),
    ]),
}


def bcrows_fixture(name, kind):
    """K, R of one assembly, then BCSystem::applyBCs at t = 0.2.  Some
    constrained rows get a zero diagonal or are zeroed whole beforehand, so
    that the row-maximum and unit fallbacks of the row scale are taken."""
    dofs, specs = BCROW_SPECS[kind]
    if kind == "poisson":
        d = deck_poisson(2, 20, 12, 1, (1.0, 0.6, 1.0))
    else:
        d = {"Mesh": refpy.mesh_block(2, 16, 16),
             "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2}, "DOFs": dofs,
             "ElmtSystem": {"ch": {"type": "cahnhilliard",
                                   "params": {"chi": 3.0, "kappa": 1e-2 / 256, "M": 1.0}}},
             "JobSystem": {"type": "transient", "assemble": "serial"},
             "TimeStepping": {"dt": 1e-3, "total_time": 2e-2}}
    d["BCSystem"] = {}
    for nm, typ, grp, dd, val, vel, direct, box in specs:
        e = {"type": typ, "phygroup": grp + "nodes_ghost", "dofs": dd,
             "value": val if len(val) > 1 else val[0]}
        if vel is not None:
            e["velocity"] = vel if len(vel) > 1 else vel[0]
        if direct:
            e["direct"] = True
        if box is not None:
            e["box"] = box
        d["BCSystem"][nm] = e
    c = refpy.Case(d)
    coords, vols, mirror = c.mesh()
    rowptr, cols = c.csr()
    rs = np.random.RandomState(5)
    u = rs.rand(c.n * c.ndof)
    k, r = c.assemble(u, u, 1e-3 if kind == "ch" else 0.0)
    groups = {grp: c.group(grp + "nodes_ghost") for _, _, grp, *_ in specs}
    # zero the diagonal of a few constrained rows and a whole row of others
    rows = sorted({(i - 1) * c.ndof for g in groups.values() for i in g})
    pick = rs.choice(len(rows), 8, replace=False)
    for q in pick[:4]:
        g = rows[q]
        k[rowptr[g] + np.searchsorted(cols[rowptr[g]:rowptr[g + 1]], g)] = 0.0
    for q in pick[4:]:
        g = rows[q]
        k[rowptr[g]:rowptr[g + 1]] = 0.0
    t = 0.2
    k2, r2 = c.apply_bcs(u, t, k, r)
    spec_arr = np.array([(nm, typ, grp, ",".join(dd), ",".join(map(repr, val)),
                          ",".join(map(repr, vel)) if vel else "", int(direct),
                          repr(sorted(box.items())) if box else "")
                         for nm, typ, grp, dd, val, vel, direct, box in specs])
    np.savez_compressed(os.path.join(HERE, name), kind=kind, ndof=c.ndof, dofs=np.array(dofs),
                        n=c.n, coords=coords, vols=vols, mirror=mirror, delta=c.delta,
                        dxyz=c.dxyz(), rowptr=rowptr, cols=cols, u=u, t=t,
                        K=k, R=r, K_bc=k2, R_bc=r2, specs=spec_arr,
                        **{"group_" + g: ids for g, ids in groups.items()})
    c.close()


def main():
    if not refpy.available():
        raise SystemExit("oracle/_ref/libperix_ref.so is missing: run `make -C oracle/ref` first")
    mesh_fixture("mesh_fam_2d.npz", 2, 20, 12, 1, (1.0, 0.6, 1.0))
    mesh_fixture("mesh_fam_3d.npz", 3, 6, 5, 4, (0.6, 0.5, 0.4))
    explicit_fixture("explicit_2d.npz", 2, 40, 20, 1, (0.02, 0.01, 1.0),
                     [(0.005, 0.0, 0.005, 0.005), (0.015, 0.0, 0.015, 0.005)],
                     [0.0075, 0.0125, -1.0, 1.0, -1.0, 1.0], 400)
    explicit_fixture("explicit_3d.npz", 3, 16, 8, 4, (0.016, 0.008, 0.004),
                     [(0.008, 0.0, 0.008, 0.0008)], [0.004, 0.006, -1.0, 1.0, -1.0, 1.0], 200)
    ch_fixture("ch_2d.npz", 16)
    diffusion_fixture("diffusion_2d.npz", 2, 20, 12, 1, (1.0, 0.6, 1.0))
    diffusion_fixture("diffusion_3d.npz", 3, 6, 5, 4, (0.6, 0.5, 0.4))
    dynfrac_fixture("dynfrac_2d.npz", 2, 24, 12, 1, (0.024, 0.012, 1.0), (0.012, 0.0, 0.012, 0.005), 1.5e-7)
    dynfrac_fixture("dynfrac_3d.npz", 3, 6, 4, 3, (0.006, 0.004, 0.003), (0.003, 0.0, 0.003, 0.002), 1e-7)
    stressch_fixture("stressch_2d.npz", 2, 14, 10, 1, (1.4, 1.0, 1.0))
    stressch_fixture("stressch_2d_ps.npz", 2, 10, 8, 1, (1.0, 0.8, 1.0), state="plane_stress")
    stressch_fixture("stressch_3d.npz", 3, 5, 4, 3, (0.5, 0.4, 0.3))
    print("fsch_2d", fsch_fixture("fsch_2d.npz", 2, 14, 10, 1, (1.4, 1.0, 1.0), (0.7, 0.0, 0.7, 0.4),
                                  rho=2.0, tension_only=True, resid=1e-3))
    print("fsch_3d", fsch_fixture("fsch_3d.npz", 3, 6, 4, 3, (0.6, 0.4, 0.3), (0.3, 0.0, 0.3, 0.15),
                                  rho=0.0, tension_only=False))
    vtu_fixture("vtu_stressch_2d.npz")
    bcrows_fixture("bcrows_poisson_2d.npz", "poisson")
    bcrows_fixture("bcrows_ch_2d.npz", "ch")
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print("%-20s %8d bytes" % (f, os.path.getsize(os.path.join(HERE, f))))


if __name__ == "__main__":
    main()
