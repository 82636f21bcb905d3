"""Problem set-ups of the benchmark configurations, built on the device.

Each builder follows the reference's start-up sequence (PDProblem::init,
PDProblem.cpp:21-162): point cloud -> families -> operators -> element
state -> BC presets, with the geometry of the corresponding deck.
"""
import math
from dataclasses import dataclass

import numpy as np

from . import pxg
from .mesh import StructuredGrid, build_structured_cloud, horizon_radius

E_STEEL = 191.0e9
NU = 0.25
RHO = 8000.0


def elastic_constants(dim, plane_stress, E, nu):
    """ExplicitPDDOFracElement::recalcElasticConstants (ExplicitPDDOFracElement.cpp:37-52)."""
    if E <= 0.0:
        return 0.0, 0.0, 0.0
    if dim >= 3 or not plane_stress:
        d = (1.0 + nu) * (1.0 - 2.0 * nu)
        return E * (1.0 - nu) / d, E * nu / d, E / (2.0 * (1.0 + nu))
    d = 1.0 - nu * nu
    return E / d, E * nu / d, E / (2.0 * (1.0 + nu))


def g0_for_stretch(sc, E, delta, dim, plane_stress=False):
    """Critical release rate that gives critical stretch sc through
    FractureCriterion::s0SqFromGc (FractureCriterion.h:30-43)."""
    if dim >= 3:
        return sc * sc * 6.0 * E * delta / 5.0
    if plane_stress:
        return sc * sc * 9.0 * E * delta / (4.0 * math.pi)
    return sc * sc * 12.0 * E * delta / (5.0 * math.pi)


def critical_stretch(G0, E, delta, dim, plane_stress=False):
    """FractureCriterion::criticalStretchFromGc (FractureCriterion.h:30-43)."""
    if G0 <= 0.0 or E <= 0.0 or delta <= 0.0:
        return 0.0
    if dim >= 3:
        s2 = 5.0 * G0 / (6.0 * E * delta)
    elif plane_stress:
        s2 = 4.0 * math.pi * G0 / (9.0 * E * delta)
    else:
        s2 = 5.0 * math.pi * G0 / (12.0 * E * delta)
    return math.sqrt(s2)


def stable_dt(dim, plane_stress, E, nu, rho, dxyz):
    """ExplicitPDDOFracElement::estimateStableDt (ExplicitPDDOFracElement.cpp:54-77)."""
    dx = min(s for s in dxyz if s > 0.0)
    if dim >= 3 or not plane_stress:
        d = (1.0 + nu) * (1.0 - 2.0 * nu)
        c11 = E * (1.0 - nu) / d
    else:
        c11 = E / (1.0 - nu * nu)
    return dx / math.sqrt(c11 / rho)


@dataclass
class FracSetup:
    ctx: object
    cloud: object
    delta: float
    params: object
    bcs: list
    dt: float
    total: float
    nsteps: int
    cracks: list
    G0: float
    u0: object


def plate_bcs(cloud, dim, box, vel):
    """Neumann mirrors on every ghost wall, then a direct Dirichlet velocity
    on the impact patch of the bottom ghosts (kalthoff_winkler.json:43-78
    pattern; later BC entries override earlier ones, BCSystem.cpp:114-129)."""
    walls = ["left", "right", "bottom", "top"] + (["back", "front"] if dim == 3 else [])
    comps = list(range(dim))
    specs = [{"type": "neumann", "nodes": cloud.group(w), "comps": comps} for w in walls]
    v = [0.0] * dim
    v[1] = vel
    specs.append({"type": "dirichlet", "nodes": cloud.group("bottom"), "comps": comps,
                  "value": [0.0] * dim, "velocity": v, "direct": True, "box": box})
    return specs


def initial_state(cloud, dim, specs):
    """u = 0 then BC presets at t = 0 (PDProblem::run, RunPDProblem.cpp:31-40)."""
    u = np.zeros(cloud.n * dim)
    for s in specs:
        for nid in s["nodes"]:
            i = nid - 1
            if s["type"] == "neumann":
                b = cloud.mirror[i] - 1
                for c in s["comps"]:
                    u[i * dim + c] = u[b * dim + c]
            else:
                box = s.get("box")
                if box is not None:
                    x = cloud.coords[3 * i: 3 * i + 3]
                    if any(box[2 * a] <= box[2 * a + 1] and not (box[2 * a] <= x[a] <= box[2 * a + 1])
                           for a in range(3)):
                        continue
                b = cloud.mirror[i] - 1
                for k, c in enumerate(s["comps"]):
                    g = s["value"][k] + s["velocity"][k] * 0.0
                    u[i * dim + c] = g if (s.get("direct") or b < 0) else 2.0 * g - u[b * dim + c]
    return u


def explicit_plate(dim=2, nx=400, ny=200, nz=1, lo=(0.0, 0.0, 0.0), hi=(0.2, 0.1, 0.04),
                   sc=0.01, cracks=None, box=None, vel=16.5, nsteps=2000, cfl=0.9,
                   precision=pxg.FP64, node_E=None, factor=3.015, device=0, ctx=None):
    """Explicit PDDO fracture plate (2D) or block (3D) with edge notches and an impactor patch.

    Defaults are the 2D plate: 0.2 m x 0.1 m, dx = 0.5 mm, two 50 mm edge
    notches at x = 50 and 150 mm from y = 0 (force-only cracks), a 16.5 m/s
    direct velocity on the bottom ghosts with 75 mm <= x <= 125 mm,
    E = 191 GPa, nu = 0.25, rho = 8000, plane strain, s_c = 0.01."""
    g = StructuredGrid(dim, nx, ny, nz, lo, hi)
    cloud = build_structured_cloud(g)
    delta = horizon_radius(g, factor)
    if cracks is None:
        cracks = [(0.05, 0.0, 0.05, 0.05), (0.15, 0.0, 0.15, 0.05)]
    if box is None:
        box = (0.075, 0.125, 1.0, -1.0, 1.0, -1.0)
    c11, c12, c66 = elastic_constants(dim, False, E_STEEL, NU)
    G0 = g0_for_stretch(sc, E_STEEL, delta, dim)
    sc_eff = critical_stretch(G0, E_STEEL, delta, dim)
    p = pxg.FracParams(dim, c11, c12, c66, RHO, 0.0, 0.0, 0.0, sc_eff, 1, 1)
    dt = cfl * stable_dt(dim, False, E_STEEL, NU, RHO, cloud.dxyz)
    total = nsteps * dt
    if ctx is None:
        ctx = pxg.Context(device)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, dim)
    ctx.set_cracks(segs=cracks)
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.frac_setup(p, precision, node_E=node_E, nu=NU, plane_stress=False)
    specs = plate_bcs(cloud, dim, box, vel)
    ctx.frac_set_bcs(specs)
    u0 = initial_state(cloud, dim, specs)
    return FracSetup(ctx, cloud, delta, p, specs, dt, total, nsteps, cracks, G0, u0)


def plate_400x200(precision=pxg.FP64, nsteps=2000, **kw):
    """The 400 x 200 plate (80,000 cells, 81,200 PD nodes)."""
    return explicit_plate(2, 400, 200, precision=precision, nsteps=nsteps, **kw)


def block_160x80x40(precision=pxg.FP32C, nsteps=1000, **kw):
    """The 160 x 80 x 40 block, dx = 1 mm (512,000 cells, 556,800 PD
    nodes).  An 8 mm deep through-thickness notch at x = 80 mm from y = 0
    (a force-only crack segment, extruded over z as in PDMesh.cpp:125-137)
    and a 16.5 m/s direct velocity on the bottom-face patch
    40 mm <= x <= 60 mm over the full 40 mm thickness."""
    return explicit_plate(3, 160, 80, 40, hi=(0.16, 0.08, 0.04), precision=precision,
                          cracks=[(0.08, 0.0, 0.08, 0.008)],
                          box=(0.04, 0.06, 1.0, -1.0, 1.0, -1.0), nsteps=nsteps, **kw)


def poisson_square(nx=100, ny=100, sigma=1.0, factor=3.015, device=0, ctx=None):
    """Poisson on the unit square, sigma = 1, manufactured f = 2 pi^2 sin(pi x) sin(pi y),
    BC rows of poisson_bcs resolved on the context."""
    g = StructuredGrid(2, nx, ny)
    cloud = build_structured_cloud(g)
    delta = horizon_radius(g, factor)
    if ctx is None:
        ctx = pxg.Context(device)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, 2)
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(1)
    ctx.csr_set_bcs(poisson_bcs(cloud))
    x = cloud.coords[0::3]
    y = cloud.coords[1::3]
    f = 2.0 * math.pi ** 2 * sigma * np.sin(math.pi * x) * np.sin(math.pi * y)
    return ctx, cloud, delta, f


def poisson_bcs(cloud):
    """Poisson BC list: Dirichlet u = 0 by reflection on the four ghost walls."""
    return [{"type": "dirichlet", "nodes": cloud.group(w), "comps": [0], "value": [0.0]}
            for w in ("left", "right", "bottom", "top")]


def cahnhilliard_bcs(cloud):
    """Cahn-Hilliard BC list: Neumann mirrors of c and mu on the four ghost walls."""
    return [{"type": "neumann", "nodes": cloud.group(w), "comps": [0, 1]}
            for w in ("left", "right", "bottom", "top")]


def cahnhilliard_square(nx=256, ny=256, seed=12345, factor=3.015, device=0, ctx=None):
    """Cahn-Hilliard on the 256 x 256 unit square, c0 = 0.5 + U(-0.01, 0.01) by RandomIC
    over all PD nodes in id order (ICSystem.cpp:57-69), mu0 = 0, Neumann
    mirrors of c and mu on all walls (presets applied to the returned state,
    BC rows of cahnhilliard_bcs resolved on the context)."""
    g = StructuredGrid(2, nx, ny)
    cloud = build_structured_cloud(g)
    delta = horizon_radius(g, factor)
    if ctx is None:
        ctx = pxg.Context(device)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, 2)
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(2)
    ctx.ch_setup()
    ctx.csr_set_bcs(cahnhilliard_bcs(cloud))
    u = np.zeros(2 * cloud.n)
    pxg.random_ic(seed, 0.49, 0.51, np.arange(1, cloud.n + 1), 2, [1], u)
    for w in ("left", "right", "bottom", "top"):
        for nid in cloud.group(w):
            b = cloud.mirror[nid - 1] - 1
            u[2 * (nid - 1)] = u[2 * b]
            u[2 * (nid - 1) + 1] = u[2 * b + 1]
    return ctx, cloud, delta, u


# silicon anode parameters of the reference's stress_cahnhilliard example deck
# (examples/silicon.json: plane stress, nondimensional units)
SILICON = dict(E=87.9336, nu=0.29, D=1e-4, Omega=3.0, cref=0.0, chi=2.6, kappa=2.19834e-6)


def stressch_square(nx=256, ny=256, seed=7, factor=3.015, device=0, ctx=None, params=None):
    """Stress-coupled Cahn-Hilliard (c, mu, ux, uy) on the nx x ny unit square
    with the silicon parameters in plane stress; returns the context (CSR with
    4 dofs per node, element set up), the cloud and a seeded (u, uold) pair:
    c = 0.3 + U(0, 0.4), mu = 0.05 N(0, 1), u = 1e-3 N(0, 1), uold with c
    lowered by U(0, 0.01)."""
    g = StructuredGrid(2, nx, ny)
    cloud = build_structured_cloud(g)
    delta = horizon_radius(g, factor)
    if ctx is None:
        ctx = pxg.Context(device)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, 2)
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(4)
    p = dict(SILICON, **(params or {}))
    ctx.stressch_setup(pxg.StressChParams(p["E"], p["nu"], p["D"], p["Omega"], p["cref"], p["chi"],
                                          p["kappa"], 1))
    rs = np.random.RandomState(seed)
    n = cloud.n
    u = np.zeros(4 * n)
    u[0::4] = 0.3 + 0.4 * rs.rand(n)
    u[1::4] = 0.05 * rs.standard_normal(n)
    u[2::4] = 1e-3 * rs.standard_normal(n)
    u[3::4] = 1e-3 * rs.standard_normal(n)
    uold = u.copy()
    uold[0::4] -= 0.01 * rs.rand(n)
    return ctx, cloud, delta, u, uold


# fracture variant: the reference's examples/silicon_frac.json element block
SILICON_FRAC = dict(SILICON, D=0.03, rho=2.41817e-22, Gc=0.10662, tension_only=True, damage_on=False)


def fsch_square(nx=256, ny=256, seed=7, factor=3.015, device=0, ctx=None, params=None):
    """Fracture stress-coupled Cahn-Hilliard on the nx x ny unit square with
    the silicon_frac element parameters (plane stress) and two force-only edge
    notches (x = 0.25 and 0.75 from y = 0 to 0.1); the element is set up from
    u_old and returned with a seeded (u, uold) pair as in stressch_square."""
    g = StructuredGrid(2, nx, ny)
    cloud = build_structured_cloud(g)
    delta = horizon_radius(g, factor)
    if ctx is None:
        ctx = pxg.Context(device)
    ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, 2)
    ctx.set_cracks(segs=[(0.25, 0.0, 0.25, 0.1), (0.75, 0.0, 0.75, 0.1)])
    ctx.build_families(delta)
    ctx.build_operators(delta, float(max(cloud.dxyz)))
    ctx.build_csr(4)
    p = dict(SILICON_FRAC, **(params or {}))
    rs = np.random.RandomState(seed)
    n = cloud.n
    u = np.zeros(4 * n)
    u[0::4] = 0.3 + 0.4 * rs.rand(n)
    u[1::4] = 0.05 * rs.standard_normal(n)
    u[2::4] = 1e-3 * rs.standard_normal(n)
    u[3::4] = 1e-3 * rs.standard_normal(n)
    uold = u.copy()
    uold[0::4] -= 0.01 * rs.rand(n)
    base = pxg.StressChParams(p["E"], p["nu"], p["D"], p["Omega"], p["cref"], p["chi"], p["kappa"], 1)
    sc = critical_stretch(p["Gc"], p["E"], delta, 2, plane_stress=True)
    ctx.fsch_setup(pxg.FschParams(base, p["rho"], sc, p.get("residual_stiffness", 1e-8),
                                  int(p["tension_only"]), int(p["damage_on"])), uold)
    return ctx, cloud, delta, u, uold

