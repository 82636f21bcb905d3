"""ctypes bindings of the C oracle (libpxo.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by the smoke test (to compare one device step against)
and by bench.py's cpu_baseline leg.  The product (perix_amd) never imports it.
Every wrapper returns numpy arrays; node ids inside neighbour lists stay
1-based as in the reference.
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
U8 = C.POINTER(C.c_ubyte)


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
                ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double)]


class Cracks(C.Structure):
    _fields_ = [("nseg", C.c_int), ("seg", D), ("nquad", C.c_int), ("quad", D)]


class Frac(C.Structure):
    _fields_ = [("dim", C.c_int), ("C11", C.c_double), ("C12", C.c_double),
                ("C66", C.c_double), ("rho", C.c_double), ("bx", C.c_double),
                ("by", C.c_double), ("bz", C.c_double), ("sc", C.c_double),
                ("tension_only", C.c_int), ("damage_on", C.c_int),
                ("nu", C.c_double), ("plane_stress", C.c_int)]


class BC(C.Structure):
    _fields_ = [("type", C.c_int), ("direct", C.c_int), ("ncomp", C.c_int),
                ("comp", C.c_int * 3), ("value", C.c_double * 3),
                ("velocity", C.c_double * 3), ("has_box", C.c_int),
                ("box", C.c_double * 6), ("nnodes", C.c_int), ("nodes", I)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libpxo.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        _LIB = C.CDLL(path)
        L = _LIB
        L.pxo_horizon.restype = C.c_double
        L.pxo_families.restype = C.c_longlong
        L.pxo_csr_pattern.restype = C.c_longlong
        L.pxo_critical_stretch.restype = C.c_double
        L.pxo_stable_dt.restype = C.c_double
        L.pxo_frac_adjoint.restype = C.c_double
        L.pxo_frac_loss_frozen.restype = C.c_double
        L.pxo_ldu_profile.restype = C.c_longlong
        L.pxo_random_ic.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_int, I,
                                    C.c_int, C.c_int, I, D]
    return _LIB


def dp(a):
    return a.ctypes.data_as(D) if a is not None else None


def ip(a):
    return a.ctypes.data_as(I) if a is not None else None


def up(a):
    return a.ctypes.data_as(U8) if a is not None else None


def make_grid(dim, nx, ny, nz=1, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    return Grid(dim, nx, ny, nz if dim == 3 else 0, lo[0], hi[0], lo[1], hi[1],
                lo[2] if dim == 3 else 0.0, hi[2] if dim == 3 else 0.0)


class Mesh:
    """Structured PD cloud (a1) + families (a2)."""

    def __init__(self, grid, factor=3.015, families=True):
        L = lib()
        self.grid = grid
        self.dim = grid.dim
        n = C.c_int(); nbk = C.c_int()
        if L.pxo_mesh_sizes(C.byref(grid), C.byref(n), C.byref(nbk)):
            raise ValueError("bad grid")
        self.n, self.nbulk = n.value, nbk.value
        self.coords = np.zeros(3 * self.n)
        self.vols = np.zeros(self.n)
        self.mirror = np.zeros(self.n, np.int32)
        self.wall = np.zeros(self.n, np.int32)
        self.dxyz = np.zeros(3)
        rc = L.pxo_mesh_build(C.byref(grid), dp(self.coords), dp(self.vols), ip(self.mirror),
                              ip(self.wall), dp(self.dxyz))
        if rc:
            raise RuntimeError("pxo_mesh_build failed %d" % rc)
        self.delta = L.pxo_horizon(C.byref(grid), C.c_double(factor))
        self.dx_char = float(max(self.dxyz))
        if families:
            self.off = np.zeros(self.n + 1, np.int32)
            nb = L.pxo_families(self.n, dp(self.coords), int(self.dim == 3),
                                C.c_double(self.delta), ip(self.off), None)
            self.nbr = np.zeros(nb, np.int32)
            L.pxo_families(self.n, dp(self.coords), int(self.dim == 3), C.c_double(self.delta),
                           ip(self.off), ip(self.nbr))
            self.nb = nb

    def group(self, wall):
        return (np.nonzero(self.wall == wall)[0] + 1).astype(np.int32)


WALLS = {"left": 0, "right": 1, "bottom": 2, "top": 3, "back": 4, "front": 5}


def operators(m, order=2, vol_corr=True, nthreads=0):
    L = lib()
    nop = L.pxo_nop(m.dim, order)
    ops = np.zeros(m.nb * nop)
    flags = np.zeros(m.n, np.uint8)
    L.pxo_operators(m.n, m.dim, order, dp(m.coords), dp(m.vols), ip(m.off), ip(m.nbr),
                    C.c_double(m.delta), C.c_double(m.dx_char), int(vol_corr), dp(ops),
                    up(flags), nthreads)
    return ops.reshape(m.nb, nop), flags


def seed_health(m, segs=(), quads=()):
    L = lib()
    seg = np.ascontiguousarray(np.array(segs, float).reshape(-1))
    quad = np.ascontiguousarray(np.array(quads, float).reshape(-1))
    cr = Cracks(len(segs), dp(seg) if len(segs) else None, len(quads),
                dp(quad) if len(quads) else None)
    h = np.ones(m.nb)
    L.pxo_seed_health(C.byref(cr), m.n, dp(m.coords), ip(m.mirror), ip(m.off), ip(m.nbr), dp(h))
    return h


def csr_pattern(m, ndof):
    L = lib()
    rows = m.n * ndof
    rowptr = np.zeros(rows + 1, np.int32)
    nnz = L.pxo_csr_pattern(m.n, ndof, ip(m.off), ip(m.nbr), ip(m.mirror), ip(rowptr), None)
    cols = np.zeros(nnz, np.int32)
    L.pxo_csr_pattern(m.n, ndof, ip(m.off), ip(m.nbr), ip(m.mirror), ip(rowptr), ip(cols))
    return rowptr, cols


def assemble_poisson(m, ops, rowptr, cols, sigma, u, f_node=None, f_const=0.0):
    L = lib()
    k = np.zeros(len(cols)); r = np.zeros(m.n)
    fn = np.ascontiguousarray(f_node, float) if f_node is not None else None
    L.pxo_assemble_poisson(m.n, m.dim, ops.shape[1], ip(m.off), ip(m.nbr), dp(ops), ip(rowptr),
                           ip(cols), C.c_double(sigma), dp(fn), C.c_double(f_const),
                           dp(np.ascontiguousarray(u, float)), dp(k), dp(r))
    return k, r


def dynfrac_commit(m, sc, tension_only, damage_on, dt_prev, uold, uprev, vel, health):
    """Per-step commit of the dynamic fracture element (in place)."""
    lib().pxo_dynfrac_commit(m.n, m.dim, dp(m.coords), ip(m.off), ip(m.nbr), C.c_double(sc),
                             int(tension_only), int(damage_on), C.c_double(dt_prev),
                             dp(np.ascontiguousarray(uold, float)), dp(uprev), dp(vel), dp(health))


def assemble_dynfrac(m, ops, rowptr, cols, C11, C12, C66, rho, b, dt, vel, health, u, uold):
    """Dynamic fracture element through the CSR assembler: (K values, residual)."""
    k = np.zeros(int(rowptr[-1])); r = np.zeros(m.n * m.dim)
    lib().pxo_assemble_dynfrac(m.n, m.dim, ops.shape[1], ip(m.off), ip(m.nbr), dp(np.ascontiguousarray(ops)),
                               ip(np.ascontiguousarray(rowptr, np.int32)), ip(np.ascontiguousarray(cols, np.int32)),
                               C.c_double(C11), C.c_double(C12), C.c_double(C66), C.c_double(rho),
                               dp(np.ascontiguousarray(b, float)), C.c_double(dt),
                               dp(np.ascontiguousarray(vel, float)), dp(np.ascontiguousarray(health, float)),
                               dp(np.ascontiguousarray(u, float)), dp(np.ascontiguousarray(uold, float)),
                               dp(k), dp(r))
    return k, r


def assemble_diffusion(m, ops, rowptr, cols, D, f, dt, u, uold):
    """DiffusionElement through the CSR assembler: (K values, residual)."""
    nop = ops.shape[1] if ops.ndim == 2 else (9 if m.dim == 3 else 5)
    k = np.zeros(int(rowptr[-1])); r = np.zeros(m.n)
    lib().pxo_assemble_diffusion(m.n, m.dim, nop, ip(m.off), ip(m.nbr),
                                 dp(np.ascontiguousarray(ops)), ip(np.ascontiguousarray(rowptr, np.int32)),
                                 ip(np.ascontiguousarray(cols, np.int32)), C.c_double(D), C.c_double(f),
                                 C.c_double(dt), dp(np.ascontiguousarray(u, float)),
                                 dp(np.ascontiguousarray(uold, float)), dp(k), dp(r))
    return k, r


def ch_geometry(m, ops):
    L = lib()
    lap = np.zeros(m.nb); rev = np.zeros(m.nb, np.int32); isg = np.zeros(m.n, np.uint8)
    L.pxo_ch_geometry(m.n, m.dim, ops.shape[1], dp(m.coords), ip(m.off), ip(m.nbr), dp(ops),
                      ip(m.mirror), dp(lap), ip(rev), up(isg))
    return lap, rev, isg


def ch_mobility(m, u, M0):
    mn = np.zeros(m.n)
    lib().pxo_ch_mobility(m.n, dp(np.ascontiguousarray(u, float)), C.c_double(M0), dp(mn))
    return mn


def assemble_ch(m, lap, rev, isg, mnode, kappa, chi, dt, rowptr, cols, u, uold):
    L = lib()
    k = np.zeros(len(cols)); r = np.zeros(2 * m.n)
    L.pxo_assemble_ch(m.n, ip(m.off), ip(m.nbr), ip(rev), dp(lap), up(isg), dp(m.vols),
                      dp(mnode), C.c_double(kappa), C.c_double(chi), C.c_double(dt),
                      ip(rowptr), ip(cols), dp(np.ascontiguousarray(u, float)),
                      dp(np.ascontiguousarray(uold, float)), dp(k), dp(r))
    return k, r


def stressch_consts(dim, plane_stress, E, nu, Omega):
    """C11, C12, C66, A, Kh, Ch of the stress-coupled Cahn-Hilliard element."""
    out = np.zeros(6)
    if lib().pxo_stressch_consts(dim, int(plane_stress), C.c_double(E), C.c_double(nu),
                                 C.c_double(Omega), dp(out)) != 0:
        raise ValueError("stress Cahn-Hilliard: Poisson ratio out of range")
    return out


def assemble_stressch(m, ops, lap, rev, isg, consts, dt, rowptr, cols, u, uold):
    """StressCahnHilliardElement through the CSR assembler: (K values,
    residual); consts = C11, C12, C66, A, Kh, Ch, D, kappa, chi, Omega, cref."""
    nd = m.dim + 2
    k = np.zeros(int(rowptr[-1])); r = np.zeros(nd * m.n)
    kc = np.ascontiguousarray(consts, float)
    assert kc.shape == (11,)
    lib().pxo_assemble_stressch(m.n, m.dim, ops.shape[1], ip(m.off), ip(m.nbr), dp(np.ascontiguousarray(ops)),
                                ip(rev), dp(lap), up(isg), dp(m.vols), dp(kc), C.c_double(dt),
                                ip(np.ascontiguousarray(rowptr, np.int32)), ip(np.ascontiguousarray(cols, np.int32)),
                                dp(np.ascontiguousarray(u, float)), dp(np.ascontiguousarray(uold, float)),
                                dp(k), dp(r))
    return k, r


def fsch_lapc(m, ops):
    """Species Laplacian of the fracture stress-CH element: per bond, zero
    on bonds to ghosts."""
    lc = np.zeros(m.nb)
    lib().pxo_fsch_lapc(m.n, m.dim, ops.shape[1], ip(m.off), ip(m.nbr), dp(np.ascontiguousarray(ops)),
                        ip(m.mirror), dp(lc))
    return lc


def fsch_commit(m, sc, tension_only, damage_on, Omega, cref, rho, dt_prev, uold, uprev, vel, health):
    """Per-step commit, in place on uprev, vel (n*dim) and health (nb, 0/1)."""
    for a in (uprev, vel, health):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    lib().pxo_fsch_commit(m.n, m.dim, dp(m.coords), ip(m.off), ip(m.nbr), C.c_double(sc), int(tension_only),
                          int(damage_on), C.c_double(Omega), C.c_double(cref), C.c_double(rho),
                          C.c_double(dt_prev), dp(np.ascontiguousarray(uold, float)), dp(uprev), dp(vel),
                          dp(health))


def assemble_fsch(m, ops, lapc, rev, isg, consts, dt, vel, health, rowptr, cols, u, uold):
    """FracStressCahnHilliardElement through the CSR assembler: (K values,
    residual); consts = C11, C12, C66, A, Kh, Ch, D, kappa, chi, Omega, cref,
    rho, residual stiffness."""
    nd = m.dim + 2
    k = np.zeros(int(rowptr[-1])); r = np.zeros(nd * m.n)
    kc = np.ascontiguousarray(consts, float)
    assert kc.shape == (13,)
    lib().pxo_assemble_fsch(m.n, m.dim, ops.shape[1], ip(m.off), ip(m.nbr), dp(np.ascontiguousarray(ops)),
                            ip(rev), dp(lapc), up(isg), dp(m.vols), dp(kc), C.c_double(dt),
                            dp(np.ascontiguousarray(vel, float)), dp(np.ascontiguousarray(health, float)),
                            ip(np.ascontiguousarray(rowptr, np.int32)), ip(np.ascontiguousarray(cols, np.int32)),
                            dp(np.ascontiguousarray(u, float)), dp(np.ascontiguousarray(uold, float)),
                            dp(k), dp(r))
    return k, r


# projections of the stress-coupled Cahn-Hilliard elements: name -> (index, components)
SCH_PROJECTIONS = {"strain": (0, 9), "stress": (1, 9), "vonMises": (2, 1), "sigmaH": (3, 1),
                   "fPrime": (4, 1), "fDoublePrime": (5, 1), "gradC": (6, 3), "damage": (7, 1),
                   "DamageI": (8, 1)}


def sch_project(m, ops, consts, frac, plane_strain, u, name, health=None, seed=None):
    """A stress-CH element's output projection (n x components); consts =
    C11, C12, C66, A, cref, chi; health / seed (nb, 0/1) for damage, DamageI."""
    which, nc = SCH_PROJECTIONS[name]
    out = np.zeros(m.n * nc)
    kc = np.ascontiguousarray(consts, float)
    assert kc.shape == (6,)
    h = None if health is None else np.ascontiguousarray(health, float)
    sd = None if seed is None else np.ascontiguousarray(seed, float)
    lib().pxo_sch_project(m.n, m.dim, ops.shape[1], ip(m.off), ip(m.nbr), dp(np.ascontiguousarray(ops)),
                          dp(kc), int(frac), int(plane_strain), dp(h), dp(sd),
                          dp(np.ascontiguousarray(u, float)), which, dp(out))
    return out.reshape(m.n, nc)


def elastic_constants(dim, plane_stress, E, nu):
    c = [C.c_double(), C.c_double(), C.c_double()]
    lib().pxo_elastic_constants(dim, int(plane_stress), C.c_double(E), C.c_double(nu),
                                C.byref(c[0]), C.byref(c[1]), C.byref(c[2]))
    return tuple(x.value for x in c)


def critical_stretch(G0, E, delta, dim, plane_stress):
    return lib().pxo_critical_stretch(C.c_double(G0), C.c_double(E), C.c_double(delta), dim,
                                      int(plane_stress))


def stable_dt(dim, plane_stress, E, nu, rho, dxyz):
    return lib().pxo_stable_dt(dim, int(plane_stress), C.c_double(E), C.c_double(nu),
                               C.c_double(rho), dp(np.ascontiguousarray(dxyz, float)))


def frac_params(dim, E, nu, rho, sc, plane_stress=False, tension_only=True, damage=True,
                body=(0.0, 0.0, 0.0)):
    c11, c12, c66 = elastic_constants(dim, plane_stress, E, nu)
    return Frac(dim, c11, c12, c66, rho, body[0], body[1], body[2], sc, int(tension_only),
                int(damage), nu, int(plane_stress))


def frac_opcache(m, ops):
    nops = 9 if m.dim == 3 else 5
    opc = np.zeros(m.nb * nops)
    lib().pxo_frac_opcache(m.n, m.dim, ip(m.off), dp(np.ascontiguousarray(ops)), dp(opc))
    return opc.reshape(m.nb, nops)


def frac_accel(m, p, opc, health, u, node_E=None, nthreads=0):
    a = np.zeros(m.n * m.dim)
    ne = np.ascontiguousarray(node_E, float) if node_E is not None else None
    lib().pxo_frac_accel(m.n, C.byref(p), dp(m.coords), ip(m.off), ip(m.nbr),
                         dp(np.ascontiguousarray(opc)), dp(health), dp(u), dp(ne), dp(a), nthreads)
    return a


def make_bcs(specs):
    """specs: list of dicts {type:'neumann'|'dirichlet', nodes, comps, value, velocity,
    direct, box}.  Returns (ctypes array, keepalive)."""
    arr = (BC * max(1, len(specs)))()
    keep = []
    for e, s in enumerate(specs):
        b = arr[e]
        b.type = 0 if s["type"] == "neumann" else 1
        b.direct = int(s.get("direct", False))
        comps = list(s["comps"])
        b.ncomp = len(comps)
        val = list(s.get("value", [0.0] * len(comps)))
        vel = list(s.get("velocity", [0.0] * len(comps)))
        if len(val) == 1: val = val * len(comps)
        if len(vel) == 1: vel = vel * len(comps)
        for k, c in enumerate(comps):
            b.comp[k] = c; b.value[k] = val[k]; b.velocity[k] = vel[k]
        box = s.get("box")
        b.has_box = int(box is not None)
        for k in range(6):
            b.box[k] = box[k] if box is not None else (1.0 if k % 2 == 0 else -1.0)
        nodes = np.ascontiguousarray(s["nodes"], np.int32)
        keep.append(nodes)
        b.nnodes = len(nodes); b.nodes = ip(nodes)
    return arr, keep


def bc_preset(m, ndof, bcs, u, t):
    arr, keep = bcs
    lib().pxo_bc_preset(m.n, ndof, dp(m.coords), ip(m.mirror), len(keep), arr, dp(u),
                        C.c_double(t))


PROJECTIONS = ("damage", "vonMises", "stress", "strain")


def frac_project(m, p, opc, health, u, name):
    """Nodal projection `name` of the explicit element (n x 1 or n x 9)."""
    which = PROJECTIONS.index(name)
    nc = 9 if which >= 2 else 1
    out = np.zeros(m.n * nc)
    lib().pxo_frac_project(m.n, m.dim, int(p.plane_stress), C.c_double(p.C12), C.c_double(p.C66),
                           ip(m.off), ip(m.nbr), dp(np.ascontiguousarray(opc)),
                           dp(np.ascontiguousarray(health, float)), dp(np.ascontiguousarray(u, float)),
                           which, dp(out))
    return out.reshape(m.n, nc)


def bc_rows(m, ndof, bcs, t, U, rowptr, cols, K, R):
    """BC row elimination on copies of K / R; returns (rc, K', R')."""
    arr, keep = bcs
    K = np.ascontiguousarray(K, float).copy(); R = np.ascontiguousarray(R, float).copy()
    rc = lib().pxo_bc_rows(m.n, ndof, dp(m.coords), ip(m.mirror), len(keep), arr, C.c_double(t),
                           dp(np.ascontiguousarray(U, float)),
                           ip(np.ascontiguousarray(rowptr, np.int32)),
                           ip(np.ascontiguousarray(cols, np.int32)), dp(K), dp(R))
    return rc, K, R


def schedule(dt, total):
    L = lib()
    n = L.pxo_schedule(C.c_double(dt), C.c_double(total), 0, None, None)
    dts = np.zeros(n); ts = np.zeros(n)
    L.pxo_schedule(C.c_double(dt), C.c_double(total), n, dp(dts), dp(ts))
    return dts, ts


def explicit_run(m, p, opc, bcs, dt, total, u, uold, health, v0=None, node_E=None,
                 nthreads=0, record_breaks=True):
    arr, keep = bcs
    bs = np.zeros(m.nb, np.int32) if record_breaks else None
    ne = np.ascontiguousarray(node_E, float) if node_E is not None else None
    v = np.ascontiguousarray(v0, float) if v0 is not None else None
    steps = lib().pxo_explicit_run(m.n, C.byref(p), dp(m.coords), ip(m.off), ip(m.nbr),
                                   dp(np.ascontiguousarray(opc)), ip(m.mirror), len(keep), arr,
                                   C.c_double(dt), C.c_double(total), dp(u), dp(uold), dp(v),
                                   dp(health), ip(bs), dp(ne), nthreads)
    return steps, bs


def frac_adjoint(m, p, opc, bcs, dt, nsteps, u0, break_step, node_E, sens, comp, v0=None,
                 nthreads=0, want_b=False):
    """(J, dJ/dE) or, with want_b, (J, dJ/dE, dJ/db)."""
    arr, keep = bcs
    g = np.zeros(m.n)
    gb = np.zeros(3)
    sens = np.ascontiguousarray(sens, np.int32)
    v = np.ascontiguousarray(v0, float) if v0 is not None else None
    J = lib().pxo_frac_adjoint(m.n, C.byref(p), dp(m.coords), ip(m.off), ip(m.nbr),
                               dp(np.ascontiguousarray(opc)), ip(m.mirror), len(keep), arr,
                               C.c_double(dt), nsteps, dp(np.ascontiguousarray(u0, float)),
                               dp(v), ip(np.ascontiguousarray(break_step, np.int32)),
                               dp(np.ascontiguousarray(node_E, float)), len(sens), ip(sens),
                               comp, dp(g), dp(gb), nthreads)
    return (J, g, gb[:m.dim]) if want_b else (J, g)


def frac_loss_frozen(m, p, opc, bcs, dt, nsteps, u0, break_step, node_E, sens, comp,
                     v0=None, nthreads=0):
    arr, keep = bcs
    sens = np.ascontiguousarray(sens, np.int32)
    v = np.ascontiguousarray(v0, float) if v0 is not None else None
    return lib().pxo_frac_loss_frozen(m.n, C.byref(p), dp(m.coords), ip(m.off), ip(m.nbr),
                                      dp(np.ascontiguousarray(opc)), ip(m.mirror), len(keep),
                                      arr, C.c_double(dt), nsteps,
                                      dp(np.ascontiguousarray(u0, float)), dp(v),
                                      ip(np.ascontiguousarray(break_step, np.int32)),
                                      dp(np.ascontiguousarray(node_E, float)), len(sens),
                                      ip(sens), comp, nthreads)


def random_ic(seed, lo, hi, nodes, ndof, dofs, values):
    nodes = np.ascontiguousarray(nodes, np.int32)
    dofs = np.ascontiguousarray(dofs, np.int32)
    lib().pxo_random_ic(C.c_uint64(seed), C.c_double(lo), C.c_double(hi), len(nodes),
                        ip(nodes), ndof, len(dofs), ip(dofs), dp(values))
    return values


def ldu_profile(rowptr, cols, use_rcm=True):
    """The default linear solver's ordering and skyline pointer: (perm, jp)."""
    n = len(rowptr) - 1
    perm = np.zeros(n, np.int32); jp = np.zeros(n + 1, np.int64)
    lib().pxo_ldu_profile(n, ip(np.ascontiguousarray(rowptr, np.int32)), ip(np.ascontiguousarray(cols, np.int32)),
                          int(use_rcm), ip(perm), jp.ctypes.data_as(C.POINTER(C.c_longlong)))
    return perm, jp


def ldu_solve(rowptr, cols, vals, b, use_rcm=True):
    """The default linear solver on K x = b: (rc, x, perm, jp); rc 0, or
    1 + the equation (new order) of a zero pivot."""
    n = len(rowptr) - 1
    x = np.zeros(n); perm = np.zeros(n, np.int32); jp = np.zeros(n + 1, np.int64)
    rc = lib().pxo_ldu_solve(n, ip(np.ascontiguousarray(rowptr, np.int32)),
                             ip(np.ascontiguousarray(cols, np.int32)), dp(np.ascontiguousarray(vals, float)),
                             dp(np.ascontiguousarray(b, float)), int(use_rcm), dp(x), ip(perm),
                             jp.ctypes.data_as(C.POINTER(C.c_longlong)))
    return rc, x, perm, jp
