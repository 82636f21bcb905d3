"""The C oracle against what the mathematics and the reference's own unit
tests fix, independently of any run of the reference.  CPU only.

Each check is chosen so that a plausible slip in the oracle (a dropped term,
a wrong index or sign, a transposed operand) fails it:
  cloud        closed-form cell centres, volumes and ghost counts
  families     brute force O(N^2) on small grids; interior family sizes 28
               (2D) and 122 (3D) at 3.015 dx; symmetry j in H_i <=> i in H_j
  crack cut    hand cases of tests/MeshModifyTests.cpp:46-73
  g-weights    polynomial reproduction sum_j g^q(xi_j) xi_j^p = p! delta_pq
  Poisson      per-bond closed form of tests/PoissonElementTests.cpp:71-88,
               row sums, O(dx^2) residual of a manufactured solution
  explicit     tests/ExplicitPDDOFracElementTests.cpp:86-154 (body force,
               CFL, 2 % expansion breaks every bond, irreversibility, severed
               family), rigid translation and linear fields give no force,
               tests/InitialVelocityTests.cpp:51-89 (v0 = 2, dt = 0.1)
  presets      tests/BCSystemTests.cpp:155-250 (reflection, direct, ramp,
               bulk pin, box, Neumann mirror)
  RandomIC     the C++ standard's 10000th std::mt19937_64 output
  adjoint      central finite differences of the frozen-health loss
"""
import ctypes as C
import math
import types

import numpy as np
import pytest

from oracle import pxo


# ---------------------------------------------------------------- cloud ----
@pytest.mark.parametrize("dim,grid,hi", [(2, (7, 5, 1), (0.7, 0.5, 1.0)),
                                         (3, (4, 3, 5), (0.4, 0.3, 0.5))])
def test_cloud_closed_form(dim, grid, hi):
    nx, ny, nz = grid
    m = pxo.Mesh(pxo.make_grid(dim, nx, ny, nz, hi=hi), families=False)
    nbulk = nx * ny * (nz if dim == 3 else 1)
    ghosts = 2 * (nx + ny) if dim == 2 else 2 * (nx * ny + nx * nz + ny * nz)
    assert m.nbulk == nbulk and m.n == nbulk + ghosts
    d = [hi[0] / nx, hi[1] / ny, hi[2] / nz if dim == 3 else 0.0]
    X = m.coords.reshape(-1, 3)[:nbulk]
    # bulk nodes: cell centres (k + 1/2) dx, x fastest
    k = np.arange(nbulk)
    ix, iy, iz = k % nx, (k // nx) % ny, k // (nx * ny)
    assert np.allclose(X[:, 0], (ix + 0.5) * d[0], rtol=1e-14, atol=0)
    assert np.allclose(X[:, 1], (iy + 0.5) * d[1], rtol=1e-14, atol=0)
    if dim == 3:
        assert np.allclose(X[:, 2], (iz + 0.5) * d[2], rtol=1e-14, atol=0)
    vol = d[0] * d[1] * (d[2] if dim == 3 else 1.0)
    assert np.allclose(m.vols[:nbulk], vol, rtol=1e-14)
    # every ghost mirrors a bulk node across its wall
    g = m.mirror[nbulk:]
    assert (g >= 1).all() and (g <= nbulk).all() and (m.mirror[:nbulk] == 0).all()


# ------------------------------------------------------------- families ----
@pytest.mark.parametrize("dim,grid,hi", [(2, (9, 7, 1), (0.9, 0.7, 1.0)),
                                         (3, (5, 4, 4), (0.5, 0.4, 0.4))])
def test_families_brute_force(dim, grid, hi):
    nx, ny, nz = grid
    m = pxo.Mesh(pxo.make_grid(dim, nx, ny, nz, hi=hi))
    X = m.coords.reshape(-1, 3)
    r2 = m.delta * m.delta
    d2 = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    for i in range(m.n):
        fam = set((m.nbr[m.off[i]:m.off[i + 1]] - 1).tolist())
        assert len(fam) == m.off[i + 1] - m.off[i], "duplicate family member"
        brute = set(np.nonzero(d2[i] <= r2 * (1 + 1e-12))[0].tolist())
        assert fam == brute, "node %d" % i
    # symmetry of the horizon relation
    pairs = set()
    for i in range(m.n):
        for j in m.nbr[m.off[i]:m.off[i + 1]] - 1:
            pairs.add((i, int(j)))
    assert all((j, i) in pairs for (i, j) in pairs)


@pytest.mark.parametrize("dim,grid,full", [(2, (15, 15, 1), 28), (3, (9, 9, 9), 122)])
def test_interior_family_size(dim, grid, full):
    nx, ny, nz = grid
    m = pxo.Mesh(pxo.make_grid(dim, nx, ny, nz))
    centre = (nx // 2) + nx * (ny // 2) + (nx * ny * (nz // 2) if dim == 3 else 0)
    assert m.off[centre + 1] - m.off[centre] == full


def test_crack_cut_hand_cases():
    """tests/MeshModifyTests.cpp:46-73: a vertical notch x = 0.5."""
    L = pxo.lib()
    L.pxo_crack_cuts_bond.argtypes = [C.POINTER(pxo.Cracks), C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_int), C.c_int, C.c_int]
    seg = np.array([0.5, 0.0, 0.5, 1.0, 0.0, 0.8, 0.25, 0.8])
    cr = pxo.Cracks(2, pxo.dp(seg), 0, None)
    xy = np.array([0.25, 0.50, 0.0, 0.75, 0.50, 0.0, 0.25, 0.25, 0.0, 0.25, 0.75, 0.0])
    assert L.pxo_crack_cuts_bond(C.byref(cr), 4, pxo.dp(xy), None, 1, 2) == 1, "crossing bond"
    assert L.pxo_crack_cuts_bond(C.byref(cr), 4, pxo.dp(xy), None, 3, 4) == 0, "same-side bond"
    # the segment is through-thickness in 3D
    seg3 = np.array([0.5, 0.0, 0.5, 1.0])
    cr3 = pxo.Cracks(1, pxo.dp(seg3), 0, None)
    xyz = np.array([0.25, 0.50, -0.5, 0.75, 0.50, 0.5])
    assert L.pxo_crack_cuts_bond(C.byref(cr3), 2, pxo.dp(xyz), None, 1, 2) == 1


# ------------------------------------------------------------ g-weights ----
def multi_indices(dim):
    L = pxo.lib()
    nop = L.pxo_nop(dim, 2)
    q = [np.zeros(nop, np.int32) for _ in range(3)]
    L.pxo_multi_indices(dim, 2, pxo.ip(q[0]), pxo.ip(q[1]), pxo.ip(q[2]))
    return list(zip(*(x.tolist() for x in q)))


@pytest.mark.parametrize("dim,grid,hi", [(2, (12, 10, 1), (1.2, 1.0, 1.0)),
                                         (3, (7, 6, 6), (0.7, 0.6, 0.6))])
def test_gweights_reproduce_polynomials(dim, grid, hi):
    nx, ny, nz = grid
    m = pxo.Mesh(pxo.make_grid(dim, nx, ny, nz, hi=hi))
    ops, flags = pxo.operators(m)
    mi = multi_indices(dim)
    X = m.coords.reshape(-1, 3)
    checked = 0
    for i in range(m.n):
        if flags[i]:
            continue
        t0, t1 = m.off[i], m.off[i + 1]
        xi = X[m.nbr[t0:t1] - 1] - X[i]
        for a, p in enumerate(mi):
            mono = xi[:, 0] ** p[0] * xi[:, 1] ** p[1] * xi[:, 2] ** p[2]
            for b in range(len(mi)):
                terms = ops[t0:t1, b] * mono
                want = math.factorial(p[0]) * math.factorial(p[1]) * math.factorial(p[2]) if a == b else 0.0
                assert abs(terms.sum() - want) <= 1e-12 * max(1.0, np.abs(terms).sum()), (i, p, b)
        checked += 1
    assert checked == m.n


# -------------------------------------------------------------- Poisson ----
def test_poisson_bond_closed_form_and_row_sums():
    """Per bond c = sigma (gxx + gyy)_ij: R_i = f + sum_j c (u_j - u_i),
    K(i,j) = -c (the Jacobian of -R), K(i,i) = -sum_j K(i,j)
    (PoissonElementTests.cpp:71-88; CSR columns are 0-based node ids)."""
    m = pxo.Mesh(pxo.make_grid(2, 10, 8))
    ops, _ = pxo.operators(m)
    rp, cols = pxo.csr_pattern(m, 1)
    sigma, f = 1.3, 0.7
    u = np.random.RandomState(1).rand(m.n)
    k, r = pxo.assemble_poisson(m, ops, rp, cols, sigma, u, f_const=f)
    for i in range(m.n):
        row = dict(zip(cols[rp[i]:rp[i + 1]].tolist(), k[rp[i]:rp[i + 1]].tolist()))
        res = f
        want = {}
        for t in range(m.off[i], m.off[i + 1]):
            j = int(m.nbr[t]) - 1
            coef = sigma * (ops[t, 2] + ops[t, 4])
            res += coef * (u[j] - u[i])
            want[j] = want.get(j, 0.0) - coef
        diag = row.pop(i)
        scale = sum(abs(v) for v in row.values())
        for j, v in want.items():
            assert abs(row[j] - v) <= 1e-14 * scale, "bond coefficient"
        assert abs(diag + sum(row.values())) <= 1e-12 * scale, "row sum"
        assert abs(r[i] - res) <= 1e-12 * (abs(f) + scale * np.abs(u).max()), "residual"


def test_poisson_manufactured_residual_is_second_order():
    """u = sin(pi x) sin(pi y), f = 2 pi^2 sin sin: the PDDO residual of the
    exact field at interior nodes falls as dx^2."""
    errs = []
    for n in (20, 40):
        m = pxo.Mesh(pxo.make_grid(2, n, n))
        ops, _ = pxo.operators(m)
        rp, cols = pxo.csr_pattern(m, 1)
        X = m.coords.reshape(-1, 3)
        u = np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1])
        fn = 2 * np.pi ** 2 * u
        _, r = pxo.assemble_poisson(m, ops, rp, cols, 1.0, u, f_node=fn)
        inner = (X[:, 0] > 0.2) & (X[:, 0] < 0.8) & (X[:, 1] > 0.2) & (X[:, 1] < 0.8)
        errs.append(np.abs(r[inner]).max())
    rate = math.log(errs[0] / errs[1], 2)
    assert rate > 1.8, (errs, rate)


# ------------------------------------------------------------- explicit ----
def small_frac(dim=2, n=8, E=100.0, nu=0.25, rho=4.0, plane_stress=True, sc=0.0, body=(0, 0, 0),
               damage=False):
    m = pxo.Mesh(pxo.make_grid(dim, n, n, n if dim == 3 else 1, hi=(float(n), float(n), float(n))))
    ops, _ = pxo.operators(m)
    opc = pxo.frac_opcache(m, ops)
    p = pxo.frac_params(dim, E, nu, rho, sc, plane_stress=plane_stress, damage=damage, body=body)
    return m, opc, p


def test_body_force_gives_b_over_rho():
    m, opc, p = small_frac(body=(8.0, -12.0, 0.0))
    a = pxo.frac_accel(m, p, opc, np.ones(m.nb), np.zeros(2 * m.n))
    assert np.allclose(a[0::2], 2.0, rtol=0, atol=1e-14) and np.allclose(a[1::2], -3.0, rtol=0, atol=1e-14)


def test_cfl_estimate():
    E, nu, rho = 100.0, 0.25, 4.0
    c11 = E / (1 - nu * nu)
    assert abs(pxo.stable_dt(2, True, E, nu, rho, np.array([1.0, 1.0, 0.0])) - 1 / math.sqrt(c11 / rho)) < 1e-15


def test_rigid_translation_and_linear_field_give_no_force():
    m, opc, p = small_frac()
    X = m.coords.reshape(-1, 3)
    scale = np.abs(opc).max() * 1.0
    a = pxo.frac_accel(m, p, opc, np.ones(m.nb), np.tile([0.3, -0.7], m.n))
    assert np.abs(a).max() <= 1e-12 * scale
    u = np.empty(2 * m.n)
    u[0::2] = 1e-3 * X[:, 0] - 2e-3 * X[:, 1]
    u[1::2] = 3e-3 * X[:, 0] + 1e-3 * X[:, 1]
    a = pxo.frac_accel(m, p, opc, np.ones(m.nb), u)
    # second-order PDDO operators annihilate linear fields at every node
    assert np.abs(a).max() <= 1e-10 * scale * np.abs(u).max() * 10


def test_expansion_breaks_every_bond_and_damage_is_irreversible():
    """2 % expansion with G0 = 1e-5 breaks every bond; restoring u = 0 keeps
    them broken; a severed family feels no force (ExplicitPDDOFracElementTests.cpp:115-154)."""
    E, nu, rho = 100.0, 0.25, 4.0
    m0 = pxo.Mesh(pxo.make_grid(2, 6, 6, hi=(6.0, 6.0, 1.0)))
    sc = pxo.critical_stretch(1e-5, E, m0.delta, 2, True)
    assert sc < 0.02
    m, opc, p = small_frac(n=6, sc=sc, damage=True)
    X = m.coords.reshape(-1, 3)
    u = np.empty(2 * m.n)
    u[0::2] = 0.02 * X[:, 0]
    u[1::2] = 0.02 * X[:, 1]
    h = np.ones(m.nb)
    pxo.frac_accel(m, p, opc, h, u)
    assert (h == 0).all(), "every bond broken"
    a = pxo.frac_accel(m, p, opc, h, np.zeros(2 * m.n))
    assert (h == 0).all() and np.abs(a).max() == 0.0


def test_initial_velocity_one_step():
    """v0 = 2, a = 0, dt = 0.1 gives u = 0.2 (InitialVelocityTests.cpp:51-89)."""
    m, opc, p = small_frac(n=5)
    bcs = pxo.make_bcs([])
    u = np.zeros(2 * m.n); uo = np.zeros(2 * m.n)
    v0 = np.full(2 * m.n, 2.0)
    steps, _ = pxo.explicit_run(m, p, opc, bcs, 0.1, 0.1, u, uo, np.ones(m.nb), v0=v0)
    assert steps == 1
    assert np.allclose(u, 0.2, rtol=0, atol=1e-14)


def test_schedule_closed_form():
    dts, ts = pxo.schedule(0.3, 1.0)
    assert len(dts) == 4
    assert (dts[:3] == 0.3).all() and abs(dts[3] - (1.0 - ts[2])) == 0.0
    assert abs(ts[-1] - 1.0) < 1e-15
    assert all(ts[k] == ts[k - 1] + dts[k] for k in range(1, len(ts)))


# -------------------------------------------------------------- presets ----
def mirror_pair():
    """BCSystemTests.cpp:66-82: node 1 bulk at x = 0, node 2 its ghost at x = 1."""
    return types.SimpleNamespace(n=2, coords=np.array([0.0, 0, 0, 1.0, 0, 0]),
                                 mirror=np.array([0, 1], np.int32))


def preset(spec, u, t=0.0):
    m = mirror_pair()
    u = np.array(u, float)
    pxo.bc_preset(m, 1, pxo.make_bcs([spec]), u, t)
    return u


def test_preset_hand_cases():
    d = {"type": "dirichlet", "comps": [0]}
    assert preset(dict(d, nodes=[2], value=[5.0]), [2.0, 0.0])[1] == 8.0, "reflection 2g - u_b"
    assert preset(dict(d, nodes=[2], value=[3.0], direct=True), [2.0, 0.0])[1] == 3.0, "direct"
    ramp = dict(d, nodes=[2], value=[2.0], velocity=[3.0])
    assert preset(ramp, [1.0, 5.0], t=0.5)[1] == 6.0, "ramp g = v0 + v t"
    assert preset(ramp, [1.0, 5.0], t=0.0)[1] == 3.0, "ramp at rest"
    assert preset(dict(d, nodes=[1], value=[3.0]), [1.0, 0.0])[0] == 3.0, "bulk node pinned directly"
    boxed = preset(dict(d, nodes=[1, 2], value=[9.0], box=[0.5, 1.5, -1.0, 1.0, 1.0, -1.0]), [1.0, 0.0])
    assert boxed[0] == 1.0 and boxed[1] == 17.0, "box"
    assert preset({"type": "neumann", "nodes": [2], "comps": [0]}, [2.0, 9.0])[1] == 2.0, "Neumann mirror"


# ------------------------------------------------------------- RandomIC ----
def test_mt19937_64_standard_value():
    """std::mt19937_64 default seed 5489: the 10000th output is
    9981545732273789042 ([rand.predef] of the C++ standard); RandomIC maps
    it to [0,1) with its top 53 bits."""
    vals = np.zeros(10000)
    pxo.random_ic(5489, 0.0, 1.0, np.arange(1, 10001), 1, [1], vals)
    assert int(round(vals[-1] * 2.0 ** 53)) == (9981545732273789042 >> 11)


# -------------------------------------------------------------- adjoint ----
def test_adjoint_matches_finite_differences():
    """Gradient of the frozen-health loss w.r.t. per-node E against central
    differences (h = 1e-6 relative) along random directions."""
    E, nu, rho = 191e9, 0.25, 8000.0
    m = pxo.Mesh(pxo.make_grid(2, 10, 6, hi=(0.01, 0.006, 1.0)))
    ops, _ = pxo.operators(m)
    opc = pxo.frac_opcache(m, ops)
    sc = 0.003
    p = pxo.frac_params(2, E, nu, rho, sc)
    comps = [0, 1]
    specs = [{"type": "neumann", "nodes": m.group(w), "comps": comps} for w in range(4)]
    specs.append({"type": "dirichlet", "nodes": m.group(2), "comps": comps, "value": [0.0, 0.0],
                  "velocity": [0.0, 16.5], "direct": True})
    bcs = pxo.make_bcs(specs)
    dt = 0.5 * pxo.stable_dt(2, False, E, nu, rho, m.dxyz)
    nsteps = 30
    rs = np.random.RandomState(5)
    nodeE = E * (1 + 0.05 * rs.randn(m.n))
    u0 = np.zeros(2 * m.n)
    pxo.bc_preset(m, 2, bcs, u0, 0.0)
    # record the break steps of the forward run
    u = u0.copy(); uo = np.zeros_like(u); h = np.ones(m.nb)
    _, bs = pxo.explicit_run(m, p, opc, bcs, dt, nsteps * dt, u, uo, h, node_E=nodeE)
    sens = rs.choice(m.nbulk, 8, replace=False) + 1
    J, g = pxo.frac_adjoint(m, p, opc, bcs, dt, nsteps, u0, bs, nodeE, sens, 1)
    assert J > 0
    for _ in range(4):
        v = rs.randn(m.n)
        hE = 1e-6 * nodeE * v
        jp = pxo.frac_loss_frozen(m, p, opc, bcs, dt, nsteps, u0, bs, nodeE + hE, sens, 1)
        jm = pxo.frac_loss_frozen(m, p, opc, bcs, dt, nsteps, u0, bs, nodeE - hE, sens, 1)
        fd = (jp - jm) / 2.0
        ad = float(g @ hE)
        assert abs(fd - ad) <= 1e-4 * max(abs(fd), abs(ad)), (fd, ad)
    # dJ/db of the uniform body force: central differences per component
    # (J is quadratic in b at frozen health, so they are exact to rounding)
    _, _, gb = pxo.frac_adjoint(m, p, opc, bcs, dt, nsteps, u0, bs, nodeE, sens, 1, want_b=True)
    for a in range(2):
        hb = 1e3 * rho                       # 1000 m/s^2 of body acceleration
        bp = [0.0, 0.0, 0.0]; bp[a] = hb
        bm = [0.0, 0.0, 0.0]; bm[a] = -hb
        jp = pxo.frac_loss_frozen(m, pxo.frac_params(2, E, nu, rho, sc, body=bp), opc, bcs, dt,
                                  nsteps, u0, bs, nodeE, sens, 1)
        jm = pxo.frac_loss_frozen(m, pxo.frac_params(2, E, nu, rho, sc, body=bm), opc, bcs, dt,
                                  nsteps, u0, bs, nodeE, sens, 1)
        fd = (jp - jm) / (2 * hb)
        assert abs(fd - gb[a]) <= 1e-6 * max(abs(fd), abs(gb[a])), (a, fd, gb[a])
    assert abs(gb[1]) > 0
