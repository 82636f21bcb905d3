"""BC row elimination of the implicit CSR system on the device
(pxg_csr_set_bcs / pxg_csr_apply_bcs[_dev]) against the reference's own
output (golden fixtures) and the oracle's sequential restatement: bit for
bit.  At full size (Poisson 100 x 100, Cahn-Hilliard 256 x 256) the size-independent properties are
checked: unconstrained rows untouched, constrained rows hold exactly their
diagonal and mirror entry, and a second application changes nothing."""
import os

import numpy as np
import pytest

from conftest import bcrow_specs
from oracle import pxo
from perix_amd import pxg, problems

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


@pytest.mark.parametrize("name", ["bcrows_poisson_2d.npz", "bcrows_ch_2d.npz"])
def test_bc_rows_vs_reference(name):
    z = np.load(os.path.join(GOLD, name), allow_pickle=False)
    ndof = int(z["ndof"])
    ctx = pxg.Context(0)
    ctx.set_points(z["coords"], z["vols"], z["mirror"], z["dxyz"], 2)
    ctx.build_families(float(z["delta"]))
    ctx.build_csr(ndof)
    rp, cols = ctx.csr()
    assert np.array_equal(rp, z["rowptr"]) and np.array_equal(cols, z["cols"])
    nrows = ctx.csr_set_bcs(bcrow_specs(z))
    print("%s: %d constrained rows" % (name, nrows), flush=True)
    k, r = ctx.csr_apply_bcs(float(z["t"]), z["u"], z["K"], z["R"])
    assert np.array_equal(bits(k), bits(z["K_bc"])) and np.array_equal(bits(r), bits(z["R_bc"]))
    # device pointers give the same bits
    import torch
    ud = torch.from_numpy(np.ascontiguousarray(z["u"])).cuda()
    kd = torch.from_numpy(np.ascontiguousarray(z["K"])).cuda()
    rd = torch.from_numpy(np.ascontiguousarray(z["R"])).cuda()
    torch.cuda.synchronize()
    ctx.csr_apply_bcs_dev(float(z["t"]), ud.data_ptr(), kd.data_ptr(), rd.data_ptr())
    ctx.sync()
    assert np.array_equal(bits(kd.cpu().numpy()), bits(z["K_bc"]))
    assert np.array_equal(bits(rd.cpu().numpy()), bits(z["R_bc"]))
    ctx.close()


def _random_list(cloud, rs, ndof):
    """Overlapping Dirichlet (reflect / direct / velocity / box) and Neumann
    entries over the four ghost walls in a seeded order."""
    specs = []
    for w in ("left", "right", "bottom", "top"):
        ids = cloud.group(w)
        for _ in range(2):
            comps = sorted(rs.choice(ndof, rs.randint(1, ndof + 1), replace=False).tolist())
            if rs.rand() < 0.4:
                specs.append({"type": "neumann", "nodes": ids, "comps": comps})
            else:
                box = None
                if rs.rand() < 0.5:
                    lo = rs.rand() * 0.5
                    box = (lo, lo + 0.4, 1.0, -1.0, 1.0, -1.0)
                specs.append({"type": "dirichlet", "nodes": ids, "comps": comps,
                              "value": rs.standard_normal(len(comps)).tolist(),
                              "velocity": rs.standard_normal(len(comps)).tolist(),
                              "direct": bool(rs.rand() < 0.3), "box": box})
    rs.shuffle(specs)
    return specs


@pytest.mark.parametrize("kind,nx", [("poisson", 40), ("ch", 40), ("poisson", 100), ("ch", 256)])
def test_bc_rows_vs_oracle_and_properties(kind, nx):
    rs = np.random.RandomState(nx + (7 if kind == "ch" else 0))
    if kind == "poisson":
        ctx, cloud, _, f = problems.poisson_square(nx=nx, ny=nx)
        ndof = 1
        u = rs.standard_normal(cloud.n)
        k, r = ctx.assemble_poisson(1.0, u, f_node=f)
    else:
        ctx, cloud, _, u0 = problems.cahnhilliard_square(nx=nx, ny=nx)
        ndof = 2
        u = u0.copy()
        u[1::2] = 0.1 * rs.standard_normal(cloud.n)
        k, r = ctx.assemble_ch(1e-2 / nx ** 2, 3.0, 1.0, 1e-3, u, u0)
    rp, cols = ctx.csr()
    specs = _random_list(cloud, rs, ndof)
    # zero some diagonals and some whole constrained rows: all three scale rules
    ghosts = np.concatenate([cloud.group(w) for w in ("left", "right", "bottom", "top")]) - 1
    for g in rs.choice(ghosts, 12, replace=False) * ndof:
        s, e = rp[g], rp[g + 1]
        d = s + int(np.searchsorted(cols[s:e], g))
        if rs.rand() < 0.5:
            k[d] = 0.0
        else:
            k[s:e] = 0.0
    t = 0.37
    ctx.csr_set_bcs(specs)
    k1, r1 = ctx.csr_apply_bcs(t, u, k, r)

    class M:
        n = cloud.n; coords = cloud.coords; mirror = cloud.mirror
    rc, ko, ro = pxo.bc_rows(M, ndof, pxo.make_bcs(specs), t, u, rp, cols, k, r)
    assert rc == 0
    assert np.array_equal(bits(k1), bits(ko)) and np.array_equal(bits(r1), bits(ro))
    # properties: untouched rows, the structure of the constrained rows, idempotence
    changed = np.zeros(len(rp) - 1, bool)
    for g in range(len(rp) - 1):
        s, e = rp[g], rp[g + 1]
        if not np.array_equal(bits(k1[s:e]), bits(k[s:e])) or bits(r1[g:g + 1]) != bits(r[g:g + 1]):
            changed[g] = True
            assert np.count_nonzero(k1[s:e]) in (1, 2)
    assert changed.any()
    k2, r2 = ctx.csr_apply_bcs(t, u, k1, r1)
    assert np.array_equal(bits(k2), bits(k1)) and np.array_equal(bits(r2), bits(r1))
    ctx.close()


def test_bc_rows_errors():
    ctx, cloud, _, _ = problems.poisson_square(nx=20, ny=20)
    bulk = np.array([1, 2, 3], np.int32)          # bulk nodes have no mirror
    with pytest.raises(pxg.PxgError):
        ctx.csr_set_bcs([{"type": "neumann", "nodes": bulk, "comps": [0]}])
    with pytest.raises(pxg.PxgError):
        ctx.csr_set_bcs([{"type": "dirichlet", "nodes": cloud.group("left"), "comps": [1],
                          "value": [0.0]}])
    # a Dirichlet pin on a bulk node is a direct row (no mirror)
    assert ctx.csr_set_bcs([{"type": "dirichlet", "nodes": bulk, "comps": [0], "value": [1.0]}]) == 3
    ctx.close()
