"""The oracle's restatement of the reference's default linear solver
(oracle/pxo_solve.cpp) against the reference itself (DefaultSolver compiled
from its sources, oracle/_ref) and against a textbook direct solve.

The ordering (RCM), the skyline pointer and the solution must equal the
reference bit for bit: the oracle sums in the reference's order.  scipy's
SuperLU (an independent pivoting LU) pins what the reference does not: that
the factorisation solves the system at all, to round-off.  A planted mistake
in the oracle (a dropped envelope entry, a wrong sign in the update) changes
the solution far beyond either bound."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle import pxo


def _ref():
    from oracle import refpy
    if not refpy.available():
        pytest.skip("oracle/_ref/libperix_ref.so is not built here (make -C oracle/ref)")
    return refpy


def _system(refpy, kind, nx):
    deck = refpy.poisson_deck(nx, "openmp") if kind == "poisson" else refpy.ch_deck(nx, 1, "openmp")
    c = refpy.Case(deck)
    u, _ = c.initial_state()
    u = c.preset(u, 0.0)
    k, r = c.assemble(u, u, 1e-3)
    k, r = c.apply_bcs(u, 1e-3, k, r)
    rp, cols = c.csr()
    return c, rp, cols, k, r


@pytest.mark.parametrize("kind,nx", [("poisson", 24), ("poisson", 60), ("ch", 16)])
@pytest.mark.parametrize("use_rcm", [True, False])
def test_oracle_ldu_equals_reference(kind, nx, use_rcm):
    refpy = _ref()
    c, rp, cols, k, r = _system(refpy, kind, nx)
    rc, x, perm, jp = c.default_solve(k, r, use_rcm)
    c.close()
    orc, ox, operm, ojp = pxo.ldu_solve(rp, cols, k, r, use_rcm)
    assert rc == 0 and orc == 0
    assert np.array_equal(perm, operm) and np.array_equal(jp, ojp)
    assert np.array_equal(x.view(np.uint64), ox.view(np.uint64))
    if not use_rcm:
        assert np.array_equal(operm, np.arange(len(r)))
    A = sp.csr_matrix((k, cols, rp), shape=(len(r), len(r)))
    xd = spl.spsolve(A.tocsc(), r)
    err = np.abs(ox - xd).max() / np.abs(xd).max()
    res = np.linalg.norm(A @ ox - r) / np.linalg.norm(r)
    print("%s %d rcm=%s profile %d: vs SuperLU %.2e, residual %.2e" % (kind, nx, use_rcm, ojp[-1], err, res))
    assert err <= 1e-11 and res <= 1e-13


def test_oracle_rcm_shrinks_profile():
    """RCM shrinks the skyline: on the 60 x 60 Poisson pattern (ghosts
    numbered after the bulk, so the natural order couples far rows) the
    profile is well below the natural order's."""
    refpy = _ref()
    c, rp, cols, k, r = _system(refpy, "poisson", 60)
    c.close()
    _, jp_rcm = pxo.ldu_profile(rp, cols, True)
    _, jp_nat = pxo.ldu_profile(rp, cols, False)
    assert jp_rcm[-1] * 3 < jp_nat[-1] * 2


def test_oracle_ldu_zero_pivot():
    """An equation whose row and column are empty meets an exact zero pivot:
    the reference and the oracle report the same equation (new order)."""
    refpy = _ref()
    c, rp, cols, k, r = _system(refpy, "poisson", 24)
    e = 300
    k = k.copy()
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    k[(rows == e) | (cols == e)] = 0.0
    rc, x, perm, jp = c.default_solve(k, r)
    c.close()
    orc, ox, operm, ojp = pxo.ldu_solve(rp, cols, k, r)
    assert rc == 1 and orc >= 1
    inv = np.empty_like(operm)
    inv[operm] = np.arange(len(operm))
    assert orc - 1 == inv[e]


def test_oracle_ldu_closed_form():
    """A tridiagonal 1D Laplacian with u = 0 ends: the exact solution of
    -u'' = 1 sampled at the nodes, x_i = i (n + 1 - i) h^2 / 2, is what the
    difference equations give exactly."""
    n = 50
    h = 1.0 / (n + 1)
    rp = [0]
    cols, vals = [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                cols.append(j)
                vals.append(2.0 / h ** 2 if i == j else -1.0 / h ** 2)
        rp.append(len(cols))
    b = np.ones(n)
    rc, x, perm, jp = pxo.ldu_solve(np.array(rp, np.int32), np.array(cols, np.int32), np.array(vals), b)
    i = np.arange(1, n + 1)
    exact = i * (n + 1 - i) * h * h / 2.0
    assert rc == 0
    assert np.abs(x - exact).max() <= 1e-12 * exact.max()
    assert jp[-1] == n - 1                 # a tridiagonal matrix keeps its one-entry envelope
    assert math.isclose(x.max(), exact.max(), rel_tol=1e-12)
