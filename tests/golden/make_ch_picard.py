"""Writes tests/golden/ch_picard_256.npz: the 256 x 256 Cahn-Hilliard state
after one Picard-frozen backward-Euler iteration from the RandomIC state.

The 256 x 256 unit square (66,560 PD nodes, c and mu per node), c0 = 0.5 +
U(-0.01, 0.01) by RandomIC with seed 12345 over all nodes, mu0 = 0, Neumann
mirrors of c and mu on the four ghost walls, chi = 3, kappa = 1e-2 dx^2,
M0 = 1, dt = 1e-3.  One iteration of NonlinearSolver's Newton loop
(NonlinearSolver.cpp:71-146): assemble K = -dF/dU and R = F at u0 with the
mobility frozen at u0 (CahnHilliardElement.cpp:155-292), eliminate the BC
rows (BCSystem::applyBCs), solve K du = R with a sparse direct LU, u1 = u0 +
du.  The assembly is the C oracle's; the solve is SciPy's SuperLU.  The
fixture is an input state for the device-vs-oracle assembly test at full
size, nothing more: u1 (133,120 values) and the SHA-256 of u0's bytes, so
the test can confirm it starts from the same RandomIC field.

    make -C oracle && python tests/golden/make_ch_picard.py    (about 2 minutes)
"""
import hashlib
import os
import sys
import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pxo  # noqa: E402


def randomic_state(m, seed=12345):
    """RandomIC over every PD node in id order, then the Neumann presets."""
    u = np.zeros(2 * m.n)
    pxo.random_ic(seed, 0.49, 0.51, np.arange(1, m.n + 1), 2, [1], u)
    for w in range(4):
        for nid in m.group(w):
            b = m.mirror[nid - 1] - 1
            u[2 * (nid - 1)] = u[2 * b]
            u[2 * (nid - 1) + 1] = u[2 * b + 1]
    return u


def main(nx=256, out=os.path.join(HERE, "ch_picard_256.npz")):
    m = pxo.Mesh(pxo.make_grid(2, nx, nx))
    ops, _ = pxo.operators(m)
    lap, rev, isg = pxo.ch_geometry(m, ops)
    rp, cols = pxo.csr_pattern(m, 2)
    u0 = randomic_state(m)
    kappa = 1e-2 * (1.0 / nx) ** 2
    mn = pxo.ch_mobility(m, u0, 1.0)
    K, R = pxo.assemble_ch(m, lap, rev, isg, mn, kappa, 3.0, 1e-3, rp, cols, u0, u0)
    bcs = pxo.make_bcs([{"type": "neumann", "nodes": m.group(w), "comps": [0, 1]} for w in range(4)])
    rc, Kb, Rb = pxo.bc_rows(m, 2, bcs, 0.0, u0, rp, cols, K, R)
    assert rc == 0
    A = sp.csr_matrix((Kb, cols, rp), shape=(2 * m.n, 2 * m.n)).tocsc()
    t0 = time.time()
    du = sla.spsolve(A, Rb, permc_spec="COLAMD")
    res = np.abs(A @ du - Rb).max() / np.abs(Rb).max()
    print("solve %.1f s, relative residual %.2e, max |du| %.3e" % (time.time() - t0, res, np.abs(du).max()))
    assert res < 1e-8
    u1 = u0 + du
    np.savez_compressed(out, u1=u1, u0_sha256=np.frombuffer(hashlib.sha256(u0.tobytes()).digest(), np.uint8))
    print("wrote", out)


if __name__ == "__main__":
    main()
