"""torch.autograd wrapper of the adjoint: E.grad after backward() equals the
device adjoint gradient and agrees with central finite differences of the
frozen-health loss (h = 1e-6 relative) to 1e-4."""
import numpy as np
import pytest
import torch

from perix_amd import problems
from perix_amd.autograd import frac_loss

pytestmark = pytest.mark.gpu


def test_autograd_wrapper_matches_adjoint_and_fd():
    nx, ny, nsteps = 40, 20, 60
    hi = (0.02, 0.01, 1.0)
    rs = np.random.RandomState(11)
    kw = dict(hi=hi, sc=0.004, cracks=[(0.005, 0.0, 0.005, 0.005)],
              box=(0.008, 0.012, 1.0, -1.0, 1.0, -1.0), nsteps=nsteps)
    probe = problems.explicit_plate(2, nx, ny, **kw)
    n, nbulk = probe.cloud.n, probe.cloud.nbulk
    probe.ctx.close()
    E0 = problems.E_STEEL * (1.0 + 0.05 * rs.standard_normal(n))
    s = problems.explicit_plate(2, nx, ny, node_E=E0, **kw)
    sens = rs.choice(nbulk, 16, replace=False).astype(np.int32)
    E = torch.tensor(E0, dtype=torch.float64, requires_grad=True)
    J = frac_loss(E, s, sens, comp=1, ckpt=16)
    J.backward()
    g = E.grad.numpy()
    # the same numbers straight from the C ABI
    s.ctx.frac_init(s.u0, s.dt)
    s.ctx.frac_run(s.dt, s.total)
    J2, g2 = s.ctx.frac_adjoint(s.u0, s.dt, s.nsteps, sens, 1, ckpt=16)
    assert abs(float(J.detach()) - J2) <= 1e-14 * abs(J2) and np.array_equal(g, g2)
    for _ in range(3):
        v = rs.standard_normal(n)
        h = 1e-6 * E0 * v
        jp = s.ctx.frac_loss(s.u0, s.dt, s.nsteps, sens, 1, node_E=E0 + h)
        jm = s.ctx.frac_loss(s.u0, s.dt, s.nsteps, sens, 1, node_E=E0 - h)
        fd = (jp - jm) / 2.0
        ad = float(g @ h)
        print("fd %.6e ad %.6e" % (fd, ad))
        assert abs(fd - ad) <= 1e-4 * max(abs(fd), abs(ad))
    s.ctx.close()
