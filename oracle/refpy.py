"""ctypes driver of oracle/_ref/libperix_ref.so -- TEST INFRASTRUCTURE ONLY.

The library is the reference's own C++ built from its sources by
oracle/ref/Makefile plus a small harness.  It is used in this container to
write golden vectors (tests/golden/) and to check the C oracle; it is optional
on machines where it was not built.
"""
import ctypes as C
import json
import os
import tempfile
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "_ref", "libperix_ref.so")
_LIB = None
D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


def available():
    return os.path.exists(_PATH)


def lib():
    global _LIB
    if _LIB is None:
        _LIB = C.CDLL(_PATH)
        _LIB.ref_open.restype = C.c_void_p
        _LIB.ref_open.argtypes = [C.c_char_p]
        for f in ("ref_close", "ref_info", "ref_mesh", "ref_families", "ref_ops", "ref_csr",
                  "ref_initial_state", "ref_preset", "ref_assemble", "ref_assemble_parallel",
                  "ref_apply_bcs", "ref_projection", "ref_ch_geometry", "ref_assemble_step",
                  "ref_dynfrac_state", "ref_dynfrac_set", "ref_stressch_consts", "ref_fsch_state", "ref_projection_u", "ref_save_vtu",
                  "ref_explicit", "ref_frac_consts", "ref_dxyz"):
            getattr(_LIB, f).argtypes = None
        _LIB.ref_horizon.restype = C.c_double
        _LIB.ref_last_bc_seconds.restype = C.c_double
        _LIB.ref_explicit_timed.argtypes = [C.c_void_p, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double), D]
        _LIB.ref_group.restype = C.c_int
        _LIB.ref_default_solve.restype = C.c_int
        _LIB.ref_default_solve.argtypes = [C.c_void_p, D, D, C.c_int, D, I, C.POINTER(C.c_longlong)]
    return _LIB


_PXG_PATH = os.path.join(_HERE, "_ref", "libperix_ref_pxg.so")
_PXG_LIB = None


def pxg_available():
    return os.path.exists(_PXG_PATH)


def pxg_lib():
    """The reference built with PERIX_CUDA_ASSEMBLE over the repository's
    perix_cuda implementation (integration/perix_cuda_pxg.cpp, libpxg).  The
    HIP library is loaded through perix_amd first, so that torch and libpxg
    share one HIP runtime."""
    global _PXG_LIB
    if _PXG_LIB is None:
        from perix_amd import pxg
        pxg.load()
        _PXG_LIB = C.CDLL(_PXG_PATH)
        _PXG_LIB.ref_run.restype = C.c_longlong
        _PXG_LIB.ref_run.argtypes = [C.c_char_p, D, C.c_longlong, D, C.c_longlong,
                                     C.POINTER(C.c_longlong), C.POINTER(C.c_double)]
        _PXG_LIB.perix_cuda_pxg_stats.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        _PXG_LIB.perix_cuda_pxg_reset_stats.argtypes = [C.c_longlong]
        _PXG_LIB.ref_open.restype = C.c_void_p
        _PXG_LIB.ref_open.argtypes = [C.c_char_p]
        _PXG_LIB.ref_horizon.restype = C.c_double
        _PXG_LIB.ref_group.restype = C.c_int
        _PXG_LIB.ref_default_solve.restype = C.c_int
        _PXG_LIB.ref_default_solve.argtypes = [C.c_void_p, D, D, C.c_int, D, I, C.POINTER(C.c_longlong)]
        _PXG_LIB.ref_pxg_linear_solve.restype = C.c_int
        _PXG_LIB.ref_pxg_linear_solve.argtypes = [C.c_void_p, D, D, C.c_char_p, D]
    return _PXG_LIB


def deck_dofs(deck):
    """PD nodes x dofs of a structured deck: cells plus one ghost per boundary
    face (CreatePDMesh.cpp:679-725), times the DOFs block."""
    m = deck["Mesh"]
    if m["dim"] == 2:
        n = m["nx"] * m["ny"] + 2 * (m["nx"] + m["ny"])
    else:
        n = m["nx"] * m["ny"] * m["nz"] + 2 * (m["nx"] * m["ny"] + m["nx"] * m["nz"] + m["ny"] * m["nz"])
    return n * len(deck["DOFs"])


def run_job(deck, fault_at_call=0, timeout=900):
    """Runs a deck's whole job with the reference's own driver (PDProblem::run)
    in the PERIX_CUDA_ASSEMBLE build (JobSystem.assemble picks "openmp" or
    "cuda", the latter through the seam over libpxg), in a child process: the
    reference ends its process on a deck or solver error.  Returns U and a
    record: the explicit element's bond health (empty for other elements),
    seconds of run(), seam calls served by the device (seam_ok) and
    refused (seam_failed), and the assembly and linear-solve seconds and
    Newton iterations of the reference's own Newton log.  The deck and its
    output files live in a temporary directory."""
    import subprocess
    import sys
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "deck.json")
        with open(path, "w") as fh:
            json.dump(deck, fh)
        out = os.path.join(td, "u.npy")
        r = subprocess.run([sys.executable, "-m", "oracle.refpy", path, out, str(int(fault_at_call))],
                           cwd=os.path.dirname(_HERE), capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            tail = "\n".join(line for line in (r.stdout + r.stderr).splitlines() if "Error" in line)[-2000:]
            raise RuntimeError("reference job failed (rc %d): %s" % (r.returncode, tail))
        rec = json.loads(r.stdout.strip().splitlines()[-1])
        u = np.load(out)
        rec["health"] = np.load(out + ".health.npy")
    # the reference's Newton log: "[assemble=<s>s,solve=<s>s]" per iteration
    import re
    pairs = re.findall(r"assemble=\s*([0-9.eE+-]+)s,\s*solve=\s*([0-9.eE+-]+)s", r.stderr)
    rec["assemble_s"] = sum(float(a) for a, _ in pairs)
    rec["solve_s"] = sum(float(b) for _, b in pairs)
    rec["newton_iterations"] = len(pairs)
    return u, rec


def _job_main(deck_path, out_path, fault_at_call):
    """Child side of run_job."""
    import sys
    L = pxg_lib()
    L.perix_cuda_pxg_reset_stats(fault_at_call)
    with open(deck_path) as fh:
        cap = deck_dofs(json.load(fh))
    u = np.zeros(cap)
    hcap = 200_000_000
    health = np.zeros(0)
    sec = C.c_double()
    nh = C.c_longlong()
    with open(deck_path) as fh:
        explicit = "explicit_pddo_frac" in fh.read()
    if explicit:
        health = np.zeros(hcap // 8)
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)         # the reference's progress goes to stderr
    try:
        n = L.ref_run(deck_path.encode(), u.ctypes.data_as(D), cap,
                      health.ctypes.data_as(D) if explicit else None, health.size, C.byref(nh), C.byref(sec))
    finally:
        C.CDLL(None).fflush(None)      # the reference's buffered C stdio
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)
    if n != cap:
        raise SystemExit("ref_run: %d dofs, expected %d" % (n, cap))
    ok, failed = C.c_longlong(), C.c_longlong()
    L.perix_cuda_pxg_stats(C.byref(ok), C.byref(failed))
    np.save(out_path, u)
    np.save(out_path + ".health.npy", health[:nh.value])
    print(json.dumps({"seconds": sec.value, "seam_ok": ok.value, "seam_failed": failed.value}))


def _dp(a):
    return a.ctypes.data_as(D) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(I) if a is not None else None


def poisson_deck(nx, assemble):
    """Poisson on the nx x nx unit square, sigma = 1, scalar source
    2 pi^2 (PoissonElement.h:74), Dirichlet u = 0 on the four ghost walls,
    static Newton with the reference's default LDU solver; "assemble" is
    "openmp" (the reference's CPU assembler) or "cuda" (the seam)."""
    import math
    walls = ["left", "right", "bottom", "top"]
    return {"Mesh": mesh_block(2, nx, nx), "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
            "DOFs": ["u"], "ElmtSystem": {"p": {"type": "poisson", "params": {"sigma": 1.0, "f": 2.0 * math.pi ** 2}}},
            "BCSystem": {w: {"type": "dirichlet", "phygroup": w + "nodes_ghost", "dofs": ["u"], "value": 0.0}
                         for w in walls},
            "JobSystem": {"type": "static", "assemble": assemble}}


def ch_deck(nx, steps, assemble):
    """The Cahn-Hilliard spinodal loop on an nx x nx square: c0 = 0.5 + U(-0.01, 0.01)
    (RandomIC, seed 12345), chi = 3, kappa = 1e-2, M = 1, backward Euler
    dt = 1e-3 (fixed), Neumann c and mu on all ghost walls, Newton (the
    mobility Picard-frozen per iteration, max 20, abs 1e-9, rel 1e-8) with the
    default LDU solver."""
    walls = ["left", "right", "bottom", "top"]
    return {"Mesh": mesh_block(2, nx, nx), "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
            "DOFs": ["c", "mu"],
            "ElmtSystem": {"ch": {"type": "cahnhilliard", "params": {"chi": 3.0, "kappa": 1e-2, "M": 1.0}}},
            "BCSystem": {w: {"type": "neumann", "phygroup": w + "nodes_ghost", "dofs": ["c", "mu"],
                             "value": [0.0, 0.0]} for w in walls},
            "ICSystem": {"noisy": {"type": "random", "dof": "c", "min": 0.49, "max": 0.51, "seed": 12345},
                         "mu0": {"type": "constant", "dof": "mu", "value": 0.0}},
            "JobSystem": {"type": "transient", "assemble": assemble},
            "TimeStepping": {"dt": 1e-3, "total_time": steps * 1e-3},
            "NonlinearSolver": {"max_iters": 20, "abs_tol": 1e-9, "rel_tol": 1e-8},
            "OutputSystem": {"interval": 100000}}


def mesh_block(dim, nx, ny, nz=1, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    m = {"type": "perix", "dim": dim, "nx": nx, "ny": ny, "xmin": lo[0], "xmax": hi[0],
         "ymin": lo[1], "ymax": hi[1], "meshtype": "quad4" if dim == 2 else "hex8",
         "savemesh": False}
    if dim == 3:
        m.update({"nz": nz, "zmin": lo[2], "zmax": hi[2]})
    return m


class Case:
    """One reference problem built from a JSON deck."""

    def __init__(self, deck, pxg=False):
        """pxg: open the case in the PERIX_CUDA_ASSEMBLE build (libperix_ref_pxg.so),
        which also carries the libpxg linear-solver back-end."""
        self._L = pxg_lib() if pxg else lib()
        self._dir = tempfile.mkdtemp(prefix="pxref_")
        self.path = os.path.join(self._dir, "deck.json")
        with open(self.path, "w") as f:
            json.dump(deck, f, indent=1)
        L = self._L
        self.h = C.c_void_p(L.ref_open(self.path.encode()))
        info = np.zeros(6, np.int64)
        L.ref_info(self.h, info.ctypes.data_as(C.POINTER(C.c_longlong)))
        self.n, self.nb, self.dim, self.ndof, self.nnz, self.nbulk = (int(x) for x in info)
        self.delta = L.ref_horizon(self.h)

    def close(self):
        if self.h:
            self._L.ref_close(self.h)
            self.h = None

    def mesh(self):
        c = np.zeros(3 * self.n); v = np.zeros(self.n); m = np.zeros(self.n, np.int32)
        self._L.ref_mesh(self.h, _dp(c), _dp(v), _ip(m))
        return c, v, m

    def dxyz(self):
        d = np.zeros(3)
        self._L.ref_dxyz(self.h, _dp(d))
        return d

    def families(self):
        off = np.zeros(self.n + 1, np.int32); nbr = np.zeros(self.nb, np.int32)
        self._L.ref_families(self.h, _ip(off), _ip(nbr))
        return off, nbr

    def group(self, name):
        L = self._L
        k = L.ref_group(self.h, name.encode(), None)
        ids = np.zeros(k, np.int32)
        if k:
            L.ref_group(self.h, name.encode(), _ip(ids))
        return ids

    def ops(self, nop):
        o = np.zeros(self.nb * nop); fl = np.zeros(self.n, np.uint8)
        self._L.ref_ops(self.h, _dp(o), fl.ctypes.data_as(C.POINTER(C.c_ubyte)))
        return o.reshape(self.nb, nop), fl

    def csr(self):
        rp = np.zeros(self.n * self.ndof + 1, np.int32); cols = np.zeros(self.nnz, np.int32)
        self._L.ref_csr(self.h, _ip(rp), _ip(cols))
        return rp, cols

    def initial_state(self):
        u = np.zeros(self.n * self.ndof); v = np.zeros(self.n * self.ndof)
        self._L.ref_initial_state(self.h, _dp(u), _dp(v))
        return u, v

    def preset(self, u, t):
        u = np.ascontiguousarray(u, float).copy()
        self._L.ref_preset(self.h, _dp(u), C.c_double(t))
        return u

    def assemble(self, u, uold, dt, parallel=False):
        """formResidualAndJacobian (serial) or its OpenMP twin with `parallel`."""
        k = np.zeros(self.nnz); r = np.zeros(self.n * self.ndof)
        f = self._L.ref_assemble_parallel if parallel else self._L.ref_assemble
        f(self.h, _dp(np.ascontiguousarray(u, float)), _dp(np.ascontiguousarray(uold, float)),
          C.c_double(dt), _dp(k), _dp(r))
        return k, r

    def apply_bcs(self, u, t, k, r):
        """BCSystem::applyBCs (row elimination) on copies of K values and RHS."""
        k = np.ascontiguousarray(k, float).copy(); r = np.ascontiguousarray(r, float).copy()
        self._L.ref_apply_bcs(self.h, _dp(np.ascontiguousarray(u, float)), C.c_double(t), _dp(k), _dp(r))
        return k, r

    def default_solve(self, k, r, use_rcm=True):
        """The reference's DefaultSolver on the case's CSR pattern with values k
        and right-hand side r: returns (rc, x, perm, jp) -- rc 1 on a zero
        pivot, perm new -> old, jp the skyline pointer."""
        n = self.n * self.ndof
        x = np.zeros(n); perm = np.zeros(n, np.int32); jp = np.zeros(n + 1, np.int64)
        rc = self._L.ref_default_solve(self.h, _dp(np.ascontiguousarray(k, float)),
                                     _dp(np.ascontiguousarray(r, float)), int(use_rcm), _dp(x), _ip(perm),
                                     jp.ctypes.data_as(C.POINTER(C.c_longlong)))
        return rc, x, perm, jp

    def pxg_linear_solve(self, k, r, params="{}"):
        """PxgLinearSolver (integration/pxg_linear_solver.cpp), the reference's
        LinearSolverBase over libpxg, on the case's SparseMatrix with values k:
        returns (ok, x).  Needs a case opened with pxg=True."""
        n = self.n * self.ndof
        x = np.zeros(n)
        ok = self._L.ref_pxg_linear_solve(self.h, _dp(np.ascontiguousarray(k, float)),
                                          _dp(np.ascontiguousarray(r, float)), params.encode(), _dp(x))
        return bool(ok), x

    def last_bc_seconds(self):
        """Wall time of the applyBCs call alone in the last apply_bcs."""
        return self._L.ref_last_bc_seconds()

    def projection(self, name):
        """computeNodalProjection of every node (n x 1 or n x 9)."""
        nc = 9 if name in ("stress", "strain") else 1
        out = np.zeros(self.n * nc)
        self._L.ref_projection(self.h, name.encode(), _dp(out))
        return out.reshape(self.n, nc)

    def projection_u(self, name, u, nc):
        """computeNodalProjection of every node on the state u (n x nc)."""
        out = np.zeros(self.n * nc)
        self._L.ref_projection_u(self.h, name.encode(), _dp(np.ascontiguousarray(u, float)), nc, _dp(out))
        return out.reshape(self.n, nc)

    def save_vtu(self, u, stem, step):
        """PDProblem::savePDResults of the state u; returns the written path."""
        self._L.ref_save_vtu(self.h, _dp(np.ascontiguousarray(u, float)), stem.encode(), step)
        return "%s-PD-results-%05d.vtu" % (stem, step) if step > 0 else "%s-PD-results.vtu" % stem

    def assemble_step(self, u, uold, dt, timestep):
        """formResidualAndJacobian as time step `timestep` (per-step commits)."""
        k = np.zeros(self.nnz); r = np.zeros(self.n * self.ndof)
        self._L.ref_assemble_step(self.h, _dp(np.ascontiguousarray(u, float)),
                                _dp(np.ascontiguousarray(uold, float)), C.c_double(dt), C.c_int(timestep),
                                _dp(k), _dp(r))
        return k, r

    def dynfrac_state(self):
        h = np.zeros(self.nb); v = np.zeros(self.n * self.ndof)
        self._L.ref_dynfrac_state(self.h, _dp(h), _dp(v))
        return h, v

    def dynfrac_set(self, b, v0):
        self._L.ref_dynfrac_set(self.h, _dp(np.ascontiguousarray(b, float)), _dp(np.ascontiguousarray(v0, float)))

    def stressch_consts(self):
        """C11, C12, C66, A, Kh, Ch, D, kappa, chi, Omega, cref of the element."""
        out = np.zeros(11)
        self._L.ref_stressch_consts(self.h, _dp(out))
        return out

    def fsch_state(self, dim):
        """FracStressCahnHilliardElement: (health, vel, lapc, constants)."""
        h = np.zeros(self.nb); v = np.zeros(self.n * dim); lc = np.zeros(self.nb); k = np.zeros(14)
        self._L.ref_fsch_state(self.h, _dp(h), _dp(v), _dp(lc), _dp(k))
        return h, v, lc, k

    def ch_geometry(self):
        lap = np.zeros(self.nb); rev = np.zeros(self.nb, np.int32); mn = np.zeros(self.n)
        self._L.ref_ch_geometry(self.h, _dp(lap), _ip(rev), _dp(mn))
        return lap, rev, mn

    def frac_consts(self):
        o = np.zeros(5)
        self._L.ref_frac_consts(self.h, _dp(o))
        return o

    def explicit_timed(self):
        """(steps, one-off setup seconds, loop seconds, final u)."""
        u = np.zeros(self.n * self.ndof)
        a = C.c_double(); b = C.c_double()
        steps = self._L.ref_explicit_timed(self.h, C.byref(a), C.byref(b), _dp(u))
        return steps, a.value, b.value, u

    def explicit(self):
        nd = self.n * self.ndof
        nops = 9 if self.dim == 3 else 5
        u = np.zeros(nd); uo = np.zeros(nd); h = np.zeros(self.nb); opc = np.zeros(self.nb * nops)
        sc = C.c_double()
        steps = self._L.ref_explicit(self.h, _dp(u), _dp(uo), _dp(h), _dp(opc), C.byref(sc))
        return steps, u, uo, h, opc.reshape(self.nb, nops), sc.value


if __name__ == "__main__":
    import sys
    _job_main(sys.argv[1], sys.argv[2], int(sys.argv[3]))
