"""ctypes binding of libpxg.so (the C ABI in include/pxg.h).

Every call goes to the HIP library; a missing library or a failing device
call raises PxgError -- there is no host fallback in this package.
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
FP64 = 0
FP32C = 1
ERR_SINGULAR = -7      # zero pivot in the direct solve
SOLVE_NO_GRAPHS = 1
_LIB = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
U8 = C.POINTER(C.c_ubyte)


class PxgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pxg error %d: %s" % (code, msg))
        self.code = code


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
                ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double)]


class FracParams(C.Structure):
    _fields_ = [("dim", C.c_int), ("C11", C.c_double), ("C12", C.c_double),
                ("C66", C.c_double), ("rho", C.c_double), ("bx", C.c_double),
                ("by", C.c_double), ("bz", C.c_double), ("sc", C.c_double),
                ("tension_only", C.c_int), ("damage_on", C.c_int)]


# PXG_SCH_* projection codes of include/pxg.h
SCH_PROJECTIONS = {"strain": 0, "stress": 1, "vonMises": 2, "sigmaH": 3, "fPrime": 4, "fDoublePrime": 5,
                   "gradC": 6, "damage": 7, "DamageI": 8}


class StressChParams(C.Structure):
    """pxg_stressch_params: the stress_cahnhilliard deck parameters."""
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("D", C.c_double),
                ("Omega", C.c_double), ("cref", C.c_double), ("chi", C.c_double),
                ("kappa", C.c_double), ("plane_stress", C.c_int)]


class FschParams(C.Structure):
    """pxg_fsch_params: the frac_stress_cahnhilliard deck parameters, with
    the critical stretch evaluated from Gc by the caller."""
    _fields_ = [("base", StressChParams), ("rho", C.c_double), ("sc", C.c_double),
                ("residual_stiffness", C.c_double), ("tension_only", C.c_int),
                ("damage_on", C.c_int)]


class SchState(C.Structure):
    """pxg_sch_state: a fracture stress-CH element's host-side state."""
    _fields_ = [(k, C.c_double) for k in ("C11", "C12", "C66", "A", "Kh", "Ch", "Omega", "cref", "chi",
                                          "kappa", "rho", "kres")] + \
               [("rev", C.POINTER(C.c_int)), ("lap", C.POINTER(C.c_double)), ("isghost", C.POINTER(C.c_ubyte)),
                ("mnode", C.POINTER(C.c_double)), ("health", C.POINTER(C.c_double)),
                ("vel", C.POINTER(C.c_double))]


class DynFracState(C.Structure):
    """pxg_dynfrac_state: a dynamic fracture element's host-side state."""
    _fields_ = [(k, C.c_double) for k in ("C11", "C12", "C66", "rho", "bx", "by", "bz")] + \
               [("vel", C.POINTER(C.c_double)), ("health", C.POINTER(C.c_double))]


class BC(C.Structure):
    _fields_ = [("type", C.c_int), ("direct", C.c_int), ("ncomp", C.c_int),
                ("comp", C.c_int * 3), ("value", C.c_double * 3),
                ("velocity", C.c_double * 3), ("has_box", C.c_int),
                ("box", C.c_double * 6), ("nnodes", C.c_int), ("nodes", I)]


# every symbol of include/pxg.h (checked by the ABI test)
SYMBOLS = [
    "pxg_create", "pxg_destroy", "pxg_last_error", "pxg_set_fault", "pxg_sync",
    "pxg_structured_sizes", "pxg_structured_cloud", "pxg_horizon_radius", "pxg_random_ic",
    "pxg_schedule", "pxg_set_points", "pxg_build_families", "pxg_get_families", "pxg_set_families",
    "pxg_set_csr", "pxg_ch_set_geometry", "pxg_ch_set_mobility", "pxg_assemble_sch_state",
    "pxg_assemble_dynfrac_state",
    "pxg_set_cracks", "pxg_get_seed", "pxg_build_csr", "pxg_get_csr", "pxg_build_operators",
    "pxg_get_operators", "pxg_assemble_poisson", "pxg_assemble_poisson_dev", "pxg_assemble_poisson_noalloc",
    "pxg_assemble_diffusion", "pxg_assemble_diffusion_dev", "pxg_ch_setup",
    "pxg_ch_get_geometry", "pxg_assemble_ch", "pxg_assemble_ch_dev", "pxg_frac_setup",
    "pxg_frac_set_bcs", "pxg_csr_set_bcs", "pxg_csr_bc_rows", "pxg_csr_apply_bcs",
    "pxg_csr_apply_bcs_dev", "pxg_solve_set_pattern", "pxg_solve_options", "pxg_solve_factor_dev", "pxg_solve_apply_dev",
    "pxg_solve_dev", "pxg_solve", "pxg_solve_info", "pxg_solve_ordering", "pxg_csr_preset_dev",
    "pxg_newton_update_dev", "pxg_dynfrac_setup", "pxg_dynfrac_commit", "pxg_dynfrac_commit_dev",
    "pxg_assemble_dynfrac", "pxg_assemble_dynfrac_dev", "pxg_dynfrac_get_state",
    "pxg_stressch_setup", "pxg_stressch_constants", "pxg_assemble_stressch", "pxg_assemble_stressch_dev",
    "pxg_fsch_setup", "pxg_fsch_constants", "pxg_fsch_commit", "pxg_fsch_commit_dev",
    "pxg_assemble_fsch", "pxg_assemble_fsch_dev", "pxg_fsch_get_state",
    "pxg_sch_components", "pxg_stressch_project", "pxg_stressch_project_dev", "pxg_fsch_project",
    "pxg_fsch_project_dev",
    "pxg_write_vtu",
    "pxg_frac_options", "pxg_frac_accel", "pxg_frac_init", "pxg_frac_run",
    "pxg_frac_project", "pxg_frac_project_dev", "pxg_frac_layout", "pxg_frac_get_state",
    "pxg_frac_state_ptrs", "pxg_frac_profile", "pxg_frac_kernel_time",
    "pxg_frac_adjoint", "pxg_frac_adjoint_ex", "pxg_frac_set_body", "pxg_frac_loss",
]


def vtu_path(deck_path, step):
    """The reference's result file name for a deck and output step
    (SaveResults.cpp:17-60): <deck without .json>-PD-results[-<step:05>].vtu."""
    stem = deck_path[:-5] if deck_path.endswith(".json") else deck_path
    return stem + "-PD-results" + ("-%05d" % step if step > 0 else "") + ".vtu"


def write_vtu(path, coords, nbulk, dof_names, u, fields=()):
    """Binary VTU of PD results (pxg_write_vtu): coords (3n), the first nbulk
    nodes bulk, u (n * len(dof_names)) and fields = [(name, n x comps array)]."""
    coords = np.ascontiguousarray(coords, float)
    n = coords.size // 3
    ndof = len(dof_names)
    u = np.ascontiguousarray(u, float)
    if u.size != n * ndof:
        raise ValueError("write_vtu: u must hold n * ndof values")
    dn = (C.c_char_p * max(ndof, 1))(*[d.encode() for d in dof_names])
    arrs = [np.ascontiguousarray(a, float).reshape(n, -1) for _, a in fields]
    fn = (C.c_char_p * max(len(arrs), 1))(*[nm.encode() for nm, _ in fields])
    fc = np.array([a.shape[1] for a in arrs] or [0], np.int32)
    fd = (D * max(len(arrs), 1))(*[a.ctypes.data_as(D) for a in arrs])
    _check(load().pxg_write_vtu(path.encode(), n, int(nbulk), _dp(coords), ndof, dn, _dp(u), len(arrs), fn,
                                _ip(fc), fd))
    return path


def lib_path():
    return os.path.join(_HERE, "libpxg.so")


def load():
    """Loads libpxg.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not os.path.exists(p):
            raise PxgError(-99, "libpxg.so not built (run python perix_amd/build.py)")
        # PyTorch-ROCm ships its own HIP runtime under the same soname.  When
        # libpxg is loaded first, torch's later import brings a second runtime
        # into the process and that one finds no GPU; with torch loaded first,
        # libpxg binds to torch's runtime and both share the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(p)
        L.pxg_last_error.restype = C.c_char_p
        L.pxg_horizon_radius.restype = C.c_double
        L.pxg_destroy.restype = None
        L.pxg_random_ic.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_int, I, C.c_int,
                                    C.c_int, I, D]
        for name in ("pxg_build_operators",):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, I]
        L.pxg_build_families.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_longlong)]
        L.pxg_build_csr.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_longlong)]
        L.pxg_assemble_poisson.argtypes = [C.c_void_p, C.c_double, D, C.c_double, D, D, D]
        L.pxg_assemble_poisson_dev.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_double,
                                               C.c_void_p, C.c_void_p, C.c_void_p]
        L.pxg_assemble_sch_state.argtypes = [C.c_void_p, C.c_void_p, C.c_double, D, D, D, D]
        L.pxg_assemble_dynfrac_state.argtypes = [C.c_void_p, C.c_void_p, C.c_double, D, D, D, D]
        L.pxg_assemble_poisson_noalloc.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_double,
                                                   C.c_void_p, C.c_void_p, C.c_void_p]
        L.pxg_assemble_diffusion.argtypes = [C.c_void_p, C.c_double, D, C.c_double, C.c_double,
                                             D, D, D, D]
        L.pxg_assemble_diffusion_dev.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_double,
                                                 C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p]
        L.pxg_assemble_ch.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                      C.c_double, D, D, D, D]
        L.pxg_assemble_ch_dev.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
        L.pxg_frac_setup.argtypes = [C.c_void_p, C.POINTER(FracParams), C.c_int, D, C.c_double,
                                     C.c_int]
        L.pxg_csr_apply_bcs.argtypes = [C.c_void_p, C.c_double, D, D, D]
        L.pxg_csr_apply_bcs_dev.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                            C.c_void_p]
        L.pxg_frac_init.argtypes = [C.c_void_p, D, D, C.c_double]
        L.pxg_dynfrac_setup.argtypes = [C.c_void_p, C.POINTER(FracParams), D, D]
        L.pxg_dynfrac_commit.argtypes = [C.c_void_p, D, C.c_double]
        L.pxg_dynfrac_commit_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
        L.pxg_assemble_dynfrac.argtypes = [C.c_void_p, C.c_double, D, D, D, D]
        L.pxg_assemble_dynfrac_dev.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p]
        L.pxg_dynfrac_get_state.argtypes = [C.c_void_p, U8, D]
        L.pxg_fsch_setup.argtypes = [C.c_void_p, C.POINTER(FschParams), D]
        L.pxg_fsch_constants.argtypes = [C.c_void_p, D]
        L.pxg_fsch_commit.argtypes = [C.c_void_p, D, C.c_double]
        L.pxg_fsch_commit_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
        L.pxg_assemble_fsch.argtypes = [C.c_void_p, C.c_double, D, D, D, D]
        L.pxg_assemble_fsch_dev.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p]
        L.pxg_fsch_get_state.argtypes = [C.c_void_p, U8, D]
        for nm in ("pxg_stressch_project", "pxg_fsch_project"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_int, D, D]
            getattr(L, nm + "_dev").argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        CP = C.POINTER(C.c_char_p)
        L.pxg_write_vtu.argtypes = [C.c_char_p, C.c_int, C.c_int, D, C.c_int, CP, D, C.c_int, CP, I,
                                    C.POINTER(D)]
        L.pxg_stressch_setup.argtypes = [C.c_void_p, C.POINTER(StressChParams)]
        L.pxg_stressch_constants.argtypes = [C.c_void_p, D]
        L.pxg_assemble_stressch.argtypes = [C.c_void_p, C.c_double, D, D, D, D]
        L.pxg_assemble_stressch_dev.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_void_p]
        L.pxg_frac_run.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, I]
        L.pxg_frac_adjoint.argtypes = [C.c_void_p, D, D, C.c_double, C.c_int, C.c_int, C.c_int,
                                       I, C.c_int, D, D]
        L.pxg_frac_loss.argtypes = [C.c_void_p, D, D, C.c_double, C.c_int, D, C.c_int, I,
                                    C.c_int, D]
        L.pxg_frac_adjoint_ex.argtypes = [C.c_void_p, D, D, C.c_double, C.c_int, C.c_int, C.c_int,
                                          I, C.c_int, D, D, D]
        L.pxg_frac_set_body.argtypes = [C.c_void_p, D]
        L.pxg_frac_project.argtypes = [C.c_void_p, C.c_int, D]
        L.pxg_frac_project_dev.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.pxg_frac_state_ptrs.argtypes = [C.c_void_p, C.POINTER(C.c_void_p),
                                          C.POINTER(C.c_void_p)]
        L.pxg_solve_options.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        L.pxg_solve_set_pattern.argtypes = [C.c_void_p, C.c_int, I, I]
        L.pxg_solve_factor_dev.argtypes = [C.c_void_p, C.c_void_p]
        L.pxg_solve_apply_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.pxg_solve_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.pxg_solve.argtypes = [C.c_void_p, D, D, D]
        L.pxg_solve_info.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
        L.pxg_solve_ordering.argtypes = [C.c_void_p, I, C.POINTER(C.c_longlong)]
        L.pxg_csr_preset_dev.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
        L.pxg_newton_update_dev.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, D]
        _LIB = L
    return _LIB


def available_devices():
    """Number of HIP devices (0 when none), without creating a context."""
    try:
        import torch
        return torch.cuda.device_count()
    except Exception:
        return 0


def _dp(a):
    return None if a is None else a.ctypes.data_as(D)


def _ip(a):
    return None if a is None else a.ctypes.data_as(I)


def _up(a):
    return None if a is None else a.ctypes.data_as(U8)


def _check(rc):
    if rc != 0:
        raise PxgError(rc, load().pxg_last_error().decode(errors="replace"))


def grid_struct(g):
    return Grid(g.dim, g.nx, g.ny, g.nz if g.dim == 3 else 0, g.lo[0], g.hi[0], g.lo[1],
                g.hi[1], g.lo[2] if g.dim == 3 else 0.0, g.hi[2] if g.dim == 3 else 0.0)


def make_bcs(specs):
    """specs: dicts {type: 'neumann'|'dirichlet', nodes (1-based), comps (0-based),
    value, velocity, direct, box=(xlo,xhi,ylo,yhi,zlo,zhi)}."""
    arr = (BC * max(1, len(specs)))()
    keep = []
    for e, s in enumerate(specs):
        b = arr[e]
        b.type = 0 if s["type"] == "neumann" else 1
        b.direct = int(bool(s.get("direct", False)))
        comps = list(s["comps"])
        b.ncomp = len(comps)
        val = list(s.get("value", [0.0] * len(comps)))
        vel = list(s.get("velocity", [0.0] * len(comps)))
        if len(val) == 1:
            val = val * len(comps)
        if len(vel) == 1:
            vel = vel * len(comps)
        for k, c in enumerate(comps):
            b.comp[k] = c
            b.value[k] = val[k]
            b.velocity[k] = vel[k]
        box = s.get("box")
        b.has_box = int(box is not None)
        for k in range(6):
            b.box[k] = box[k] if box is not None else (1.0 if k % 2 == 0 else -1.0)
        nodes = np.ascontiguousarray(s["nodes"], np.int32)
        keep.append(nodes)
        b.nnodes = len(nodes)
        b.nodes = _ip(nodes)
    return arr, len(specs), keep


class Context:
    """One PD problem on one GPU (pxg_ctx)."""

    def __init__(self, device=0):
        L = load()
        h = C.c_void_p()
        _check(L.pxg_create(C.byref(h), device))
        self.h = h
        self.device = device
        self.n = 0
        self.dim = 0
        self.nb = 0
        self.nop = 0
        self.nnz = 0
        self.ndof = 0
        # bumped by every call that changes the explicit state (set-up, init,
        # run, per-step accel): lets a caller tell whether the state it
        # recorded is still the context's
        self.frac_serial = 0

    def close(self):
        if self.h:
            load().pxg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fault(self, code):
        _check(load().pxg_set_fault(self.h, code))

    def sync(self):
        _check(load().pxg_sync(self.h))

    # -- a1-a4 ---------------------------------------------------------------
    def set_points(self, coords, vols, mirror, dxyz, dim):
        coords = np.ascontiguousarray(coords, float)
        vols = np.ascontiguousarray(vols, float)
        mirror = np.ascontiguousarray(mirror, np.int32)
        dxyz = np.ascontiguousarray(dxyz, float)
        self.n = len(vols)
        self.dim = dim
        self.nb = self.nop = self.nnz = self.ndof = 0
        _check(load().pxg_set_points(self.h, self.n, dim, _dp(coords), _dp(vols), _ip(mirror),
                                     _dp(dxyz)))

    def build_families(self, r):
        nb = C.c_longlong()
        _check(load().pxg_build_families(self.h, C.c_double(r), C.byref(nb)))
        self.nb = nb.value
        return self.nb

    def set_families(self, off, nbr):
        """pxg_set_families: the caller's own flat families (1-based ids)."""
        off = np.ascontiguousarray(off, np.int32)
        nbr = np.ascontiguousarray(nbr, np.int32)
        _check(load().pxg_set_families(self.h, _ip(off), _ip(nbr)))
        self.nb = int(off[-1])
        self.nop = self.nnz = 0

    def set_csr(self, ndof, rowptr, cols):
        """pxg_set_csr: the caller's own CSR pattern."""
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        _check(load().pxg_set_csr(self.h, int(ndof), _ip(rowptr), _ip(cols)))
        self.nnz = int(rowptr[-1])
        self.ndof = ndof

    def ch_set_geometry(self, lap, rev, isghost=None):
        lap = np.ascontiguousarray(lap, float)
        rev = np.ascontiguousarray(rev, np.int32)
        g = None if isghost is None else np.ascontiguousarray(isghost, np.uint8)
        _check(load().pxg_ch_set_geometry(self.h, _dp(lap), _ip(rev), _up(g)))

    def ch_set_mobility(self, mnode):
        m = None if mnode is None else np.ascontiguousarray(mnode, float)
        _check(load().pxg_ch_set_mobility(self.h, _dp(m)))

    def assemble_sch_state(self, consts, dt, u, uold, rev, lap, mnode, isghost=None, health=None, vel=None):
        """pxg_assemble_sch_state; consts: dict of the 12 pxg_sch_state constants."""
        keep = [np.ascontiguousarray(rev, np.int32), np.ascontiguousarray(lap, float),
                None if isghost is None else np.ascontiguousarray(isghost, np.uint8),
                np.ascontiguousarray(mnode, float),
                None if health is None else np.ascontiguousarray(health, float),
                None if vel is None else np.ascontiguousarray(vel, float)]
        st = SchState(**consts)
        st.rev, st.lap, st.isghost = _ip(keep[0]), _dp(keep[1]), _up(keep[2])
        st.mnode, st.health, st.vel = _dp(keep[3]), _dp(keep[4]), _dp(keep[5])
        k = np.zeros(self.nnz)
        r = np.zeros(self.n * (self.dim + 2))
        _check(load().pxg_assemble_sch_state(self.h, C.byref(st), dt, _dp(np.ascontiguousarray(u, float)),
                                             _dp(np.ascontiguousarray(uold, float)), _dp(k), _dp(r)))
        return k, r

    def assemble_dynfrac_state(self, consts, dt, u, uold, vel, health=None):
        keep = [np.ascontiguousarray(vel, float), None if health is None else np.ascontiguousarray(health, float)]
        st = DynFracState(**consts)
        st.vel, st.health = _dp(keep[0]), _dp(keep[1])
        k = np.zeros(self.nnz)
        r = np.zeros(self.n * self.dim)
        _check(load().pxg_assemble_dynfrac_state(self.h, C.byref(st), dt, _dp(np.ascontiguousarray(u, float)),
                                                 _dp(np.ascontiguousarray(uold, float)), _dp(k), _dp(r)))
        return k, r

    def families(self):
        off = np.zeros(self.n + 1, np.int32)
        nbr = np.zeros(self.nb, np.int32)
        _check(load().pxg_get_families(self.h, _ip(off), _ip(nbr)))
        return off, nbr

    def set_cracks(self, segs=(), quads=()):
        seg = np.ascontiguousarray(np.array(segs, float).reshape(-1))
        quad = np.ascontiguousarray(np.array(quads, float).reshape(-1))
        _check(load().pxg_set_cracks(self.h, len(segs), _dp(seg) if len(segs) else None,
                                     len(quads), _dp(quad) if len(quads) else None))

    def seed(self):
        s = np.zeros(self.nb, np.uint8)
        _check(load().pxg_get_seed(self.h, _up(s)))
        return s

    def build_csr(self, ndof):
        nnz = C.c_longlong()
        _check(load().pxg_build_csr(self.h, ndof, C.byref(nnz)))
        self.nnz = nnz.value
        self.ndof = ndof
        return self.nnz

    def csr(self):
        rp = np.zeros(self.n * self.ndof + 1, np.int32)
        cols = np.zeros(self.nnz, np.int32)
        _check(load().pxg_get_csr(self.h, _ip(rp), _ip(cols)))
        return rp, cols

    # -- a5-a8 ---------------------------------------------------------------
    def build_operators(self, delta, dx_char, order=2):
        nd = C.c_int()
        _check(load().pxg_build_operators(self.h, order, C.c_double(delta), C.c_double(dx_char),
                                          C.byref(nd)))
        self.nop = 5 if self.dim == 2 else 9
        return nd.value

    def operators(self):
        ops = np.zeros(self.nb * self.nop)
        fl = np.zeros(self.n, np.uint8)
        _check(load().pxg_get_operators(self.h, _dp(ops), _up(fl)))
        return ops.reshape(self.nb, self.nop), fl

    # -- a10-a13 ---------------------------------------------------------------
    def assemble_poisson(self, sigma, u, f_node=None, f_const=0.0):
        k = np.zeros(self.nnz)
        r = np.zeros(self.n)
        fn = None if f_node is None else np.ascontiguousarray(f_node, float)
        _check(load().pxg_assemble_poisson(self.h, sigma, _dp(fn), f_const,
                                           _dp(np.ascontiguousarray(u, float)), _dp(k), _dp(r)))
        return k, r

    def assemble_poisson_dev(self, sigma, u_ptr, k_ptr, r_ptr, f_ptr=None, f_const=0.0):
        _check(load().pxg_assemble_poisson_dev(self.h, sigma, f_ptr, f_const, u_ptr, k_ptr, r_ptr))

    def assemble_poisson_noalloc(self, sigma, u_ptr, k_ptr, r_ptr, f_ptr=None, f_const=0.0):
        """pxg_assemble_poisson_noalloc: device pointers, no allocation."""
        _check(load().pxg_assemble_poisson_noalloc(self.h, sigma, f_ptr, f_const, u_ptr, k_ptr, r_ptr))

    def assemble_diffusion(self, D, dt, u, uold, f_node=None, f_const=0.0):
        k = np.zeros(self.nnz)
        r = np.zeros(self.n)
        fn = None if f_node is None else np.ascontiguousarray(f_node, float)
        _check(load().pxg_assemble_diffusion(self.h, D, _dp(fn), f_const, dt,
                                             _dp(np.ascontiguousarray(u, float)),
                                             _dp(np.ascontiguousarray(uold, float)), _dp(k), _dp(r)))
        return k, r

    def assemble_diffusion_dev(self, D, dt, u_ptr, uold_ptr, k_ptr, r_ptr, f_ptr=None, f_const=0.0):
        _check(load().pxg_assemble_diffusion_dev(self.h, D, f_ptr, f_const, dt, u_ptr, uold_ptr, k_ptr,
                                                 r_ptr))

    def ch_setup(self):
        _check(load().pxg_ch_setup(self.h))

    def ch_geometry(self):
        lap = np.zeros(self.nb)
        rev = np.zeros(self.nb, np.int32)
        isg = np.zeros(self.n, np.uint8)
        _check(load().pxg_ch_get_geometry(self.h, _dp(lap), _ip(rev), _up(isg)))
        return lap, rev, isg

    def assemble_ch(self, kappa, chi, M0, dt, u, uold):
        k = np.zeros(self.nnz)
        r = np.zeros(2 * self.n)
        _check(load().pxg_assemble_ch(self.h, kappa, chi, M0, dt,
                                      _dp(np.ascontiguousarray(u, float)),
                                      _dp(np.ascontiguousarray(uold, float)), _dp(k), _dp(r)))
        return k, r

    def assemble_ch_dev(self, kappa, chi, M0, dt, u_ptr, uold_ptr, k_ptr, r_ptr):
        _check(load().pxg_assemble_ch_dev(self.h, kappa, chi, M0, dt, u_ptr, uold_ptr, k_ptr,
                                          r_ptr))

    # -- BC row elimination of the implicit system -------------------------------
    def csr_set_bcs(self, specs):
        """Resolves a BC list (make_bcs specs, components index the CSR dofs)
        against the current CSR pattern; returns the number of constrained rows."""
        arr, n, keep = make_bcs(specs)
        _check(load().pxg_csr_set_bcs(self.h, n, arr))
        cnt = C.c_int()
        _check(load().pxg_csr_bc_rows(self.h, C.byref(cnt)))
        self.csr_set_bcs_count = cnt.value
        return cnt.value

    def csr_apply_bcs(self, t, u, k, r):
        """BCSystem::applyBCs on copies of the host arrays k (nnz), r (rows)."""
        k = np.ascontiguousarray(k, float).copy()
        r = np.ascontiguousarray(r, float).copy()
        _check(load().pxg_csr_apply_bcs(self.h, t, _dp(np.ascontiguousarray(u, float)), _dp(k),
                                        _dp(r)))
        return k, r

    def csr_apply_bcs_dev(self, t, u_ptr, k_ptr, r_ptr):
        _check(load().pxg_csr_apply_bcs_dev(self.h, t, u_ptr, k_ptr, r_ptr))

    # -- direct solve of the implicit system (DefaultSolver on the device) -------
    def solve_options(self, use_rcm=True, max_profile_gib=16.0, graphs=True):
        _check(load().pxg_solve_options(self.h, int(use_rcm), max_profile_gib, 0 if graphs else SOLVE_NO_GRAPHS))

    def solve_set_pattern(self, rowptr, cols):
        """A square CSR pattern of the solver's own (any system assembled elsewhere)."""
        rp = np.ascontiguousarray(rowptr, np.int32)
        _check(load().pxg_solve_set_pattern(self.h, len(rp) - 1, _ip(rp), _ip(np.ascontiguousarray(cols, np.int32))))

    def solve(self, k, r):
        """K x = r for host values k (nnz, BC rows applied) and r (rows)."""
        x = np.zeros(len(r))
        _check(load().pxg_solve(self.h, _dp(np.ascontiguousarray(k, float)),
                                _dp(np.ascontiguousarray(r, float)), _dp(x)))
        return x

    def solve_dev(self, k_ptr, r_ptr, x_ptr):
        _check(load().pxg_solve_dev(self.h, k_ptr, r_ptr, x_ptr))

    def solve_factor_dev(self, k_ptr):
        _check(load().pxg_solve_factor_dev(self.h, k_ptr))

    def solve_apply_dev(self, r_ptr, x_ptr):
        _check(load().pxg_solve_apply_dev(self.h, r_ptr, x_ptr))

    def solve_info(self):
        out = np.zeros(8, np.int64)
        _check(load().pxg_solve_info(self.h, out.ctypes.data_as(C.POINTER(C.c_longlong))))
        keys = ("equations", "blocks", "block", "profile", "strip_blocks", "max_active", "bytes", "block_products")
        return dict(zip(keys, (int(v) for v in out)))

    def solve_ordering(self):
        n = self.solve_info()["equations"]
        perm = np.zeros(n, np.int32); jp = np.zeros(n + 1, np.int64)
        _check(load().pxg_solve_ordering(self.h, _ip(perm), jp.ctypes.data_as(C.POINTER(C.c_longlong))))
        return perm, jp

    def csr_preset_dev(self, t, u_ptr):
        _check(load().pxg_csr_preset_dev(self.h, t, u_ptr))

    def newton_update_dev(self, t, u_ptr, du_ptr, r_ptr):
        """u += du and the presets at t when |du| is finite; returns (|r|, |du|)."""
        nrm = np.zeros(2)
        _check(load().pxg_newton_update_dev(self.h, t, u_ptr, du_ptr, r_ptr, _dp(nrm)))
        return float(nrm[0]), float(nrm[1])

    # -- implicit dynamic fracture element ---------------------------------------
    def dynfrac_setup(self, params, uold0, v0=None):
        v = None if v0 is None else np.ascontiguousarray(list(v0) + [0.0] * (3 - len(v0)), float)
        _check(load().pxg_dynfrac_setup(self.h, C.byref(params), _dp(v),
                                        _dp(np.ascontiguousarray(uold0, float))))

    def dynfrac_commit(self, uold, dt_prev):
        _check(load().pxg_dynfrac_commit(self.h, _dp(np.ascontiguousarray(uold, float)), dt_prev))

    def assemble_dynfrac(self, dt, u, uold):
        k = np.zeros(self.nnz)
        r = np.zeros(self.n * self.dim)
        _check(load().pxg_assemble_dynfrac(self.h, dt, _dp(np.ascontiguousarray(u, float)),
                                           _dp(np.ascontiguousarray(uold, float)), _dp(k), _dp(r)))
        return k, r

    def assemble_dynfrac_dev(self, dt, u_ptr, uold_ptr, k_ptr, r_ptr):
        _check(load().pxg_assemble_dynfrac_dev(self.h, dt, u_ptr, uold_ptr, k_ptr, r_ptr))

    def dynfrac_state(self):
        h = np.zeros(self.nb, np.uint8)
        v = np.zeros(self.n * self.dim)
        _check(load().pxg_dynfrac_get_state(self.h, _up(h), _dp(v)))
        return h, v

    # -- stress-coupled Cahn-Hilliard element -------------------------------------
    def stressch_setup(self, params):
        _check(load().pxg_stressch_setup(self.h, C.byref(params)))

    def stressch_constants(self):
        """C11, C12, C66, A, Kh, Ch, D, kappa, chi, Omega, cref."""
        out = np.zeros(11)
        _check(load().pxg_stressch_constants(self.h, _dp(out)))
        return out

    def assemble_stressch(self, dt, u, uold):
        k = np.zeros(self.nnz)
        r = np.zeros(self.n * (self.dim + 2))
        _check(load().pxg_assemble_stressch(self.h, dt, _dp(np.ascontiguousarray(u, float)),
                                            _dp(np.ascontiguousarray(uold, float)), _dp(k), _dp(r)))
        return k, r

    def assemble_stressch_dev(self, dt, u_ptr, uold_ptr, k_ptr, r_ptr):
        _check(load().pxg_assemble_stressch_dev(self.h, dt, u_ptr, uold_ptr, k_ptr, r_ptr))

    def stressch_project(self, name, u):
        """A projection of the stress-CH element for the state u (n x components)."""
        return self._sch_project("pxg_stressch_project", name, u)

    def stressch_project_dev(self, which, u_ptr, out_ptr):
        _check(load().pxg_stressch_project_dev(self.h, which, u_ptr, out_ptr))

    def fsch_project_dev(self, which, u_ptr, out_ptr):
        _check(load().pxg_fsch_project_dev(self.h, which, u_ptr, out_ptr))

    def _sch_project(self, fn, name, u):
        which = SCH_PROJECTIONS[name]
        nc = load().pxg_sch_components(which)
        out = np.zeros(self.n * nc)
        _check(getattr(load(), fn)(self.h, which, _dp(np.ascontiguousarray(u, float)), _dp(out)))
        return out.reshape(self.n, nc)

    # -- fracture stress-coupled Cahn-Hilliard element ----------------------------
    def fsch_setup(self, params, uold0=None):
        uo = None if uold0 is None else np.ascontiguousarray(uold0, float)
        _check(load().pxg_fsch_setup(self.h, C.byref(params), _dp(uo)))

    def fsch_constants(self):
        """C11, C12, C66, A, Kh, Ch, D, kappa, chi, Omega, cref, rho, residual
        stiffness, critical stretch."""
        out = np.zeros(14)
        _check(load().pxg_fsch_constants(self.h, _dp(out)))
        return out

    def fsch_commit(self, uold, dt_prev):
        _check(load().pxg_fsch_commit(self.h, _dp(np.ascontiguousarray(uold, float)), dt_prev))

    def fsch_commit_dev(self, uold_ptr, dt_prev):
        _check(load().pxg_fsch_commit_dev(self.h, uold_ptr, dt_prev))

    def assemble_fsch(self, dt, u, uold):
        k = np.zeros(self.nnz)
        r = np.zeros(self.n * (self.dim + 2))
        _check(load().pxg_assemble_fsch(self.h, dt, _dp(np.ascontiguousarray(u, float)),
                                        _dp(np.ascontiguousarray(uold, float)), _dp(k), _dp(r)))
        return k, r

    def assemble_fsch_dev(self, dt, u_ptr, uold_ptr, k_ptr, r_ptr):
        _check(load().pxg_assemble_fsch_dev(self.h, dt, u_ptr, uold_ptr, k_ptr, r_ptr))

    def fsch_project(self, name, u):
        """A projection of the fracture stress-CH element for the state u."""
        return self._sch_project("pxg_fsch_project", name, u)

    def fsch_state(self):
        h = np.zeros(self.nb, np.uint8)
        v = np.zeros(self.n * self.dim)
        _check(load().pxg_fsch_get_state(self.h, _up(h), _dp(v)))
        return h, v

    # -- a15-a20 ---------------------------------------------------------------
    def frac_setup(self, params, precision=FP64, node_E=None, nu=0.25, plane_stress=False):
        ne = None if node_E is None else np.ascontiguousarray(node_E, float)
        self.frac_serial += 1
        _check(load().pxg_frac_setup(self.h, C.byref(params), precision, _dp(ne), nu,
                                     int(plane_stress)))

    FRAC_NO_SHARED_RUNS = 1
    FRAC_FP32_OPERATORS = 2

    def frac_options(self, flags):
        """PXG_FRAC_* diagnostic options, applied at the next frac_setup."""
        _check(load().pxg_frac_options(self.h, int(flags)))

    def frac_set_bcs(self, specs):
        arr, n, keep = make_bcs(specs)
        _check(load().pxg_frac_set_bcs(self.h, n, arr))

    def frac_accel(self, u, health):
        a = np.zeros(self.n * self.dim)
        ab = C.c_int()
        self.frac_serial += 1
        _check(load().pxg_frac_accel(self.h, _dp(np.ascontiguousarray(u, float)), _dp(health),
                                     _dp(a), C.byref(ab)))
        return a, ab.value

    PROJECTIONS = ("damage", "vonMises", "stress", "strain")

    def frac_project(self, name):
        """Nodal projection `name` of the current state (n x 1 or n x 9)."""
        which = self.PROJECTIONS.index(name)
        nc = 9 if which >= 2 else 1
        out = np.zeros(self.n * nc)
        _check(load().pxg_frac_project(self.h, which, _dp(out)))
        return out.reshape(self.n, nc)

    def frac_project_dev(self, which, out_ptr):
        _check(load().pxg_frac_project_dev(self.h, which, out_ptr))

    def frac_layout(self):
        """(bond slots per component, tiled, shared runs, nodes on shared runs)."""
        info = (C.c_longlong * 4)()
        _check(load().pxg_frac_layout(self.h, info))
        return tuple(int(v) for v in info)

    def frac_init(self, u0, dt, v0=None):
        v = None if v0 is None else np.ascontiguousarray(v0, float)
        self.frac_serial += 1
        _check(load().pxg_frac_init(self.h, _dp(np.ascontiguousarray(u0, float)), _dp(v), dt))

    def frac_run(self, dt, total, nsteps_max=0):
        done = C.c_int()
        self.frac_serial += 1
        rc = load().pxg_frac_run(self.h, dt, total, nsteps_max, C.byref(done))
        if rc != 0:
            err = PxgError(rc, load().pxg_last_error().decode(errors="replace"))
            err.steps = done.value
            raise err
        return done.value

    def frac_profile(self, enable):
        _check(load().pxg_frac_profile(self.h, int(enable)))

    def frac_kernel_time(self):
        """(summed step-kernel milliseconds, launches) since profiling was enabled."""
        ms = C.c_double()
        n = C.c_int()
        _check(load().pxg_frac_kernel_time(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def frac_state(self):
        nd = self.n * self.dim
        u = np.zeros(nd)
        uo = np.zeros(nd)
        h = np.zeros(self.nb, np.uint8)
        bs = np.zeros(self.nb, np.int32)
        _check(load().pxg_frac_get_state(self.h, _dp(u), _dp(uo), _up(h), _ip(bs)))
        return u, uo, h, bs

    def frac_state_ptrs(self):
        a = C.c_void_p()
        b = C.c_void_p()
        _check(load().pxg_frac_state_ptrs(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # -- a22 ---------------------------------------------------------------------
    def frac_adjoint(self, u0, dt, nsteps, sens, comp, ckpt=50, v0=None):
        g = np.zeros(self.n)
        J = C.c_double()
        sens = np.ascontiguousarray(sens, np.int32)
        v = None if v0 is None else np.ascontiguousarray(v0, float)
        _check(load().pxg_frac_adjoint(self.h, _dp(np.ascontiguousarray(u0, float)), _dp(v), dt,
                                       nsteps, ckpt, len(sens), _ip(sens), comp, C.byref(J),
                                       _dp(g)))
        return J.value, g

    def frac_adjoint_ex(self, u0, dt, nsteps, sens, comp, ckpt=50, v0=None, grad_E=True):
        """(J, dJ/dE or None, dJ/db) in one reverse sweep."""
        g = np.zeros(self.n) if grad_E else None
        gb = np.zeros(self.dim)
        J = C.c_double()
        sens = np.ascontiguousarray(sens, np.int32)
        v = None if v0 is None else np.ascontiguousarray(v0, float)
        _check(load().pxg_frac_adjoint_ex(self.h, _dp(np.ascontiguousarray(u0, float)), _dp(v), dt,
                                          nsteps, ckpt, len(sens), _ip(sens), comp, C.byref(J),
                                          _dp(g), _dp(gb)))
        return J.value, g, gb

    def frac_set_body(self, b):
        _check(load().pxg_frac_set_body(self.h, _dp(np.ascontiguousarray(b, float))))

    def frac_loss(self, u0, dt, nsteps, sens, comp, node_E=None, v0=None):
        J = C.c_double()
        sens = np.ascontiguousarray(sens, np.int32)
        v = None if v0 is None else np.ascontiguousarray(v0, float)
        ne = None if node_E is None else np.ascontiguousarray(node_E, float)
        _check(load().pxg_frac_loss(self.h, _dp(np.ascontiguousarray(u0, float)), _dp(v), dt,
                                    nsteps, _dp(ne), len(sens), _ip(sens), comp, C.byref(J)))
        return J.value


def structured_cloud(g):
    L = load()
    gs = grid_struct(g)
    n = C.c_int()
    nb = C.c_int()
    _check(L.pxg_structured_sizes(C.byref(gs), C.byref(n), C.byref(nb)))
    coords = np.zeros(3 * n.value)
    vols = np.zeros(n.value)
    mirror = np.zeros(n.value, np.int32)
    wall = np.zeros(n.value, np.int32)
    dxyz = np.zeros(3)
    _check(L.pxg_structured_cloud(C.byref(gs), _dp(coords), _dp(vols), _ip(mirror), _ip(wall),
                                  _dp(dxyz)))
    return coords, vols, mirror, wall, dxyz, nb.value


def horizon(g, factor):
    gs = grid_struct(g)
    return load().pxg_horizon_radius(C.byref(gs), C.c_double(factor))


def random_ic(seed, lo, hi, nodes, ndof, dofs, values):
    nodes = np.ascontiguousarray(nodes, np.int32)
    dofs = np.ascontiguousarray(dofs, np.int32)
    _check(load().pxg_random_ic(C.c_uint64(seed), lo, hi, len(nodes), _ip(nodes), ndof,
                                len(dofs), _ip(dofs), _dp(values)))
    return values


def schedule(dt, total):
    L = load()
    L.pxg_schedule.argtypes = [C.c_double, C.c_double, C.c_int, D, D]
    n = L.pxg_schedule(dt, total, 0, None, None)
    dts = np.zeros(n)
    ts = np.zeros(n)
    L.pxg_schedule(dt, total, n, _dp(dts), _dp(ts))
    return dts, ts
