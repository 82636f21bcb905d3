"""CPU baseline of the explicit fracture step -- TEST INFRASTRUCTURE ONLY.

Times the reference's own OpenMP explicit loop (oracle/_ref/libperix_ref.so,
built from the reference sources by oracle/ref/Makefile) on a bounded sample
of a benchmark workload, in a process of its own, so that the OpenMP
settings the reference advises (README.md:391-393: OMP_NUM_THREADS = physical
cores, OMP_PLACES=cores, OMP_PROC_BIND=close) are in the environment before
any OpenMP runtime starts.  bench.py runs it as

    python -m oracle.cpu_bench {plate|block} --steps N [--probe-steps M]

with the environment it chose; the last stdout line is a JSON record.  The
one-off element build (families, operator cache, crack seeds) is timed
separately from the step loop.  Without the reference build it falls back to
the C oracle (oracle/libpxo.so), reported as kind "port".
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

E, NU, RHO = 191.0e9, 0.25, 8000.0

# workload geometry: (dim, nx, ny, nz, hi, cracks, impact box x-range)
CASES = {
    "plate": (2, 400, 200, 1, (0.2, 0.1, 1.0), [(0.05, 0.0, 0.05, 0.05), (0.15, 0.0, 0.15, 0.05)], (0.075, 0.125)),
    "block": (3, 160, 80, 40, (0.16, 0.08, 0.04), [(0.08, 0.0, 0.08, 0.008)], (0.04, 0.06)),
}


class quiet_stdout:
    """Sends fd 1 to fd 2 while the reference prints its progress."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)


def case_constants(name):
    """Node count, the release rate G0 that gives critical stretch 0.01
    (FractureCriterion.h:30-43 inverted) and the step 0.9 x the CFL limit
    (ExplicitPDDOFracElement.cpp:54-77)."""
    import math
    from oracle import pxo
    dim, nx, ny, nz, hi, cracks, box = CASES[name]
    m = pxo.Mesh(pxo.make_grid(dim, nx, ny, nz, hi=hi), families=False)
    sc = 0.01
    G0 = sc * sc * 6.0 * E * m.delta / 5.0 if dim == 3 else sc * sc * 12.0 * E * m.delta / (5.0 * math.pi)
    dt = 0.9 * pxo.stable_dt(dim, False, E, NU, RHO, m.dxyz)
    return m.n, G0, dt


def deck(name, G0, dt, nsteps):
    from oracle import refpy
    dim, nx, ny, nz, hi, cracks, box = CASES[name]
    dofs = ["ux", "uy", "uz"][:dim]
    params = {"E": E, "nu": NU, "rho": RHO, "G0": G0, "tension_only": True, "damage": True}
    if dim == 2:
        params["state"] = "plane_strain"
    d = {"Mesh": refpy.mesh_block(dim, nx, ny, nz, hi=hi),
         "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
         "MeshModify": {"type": "pre_existing_cracks", "treatment": "force_only",
                        "Cracks": [{"x1": c[0], "y1": c[1], "x2": c[2], "y2": c[3]} for c in cracks]},
         "DOFs": dofs,
         "ElmtSystem": {"f": {"type": "explicit_pddo_frac", "params": params}},
         "BCSystem": {}, "JobSystem": {"type": "transient", "assemble": "openmp"},
         "TimeStepping": {"dt": dt, "total_time": nsteps * dt, "verbose": False}}
    walls = ["left", "right", "bottom", "top"] + (["back", "front"] if dim == 3 else [])
    for w in walls:
        d["BCSystem"]["n_" + w] = {"type": "neumann", "phygroup": w + "nodes_ghost", "dofs": dofs,
                                  "value": [0.0] * dim}
    vel = [0.0] * dim
    vel[1] = 16.5
    d["BCSystem"]["impactor"] = {"type": "dirichlet", "phygroup": "bottomnodes_ghost", "dofs": dofs,
                                 "value": [0.0] * dim, "velocity": vel, "direct": True,
                                 "box": {"xmin": box[0], "xmax": box[1]}}
    return d


def run_reference(name, nsteps):
    from oracle import refpy
    n, G0, dt = case_constants(name)
    with quiet_stdout():
        c = refpy.Case(deck(name, G0, dt, nsteps))
        steps, setup_s, loop_s, _ = c.explicit_timed()
        c.close()
    return n, steps, setup_s, loop_s


def run_port(name, nsteps):
    from oracle import pxo
    dim, nx, ny, nz, hi, cracks, box = CASES[name]
    n, G0, dt = case_constants(name)
    t0 = time.perf_counter()
    m = pxo.Mesh(pxo.make_grid(dim, nx, ny, nz, hi=hi))
    ops, _ = pxo.operators(m)
    opc = pxo.frac_opcache(m, ops)
    del ops
    sc = pxo.critical_stretch(G0, E, m.delta, dim, False)
    p = pxo.frac_params(dim, E, NU, RHO, sc)
    h = pxo.seed_health(m, segs=cracks)
    comps = list(range(dim))
    specs = [{"type": "neumann", "nodes": m.group(w), "comps": comps} for w in range(2 * dim)]
    vel = [0.0] * dim
    vel[1] = 16.5
    specs.append({"type": "dirichlet", "nodes": m.group(2), "comps": comps, "value": [0.0] * dim,
                  "velocity": vel, "direct": True, "box": (box[0], box[1], 1.0, -1.0, 1.0, -1.0)})
    bcs = pxo.make_bcs(specs)
    u = np.zeros(dim * m.n)
    pxo.bc_preset(m, dim, bcs, u, 0.0)
    uo = np.zeros_like(u)
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    steps, _ = pxo.explicit_run(m, p, opc, bcs, dt, nsteps * dt, u, uo, h)
    return m.n, steps, setup_s, time.perf_counter() - t0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("case", choices=sorted(CASES))
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--budget-s", type=float, default=0.0,
                    help="with > 0: a short probe run first, then as many steps as fit the budget "
                         "(at most --steps)")
    args = ap.parse_args()
    from oracle import refpy
    kind = "reference" if refpy.available() else "port"
    run = run_reference if kind == "reference" else run_port
    nsteps = args.steps
    if args.budget_s > 0:
        _, st, _, loop = run(args.case, 5)
        nsteps = int(min(args.steps, max(5, args.budget_s / max(loop / max(st, 1), 1e-6))))
    n, steps, setup_s, loop_s = run(args.case, nsteps)
    print(json.dumps({"case": args.case, "kind": kind, "nodes": n, "steps": steps, "setup_s": setup_s,
                      "loop_s": loop_s, "node_steps_per_s": n * steps / loop_s,
                      "threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0),
                      "omp_places": os.environ.get("OMP_PLACES", ""),
                      "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""), "cpu_model": cpu_model()}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
