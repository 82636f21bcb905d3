#!/usr/bin/env python3
"""Benchmark of the explicit PDDO fracture hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): the 2D plate
0.2 m x 0.1 m at dx = 0.5 mm (400 x 200 cells, 81,200 PD nodes, 2,251,924
directed bonds), two 50 mm force-only edge notches, a 16.5 m/s impactor
patch, E = 191 GPa, nu = 0.25, rho = 8000 kg/m^3, plane strain, critical
stretch 0.01, CFL-limited dt (0.9 x dx/c), 2000 central-difference steps.

A "step" is one explicit time step (damage, Navier force, central-difference
update, BC presets, blow-up guard) over all nodes.  Inputs are resident in
HBM; the one-off family / g-weight build is timed separately.  `value` is the
FP32C form's node-steps/s (operators stored in FP32, bond products and family
sums formed in FP64, which is at least as accurate as FP32-pair compensated
sums of the same FP32 operators); the FP64 form, the FP32C error against
FP64, the adjoint/forward cost (200 x 100 plate, 500 steps) and the 3D block
(160 x 80 x 40, 556,800 nodes) are reported beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "explicit PDDO node-steps/s FP32-comp vs FP64 at stated err; adjoint/fwd cost"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class quiet_stdout:
    """Redirects fd 1 to fd 2 (the reference prints progress on stdout),
    flushing the C library's buffered output before fd 1 is restored."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        sys.stdout.flush()
        ctypes.CDLL(None).fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)


def frac_bytes(nb, n, dim, precision):
    """Algorithmic HBM bytes of one step launch: per directed bond the
    neighbour id (4 B), the Navier operators (3 or 6 values) and the health
    byte; per node u^n, u^{n-1} read, u^{n+1} written and x read."""
    ncomp = 3 if dim == 2 else 6
    opb = 4 if precision == 1 else 8
    per_bond = 4 + ncomp * opb + 1
    per_node = 4 * dim * 8
    return nb * per_bond + n * per_node


def timed_run(ctx, s, steps, warmup, profile=False):
    """Warm-up pass, then a fresh run of `steps` steps timed on the host
    clock (device synchronised on both sides).  Returns seconds."""
    total = steps * s.dt
    if warmup > 0:
        ctx.frac_init(s.u0, s.dt)
        ctx.frac_run(s.dt, total, nsteps_max=warmup)
    ctx.frac_init(s.u0, s.dt)
    ctx.sync()
    if profile:
        ctx.frac_profile(True)
    t0 = time.perf_counter()
    done = ctx.frac_run(s.dt, total, nsteps_max=steps)
    ctx.sync()
    t1 = time.perf_counter()
    if profile:
        ms, launches = ctx.frac_kernel_time()
        ctx.frac_profile(False)
        return t1 - t0, done, ms, launches
    return t1 - t0, done


def plate_case(problems, pxg, precision, steps, opts=0):
    t0 = time.perf_counter()
    ctx = pxg.Context(0)
    ctx.frac_options(opts)
    s = problems.plate_400x200(precision=precision, nsteps=steps, ctx=ctx)
    s.ctx.sync()
    return s, time.perf_counter() - t0


def plate_throughput(problems, pxg, K, W):
    """The plate's timed runs: W warm-up steps, then K steps timed on the host
    clock with the step kernel's HIP events recorded in the same run, for
    FP32C and FP64.  The event spans cover the step launches only (no guard
    polls), so the kernel time is at most the wall time."""
    out = {}
    for prec, name in ((pxg.FP32C, "fp32c"), (pxg.FP64, "fp64")):
        log("plate %s: setup" % name)
        s, setup_s = plate_case(problems, pxg, prec, K)
        sec, done, ms, launches = timed_run(s.ctx, s, K, W, profile=True)
        assert launches == done and ms <= sec * 1e3 * 1.0001, (ms, sec, launches, done)
        log("plate %s: %d steps in %.3f s (step kernel %.3f ms)" % (name, done, sec, ms))
        out[name] = {"sec": sec, "done": done, "setup_s": setup_s, "kernel_ms": ms, "launches": launches,
                     "n": s.cloud.n, "nb": s.ctx.nb, "dt": s.dt}
        s.ctx.close()
    return out


def plate_error(problems, pxg, nsteps=2000):
    """FP32C against FP64 over the workload's full 2,000 steps (independent of
    the timed step count): relative-L2 and max-abs displacement error and the
    broken-bond mismatch count."""
    st = {}
    for prec, name in ((pxg.FP32C, "fp32c"), (pxg.FP64, "fp64")):
        s = problems.plate_400x200(precision=prec, nsteps=nsteps)
        s.ctx.frac_init(s.u0, s.dt)
        done = s.ctx.frac_run(s.dt, s.total)
        u, _, h, _ = s.ctx.frac_state()
        st[name] = (u, h, done)
        s.ctx.close()
    (u32, h32, d32), (u64, h64, d64) = st["fp32c"], st["fp64"]
    assert d32 == d64 == nsteps
    return {"steps": nsteps, "rel_l2": float(np.linalg.norm(u32 - u64) / max(np.linalg.norm(u64), 1e-300)),
            "max_abs_m": float(np.abs(u32 - u64).max()), "broken_bond_mismatch": int((h32 != h64).sum()),
            "broken_bonds_fp64": int((h64 == 0).sum())}


def physical_cores():
    """Distinct physical cores among the CPUs this process may run on."""
    cores = set()
    for c in sorted(os.sched_getaffinity(0)):
        base = "/sys/devices/system/cpu/cpu%d/topology/" % c
        try:
            with open(base + "core_id") as a, open(base + "physical_package_id") as b:
                cores.add((int(b.read()), int(a.read())))
        except OSError:
            cores.add((0, c))
    return len(cores)


def cpu_run(case, threads, steps, budget_s=0.0, timeout=300):
    """oracle/cpu_bench.py in a process of its own with the reference's
    advised OpenMP placement set before its runtime starts."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores", OMP_PROC_BIND="close")
    cmd = [sys.executable, "-m", "oracle.cpu_bench", case, "--steps", str(steps)]
    if budget_s > 0:
        cmd += ["--budget-s", str(budget_s)]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("cpu_bench %s failed: %s" % (case, r.stderr[-400:]))
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baselines():
    """The reference's OpenMP explicit loop on this host (bounded samples):
    the plate on the physical cores this process may use (at most
    OMP_NUM_THREADS when that is set), the plate on one thread, and the block."""
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    phys = physical_cores()
    threads = min(phys, cap) if cap > 0 else phys
    plate = cpu_run("plate", threads, 2000, budget_s=12.0)
    one = cpu_run("plate", 1, 400, budget_s=6.0)
    block = cpu_run("block", threads, 20)
    return {"value": plate["node_steps_per_s"], "unit": "node-steps/s", "cores": threads, "kind": plate["kind"],
            "sample": "reference OpenMP explicit loop (oracle/_ref, built from the reference sources), "
                      "first %d steps of the same 400x200 plate; one-off element build (%.2f s) excluded"
                      % (plate["steps"], plate["setup_s"]),
            "cpu_model": plate["cpu_model"], "physical_cores_available": phys,
            "omp": "OMP_NUM_THREADS=%d OMP_PLACES=cores OMP_PROC_BIND=close" % threads,
            "one_thread_node_steps_per_s": one["node_steps_per_s"], "one_thread_steps": one["steps"],
            "block3d_node_steps_per_s": block["node_steps_per_s"], "block3d_steps": block["steps"],
            "block3d_setup_s": block["setup_s"], "block3d_kind": block["kind"],
            "implicit_poisson_100x100": implicit_solve_split(),
            "implicit_ch_128x128_ldu": reference_ldu_ch(128)}


def implicit_solve_split(nx=100):
    """The Poisson 100 x 100 job run whole by the reference's own driver (oracle/_ref, Newton
    with its default LDU solver), once on its OpenMP assembler and once with
    "assemble": "cuda" through the seam over libpxg: the reference's own
    timers split each run into assembly and linear solve, the part that
    stays on the CPU.  None when the seam build is absent."""
    from oracle import refpy
    if not refpy.pxg_available():
        return None
    out = {}
    for assemble in ("openmp", "cuda"):
        _, rec = refpy.run_job(refpy.poisson_deck(nx, assemble), timeout=300)
        out[assemble] = {k: rec[k] for k in ("seconds", "assemble_s", "solve_s", "newton_iterations", "seam_ok")}
    out["deck"] = "Poisson %dx%d, static Newton, default LDU (reference driver)" % (nx, nx)
    return out


def reference_ldu_ch(nx=128):
    """The reference's DefaultSolver (oracle/_ref, RCM + skyline LDU, one
    thread) on the Cahn-Hilliard spinodal loop's first Newton system at nx x nx: seconds of one
    factorisation and solve.  None when the reference build is absent.  The
    reference prints its progress to stdout; it goes to stderr here, so that
    the bench's stdout stays one JSON line."""
    from oracle import refpy
    if not refpy.available():
        return None
    with quiet_stdout():
        c = refpy.Case(refpy.ch_deck(nx, 1, "openmp"))
        u, _ = c.initial_state()
        u = c.preset(u, 0.0)
        k, r = c.assemble(u, u, 1e-3)
        k, r = c.apply_bcs(u, 1e-3, k, r)
        t0 = time.perf_counter()
        rc, x, perm, jp = c.default_solve(k, r)
        sec = time.perf_counter() - t0
        c.close()
    return {"seconds": sec, "equations": len(r), "profile": int(jp[-1]), "rc": rc,
            "gflop": 2.0 * float(np.sum(np.diff(jp).astype(float) ** 2)) / 1e9,
            "system": "Cahn-Hilliard %dx%d, first Newton system of the spinodal loop (chi 3, kappa 1e-2, dt 1e-3)" % (nx, nx)}


def implicit_case(problems, pxg, ch_nx=256, ch_steps=20):
    """The implicit jobs with the linear solve on the device (the reference's
    default LDU as a block-skyline LU): Poisson 100 x 100 (static Newton)
    and the Cahn-Hilliard spinodal job (ch_nx^2, ch_steps backward-Euler
    steps, Newton max 20 / abs 1e-9 / rel 1e-8), wall time of the whole job
    after a warm-up solve that builds the ordering and captures the launch
    graphs, and the factorisation / substitution split of one Newton system."""
    import math
    import statistics
    import torch
    from perix_amd import implicit
    dev = torch.device("cuda", 0)
    out = {}

    def split(ctx, asm_sys, reps=5):
        ws = implicit.Workspace(ctx)
        asm_sys(ws)
        ctx.sync()
        tf, ta = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            ctx.solve_factor_dev(ws.K.data_ptr())
            ctx.sync()
            t1 = time.perf_counter()
            ctx.solve_apply_dev(ws.R.data_ptr(), ws.dU.data_ptr())
            ctx.sync()
            t2 = time.perf_counter()
            tf.append(t1 - t0)
            ta.append(t2 - t1)
        info = ctx.solve_info()
        fms, ams = statistics.median(tf) * 1e3, statistics.median(ta) * 1e3
        flop = info["block_products"] * 2.0 * info["block"] ** 3
        return {"factor_ms": fms, "substitution_ms": ams, "equations": info["equations"],
                "blocks": info["blocks"], "profile": info["profile"], "block_products": info["block_products"],
                "factor_tflops": flop / (fms * 1e-3) / 1e12, "block_storage_gb": info["bytes"] / 1e9}

    ctx, cloud, delta, f = problems.poisson_square(nx=100, ny=100)
    f0 = 2.0 * math.pi ** 2

    def asm_p(up, uop, kp, rp):
        ctx.assemble_poisson_dev(1.0, up, kp, rp, f_const=f0)

    u = torch.zeros(cloud.n, dtype=torch.float64, device=dev)
    implicit.static_solve(ctx, asm_p, u)
    u.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = implicit.static_solve(ctx, asm_p, u)
    job = time.perf_counter() - t0

    def sys_p(ws):
        asm_p(u.data_ptr(), u.data_ptr(), ws.K.data_ptr(), ws.R.data_ptr())
        ctx.csr_apply_bcs_dev(0.0, u.data_ptr(), ws.K.data_ptr(), ws.R.data_ptr())

    out["poisson_100x100"] = dict(job_s=job, newton_iterations=res.iters, converged=res.converged,
                                  **split(ctx, sys_p))
    ctx.close()

    ctx, cloud, delta, u0 = problems.cahnhilliard_square(nx=ch_nx, ny=ch_nx)

    def asm_c(dt, up, uop, kp, rp):
        ctx.assemble_ch_dev(1e-2, 3.0, 1.0, dt, up, uop, kp, rp)

    u = torch.tensor(u0, device=dev)
    torch.cuda.synchronize()
    # warm-up: ordering, block storage and graph capture on the first system
    wsw = implicit.Workspace(ctx)
    asm_c(1e-3, u.data_ptr(), u.data_ptr(), wsw.K.data_ptr(), wsw.R.data_ptr())
    ctx.csr_apply_bcs_dev(1e-3, u.data_ptr(), wsw.K.data_ptr(), wsw.R.data_ptr())
    ctx.solve_dev(wsw.K.data_ptr(), wsw.R.data_ptr(), wsw.dU.data_ptr())
    ctx.sync()
    del wsw
    t0 = time.perf_counter()
    iters = implicit.transient(ctx, asm_c, u, 1e-3, ch_steps * 1e-3, max_iters=20, abs_tol=1e-9, rel_tol=1e-8)
    job = time.perf_counter() - t0
    torch.cuda.synchronize()
    uo = u.clone()
    torch.cuda.synchronize()

    def sys_c(ws):
        asm_c(1e-3, u.data_ptr(), uo.data_ptr(), ws.K.data_ptr(), ws.R.data_ptr())
        ctx.csr_apply_bcs_dev(1e-3, u.data_ptr(), ws.K.data_ptr(), ws.R.data_ptr())

    sp = split(ctx, sys_c)
    out["ch_%dx%d_%d_steps" % (ch_nx, ch_nx, ch_steps)] = dict(
        job_s=job, newton_iterations=sum(iters), per_step_iterations=iters,
        ms_per_newton_iteration=job * 1e3 / max(sum(iters), 1), **sp)
    ctx.close()
    return out


def adjoint_case(problems, pxg, nsteps=500, nsens=64, ndirs=20, seed=2024):
    """Adjoint of the 200 x 100 plate (dx = 1 mm), E_i = 191 GPa (1 + 0.05 N(0,1)),
    J = sum over 64 sensors of u_y^2 at step 500, checkpoints every 50 steps."""
    rs = np.random.RandomState(seed)
    probe = problems.explicit_plate(2, 200, 100, nsteps=nsteps)
    n, nbulk = probe.cloud.n, probe.cloud.nbulk
    probe.ctx.close()
    E = problems.E_STEEL * (1.0 + 0.05 * rs.standard_normal(n))
    sens = rs.choice(nbulk, nsens, replace=False).astype(np.int32)
    s = problems.explicit_plate(2, 200, 100, nsteps=nsteps, node_E=E)
    ctx = s.ctx
    ctx.frac_init(s.u0, s.dt)
    ctx.frac_run(s.dt, s.total)          # records break steps
    ctx.sync()
    # warm, then time: one forward (frozen health) and one adjoint
    ctx.frac_loss(s.u0, s.dt, nsteps, sens, 1)
    ctx.frac_adjoint(s.u0, s.dt, nsteps, sens, 1, ckpt=50)
    t0 = time.perf_counter()
    ctx.frac_init(s.u0, s.dt)
    ctx.frac_run(s.dt, s.total)
    ctx.sync()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J0 = ctx.frac_loss(s.u0, s.dt, nsteps, sens, 1)
    t_fwd_frozen = time.perf_counter() - t0
    t0 = time.perf_counter()
    J, g = ctx.frac_adjoint(s.u0, s.dt, nsteps, sens, 1, ckpt=50)
    t_adj = time.perf_counter() - t0
    worst = 0.0
    hh = 1e-6
    for _ in range(ndirs):
        v = rs.standard_normal(n) * E
        Jp = ctx.frac_loss(s.u0, s.dt, nsteps, sens, 1, node_E=E + hh * v)
        Jm = ctx.frac_loss(s.u0, s.dt, nsteps, sens, 1, node_E=E - hh * v)
        fd = (Jp - Jm) / (2 * hh)
        ad = float(g @ v)
        worst = max(worst, abs(fd - ad) / max(abs(fd), 1e-300))
    ctx.close()
    return {"workload": "200x100 plate, 500 steps, 64 sensors, checkpoint every 50",
            "forward_ms": t_fwd * 1e3, "adjoint_ms": t_adj * 1e3,
            "adjoint_over_forward": t_adj / t_fwd,
            "reverse_over_forward": (t_adj - t_fwd_frozen) / t_fwd,
            "fd_max_rel_err_20dirs": worst, "J": J, "J_forward": J0}


def block_adjoint_case(problems, pxg, nsteps=200, nsens=64, ckpt=20, seed=2025):
    """Adjoint of the 160 x 80 x 40 block over 200 steps, J = sum over 64 sensors
    of u_y^2, gradient w.r.t. the uniform body force (3 scalars) from one
    reverse sweep with checkpoints every 20 steps (2 x 10 states of 13 MB),
    checked per component against central differences."""
    rs = np.random.RandomState(seed)
    probe = problems.block_160x80x40(precision=pxg.FP64, nsteps=nsteps)
    n, nbulk = probe.cloud.n, probe.cloud.nbulk
    probe.ctx.close()
    E = np.full(n, problems.E_STEEL)
    s = problems.block_160x80x40(precision=pxg.FP64, nsteps=nsteps, node_E=E)
    sens = rs.choice(nbulk, nsens, replace=False).astype(np.int32)
    ctx = s.ctx
    ctx.frac_init(s.u0, s.dt)
    ctx.frac_run(s.dt, s.total)          # records break steps
    ctx.sync()
    ctx.frac_adjoint_ex(s.u0, s.dt, nsteps, sens, 1, ckpt=ckpt, grad_E=False)   # warm
    t0 = time.perf_counter()
    ctx.frac_init(s.u0, s.dt)
    ctx.frac_run(s.dt, s.total)
    ctx.sync()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J, _, gb = ctx.frac_adjoint_ex(s.u0, s.dt, nsteps, sens, 1, ckpt=ckpt, grad_E=False)
    t_adj = time.perf_counter() - t0
    worst = 0.0
    hb = 1e3 * problems.RHO
    for a in range(3):
        b = np.zeros(3); b[a] = hb
        ctx.frac_set_body(b)
        jp = ctx.frac_loss(s.u0, s.dt, nsteps, sens, 1)
        ctx.frac_set_body(-b)
        jm = ctx.frac_loss(s.u0, s.dt, nsteps, sens, 1)
        fd = (jp - jm) / (2 * hb)
        worst = max(worst, abs(fd - gb[a]) / max(abs(fd), 1e-300))
    ctx.close()
    return {"workload": "160x80x40 block, 200 steps, 64 sensors, dJ/d(body force), checkpoint every 20",
            "forward_ms": t_fwd * 1e3, "adjoint_ms": t_adj * 1e3,
            "adjoint_over_forward": t_adj / t_fwd, "dJ_db": [float(v) for v in gb],
            "fd_max_rel_err_3comps": worst, "J": J}


def block_case(problems, pxg, steps=100):
    """3D block 160 x 80 x 40 (556,800 PD nodes): FP32C (shared operator
    runs), FP32C with every node on its own operators, and FP64.  Each form
    is timed in one run: host clock around `steps` steps and the step
    kernel's HIP events of the same run."""
    out = {}
    states = {}
    forms = (("fp32c", pxg.FP32C, 0), ("fp32c_own_ops", pxg.FP32C, 1), ("fp64", pxg.FP64, 0))
    for name, prec, opts in forms:
        t0 = time.perf_counter()
        ctx = pxg.Context(0)
        ctx.frac_options(opts)
        s = problems.block_160x80x40(precision=prec, nsteps=steps, ctx=ctx)
        s.ctx.sync()
        setup = time.perf_counter() - t0
        sec, done, ms, launches = timed_run(s.ctx, s, steps, 5, profile=True)
        assert launches == done and ms <= sec * 1e3 * 1.0001, (ms, sec, launches, done)
        u, _, h, _ = s.ctx.frac_state()
        states[name] = (u, h)
        nb, n = s.ctx.nb, s.cloud.n
        avg_ms = ms / max(launches, 1)
        alg = frac_bytes(nb, n, 3, prec)
        gbs = alg / (avg_ms * 1e-3) / 1e9
        traffic = profiled_traffic("frac_step_block160x80x40", name)
        out[name] = {"node_steps_per_s": n * done / sec, "ms_per_step": sec * 1e3 / done,
                     "step_kernel_ms": avg_ms, "setup_s": setup,
                     "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": gbs / HBM_PEAK_GBS, "traffic": traffic,
                                  "counter_frac": (traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
                                                   if traffic else None),
                                  "kernel": "k_frac_step<3,%s>" % ("fp64" if prec == pxg.FP64 else "fp32c"),
                                  "algorithmic_bytes_per_launch": alg, "avg_launch_ms": avg_ms}}
        if name == "fp64":
            # output projections of the final state (device, n or 9n doubles)
            import torch
            proj = {}
            buf = torch.empty(9 * n, dtype=torch.float64, device="cuda")
            for which, pname in enumerate(("damage", "vonMises", "stress", "strain")):
                s.ctx.frac_project_dev(which, buf.data_ptr())
                s.ctx.sync()
                t0 = time.perf_counter()
                for _ in range(5):
                    s.ctx.frac_project_dev(which, buf.data_ptr())
                s.ctx.sync()
                proj[pname + "_ms"] = (time.perf_counter() - t0) * 1e3 / 5
            out["projections"] = proj
        pst, tiled, nclasses, nshared = s.ctx.frac_layout()
        out[name]["layout"] = {"bond_slots": pst, "tiled": bool(tiled), "shared_operator_runs": nclasses,
                               "nodes_on_shared_runs": nshared}
        out["nodes"] = n
        out["bonds"] = nb
        s.ctx.close()
    u64, h64 = states["fp64"]
    for name in ("fp32c", "fp32c_own_ops"):
        u32, h32 = states[name]
        out[name]["rel_l2_err_vs_fp64"] = float(np.linalg.norm(u32 - u64) / max(np.linalg.norm(u64), 1e-300))
        out[name]["broken_bond_mismatch"] = int((h32 != h64).sum())
        out[name]["over_fp64"] = out[name]["node_steps_per_s"] / out["fp64"]["node_steps_per_s"]
    out["steps"] = steps
    return out


def _observables(s, pxg):
    """Pointwise state and non-pointwise observables of a block run: broken
    bonds, the L2 norm of the nodal damage field, kinetic energy of the bulk
    nodes with v = (u^n - u^{n-1}) / dt."""
    u, uo, h, _ = s.ctx.frac_state()
    dmg = s.ctx.frac_project("damage").ravel()
    nb = s.cloud.nbulk
    v = (u - uo).reshape(-1, 3)[:nb] / s.dt
    ke = 0.5 * RHO_BLOCK * float((s.cloud.vols[:nb] * (v * v).sum(axis=1)).sum())
    return u, h, {"broken": int((h == 0).sum()), "damage_l2": float(np.linalg.norm(dmg[:nb])), "kinetic_J": ke}


RHO_BLOCK = 8000.0


def stability_probe(problems, pxg, dim, iters=60, seed=5):
    """Largest eigenvalue of the intact discrete operator u -> -a(u) (a =
    L u / rho, the step's acceleration) by power iteration through the
    per-step accel entry point, and the central-difference limit
    dt_crit = 2 / sqrt(lambda_max) beside the step the workload uses
    (0.9 x the reference's estimate dx / sqrt(C11 / rho),
    ExplicitPDDOFracElement.cpp:54-77)."""
    s = problems.block_160x80x40(precision=pxg.FP64, nsteps=1) if dim == 3 else \
        problems.plate_400x200(precision=pxg.FP64, nsteps=1)
    rs = np.random.RandomState(seed)
    nd = s.cloud.n * dim
    h = np.ones(s.ctx.nb)
    v = rs.standard_normal(nd)
    v /= np.linalg.norm(v)
    lam = 0.0
    for _ in range(iters):
        a, _ = s.ctx.frac_accel(1e-12 * v, h)
        w = -a / 1e-12
        lam = float(v @ w)
        v = w / np.linalg.norm(w)
    dt_crit = 2.0 / math.sqrt(abs(lam))
    s.ctx.close()
    return {"lambda_max": lam, "dt_crit": dt_crit, "dt": s.dt, "dt_over_dt_crit": s.dt / dt_crit,
            "iterations": iters}


def block_long_run(problems, pxg, steps=1000, chunk=100, cfl=0.9):
    """FP32C against FP64 over the block's full 1,000-step run, every 100
    steps, beside controls that separate the causes of a divergence:
      fp32c_own_ops  FP32C with every node on its own FP32 operators (no
                     shared runs);
      fp64_f32ops    FP64 arithmetic on operators rounded to FP32 (FP32C's
                     operator rounding without its FP32 family sums);
      fp64_eps15     FP64 started 1e-15 m away (random per component);
      fp64_epsc      FP64 started away by FP32C's own error after the first
                     chunk (max |u_fp32c - u_fp64| at step 100, random per
                     component).
    Each row holds the relative-L2 displacement error and the broken-bond
    mismatch count against FP64, and the broken-bond count, damage-field L2
    and kinetic energy of every run."""
    def make(prec, opts, eps):
        ctx = pxg.Context(0)
        ctx.frac_options(opts)
        s = problems.block_160x80x40(precision=prec, nsteps=steps, ctx=ctx, cfl=cfl)
        u0 = s.u0.copy()
        if eps:
            u0 += eps * np.random.RandomState(1).standard_normal(u0.size)
        s.ctx.frac_init(u0, s.dt)
        return s

    # first chunk of FP32C vs FP64 fixes the control's perturbation size
    a64, a32 = make(pxg.FP64, 0, 0.0), make(pxg.FP32C, 0, 0.0)
    for s in (a64, a32):
        s.ctx.frac_run(s.dt, s.total, nsteps_max=chunk)
    epsc = float(np.abs(a32.ctx.frac_state()[0] - a64.ctx.frac_state()[0]).max())
    for s in (a64, a32):
        s.ctx.close()
    runs = {"fp64": make(pxg.FP64, 0, 0.0), "fp32c": make(pxg.FP32C, 0, 0.0),
            "fp32c_own_ops": make(pxg.FP32C, pxg.Context.FRAC_NO_SHARED_RUNS, 0.0),
            "fp64_f32ops": make(pxg.FP64, pxg.Context.FRAC_FP32_OPERATORS, 0.0),
            "fp64_eps15": make(pxg.FP64, 0, 1e-15), "fp64_epsc": make(pxg.FP64, 0, epsc)}
    rows = []
    for c in range(steps // chunk):
        st = {}
        for name, s in runs.items():
            s.ctx.frac_run(s.dt, s.total, nsteps_max=chunk)
            st[name] = _observables(s, pxg)
        u64, h64, ob64 = st["fp64"]
        nrm = max(np.linalg.norm(u64), 1e-300)
        row = {"step": (c + 1) * chunk, "fp64": ob64}
        for name, (u, h, ob) in st.items():
            if name == "fp64":
                continue
            row[name] = dict(ob, rel_l2=float(np.linalg.norm(u - u64) / nrm), mismatch=int((h != h64).sum()))
        rows.append(row)
        log("block long run", json.dumps(row))
    for s in runs.values():
        s.ctx.close()
    return {"workload": "160x80x40 block, 1000 steps, FP32C and controls against FP64",
            "cfl": cfl, "control_eps_m": epsc, "chunks": rows}


def gweight_model(n, nb, dim, nop):
    """Algorithmic cost of one g-weight build (k_operators): (bytes, flops).

    Bytes: per directed bond the neighbour id (4 B) and the nop FP64 weights
    written; per node the family offset (4 B), x (dim x 8 B), V (8 B) and the
    flag byte.  Flops (sqrt and exp counted as one each): per bond the
    geometry twice (moment pass and weight pass: 21 + dim(dim+1)/2), the
    moment products nop^2 x 3 and the weights nop x (2 nop + 2); per node the
    complete-pivot LU (2/3 nop^3), the nop triangular solves (2 nop^3) and the
    A X = B residual (2 nop^3)."""
    geom = 21 + dim * (dim + 1) // 2
    per_bond_f = 2 * geom + 3 * nop * nop + nop * (2 * nop + 2)
    per_node_f = (2.0 / 3.0 + 2.0 + 2.0) * nop ** 3
    per_bond_b = 4 + 8 * nop
    per_node_b = 4 + 8 * dim + 8 + 1
    return nb * per_bond_b + n * per_node_b, nb * per_bond_f + n * per_node_f, 3 * nop * nop * nb


def gweight_case(problems, pxg, reps=3):
    """Family search and g-weight build of the 2D plate and the 3D block.

    Each build runs once to allocate, then `reps` rebuilds are timed on the
    host clock around the synchronous call (buffers already allocated, so the
    figure is the kernel plus a few microseconds of launch and copy-back)."""
    from perix_amd.mesh import StructuredGrid, build_structured_cloud, horizon_radius
    out = {}
    for name, dim, nx, ny, nz, hi in (("plate400x200", 2, 400, 200, 1, (0.2, 0.1, 0.04)),
                                      ("block160x80x40", 3, 160, 80, 40, (0.16, 0.08, 0.04))):
        g = StructuredGrid(dim, nx, ny, nz, (0.0, 0.0, 0.0), hi)
        cloud = build_structured_cloud(g)
        delta = horizon_radius(g, 3.015)
        ctx = pxg.Context(0)
        ctx.set_points(cloud.coords, cloud.vols, cloud.mirror, cloud.dxyz, dim)
        ctx.build_families(delta)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.build_families(delta)
        ctx.sync()
        fam_ms = (time.perf_counter() - t0) * 1e3 / reps
        dxc = float(max(cloud.dxyz))
        ndeg = ctx.build_operators(delta, dxc)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.build_operators(delta, dxc)
        ctx.sync()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        n, nb, nop = cloud.n, ctx.nb, ctx.nop
        b, f, fmom = gweight_model(n, nb, dim, nop)
        gbs = b / (ms * 1e-3) / 1e9
        tfl = f / (ms * 1e-3) / 1e12
        out[name] = {"nodes": n, "bonds": nb, "nop": nop, "degenerate_nodes": ndeg,
                     "families_ms": fam_ms, "gweight_ms": ms,
                     "gweight_hbm": {"algorithmic_bytes": b, "achieved_GBs": gbs,
                                     "frac": gbs / HBM_PEAK_GBS},
                     "gweight_fp64": {"flops": f, "achieved_TFLOPs": tfl,
                                      "frac": tfl / FP64_PEAK_TFLOPS,
                                      "moment_accumulation_share": fmom / f}}
        ctx.close()
    return out


def assembly_case(problems, pxg, reps=20):
    """Poisson and diffusion (100x100), dynamic fracture (400x200 plate),
    Cahn-Hilliard and stress-coupled Cahn-Hilliard (256x256, one Picard
    iteration) CSR assembly on the device with inputs and outputs
    resident in HBM: nnz/s, beside the reference's own CPU assembly (oracle/_ref)
    of the same system when that build is present."""
    import torch
    out = {}
    ctx, cloud, delta, f = problems.poisson_square()
    n, nnz = cloud.n, ctx.nnz
    X = cloud.coords.reshape(-1, 3)
    u_h = np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1])
    u_h_poisson = u_h
    u = torch.from_numpy(u_h).cuda()
    fd = torch.from_numpy(np.ascontiguousarray(f)).cuda()
    k = torch.empty(nnz, dtype=torch.float64, device="cuda")
    r = torch.empty(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    def poisson():
        ctx.assemble_poisson_dev(1.0, u.data_ptr(), k.data_ptr(), r.data_ptr(), f_ptr=fd.data_ptr())

    poisson()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        poisson()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    out["poisson_100x100"] = {"nnz": nnz, "ms_per_assembly": ms, "nnz_per_s": nnz / (ms * 1e-3),
                              "roofline": assembly_roofline("poisson", ctx, ms)}
    out["poisson_100x100"].update(bc_rows_timing(ctx, u, k, r, poisson, reps))
    ctx.close()
    # diffusion (backward Euler) on the same 100x100 square
    ctx, cloud, delta, f = problems.poisson_square()
    n, nnz = cloud.n, ctx.nnz
    uo = torch.from_numpy(u_h * 0.99).cuda()
    u = torch.from_numpy(u_h).cuda()
    k = torch.empty(nnz, dtype=torch.float64, device="cuda")
    r = torch.empty(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    def diffusion():
        ctx.assemble_diffusion_dev(0.7, 1e-3, u.data_ptr(), uo.data_ptr(), k.data_ptr(), r.data_ptr(),
                                   f_const=0.3)

    diffusion()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        diffusion()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    out["diffusion_100x100"] = {"nnz": nnz, "ms_per_assembly": ms, "nnz_per_s": nnz / (ms * 1e-3),
                                "roofline": assembly_roofline("diffusion", ctx, ms)}
    ctx.close()
    # implicit dynamic fracture element on the 400x200 plate (2 dofs per node)
    import perix_amd
    g = perix_amd.StructuredGrid(2, 400, 200, hi=(0.2, 0.1, 1.0))
    pc = perix_amd.build_structured_cloud(g)
    pdelta = perix_amd.horizon_radius(g, 3.015)
    ctx = pxg.Context(0)
    ctx.set_points(pc.coords, pc.vols, pc.mirror, pc.dxyz, 2)
    ctx.set_cracks(segs=[(0.05, 0.0, 0.05, 0.05), (0.15, 0.0, 0.15, 0.05)])
    ctx.build_families(pdelta)
    ctx.build_operators(pdelta, float(max(pc.dxyz)))
    ctx.build_csr(2)
    c11, c12, c66 = problems.elastic_constants(2, False, problems.E_STEEL, problems.NU)
    G0 = problems.g0_for_stretch(0.01, problems.E_STEEL, pdelta, 2)
    sc = problems.critical_stretch(G0, problems.E_STEEL, pdelta, 2)
    nd, nnz = 2 * pc.n, ctx.nnz
    rs = np.random.RandomState(4)
    uold_h = 1e-7 * rs.standard_normal(nd)
    u_h = uold_h + 1e-8 * rs.standard_normal(nd)
    ctx.dynfrac_setup(pxg.FracParams(2, c11, c12, c66, problems.RHO, 0.0, 0.0, 0.0, sc, 1, 1), uold_h)
    ud, uod = torch.from_numpy(u_h).cuda(), torch.from_numpy(uold_h).cuda()
    k = torch.empty(nnz, dtype=torch.float64, device="cuda")
    r = torch.empty(nd, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    dt_dyn = 1e-7

    def dynfrac():
        ctx.assemble_dynfrac_dev(dt_dyn, ud.data_ptr(), uod.data_ptr(), k.data_ptr(), r.data_ptr())

    dynfrac()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        dynfrac()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    out["dynamic_fracture_400x200"] = {"nnz": nnz, "ms_per_assembly": ms, "nnz_per_s": nnz / (ms * 1e-3),
                                       "roofline": assembly_roofline("dynfrac", ctx, ms)}
    ctx.close()
    ctx, cloud, delta, u0 = problems.cahnhilliard_square()
    n, nnz = cloud.n, ctx.nnz
    u1 = u0.copy()
    u1[1::2] = 0.01 * np.random.RandomState(3).rand(n)
    ud, uod = torch.from_numpy(u1).cuda(), torch.from_numpy(u0).cuda()
    k = torch.empty(nnz, dtype=torch.float64, device="cuda")
    r = torch.empty(2 * n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    kappa = 1e-2 * (1.0 / 256) ** 2

    def ch():
        ctx.assemble_ch_dev(kappa, 3.0, 1.0, 1e-3, ud.data_ptr(), uod.data_ptr(), k.data_ptr(), r.data_ptr())

    ch()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        ch()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    out["cahnhilliard_256x256"] = {"nnz": nnz, "ms_per_picard_assembly": ms, "nnz_per_s": nnz / (ms * 1e-3),
                                   "roofline": assembly_roofline("ch", ctx, ms)}
    out["cahnhilliard_256x256"].update(bc_rows_timing(ctx, ud, k, r, ch, reps))
    ctx.close()
    # stress-coupled Cahn-Hilliard (c, mu, ux, uy), silicon parameters, 256x256
    ctx, cloud, delta, us_h, usold_h = problems.stressch_square()
    n, nnz = cloud.n, ctx.nnz
    ud, uod = torch.from_numpy(us_h).cuda(), torch.from_numpy(usold_h).cuda()
    k = torch.empty(nnz, dtype=torch.float64, device="cuda")
    r = torch.empty(4 * n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    def stressch():
        ctx.assemble_stressch_dev(1e-4, ud.data_ptr(), uod.data_ptr(), k.data_ptr(), r.data_ptr())

    stressch()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        stressch()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    out["stress_cahnhilliard_256x256"] = {"nnz": nnz, "ms_per_picard_assembly": ms,
                                          "nnz_per_s": nnz / (ms * 1e-3),
                                          "roofline": assembly_roofline("stressch", ctx, ms)}
    ctx.close()
    # fracture variant (silicon_frac parameters, two edge notches), with the
    # per-step commit timed on its own
    ctx, cloud, delta, uf_h, ufold_h = problems.fsch_square()
    n, nnz = cloud.n, ctx.nnz
    ud, uod = torch.from_numpy(uf_h).cuda(), torch.from_numpy(ufold_h).cuda()
    k = torch.empty(nnz, dtype=torch.float64, device="cuda")
    r = torch.empty(4 * n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    def fsch():
        ctx.assemble_fsch_dev(1e-4, ud.data_ptr(), uod.data_ptr(), k.data_ptr(), r.data_ptr())

    def fsch_commit():
        ctx.fsch_commit_dev(uod.data_ptr(), 1e-4)

    fsch()
    fsch_commit()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fsch()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        fsch_commit()
    ctx.sync()
    ms_c = (time.perf_counter() - t0) * 1e3 / reps
    out["frac_stress_cahnhilliard_256x256"] = {"nnz": nnz, "ms_per_picard_assembly": ms,
                                               "nnz_per_s": nnz / (ms * 1e-3), "ms_per_commit": ms_c,
                                               "roofline": assembly_roofline("fsch", ctx, ms)}
    ctx.close()
    try:
        out["cpu_baseline"] = cpu_assembly_baseline(u_h_poisson, u0, u1, kappa, (u_h, uold_h, G0),
                                                    (us_h, usold_h), (uf_h, ufold_h))
    except Exception as e:       # reported, never required
        out["cpu_baseline"] = {"value": None, "sample": "failed: %s" % e}
    return out


def assembly_bytes(kind, n, nb, nnz, dim, nop):
    """Algorithmic HBM bytes of one assembly launch: the K values and the
    residual written once; per directed bond the element's bond data read
    once (neighbour id and CSR slot 4 B each; the per-bond Laplacian 8 B and
    reverse-bond index 4 B for the Laplacian elements; the operators nop x 8 B
    for the Navier elements; the health byte for the fracture ones); per node
    the state, old state and node data.  Neighbour gathers are not counted
    again (each node's data is fetched once from HBM)."""
    if kind in ("poisson", "diffusion"):
        nd = 1
        bond = 4 + 4 + 8
        node = 8 + 8 + 12 + (8 if kind == "diffusion" else 0)
    elif kind == "ch":
        nd = 2
        bond = 4 + 4 + 4 + 8
        node = 2 * 8 + 8 + 8 + 8 + 1 + 12
    elif kind == "dynfrac":
        nd = dim
        bond = 4 + 4 + 1 + nop * 8
        node = 3 * dim * 8 + 12
    else:   # stressch, fsch
        nd = dim + 2
        bond = 4 + 4 + 4 + 8 + nop * 8 + (1 if kind == "fsch" else 0)
        node = nd * 8 + 8 + 8 + 1 + 12 + (dim * 8 if kind == "fsch" else 0)
    return nnz * 8 + n * nd * 8 + nb * bond + n * node


def assembly_roofline(kind, ctx, ms):
    b = assembly_bytes(kind, ctx.n, ctx.nb, ctx.nnz, ctx.dim, 9 if ctx.dim == 3 else 5)
    gbs = b / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "algorithmic_bytes": b, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
            "timing": "host clock over back-to-back launches (an upper bound on the kernel time)"}


def bc_rows_timing(ctx, u, k, r, assemble, reps):
    """BC row elimination (BCSystem::applyBCs) of the assembled system on the
    device, alone and with the assembly (the whole linear-system build of one
    Newton/Picard iteration), inputs and outputs resident in HBM."""
    nrows = ctx.csr_set_bcs_count
    assemble()
    ctx.csr_apply_bcs_dev(0.0, u.data_ptr(), k.data_ptr(), r.data_ptr())
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.csr_apply_bcs_dev(0.0, u.data_ptr(), k.data_ptr(), r.data_ptr())
    ctx.sync()
    ms_bc = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        assemble()
        ctx.csr_apply_bcs_dev(0.0, u.data_ptr(), k.data_ptr(), r.data_ptr())
    ctx.sync()
    ms_sys = (time.perf_counter() - t0) * 1e3 / reps
    return {"bc_rows": nrows, "ms_bc_rows": ms_bc, "ms_assembly_plus_bc_rows": ms_sys}


def cpu_assembly_baseline(u_poisson, u0_ch, u1_ch, kappa, dyn=None, stressch=None, fsch=None):
    """The reference's own OpenMP assembly (PDSystem::formResidualAndJacobianParallel
    through oracle/_ref) of the same systems on this host, one call each."""
    from oracle import refpy
    if not refpy.available():
        return {"value": None, "sample": "reference build not present"}
    res = {}
    with quiet_stdout():
        d = {"Mesh": refpy.mesh_block(2, 100, 100), "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
             "DOFs": ["u"], "ElmtSystem": {"p": {"type": "poisson", "params": {"sigma": 1.0, "f": 0.0}}},
             "BCSystem": {w: {"type": "dirichlet", "phygroup": w + "nodes_ghost", "dofs": ["u"], "value": 0.0}
                          for w in ("left", "right", "bottom", "top")},
             "JobSystem": {"type": "static", "assemble": "openmp"}}
        c = refpy.Case(d)
        c.assemble(u_poisson, u_poisson, 0.0, parallel=True)
        t0 = time.perf_counter()
        k, r = c.assemble(u_poisson, u_poisson, 0.0, parallel=True)
        ms = (time.perf_counter() - t0) * 1e3
        c.apply_bcs(u_poisson, 0.0, k, r)
        res["poisson_100x100"] = {"nnz": c.nnz, "ms_per_assembly": ms, "nnz_per_s": c.nnz / (ms * 1e-3),
                                  "ms_bc_rows": c.last_bc_seconds() * 1e3}
        c.close()
        d = {"Mesh": refpy.mesh_block(2, 256, 256), "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
             "DOFs": ["c", "mu"],
             "ElmtSystem": {"ch": {"type": "cahnhilliard", "params": {"chi": 3.0, "kappa": kappa, "M": 1.0}}},
             "BCSystem": {w: {"type": "neumann", "phygroup": w + "nodes_ghost", "dofs": ["c", "mu"],
                              "value": [0.0, 0.0]} for w in ("left", "right", "bottom", "top")},
             "JobSystem": {"type": "transient", "assemble": "openmp"},
             "TimeStepping": {"dt": 1e-3, "total_time": 2e-2}}
        c = refpy.Case(d)
        c.assemble(u1_ch, u0_ch, 1e-3, parallel=True)
        t0 = time.perf_counter()
        k, r = c.assemble(u1_ch, u0_ch, 1e-3, parallel=True)
        ms = (time.perf_counter() - t0) * 1e3
        c.apply_bcs(u1_ch, 0.0, k, r)
        res["cahnhilliard_256x256"] = {"nnz": c.nnz, "ms_per_picard_assembly": ms, "nnz_per_s": c.nnz / (ms * 1e-3),
                                       "ms_bc_rows": c.last_bc_seconds() * 1e3}
        c.close()
        if dyn is not None:
            u_d, uold_d, G0 = dyn
            d = {"Mesh": refpy.mesh_block(2, 400, 200, hi=(0.2, 0.1, 1.0)),
                 "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
                 "MeshModify": {"type": "pre_existing_cracks", "treatment": "force_only",
                                "Cracks": [{"x1": 0.05, "y1": 0.0, "x2": 0.05, "y2": 0.05},
                                           {"x1": 0.15, "y1": 0.0, "x2": 0.15, "y2": 0.05}]},
                 "DOFs": ["ux", "uy"],
                 "ElmtSystem": {"f": {"type": "pddo_dynamic_frac", "params": {
                     "E": 191e9, "nu": 0.25, "rho": 8000.0, "G0": G0, "state": "plane_strain",
                     "tension_only": True, "damage": True}}},
                 "BCSystem": {"l": {"type": "neumann", "phygroup": "leftnodes_ghost", "dofs": ["ux", "uy"],
                                    "value": [0.0, 0.0]}},
                 "JobSystem": {"type": "transient", "assemble": "openmp"},
                 "TimeStepping": {"dt": 1e-7, "total_time": 1e-6}}
            c = refpy.Case(d)
            c.assemble(u_d, uold_d, 1e-7, parallel=True)
            t0 = time.perf_counter()
            c.assemble(u_d, uold_d, 1e-7, parallel=True)
            ms = (time.perf_counter() - t0) * 1e3
            res["dynamic_fracture_400x200"] = {"nnz": c.nnz, "ms_per_assembly": ms,
                                               "nnz_per_s": c.nnz / (ms * 1e-3)}
            c.close()
        if stressch is not None:
            from perix_amd.problems import SILICON
            us, usold = stressch
            d = {"Mesh": refpy.mesh_block(2, 256, 256), "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
                 "DOFs": ["c", "mu", "ux", "uy"],
                 "ElmtSystem": {"s": {"type": "stress_cahnhilliard",
                                      "params": dict(SILICON, state="plane_stress")}},
                 "BCSystem": {"n": {"type": "neumann", "phygroup": "leftnodes_ghost", "dofs": ["c", "mu"],
                                    "value": [0.0, 0.0]}},
                 "JobSystem": {"type": "transient", "assemble": "openmp"},
                 "TimeStepping": {"dt": 1e-4, "total_time": 1e-3}}
            c = refpy.Case(d)
            c.assemble(us, usold, 1e-4, parallel=True)
            t0 = time.perf_counter()
            c.assemble(us, usold, 1e-4, parallel=True)
            ms = (time.perf_counter() - t0) * 1e3
            res["stress_cahnhilliard_256x256"] = {"nnz": c.nnz, "ms_per_picard_assembly": ms,
                                                  "nnz_per_s": c.nnz / (ms * 1e-3)}
            c.close()
        if fsch is not None:
            from perix_amd.problems import SILICON_FRAC
            uf, ufold = fsch
            prm = dict(SILICON_FRAC, state="plane_stress")
            d = {"Mesh": refpy.mesh_block(2, 256, 256), "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
                 "MeshModify": {"type": "pre_existing_cracks", "treatment": "force_only",
                                "Cracks": [{"x1": 0.25, "y1": 0.0, "x2": 0.25, "y2": 0.1},
                                           {"x1": 0.75, "y1": 0.0, "x2": 0.75, "y2": 0.1}]},
                 "DOFs": ["c", "mu", "ux", "uy"],
                 "ElmtSystem": {"s": {"type": "frac_stress_cahnhilliard", "params": prm}},
                 "BCSystem": {"n": {"type": "neumann", "phygroup": "leftnodes_ghost", "dofs": ["c", "mu"],
                                    "value": [0.0, 0.0]}},
                 "JobSystem": {"type": "transient", "assemble": "openmp"},
                 "TimeStepping": {"dt": 1e-4, "total_time": 1e-3}}
            c = refpy.Case(d)
            c.assemble(uf, ufold, 1e-4, parallel=True)
            t0 = time.perf_counter()
            c.assemble(uf, ufold, 1e-4, parallel=True)
            ms = (time.perf_counter() - t0) * 1e3
            res["frac_stress_cahnhilliard_256x256"] = {"nnz": c.nnz, "ms_per_picard_assembly": ms,
                                                       "nnz_per_s": c.nnz / (ms * 1e-3)}
            c.close()
    res["kind"] = "reference"
    res["cores"] = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    res["sample"] = ("one assembly call of each system after a warm-up call (OpenMP assembler), "
                     "then one BCSystem::applyBCs (serial, timed alone inside the harness)")
    return res


def profiled_traffic(kernel_case, prec):
    """HBM bytes per launch of a step kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (profiles/traffic.json), or None."""
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(tpath):
        return None
    with open(tpath) as fh:
        t = json.load(fh)
    return t.get(kernel_case, {}).get(prec, {}).get("traffic_bytes")


def stream_ceiling(nbytes=1_750_000_000, reps=10):
    """Device-wide read bandwidth of a plain reduction over nbytes (the volume
    of the 3D block's bond streams): the practical ceiling for the step."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    x.sum()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        x.sum()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    print(json.dumps({"stream_read_bytes": x.numel() * 4, "ms": ms,
                      "GBs": x.numel() * 4 / (ms * 1e-3) / 1e9}))


def copy_calibration():
    """A plain device copy of known size for calibrating rocprofv3 FETCH_SIZE /
    WRITE_SIZE counters (64 MiB read, 64 MiB written, 10 launches)."""
    import torch
    a = torch.empty(8 * 1024 * 1024, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8,
                      "launches": 10}))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the adjoint and 3D block figures")
    ap.add_argument("--only-plate", action="store_true", help="plate FP32C/FP64 only (profiling)")
    ap.add_argument("--only-block", action="store_true", help="3D block only (profiling)")
    ap.add_argument("--long-run", action="store_true", help="with --only-block: the 1000-step comparison")
    ap.add_argument("--cfl", type=float, default=0.9, help="with --long-run: step as a fraction of the CFL estimate")
    ap.add_argument("--long-steps", type=int, default=1000, help="with --long-run: steps of the comparison")
    ap.add_argument("--stability", action="store_true",
                    help="largest eigenvalue of the plate's and the block's step operator (power iteration)")
    ap.add_argument("--only-adjoint", action="store_true", help="adjoint case only (profiling)")
    ap.add_argument("--only-assembly", action="store_true", help="CSR assembly only (profiling)")
    ap.add_argument("--only-gweight", action="store_true", help="family + g-weight build only (profiling)")
    ap.add_argument("--only-implicit", action="store_true", help="implicit jobs with the device solve only")
    ap.add_argument("--ch-nx", type=int, default=256, help="with --only-implicit: Cahn-Hilliard grid")
    ap.add_argument("--ch-steps", type=int, default=20, help="with --only-implicit: Cahn-Hilliard steps")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--stream-ceiling", action="store_true")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return 0
    if args.stream_ceiling:
        stream_ceiling()
        return 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: perix_amd runs on one GPU only; the multi-GPU slab "
              "decomposition is a design, not a measured path. Run with --gpus 1.", file=sys.stderr)
        return 2

    import perix_amd
    from perix_amd import pxg, problems

    K, W = args.steps, args.warmup
    if args.stability:
        print(json.dumps({"plate": stability_probe(problems, pxg, 2), "block": stability_probe(problems, pxg, 3)}))
        return 0
    if args.only_block:
        blk = {} if args.long_run else block_case(problems, pxg, steps=K)
        if args.long_run:
            blk["long_run"] = block_long_run(problems, pxg, steps=args.long_steps, cfl=args.cfl)
        print(json.dumps({"block3d": blk}))
        return 0
    if args.only_adjoint:
        print(json.dumps({"adjoint": adjoint_case(problems, pxg),
                          "adjoint_block3d_body_force": block_adjoint_case(problems, pxg)}))
        return 0
    if args.only_assembly:
        print(json.dumps({"assembly": assembly_case(problems, pxg)}))
        return 0
    if args.only_gweight:
        print(json.dumps({"gweight_build": gweight_case(problems, pxg)}))
        return 0
    if args.only_implicit:
        print(json.dumps({"implicit": implicit_case(problems, pxg, args.ch_nx, args.ch_steps)}))
        return 0
    pt = plate_throughput(problems, pxg, K, W)
    f32, f64 = pt["fp32c"], pt["fp64"]
    n, nb = f32["n"], f32["nb"]
    value = n * f32["done"] / f32["sec"]
    v64 = n * f64["done"] / f64["sec"]
    avg_ms = f32["kernel_ms"] / max(f32["launches"], 1)
    avg64 = f64["kernel_ms"] / max(f64["launches"], 1)
    bytes32 = frac_bytes(nb, n, 2, 1)
    achieved = bytes32 / (avg_ms * 1e-3) / 1e9
    achieved64 = frac_bytes(nb, n, 2, 0) / (avg64 * 1e-3) / 1e9
    traffic = profiled_traffic("frac_step_plate400x200", "fp32c")
    if args.only_plate:
        print(json.dumps({"value": value, "fp64": v64, "kernel_ms": avg_ms, "kernel64_ms": avg64}))
        return 0
    log("plate FP32C vs FP64 over 2000 steps")
    err = plate_error(problems, pxg)
    extra = {}
    if not args.no_cpu:
        log("cpu baseline (reference OpenMP loop in a process of its own)")
        try:
            extra["cpu_baseline"] = cpu_baselines()
        except Exception as e:   # the baseline is reported, never required
            extra["cpu_baseline"] = {"value": None, "unit": "node-steps/s", "cores": physical_cores(),
                                     "kind": "reference", "sample": "failed: %s" % e}
    if not args.no_extra:
        log("adjoint (200x100, 500 steps)")
        adj = adjoint_case(problems, pxg)
        log("adjoint (160x80x40 block, body force, 200 steps)")
        adjb = block_adjoint_case(problems, pxg)
        log("3D block (160x80x40)")
        blk = block_case(problems, pxg)
        log("assembly")
        asm = assembly_case(problems, pxg)
        log("family search and g-weight build")
        gw = gweight_case(problems, pxg)
        log("implicit jobs with the device solve")
        imp = implicit_case(problems, pxg)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": f32["sec"] * 1e3 / f32["done"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 operators; f64 bond products, family sums, damage and update",
        "data": "synthetic",
        "config": {"workload": "explicit_pddo_frac 2D plate 0.2x0.1 m, 400x200 cells (81,200 PD nodes, "
                               "%d directed bonds), two force-only edge notches, 16.5 m/s impactor; "
                               "throughput over %d CFL-limited steps, FP32C error over the workload's "
                               "2000 steps" % (nb, K),
                   "nodes": n, "bonds": nb, "horizon": "3.015 dx", "order": 2},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "counter_frac": (traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None),
                     "kernel": "k_frac_step<2,fp32c>", "algorithmic_bytes_per_launch": bytes32,
                     "avg_launch_ms": avg_ms,
                     "note": "achieved prices each launch at the per-bond model (id, 3 FP32 operators, "
                             "health) and time from the HIP events of the timed run; nodes on shared "
                             "operator runs (bit-identical FP32 values) read them from L2, so the "
                             "measured traffic is below the model (counter_frac prices the counters). "
                             "The 43.5 MB working set stays in the 256 MB Infinity Cache and the launch "
                             "is latency-bound; the HBM claim is block3d's own-operator form"},
    }
    out.update(extra)
    # the metric's components, last on the line
    out["fp64"] = {"node_steps_per_s": v64, "ms_per_step": f64["sec"] * 1e3 / f64["done"],
                   "step_kernel_ms": avg64, "step_kernel_GBs": achieved64}
    out["fp32c_over_fp64"] = value / v64
    out["fp32c_error_vs_fp64"] = err
    if not args.no_extra:
        out["adjoint_over_forward"] = {
            "plate200x100_dJdE_500_steps": adj["adjoint_over_forward"],
            "plate200x100_fd_max_rel_20dirs": adj["fd_max_rel_err_20dirs"],
            "block3d_dJdb_200_steps": adjb["adjoint_over_forward"],
            "block3d_fd_max_rel": adjb["fd_max_rel_err_3comps"]}
        b32, bown, b64 = blk["fp32c"], blk["fp32c_own_ops"], blk["fp64"]
        out["block3d"] = {
            "fp32c_ms": b32["ms_per_step"], "fp32c_kernel_ms": b32["step_kernel_ms"],
            "fp32c_frac_counters": b32["roofline"]["counter_frac"],
            "fp32c_own_ops_kernel_ms": bown["step_kernel_ms"], "fp32c_own_ops_frac": bown["roofline"]["frac"],
            "fp32c_own_ops_frac_counters": bown["roofline"]["counter_frac"],
            "fp64_ms": b64["ms_per_step"], "fp64_kernel_ms": b64["step_kernel_ms"],
            "fp64_frac": b64["roofline"]["frac"], "fp32c_over_fp64": b32["over_fp64"],
            "fp32c_own_ops_over_fp64": bown["over_fp64"],
            "fp32c_err_100_steps": [b32["rel_l2_err_vs_fp64"], b32["broken_bond_mismatch"]],
            "node_steps_per_s_fp32c": b32["node_steps_per_s"]}
        out["assembly_nnz_per_s"] = {k: round(v["nnz_per_s"] / 1e9, 2) for k, v in asm.items()
                                     if isinstance(v, dict) and "nnz_per_s" in v}
        out["assembly_nnz_per_s"]["unit"] = "1e9 nnz/s"
        out["gweight_build_ms"] = {k: round(v["gweight_ms"], 3) for k, v in gw.items()}
        ip, ic = imp["poisson_100x100"], imp["ch_256x256_20_steps"]
        ref_ch = (extra.get("cpu_baseline") or {}).get("implicit_ch_128x128_ldu")
        out["implicit_device_solve"] = {
            "poisson_100x100_job_s": ip["job_s"], "poisson_newton_iterations": ip["newton_iterations"],
            "poisson_factor_ms": ip["factor_ms"], "poisson_substitution_ms": ip["substitution_ms"],
            "ch256_20_steps_job_s": ic["job_s"], "ch256_newton_iterations": ic["newton_iterations"],
            "ch256_factor_ms": ic["factor_ms"], "ch256_substitution_ms": ic["substitution_ms"],
            "ch256_factor_tflops": ic["factor_tflops"],
            "roofline": {"bound": "fp64", "achieved": ic["factor_tflops"], "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": ic["factor_tflops"] / FP64_PEAK_TFLOPS,
                         "kernel": "pxg_solve_factor_dev (k_diag, k_panel, k_update) on the CH 256x256 system",
                         "note": "block products over the whole factorisation; the chain of n pivots through "
                                 "the diagonal blocks bounds it, not the FP64 rate"},
            "reference_ldu_ch128_s": ref_ch["seconds"] if ref_ch else None}
    sys.stderr.flush()
    print(json.dumps(out))
    sys.stdout.flush()
    return 0


if __name__ == "__main__":
    sys.exit(main())
