"""The drop-in at the reference's own seam: the reference's host code
(PDProblem::run and everything under it, compiled from its sources with
PERIX_CUDA_ASSEMBLE, its device dispatch FormResidualAndJacobianCUDA.cpp and
ExplicitPDDOFracElement.cpp:174-235 unchanged) runs whole decks with
"assemble": "cuda" through integration/perix_cuda_pxg.cpp over libpxg, and
with "assemble": "openmp" on its own CPU assembler.  Each job runs in a child
process (oracle/refpy.run_job).

The device rebuilds the g-weights from the points (as the reference's own
device path does, CudaAssembleKernels.cu:208-243), so the two runs differ by
the last bit of exp() in the weights: solutions must agree to 1e-10
(implicit, direct LDU solves) and 1e-9 (explicit), and the broken-bond set
of the explicit run must be the same.  Every assembly of a "cuda" run must
have been served by the device: the seam's counter of served calls equals
the Newton iterations of the reference's log (or the explicit steps) and no
call was refused, so a silent CPU fallback fails the test.  The failure
contract: a fault injected into one explicit call makes the element fall back
to its CPU loop for the rest of the run (ExplicitPDDOFracElement.cpp:
225-234), and the result still agrees."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E_STEEL, NU, RHO = 191.0e9, 0.25, 8000.0


@pytest.fixture(scope="module")
def refpy():
    from oracle import refpy as r
    if not r.pxg_available():
        pytest.fail("oracle/_ref/libperix_ref_pxg.so is not built (make -C oracle/ref)")
    return r


def _walls(dim):
    return ["left", "right", "bottom", "top"] + (["back", "front"] if dim == 3 else [])


def poisson_deck(refpy, nx, assemble):
    """oracle/refpy.poisson_deck: Poisson on the 100 x 100 unit
    square, Dirichlet ghost walls, static Newton with the default LDU."""
    return refpy.poisson_deck(nx, assemble)


def ch_deck(refpy, nx, steps, assemble):
    """The Cahn-Hilliard spinodal loop (oracle/refpy.ch_deck): RandomIC seed 12345, chi = 3,
    kappa = 1e-2, M = 1, backward Euler dt = 1e-3, Neumann c and mu ghosts,
    Newton with the default LDU solver."""
    return refpy.ch_deck(nx, steps, assemble)


def plate_deck(refpy, nx, ny, steps, assemble):
    """The 0.2 x 0.1 m plate geometry: two force-only edge notches,
    the 16.5 m/s impactor on the bottom ghosts with 75 <= x <= 125 mm,
    Neumann ghosts elsewhere, critical stretch 0.01 (G0 from
    FractureCriterion.h:30-43), 0.9 of the CFL step."""
    dx = 0.2 / nx
    delta = 3.015 * dx * (1.0 + 1e-6)
    G0 = 0.01 ** 2 * 12.0 * E_STEEL * delta / (5.0 * math.pi)
    c11 = E_STEEL * (1.0 - NU) / ((1.0 + NU) * (1.0 - 2.0 * NU))
    dt = 0.9 * dx / math.sqrt(c11 / RHO)
    d = {"Mesh": refpy.mesh_block(2, nx, ny, hi=(0.2, 0.1, 1.0)),
         "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
         "MeshModify": {"type": "pre_existing_cracks", "treatment": "force_only",
                        "Cracks": [{"x1": 0.05, "y1": 0.0, "x2": 0.05, "y2": 0.05},
                                   {"x1": 0.15, "y1": 0.0, "x2": 0.15, "y2": 0.05}]},
         "DOFs": ["ux", "uy"],
         "ElmtSystem": {"f": {"type": "explicit_pddo_frac", "params": {
             "E": E_STEEL, "nu": NU, "rho": RHO, "G0": G0, "state": "plane_strain",
             "tension_only": True, "damage": True}}},
         "BCSystem": {}, "JobSystem": {"type": "transient", "assemble": assemble},
         "TimeStepping": {"dt": dt, "total_time": steps * dt, "verbose": False},
         "OutputSystem": {"interval": 1000000}}
    for w in _walls(2):
        d["BCSystem"]["n_" + w] = {"type": "neumann", "phygroup": w + "nodes_ghost", "dofs": ["ux", "uy"],
                                  "value": [0.0, 0.0]}
    d["BCSystem"]["impactor"] = {"type": "dirichlet", "phygroup": "bottomnodes_ghost", "dofs": ["ux", "uy"],
                                 "value": [0.0, 0.0], "velocity": [0.0, 16.5], "direct": True,
                                 "box": {"xmin": 0.075, "xmax": 0.125}}
    return d


def diffusion_deck(refpy, nx, steps, assemble):
    """The reference's diffusion example (examples/diffusion.json) at nx x nx."""
    return {"Mesh": refpy.mesh_block(2, nx, nx), "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
            "DOFs": ["c"], "ElmtSystem": {"d": {"type": "diffusion", "params": {"D": 0.01, "f": 0.3}}},
            "BCSystem": {"bottom": {"type": "dirichlet", "phygroup": "bottomnodes_ghost", "dofs": ["c"], "value": 1.0},
                         "top": {"type": "dirichlet", "phygroup": "topnodes_ghost", "dofs": ["c"], "value": 0.0},
                         "left": {"type": "neumann", "phygroup": "leftnodes_ghost", "dofs": ["c"], "value": 0.0},
                         "right": {"type": "neumann", "phygroup": "rightnodes_ghost", "dofs": ["c"], "value": 0.0}},
            "JobSystem": {"type": "transient", "assemble": assemble},
            "TimeStepping": {"dt": 0.5, "total_time": steps * 0.5},
            "NonlinearSolver": {"max_iters": 20, "abs_tol": 1e-9, "rel_tol": 1e-12},
            "OutputSystem": {"interval": 100000}}


def dynfrac_deck(refpy, nx, ny, steps, assemble):
    """The reference's tensile plate (examples/tensile_plate.json: implicit
    dynamic fracture under PD tractions) at nx x ny, fixed steps."""
    prm = {"E": 79.1403, "nu": 0.333, "state": "plane_stress"}
    bc = {"left_free": ("leftnodes", [0.0, 0.0]), "right_free": ("rightnodes", [0.0, 0.0]),
          "top_traction": ("topnodes", [0.0, 0.0153884]), "bottom_traction": ("bottomnodes", [0.0, -0.0153884])}
    return {"Mesh": refpy.mesh_block(2, nx, ny, hi=(1.0, 0.4, 1.0)),
            "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
            "MeshModify": {"type": "pre_existing_cracks", "treatment": "force_only",
                           "Cracks": [{"x1": 0.0, "y1": 0.2, "x2": 0.5, "y2": 0.2}]},
            "DOFs": ["ux", "uy"],
            "ElmtSystem": {"frac": {"type": "pddo_dynamic_frac", "params": {
                "E": 79.1403, "nu": 0.333, "rho": 268.198, "Gc": 1.48388e-06, "state": "plane_stress",
                "tension_only": True, "damage": True}}},
            "BCSystem": {k: dict(prm, type="pdtraction", phygroup=g, value=v) for k, (g, v) in bc.items()},
            "JobSystem": {"type": "transient", "assemble": assemble},
            "TimeStepping": {"dt": 0.01, "total_time": steps * 0.01},
            "NonlinearSolver": {"max_iters": 20, "abs_tol": 1e-9, "rel_tol": 1e-9},
            "OutputSystem": {"interval": 100000}}


def fsch_deck(refpy, nx, steps, assemble):
    """The reference's silicon_frac element (frac_stress_cahnhilliard, plane
    stress, with inertia) on a structured nx x nx square with a notch: a
    species flux into the bottom ghosts, Neumann mu, a free surface by
    sdtraction on every wall, the centre pinned; bond failure on."""
    prm = {"E": 87.9336, "nu": 0.29, "D": 0.03, "Omega": 3, "cref": 0.0, "chi": 2.6, "rho": 2.41817e-22,
           "kappa": 2.19834e-06, "Gc": 0.10662, "tension_only": True, "damage_on": True, "state": "plane_stress"}
    bcs = {"li_flux": {"type": "speciesflux", "phygroup": "bottomnodes_ghost", "dofs": ["c"], "value": 0.013889}}
    for w in _walls(2):
        bcs["mu_" + w] = {"type": "neumann", "phygroup": w + "nodes_ghost", "dofs": ["mu"], "value": 0.0}
        bcs["free_" + w] = {"type": "sdtraction", "phygroup": w + "nodes_ghost", "dofs": ["ux", "uy"],
                            "value": [0.0, 0.0], "E": 87.9336, "nu": 0.29, "Omega": 3, "cref": 0.0, "c_dof": "c",
                            "state": "plane_stress"}
    bcs["pin_center"] = {"type": "dirichlet", "phygroup": "alldomain", "dofs": ["ux", "uy"], "value": 0.0,
                         "box": {"xmin": 0.45, "xmax": 0.55, "ymin": 0.45, "ymax": 0.55}}
    return {"Mesh": refpy.mesh_block(2, nx, nx), "PDMesh": {"HorizonRadiusFactor": 3.015, "Order": 2},
            "MeshModify": {"type": "pre_existing_cracks", "treatment": "force_only",
                           "Cracks": [{"x1": 0.5, "y1": 0.0, "x2": 0.5, "y2": 0.2}]},
            "DOFs": ["c", "mu", "ux", "uy"],
            "ElmtSystem": {"si": {"type": "frac_stress_cahnhilliard", "params": prm}},
            "BCSystem": bcs, "ICSystem": {"c0": {"type": "constant", "dof": "c", "value": 0.0}},
            "JobSystem": {"type": "transient", "assemble": assemble},
            "TimeStepping": {"dt": 1e-5, "total_time": steps * 1e-5},
            "NonlinearSolver": {"max_iters": 30, "abs_tol": 1e-6, "rel_tol": 1e-6},
            "OutputSystem": {"interval": 100000}}


def _pair(refpy, make, *args):
    cpu, rc = refpy.run_job(make(refpy, *args, "openmp"))
    dev, rd = refpy.run_job(make(refpy, *args, "cuda"))
    rel = np.linalg.norm(dev - cpu) / max(np.linalg.norm(cpu), 1e-300)
    print("%s%s: rel %.2e; openmp %.2f s (assembly %.3f s, solve %.3f s, %d Newton iterations); "
          "cuda %.2f s (assembly %.3f s, solve %.3f s, %d iterations); seam calls %d served, %d refused"
          % (make.__name__, args, rel, rc["seconds"], rc["assemble_s"], rc["solve_s"], rc["newton_iterations"],
             rd["seconds"], rd["assemble_s"], rd["solve_s"], rd["newton_iterations"], rd["seam_ok"],
             rd["seam_failed"]))
    assert rc["seam_ok"] == 0 and rc["seam_failed"] == 0
    assert rd["seam_failed"] == 0, "a seam call fell back to the CPU"
    return cpu, dev, rel, rc, rd


def test_poisson_100x100_static(refpy):
    cpu, dev, rel, rc, rd = _pair(refpy, poisson_deck, 100)
    assert rd["newton_iterations"] >= 1 and rd["seam_ok"] == rd["newton_iterations"]
    assert rel <= 1e-10


def test_cahnhilliard_20_steps(refpy):
    cpu, dev, rel, rc, rd = _pair(refpy, ch_deck, 48, 20)
    assert rd["seam_ok"] == rd["newton_iterations"] >= 20
    assert rel <= 1e-10


def test_explicit_plate_200_steps(refpy):
    cpu, rc = refpy.run_job(plate_deck(refpy, 400, 200, 200, "openmp"))
    dev, rd = refpy.run_job(plate_deck(refpy, 400, 200, 200, "cuda"))
    rel = np.linalg.norm(dev - cpu) / np.linalg.norm(cpu)
    broken = int((rc["health"] == 0).sum())
    print("explicit plate 200 steps: rel %.2e, broken %d, mismatches %d; openmp %.2f s, cuda %.2f s, "
          "seam %d served %d refused" % (rel, broken, int((rd["health"] != rc["health"]).sum()), rc["seconds"],
                                         rd["seconds"], rd["seam_ok"], rd["seam_failed"]))
    assert rd["seam_failed"] == 0 and rd["seam_ok"] >= 200
    assert broken > 0
    assert np.array_equal(rd["health"], rc["health"])
    assert rel <= 1e-9


def test_explicit_fault_latches_cpu_path(refpy):
    """A fault in the 50th device call: the element warns, shuts the seam
    down and finishes on its CPU loop; no further device call happens."""
    cpu, rc = refpy.run_job(plate_deck(refpy, 200, 100, 150, "openmp"))
    dev, rd = refpy.run_job(plate_deck(refpy, 200, 100, 150, "cuda"), fault_at_call=50)
    assert rd["seam_ok"] == 49 and rd["seam_failed"] == 1
    assert np.array_equal(rd["health"], rc["health"])
    assert np.linalg.norm(dev - cpu) <= 1e-9 * np.linalg.norm(cpu)


def test_other_elements_through_the_seam(refpy):
    for make, args in ((diffusion_deck, (40, 5)), (dynfrac_deck, (60, 24, 5)), (fsch_deck, (32, 5))):
        cpu, dev, rel, rc, rd = _pair(refpy, make, *args)
        assert rd["seam_ok"] == rd["newton_iterations"] >= 1
        assert rel <= 1e-10
