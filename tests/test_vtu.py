"""Result output (pxg_write_vtu) against the reference's own VTU writer
(PDProblem::savePDResults, fixture tests/golden/vtu_stressch_2d.npz): the
same arrays in the same order, names and component counts; ElmtID, cells
and the dof values exactly; coordinates and projections to the 16 digits the
reference's ASCII keeps.  Runs without a GPU (the writer is host code)."""
import os

import numpy as np

from perix_amd import pxg
from vtu_io import read_vtu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_vtu_matches_reference_writer(oracle, tmp_path):
    z = np.load(os.path.join(GOLD, "vtu_stressch_2d.npz"), allow_pickle=False)
    s = np.load(os.path.join(GOLD, "stressch_2d.npz"), allow_pickle=False)
    m = oracle.Mesh(oracle.make_grid(2, 14, 10, 1, hi=(1.4, 1.0, 1.0)))
    ops, _ = oracle.operators(m)
    k = s["consts"]
    consts = [k[0], k[1], k[2], k[3], k[10], k[8]]
    names = [str(x) for x in z["names"]]
    dofs = ["c", "mu", "ux", "uy"]
    assert names[:5] == ["ElmtID"] + dofs
    fields = [(nm, oracle.sch_project(m, ops, consts, False, True, z["u"], nm)) for nm in names[5:]]
    assert pxg.vtu_path("/x/case.json", 3) == "/x/" + str(z["vtu_name"])
    path = pxg.write_vtu(str(tmp_path / "ours.vtu"), m.coords, m.nbulk, dofs, z["u"], fields)
    v = read_vtu(path)
    assert v["npoints"] == v["ncells"] == m.n
    assert [a[0] for a in v["point_data"]] == names
    assert [a[1] for a in v["point_data"]] == [int(c) for c in z["comps"]]
    for q, (nm, nc, arr) in enumerate(v["point_data"]):
        ref = z["pd_%d" % q]
        assert arr.shape == ref.shape, nm
        if nm == "ElmtID":
            assert np.array_equal(arr, ref)
        else:
            scale = max(np.abs(ref).max(), 1e-300)
            assert np.abs(arr - ref).max() <= 1e-15 * scale, nm
        if nm in dofs:
            assert np.array_equal(arr, z["u"][dofs.index(nm)::4])     # exact, not 16 digits
    assert np.abs(v["points"] - z["points"]).max() <= 1e-15 * np.abs(z["points"]).max()
    assert np.array_equal(v["points"].ravel(), m.coords)
    cells = {nm: arr for nm, _, arr in v["cells"]}
    assert np.array_equal(cells["connectivity"], z["connectivity"])
    assert np.array_equal(cells["offsets"], z["offsets"])
    assert np.array_equal(cells["types"], z["types"])


def test_vtu_rejects_bad_input(tmp_path):
    """A field of 2 components (VTK point data here is 1, 3 or 9) and an
    unwritable path are library errors; a mis-sized state a host error."""
    coords = np.zeros(6)
    import pytest
    with pytest.raises(ValueError):
        pxg.write_vtu(str(tmp_path / "a.vtu"), coords, 2, ["u"], np.zeros(3))     # u size
    with pytest.raises(pxg.PxgError):
        pxg.write_vtu(str(tmp_path / "b.vtu"), coords, 2, ["u"], np.zeros(2), [("bad", np.zeros((2, 2)))])
    with pytest.raises(pxg.PxgError):
        pxg.write_vtu(str(tmp_path / "no" / "dir" / "c.vtu"), coords, 2, ["u"], np.zeros(2))
