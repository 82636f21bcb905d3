"""The C-ABI library without a GPU: it loads, exports every entry point that
include/pxg.h declares, its host-only helpers agree with the oracle bit for
bit, and its failure paths report errors instead of falling back.  CPU only.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import pxo
from perix_amd import pxg
from perix_amd.mesh import StructuredGrid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pxg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(pxg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 30
    lib = C.CDLL(pxg.lib_path())
    missing = [s for s in names if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(pxg.SYMBOLS) == names, "Python mirror lists the same entry points"


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(pxg, "_LIB", None)
    monkeypatch.setattr(pxg, "lib_path", lambda: os.path.join(ROOT, "no_such_libpxg.so"))
    with pytest.raises(pxg.PxgError):
        pxg.load()


@pytest.mark.parametrize("dim,grid,hi", [(2, (20, 12, 1), (1.0, 0.6, 1.0)),
                                         (3, (6, 5, 4), (0.6, 0.5, 0.4))])
def test_structured_cloud_matches_oracle(dim, grid, hi):
    nx, ny, nz = grid
    g = StructuredGrid(dim, nx, ny, nz, hi=hi)
    coords, vols, mirror, wall, dxyz, nbulk = pxg.structured_cloud(g)
    m = pxo.Mesh(pxo.make_grid(dim, nx, ny, nz, hi=hi))
    assert nbulk == m.nbulk
    for a, b in ((coords, m.coords), (vols, m.vols), (dxyz, m.dxyz)):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert np.array_equal(mirror, m.mirror) and np.array_equal(wall, m.wall)
    assert pxg.horizon(g, 3.015) == m.delta


def test_random_ic_and_schedule_match_oracle():
    nodes = np.arange(1, 501)
    a = np.zeros(2 * 500); b = np.zeros(2 * 500)
    pxg.random_ic(12345, 0.49, 0.51, nodes, 2, [1], a)
    pxo.random_ic(12345, 0.49, 0.51, nodes, 2, [1], b)
    assert np.array_equal(a, b) and (a[1::2] == 0).all()
    for dt, total in ((0.3, 1.0), (1e-7, 2.3e-6), (0.25, 1.0)):
        d1, t1 = pxg.schedule(dt, total)
        d2, t2 = pxo.schedule(dt, total)
        assert np.array_equal(d1, d2) and np.array_equal(t1, t2)


def test_bad_arguments_are_rejected():
    L = pxg.load()
    bad = pxg.Grid(4, 10, 10, 0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0)
    n = C.c_int(); nb = C.c_int()
    assert L.pxg_structured_sizes(C.byref(bad), C.byref(n), C.byref(nb)) == -1
    assert L.pxg_last_error()


def test_context_without_gpu_reports_error():
    L = pxg.load()
    h = C.c_void_p()
    rc = L.pxg_create(C.byref(h), 0)
    if rc == 0:
        L.pxg_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == -2, "HIP error code"
    assert b"HIP" in L.pxg_last_error()
