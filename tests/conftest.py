import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU case")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pxo
    pxo.lib()
    return pxo


def bcrow_specs(z):
    """BC list of a bcrows_*.npz fixture as make_bcs specs (0-based
    components), in deck order.  The box starts from the (1, -1) sentinel per
    axis and takes the bounds the deck gives, as ReadBCSystemBlock.cpp:43-55
    does, so an axis is only constrained when both bounds are given."""
    import ast
    dofs = [str(d) for d in z["dofs"]]
    out = []
    for nm, typ, grp, dd, val, vel, direct, box in z["specs"]:
        comps = [dofs.index(d) for d in str(dd).split(",")]
        vals = [float(v) for v in str(val).split(",")]
        vels = [float(v) for v in str(vel).split(",")] if str(vel) else [0.0]
        b = None
        if str(box):
            items = dict(ast.literal_eval(str(box)))
            b = [1.0, -1.0, 1.0, -1.0, 1.0, -1.0]
            for a, ax in enumerate("xyz"):
                if ax + "min" in items:
                    b[2 * a] = items[ax + "min"]
                if ax + "max" in items:
                    b[2 * a + 1] = items[ax + "max"]
        out.append({"type": str(typ), "nodes": z["group_" + str(grp)], "comps": comps,
                    "value": vals * (len(comps) if len(vals) == 1 else 1),
                    "velocity": vels * (len(comps) if len(vels) == 1 else 1),
                    "direct": bool(int(direct)), "box": b})
    return out
