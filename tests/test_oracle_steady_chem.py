"""The oracle's inline (operator-split) chemistry on the steady path: the
chemistry step that follows the flow update of a pseudo-iteration is the
batch point-implicit integrator (orc_chem_batch, pinned in
test_oracle_pins.py) applied to every interior cell at its density and
internal energy."""
import numpy as np

import pyoracle as po
from ostc import cases
from helpers import oracle_zone, relerr

NS = 9


def _c3(mech):
    case = cases.c3_case(24, 12, reconstruction=2)
    case["setup"]["int implicit_method"] = 1
    case["scheme"].update(implicit_method=1, DPLUR_inner_step=3, chemSrcMethod=1)
    import ostc
    import tempfile
    d = tempfile.mkdtemp()
    cases.write_case_dir(case, d)
    h = ostc.Solver(d, device=-1)
    states = {}
    for name in ("air_stream", "fuel_stream"):
        o = h.inflow_state(name, 7 + NS)
        states[name] = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=o[7:7 + NS])
    h.close()
    bv, sv = cases.c3_fields(case, states, np.array([mech.mw[i] for i in range(NS)]))
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = states[b["stream"]]
    return case, bv, sv


def test_inline_chemistry_step_equals_the_batch_integrator(mech):
    """A split pseudo-iteration with a chemistry step of 3e-8 s equals the
    same pseudo-iteration without it followed by the batch integrator on the
    interior (away from the boundaries, whose ghosts the step then updates)."""
    case, bv, sv = _c3(mech)
    flow = oracle_zone(dict(case, scheme=dict(case["scheme"], split_chem=1, steady_chem_dt=0.0)), mech=mech,
                       bv=bv, sv=sv)
    both = oracle_zone(dict(case, scheme=dict(case["scheme"], split_chem=1, steady_chem_dt=3e-8)), mech=mech,
                       bv=bv, sv=sv)
    flow.run_steady(1)
    both.run_steady(1)
    g = case["ngg"]
    shp = cases.ghost_shape(case["mx"], case["my"], 1, g)
    fb = flow.field("bv").reshape((6,) + shp)[:, g, g:-g, g:-g]
    fs = flow.field("sv").reshape((NS,) + shp)[:, g, g:-g, g:-g]
    rho, T = fb[0].ravel().copy(), fb[5].ravel().copy()
    Y = fs.reshape(NS, -1).copy()
    dY, dT = po.chem_batch(mech, rho, T, Y, 3e-8)
    bb = both.field("bv").reshape((6,) + shp)[:, g, g:-g, g:-g]
    bs = both.field("sv").reshape((NS,) + shp)[:, g, g:-g, g:-g]
    # interior cells next to a boundary see the boundary update of their
    # ghosts only after the step: compare cells two away from the edges
    inner = (slice(None), slice(2, -2), slice(2, -2))
    assert relerr(bs.reshape(NS, case["my"], case["mx"])[inner],
                  (Y + dY).reshape(NS, case["my"], case["mx"])[inner]) < 1e-14
    assert relerr(bb[5][inner[1:]], (T + dT).reshape(case["my"], case["mx"])[inner[1:]]) < 1e-14
