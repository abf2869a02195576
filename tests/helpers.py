"""Shared test helpers: oracle zones for generated cases and normwise errors."""
import numpy as np

import pyoracle as po
from ostc import cases


def oracle_zone(case, mech=None, bv=None, sv=None):
    """Oracle zone for a generated case, initialised like the solver."""
    bcs = []
    for b in case["bcs"]:
        d = dict(type=b["type"], raw=b["raw"])
        if b["type"] == 5:
            st = b.get("state", case.get("inflow"))
            d["state"] = dict(st, sv=list(st.get("sv", [])))
        elif "state" in b:
            d["state"] = dict(b["state"], sv=list(b["state"].get("sv", [])))
        if b["type"] == 2:
            d["t_wall"] = b.get("t_wall", 300.0)
            d["isothermal"] = b.get("isothermal", True)
        bcs.append(d)
    sc = dict(case["scheme"], bcs=bcs)
    z = po.Zone(case["mx"], case["my"], case["mz"], case["ngg"], sc, case["x"], case["y"], case["z"], mech=mech)
    if bv is None:
        bv, sv = cases.initial_fields(case)
    z.field("bv")[:] = np.asarray(bv).ravel()
    if sv is not None and np.asarray(sv).size:
        z.field("sv")[:] = np.asarray(sv).ravel()
    z.initialize()
    return z


def _oracle_bcs(case, bcs):
    out = []
    for b in bcs:
        d = dict(type=b["type"], raw=b["raw"])
        if b["type"] == 5:
            st = b.get("state", case.get("inflow"))
            d["state"] = dict(st, sv=list(st.get("sv", [])))
        elif "state" in b:
            d["state"] = dict(b["state"], sv=list(b["state"].get("sv", [])))
        if b["type"] == 2:
            d["t_wall"] = b.get("t_wall", 300.0)
            d["isothermal"] = b.get("isothermal", True)
        out.append(d)
    return out


def oracle_mesh(mcase, fields, mech=None):
    """Oracle mesh for a multi-block case (ostc.cases.split_case) with the
    given per-block (bv, sv), initialised like the solver."""
    blocks = [dict(b, bcs=_oracle_bcs(mcase, b["bcs"])) for b in mcase["blocks"]]
    M = po.Mesh(blocks, mcase["ngg"], mcase["scheme"], mcase["inner"], mech=mech)
    for z, (bv, sv) in zip(M.zones, fields):
        z.field("bv")[:] = np.asarray(bv).ravel()
        if np.asarray(sv).size:
            z.field("sv")[:] = np.asarray(sv).ravel()
    M.initialize()
    return M


def relerr(a, b):
    """Normwise relative error ||a-b||_inf / ||b||_inf (0 if both vanish)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if a.size else 0.0
    if den == 0.0:
        return num
    return num / den


def per_component(a, b, ncomp):
    a = np.asarray(a).reshape(ncomp, -1)
    b = np.asarray(b).reshape(ncomp, -1)
    return [relerr(a[c], b[c]) for c in range(ncomp)]


def interior(arr, case, ncomp):
    g = case["ngg"]
    a = np.asarray(arr).reshape(ncomp, case["mz"] + 2 * g, case["my"] + 2 * g, case["mx"] + 2 * g)
    return a[:, g:g + case["mz"], g:g + case["my"], g:g + case["mx"]]


def grouped(a, b, ncomp, groups):
    """Normwise relative errors per named group of components; a vector
    group (e.g. the three velocity components) is scaled by the largest
    component magnitude of the whole group, so a component that is zero in
    exact arithmetic is judged against the vector's size, not its own
    rounding noise."""
    a = np.asarray(a).reshape(ncomp, -1)
    b = np.asarray(b).reshape(ncomp, -1)
    out = {}
    for name, comps in groups.items():
        den = max(np.max(np.abs(b[c])) for c in comps)
        num = max(np.max(np.abs(a[c] - b[c])) for c in comps)
        out[name] = num / den if den > 0 else num
    return out


BV_GROUPS = {"rho": [0], "vel": [1, 2, 3], "p": [4], "T": [5]}
CV_GROUPS = {"rho": [0], "mom": [1, 2, 3], "rhoE": [4]}
