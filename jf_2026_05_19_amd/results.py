"""SOL 101 results in the reference's result layout, from the device's
static solve and result recovery (``sol105_static_recover``).

``static_results`` mirrors ``_build_results_from_state``
(JFEM/src/solver/solve_case.jl:126-186) for the supported element set:
``displacements`` (one row per GRID in ascending id order, analysis frame =
basic frame since CD != 0 is refused), ``spc_forces`` (rows of
``R = K u - F`` whose 6-vector norm exceeds 1e-20, the reference's test; at
free DOFs R is the round-off residual), ``forces`` / ``forces_bilin`` /
``stresses`` / ``strains`` for CQUAD4, CTRIA3 and CBAR/CBEAM
(JFEM/src/solver/stress_recovery.jl:163-513), and
``solver_diagnostics["equilibrium"]`` with the reaction sums.  Element rows
come in model-array order (CQUAD4 and CTRIA3 in deck order of element id,
CBARs before CBEAMs); the reference walks its element dictionaries in hash
order.  Stress entries carry ``major = minor = 0.0`` as the reference writes
them; the per-element scalar of the second return value is the major
principal stress at z1 for shells and |axial stress| for bars, as the
reference's ``stresses`` dictionary holds.

``solve_static`` runs the whole SOL 101 path on the device: upload, K
assembly, AUTOSPC and sparse Cholesky, solve, recovery.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["static_results", "solve_static", "Q4_COLS", "T3_COLS", "BAR_COLS"]

Q4_COLS, T3_COLS, BAR_COLS = 44, 20, 18
_FORCE_KEYS = ("fx", "fy", "fxy", "mx", "my", "mxy", "qx", "qy")
_BAR_KEYS = ("axial", "shear_1", "shear_2", "torque", "moment_a1", "moment_a2", "moment_b1", "moment_b2")


def _von_mises(s):
    return math.sqrt(s[0] ** 2 - s[0] * s[1] + s[1] ** 2 + 3 * s[2] ** 2)


def _stress_entry(s, z):
    return {"fiber_dist": z, "normal_x": float(s[0]), "normal_y": float(s[1]), "shear_xy": float(s[2]),
            "von_mises": _von_mises(s), "major": 0.0, "minor": 0.0}


def _strain_entry(e, z):
    return {"fiber_dist": z, "normal_x": float(e[0]), "normal_y": float(e[1]), "shear_xy": float(e[2]),
            "major": 0.0, "minor": 0.0}


def _principal_major(s):
    avg = (s[0] + s[1]) / 2.0
    return avg + math.sqrt(((s[0] - s[1]) / 2.0) ** 2 + s[2] ** 2)


def _shell_rows(rows, eids, thickness, key, out, scal, node_ids=None, nodes=None):
    forces, stresses, strains = out["forces"][key], out["stresses"][key], out["strains"][key]
    bilin = out["forces_bilin"][key]
    for r, eid, t in zip(rows, eids, thickness):
        eid = int(eid)
        f = {"eid": eid}
        f.update({k: float(v) for k, v in zip(_FORCE_KEYS, r[:8])})
        forces.append(f)
        s1, s2, e1, e2 = r[8:11], r[11:14], r[14:17], r[17:20]
        scal[eid] = _principal_major(s1)
        stresses.append({"eid": eid, "z1": _stress_entry(s1, -t / 2), "z2": _stress_entry(s2, t / 2)})
        strains.append({"eid": eid, "z1": _strain_entry((e1 + e2) / 2.0, 0.0),
                        "z2": _strain_entry((e1 - e2) / t, -1.0)})
    if nodes is not None:
        for r, eid, nd in zip(rows, eids, nodes):
            cen = {"eid": int(eid), "grid_id": "CEN/4"}
            cen.update({k: float(v) for k, v in zip(_FORCE_KEYS, r[:8])})
            bilin.append(cen)
            Nc, Mc = r[20:32].reshape(4, 3), r[32:44].reshape(4, 3)
            for k in range(4):
                row = {"eid": int(eid), "grid_id": int(node_ids[nd[k]])}
                row.update(dict(zip(_FORCE_KEYS, [float(v) for v in (*Nc[k], *Mc[k], r[6], r[7])])))
                bilin.append(row)


def static_results(arr, u, residual, q4_rows, t3_rows, bar_rows, solver_diagnostics=None):
    """(u_out, stresses, results_json) of the reference's
    ``_build_results_from_state`` from the recovered rows."""
    u = np.asarray(u, float)
    residual = np.asarray(residual, float)
    elem_keys = ("cbar", "quad4", "tria3", "crod", "conrod", "celas1")
    solid_keys = elem_keys + ("ctetra", "chexa", "cpenta")
    diag = dict(solver_diagnostics or {})
    out = {
        "displacements": [],
        "spc_forces": [],
        "forces": {k: [] for k in elem_keys},
        "forces_bilin": {"quad4": [], "tria3": []},
        "stresses": {k: [] for k in solid_keys},
        "strains": {k: [] for k in solid_keys},
        "solver_diagnostics": diag,
    }
    comps = ("t1", "t2", "t3", "r1", "r2", "r3")
    for i, gid in enumerate(np.asarray(arr.node_ids)):
        d = {"grid_id": int(gid)}
        d.update({c: float(v) for c, v in zip(comps, u[6 * i:6 * i + 6])})
        out["displacements"].append(d)
        r = residual[6 * i:6 * i + 6]
        if np.linalg.norm(r) > 1e-20:
            row = {"grid_id": int(gid)}
            row.update({c: float(v) for c, v in zip(comps, r)})
            out["spc_forces"].append(row)
    sums = {k: sum(s[c] for s in out["spc_forces"]) if out["spc_forces"] else 0.0
            for k, c in (("fx", "t1"), ("fy", "t2"), ("fz", "t3"))}
    diag["equilibrium"] = {"spc_reaction_sum": sums,
                           "residual": math.sqrt(sums["fx"] ** 2 + sums["fy"] ** 2 + sums["fz"] ** 2),
                           "relative_residual": 0.0}
    scal = {}
    node_ids = np.asarray(arr.node_ids)
    _shell_rows(np.asarray(q4_rows), arr.q4_eid, np.asarray(arr.q4prop)[:, 2], "quad4", out, scal,
                node_ids=node_ids, nodes=np.asarray(arr.q4))
    _shell_rows(np.asarray(t3_rows), arr.t3_eid, np.asarray(arr.t3prop)[:, 2], "tria3", out, scal)
    for r, eid in zip(np.asarray(bar_rows), arr.bar_eid):
        if r[17] == 0.0:
            continue
        eid = int(eid)
        f = {"eid": eid}
        f.update({k: float(v) for k, v in zip(_BAR_KEYS, r[:8])})
        out["forces"]["cbar"].append(f)
        scal[eid] = abs(float(r[8]))
        out["stresses"]["cbar"].append({
            "eid": eid, "axial": float(r[8]),
            "end_a": {"p%d" % (j + 1): float(r[9 + j]) for j in range(4)},
            "end_b": {"p%d" % (j + 1): float(r[13 + j]) for j in range(4)}})
    return u.copy(), scal, out


def solve_static(arr, ctx=None, device=0):
    """SOL 101 on the device for the model arrays: (u_out, stresses,
    results_json)."""
    from . import _native as N
    own = ctx is None
    ctx = ctx or N.Context(device)
    try:
        dm = N.DeviceModel(ctx, arr)
        try:
            u, free = dm.static_solve()
            res, q4, t3, bar = dm.static_recover()
        finally:
            dm.free()
    finally:
        if own:
            ctx.close()
    diag = {"backend": "mi355x_multifrontal_cholesky", "n_free": int(free.sum()), "ndof": int(arr.ndof)}
    return static_results(arr, u, res, q4, t3, bar, diag)
