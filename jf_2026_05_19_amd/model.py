"""Model dictionary -> typed structure-of-arrays for the SOL 105 path.

The reference keeps its model as a ``Dict{String,Any}`` (built by
``JFEM/src/ModelBuilder.jl:421-823``) and re-derives every per-element array
inside ``assemble_stiffness`` (``JFEM/src/solver/assembly.jl:1953-2530``).
Here that derivation happens once, on the host, into contiguous arrays that
the C-ABI takes by pointer:

* node order = ascending GRID id (``assembly.jl:1958-1961``); DOF ``6*rank+c``;
* per-CQUAD4 and per-CTRIA3 property row ``[E, nu, T, BEND_RATIO, TS_T,
  G_shear]`` (``assembly.jl:2393-2476``: MID3 falls back to MID1 for the
  shear modulus; a CTRIA3 uses G_shear as its drilling reference modulus,
  ``assembly.jl:2454-2460, 2515``);
* per-CBAR property row ``[A, I1, I2, J, K1, K2, E, G]`` and the resolved
  orientation vector (G0 - GA, or V; ``helpers.jl:8-18``);
* SPC1/SPC set of the static subcase (``boundary_conditions.jl:634-677``);
* FORCE vector of the static subcase LOAD (``solve_case.jl:3-57``,
  ``loads.jl:50-53``), rotated into the analysis frame (identity for CD = 0);
* subcase pairing and EIGRL ND/V1/V2 (``JFEMSolver.jl:1302-1326, 1464-1472``).
* RBE2 (and RBAR) master-slave pairs and RBE3 weight groups by node rank, in
  element-id order; grids absent from the model are dropped as
  ``assemble_constraints`` skips them (``constraints.jl:44-60, 236-266``).
  The device builds the dependent-DOF map from them.

Features outside the supported default branch raise ``UnsupportedModel``
(the C-ABI returns SOL105_E_UNSUPPORTED for the same cases).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

__all__ = ["UnsupportedModel", "Sol105Arrays", "build_arrays"]


class UnsupportedModel(ValueError):
    pass


@dataclass
class Sol105Arrays:
    node_ids: np.ndarray          # (n,) int64, ascending GRID ids
    xyz: np.ndarray               # (n, 3) float64
    q4: np.ndarray                # (n_q4, 4) int32 node ranks
    q4_eid: np.ndarray            # (n_q4,) int64
    q4_pid: np.ndarray            # (n_q4,) int64
    q4prop: np.ndarray            # (n_q4, 6) float64  E, nu, T, BR, TS/T, G_shear
    t3: np.ndarray                # (n_t3, 3) int32 node ranks
    t3_eid: np.ndarray            # (n_t3,) int64
    t3_pid: np.ndarray            # (n_t3,) int64
    t3prop: np.ndarray            # (n_t3, 6) float64, same columns as q4prop
    bar: np.ndarray               # (n_bar, 2) int32
    bar_eid: np.ndarray
    bar_vref: np.ndarray          # (n_bar, 3)
    barprop: np.ndarray           # (n_bar, 8)  A, I1, I2, J, K1, K2, E, G
    spc_mask: np.ndarray          # (ndof,) uint8 1 = fixed by SPC1/SPC
    F: np.ndarray                 # (ndof,) float64
    nd: int = 3
    v1: float = 0.0
    v2: float = 0.0
    k6rot: float = 100.0
    autospc: bool = True
    node_has_shell: np.ndarray = field(default=None)
    node_has_bar: np.ndarray = field(default=None)
    buckling_subcase: int = 1
    static_subcase: int = 1
    shell_cards: int = -1         # CSHELL cards of the deck, assembled or not (-1: n_shell)
    bar_points: np.ndarray = field(default=None)   # (n_bar, 8) stress points C1 C2 D1 D2 E1 E2 F1 F2
    # constraints (RBE2 / RBAR as master-slave pairs, RBE3 elements with
    # weight groups), node ranks; component sets as digit integers (123456)
    rbe2: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))      # master, CM, slave
    rbe3: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))      # ref, REFC, first group, groups
    rbe3_grp: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))  # comps, first node, nodes
    rbe3_wt: np.ndarray = field(default_factory=lambda: np.zeros(0))
    rbe3_node: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    mpc_eq: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))    # first term, terms
    mpc_term: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))  # node rank, component
    mpc_coef: np.ndarray = field(default_factory=lambda: np.zeros(0))
    rspline: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))   # indep first/count, dep first/count
    rspline_indep: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    rspline_dep: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))  # node rank, component
    rbe1: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))       # indep first/count, dep first/count
    rbe1_indep: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))  # node rank, component
    rbe1_dep: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))    # node rank, component

    @property
    def n_nodes(self):
        return int(self.xyz.shape[0])

    @property
    def ndof(self):
        return 6 * self.n_nodes

    @property
    def n_q4(self):
        return int(self.q4.shape[0])

    @property
    def n_t3(self):
        return int(self.t3.shape[0])

    @property
    def n_shell(self):
        return self.n_q4 + self.n_t3

    @property
    def n_shell_cards(self):
        """Shell cards the mode filters' mesh-size guards count
        (``length(model["CSHELLs"])`` in solve_case.jl:2609, 3071), including
        cards the assembly skips."""
        return self.n_shell if self.shell_cards < 0 else int(self.shell_cards)

    @property
    def has_range(self):
        """EIGRL V1/V2 range mode (solve_case.jl:1938)."""
        return (self.v1 != 0.0 or self.v2 != 0.0) and self.v2 > self.v1

    @property
    def n_bar(self):
        return int(self.bar.shape[0])

    def with_q4_thickness(self, t: np.ndarray) -> "Sol105Arrays":
        """Copy with per-element CQUAD4 thickness replaced (design variants)."""
        import copy
        out = copy.copy(self)
        out.q4prop = self.q4prop.copy()
        out.q4prop[:, 2] = t
        return out

    def shell_thickness(self) -> np.ndarray:
        """Thickness of every shell, CQUAD4 first, then CTRIA3."""
        return np.concatenate([self.q4prop[:, 2], self.t3prop[:, 2]])

    def with_shell_thickness(self, t: np.ndarray) -> "Sol105Arrays":
        """Copy with the thickness of every shell replaced (CQUAD4 first,
        then CTRIA3)."""
        import copy
        t = np.asarray(t, dtype=float)
        out = copy.copy(self)
        out.q4prop = self.q4prop.copy()
        out.t3prop = self.t3prop.copy()
        out.q4prop[:, 2] = t[:self.n_q4]
        out.t3prop[:, 2] = t[self.n_q4:]
        return out


def _flat_dev(p):
    """Max out-of-plane corner deviation and the longer diagonal of each quad
    (p: E x 4 x 3), assembly.jl:2815-2826; zero deviation where the diagonals
    are parallel."""
    d13 = p[:, 2] - p[:, 0]
    d24 = p[:, 3] - p[:, 1]
    n = np.cross(d13, d24)
    ln = np.linalg.norm(n, axis=1)
    c = p.mean(axis=1)
    L = np.maximum(np.linalg.norm(d13, axis=1), np.linalg.norm(d24, axis=1))
    ok = ln > 1e-12
    nn = np.where(ok[:, None], n / np.where(ok, ln, 1.0)[:, None], 0.0)
    dev = np.abs(np.einsum("eki,ei->ek", p - c[:, None, :], nn)).max(axis=1)
    return np.where(ok, dev, 0.0), L


def _subcase_pairing(cc):
    subs = cc["SUBCASES"]
    sids = sorted(subs)
    buck = []
    for sid in sids:
        sub = subs[sid]
        if sub.get("METHOD") is not None:
            st = sub.get("STATSUB")
            buck.append((sid, int(st) if st is not None else 0))
    if buck and all(p[1] == 0 for p in buck):
        default = None
        for sid in sids:
            if subs[sid].get("LOAD") is not None and not any(p[0] == sid for p in buck):
                default = sid
                break
        if default is None:
            default = sids[0]
        buck = [(p[0], default) for p in buck]
    if not buck:
        buck = [(sids[1] if len(sids) >= 2 else sids[0], sids[0])]
    return buck


def build_arrays(model: dict, buckling_index: int = 0) -> Sol105Arrays:
    if model.get("_UNSUPPORTED_CARDS"):
        raise UnsupportedModel(f"unsupported cards: {model['_UNSUPPORTED_CARDS']}")
    grids = model["GRIDs"]
    ids = sorted(int(k) for k in grids)
    rank = {g: i for i, g in enumerate(ids)}
    n = len(ids)
    xyz = np.zeros((n, 3))
    for g in ids:
        gr = grids[str(g)]
        if (gr.get("CD", 0) or 0) != 0 or (gr.get("CP", 0) or 0) != 0:
            raise UnsupportedModel(f"GRID {g}: CP/CD coordinate systems are not supported")
        xyz[rank[g]] = gr["X"]
    mats = model["MATs"]
    pshells = model["PSHELLs"]
    q4, q4eid, q4pid, q4prop = [], [], [], []
    t3, t3eid, t3pid, t3prop = [], [], [], []
    for key in sorted(model["CSHELLs"], key=lambda k: int(k)):
        el = model["CSHELLs"][key]
        nodes = el["NODES"]
        if len(nodes) not in (3, 4):
            raise UnsupportedModel(f"element {el['ID']}: shells need 3 or 4 nodes")
        prop = pshells.get(str(el["PID"]))
        if prop is None:
            continue
        mat = mats.get(str(prop["MID"]))
        if mat is None:
            continue
        if any(nd not in rank for nd in nodes):
            continue
        mid3 = prop.get("MID3", 0) or 0
        smat = mats.get(str(mid3), mat) if mid3 else mat
        Gs = float(smat.get("G", 0.0))
        if Gs <= 0.0:
            E_ = float(smat.get("E", 0.0)); nu_ = float(smat.get("NU", 0.0))
            Gs = E_ / (2 * (1 + nu_)) if E_ > 0 else 0.0
        row = [float(mat["E"]), float(mat["NU"]), float(prop["T"]),
               float(prop.get("BEND_RATIO", 1.0)), float(prop.get("TS_T", 5.0 / 6.0)), Gs]
        if len(nodes) == 3:
            t3.append([rank[v] for v in nodes])
            t3eid.append(int(el["ID"]))
            t3pid.append(int(el["PID"]))
            t3prop.append(row)
            continue
        q4.append([rank[v] for v in nodes])
        q4eid.append(int(el["ID"]))
        q4pid.append(int(el["PID"]))
        q4prop.append(row)
    pbarls = model.get("PBARLs", {})
    bars, bareid, vref, barprop, barpts = [], [], [], [], []
    for group in ("CBARs", "CBEAMs"):
        for key in sorted(model.get(group, {}), key=lambda k: int(k)):
            b = model[group][key]
            prop = pbarls.get(str(b["PID"]))
            if prop is None:
                continue
            if prop.get("UNSUPPORTED"):
                raise UnsupportedModel(f"PBARL {prop['PID']}: only TYPE=BAR is supported")
            mat = mats.get(str(prop["MID"]))
            if mat is None or b["GA"] not in rank or b["GB"] not in rank:
                continue
            if (b.get("PA", 0) or 0) != 0 or (b.get("PB", 0) or 0) != 0:
                raise UnsupportedModel(f"CBAR {b['ID']}: pin flags are not supported")
            if any(abs(v) > 1e-10 for v in list(b.get("WA", [0, 0, 0])) + list(b.get("WB", [0, 0, 0]))):
                raise UnsupportedModel(f"CBAR {b['ID']}: offsets are not supported")
            g0 = b.get("G0", 0) or 0
            if g0 > 0 and g0 in rank:
                v = xyz[rank[g0]] - xyz[rank[b["GA"]]]
            else:
                v = np.asarray(b.get("V", [0.0, 0.0, 0.0]), dtype=float)
            bars.append([rank[b["GA"]], rank[b["GB"]]])
            bareid.append(int(b["ID"]))
            vref.append(v)
            barprop.append([prop["A"], prop.get("I1", prop.get("I", 0.0)), prop.get("I2", prop.get("I", 0.0)),
                            prop["J"], prop.get("K1", 0.0), prop.get("K2", 0.0), float(mat["E"]), float(mat["G"])])
            # stress recovery points of the BAR section (y, z) = C (h/2, w/2),
            # D (-h/2, w/2), E (-h/2, -w/2), F (h/2, -w/2), extract_properties.jl:176-178
            w2, h2 = prop["DIMS"][0] / 2, prop["DIMS"][1] / 2
            barpts.append([h2, w2, -h2, w2, -h2, -w2, h2, -w2])
    rbe2 = []
    for key in sorted(model.get("RBE2s", {}), key=lambda k: int(k)):
        r = model["RBE2s"][key]
        if r["GN"] not in rank:
            continue
        for g in r["GM"]:
            if g in rank:
                rbe2.append([rank[r["GN"]], int(r["CM"]), rank[g]])
    rbe3, rbe3_grp, rbe3_wt, rbe3_node = [], [], [], []
    for key in sorted(model.get("RBE3s", {}), key=lambda k: int(k)):
        r = model["RBE3s"][key]
        if r["REFGRID"] not in rank:
            continue
        g0 = len(rbe3_grp)
        for grp in r["WT_GROUPS"]:
            nodes = [rank[g] for g in grp["grids"] if g in rank]
            rbe3_grp.append([int(grp["comps"]), len(rbe3_node), len(nodes)])
            rbe3_wt.append(float(grp["wt"]))
            rbe3_node.extend(nodes)
        rbe3.append([rank[r["REFGRID"]], int(r["REFC"]), g0, len(rbe3_grp) - g0])
    # explicit MPC equations in card order, filtered as constraints.jl:
    # 294-312: the dependent term needs a known GRID, a component 1..6 and
    # |A| >= 1e-30; other terms with an unknown GRID or component are
    # skipped; an equation left without independent terms is dropped
    mpc_eq, mpc_term, mpc_coef = [], [], []
    for mpc in model.get("MPCs", []):
        terms = mpc["TERMS"]
        if len(terms) < 2:
            continue
        d0 = terms[0]
        if d0["G"] not in rank or not 1 <= d0["C"] <= 6 or abs(d0["A"]) < 1e-30:
            continue
        rest = [t for t in terms[1:] if t["G"] in rank and 1 <= t["C"] <= 6]
        if not rest:
            continue
        mpc_eq.append([len(mpc_term), 1 + len(rest)])
        for t in [d0] + rest:
            mpc_term.append([rank[t["G"]], int(t["C"])])
            mpc_coef.append(float(t["A"]))
    # RSPLINE elements in element-id order: independent GRIDs known to the
    # model (constraints.jl:155-160), dependent (GRID, component) pairs on
    # known GRIDs; elements with fewer than two known independent GRIDs skipped
    rspline, rspline_indep, rspline_dep = [], [], []
    for key in sorted(model.get("RSPLINEs", {}), key=lambda k: int(k)):
        r = model["RSPLINEs"][key]
        ind = [rank[g] for g in r["INDEP_GRIDS"] if g in rank]
        if len(ind) < 2:
            continue
        dep = [[rank[g], int(c)] for g, c in r["DEP"] if g in rank]
        rspline.append([len(rspline_indep), len(ind), len(rspline_dep), len(dep)])
        rspline_indep.extend(ind)
        rspline_dep.extend(dep)
    # RBE1 elements in element-id order (constraints.jl:80-112): one to six
    # independent DOFs, all on known GRIDs (else the element is skipped);
    # dependent DOFs on unknown GRIDs skipped.  Components outside 1..6 make
    # the reference's rigid row fail, so they are refused here.
    rbe1, rbe1_indep, rbe1_dep = [], [], []
    for key in sorted(model.get("RBE1s", {}), key=lambda k: int(k)):
        r = model["RBE1s"][key]
        for g, c in list(r["INDEP"]) + list(r["DEP"]):
            if not 1 <= int(c) <= 6:
                raise UnsupportedModel(f"RBE1 {r['ID']}: component {c} of GRID {g}")
        if not 1 <= len(r["INDEP"]) <= 6 or any(g not in rank for g, _ in r["INDEP"]):
            continue
        dep = [[rank[g], int(c)] for g, c in r["DEP"] if g in rank]
        rbe1.append([len(rbe1_indep), len(r["INDEP"]), len(rbe1_dep), len(dep)])
        rbe1_indep.extend([rank[g], int(c)] for g, c in r["INDEP"])
        rbe1_dep.extend(dep)
    ndof = 6 * n
    cc = model["CASE_CONTROL"]
    pairs = _subcase_pairing(cc)
    buck_sid, stat_sid = pairs[buckling_index]
    subs = cc["SUBCASES"]
    sub_static = subs.get(stat_sid, subs[sorted(subs)[0]])
    sub_buck = subs[buck_sid]
    load_id = sub_static.get("LOAD")
    spc_id = sub_buck.get("SPC", sub_static.get("SPC"))
    spc_mask = np.zeros(ndof, dtype=np.uint8)
    if spc_id is not None:
        sets = set(model.get("SPCADDs", {}).get(int(spc_id), [int(spc_id)]))
        for spc in model["SPC1s"]:
            if int(spc["SID"]) in sets:
                if abs(float(spc.get("D", 0.0))) > 0:
                    raise UnsupportedModel("enforced displacements are not supported")
                for g in spc["NODES"]:
                    if g in rank:
                        for c in spc["C"]:
                            spc_mask[6 * rank[g] + int(c) - 1] = 1
    for g in ids:
        ps = str(grids[str(g)].get("PS", "") or "").strip()
        for c in ps:
            if c.isdigit() and c != "0":
                spc_mask[6 * rank[g] + int(c) - 1] = 1
    F = np.zeros(ndof)
    if load_id is not None:
        for f in model["FORCEs"]:
            if int(f["SID"]) == int(load_id):
                if (f.get("CID", 0) or 0) != 0:
                    raise UnsupportedModel("FORCE with CID != 0 is not supported")
                g = f["GID"]
                F[6 * rank[g]:6 * rank[g] + 3] += np.asarray(f["Dir"], dtype=float) * f["Mag"]
        for m in model.get("MOMENTs", []):
            if int(m["SID"]) == int(load_id):
                raise UnsupportedModel("MOMENT loads are not supported")
    nd, v1, v2 = 3, 0.0, 0.0
    method = sub_buck.get("METHOD")
    if method is not None:
        e = model["EIGRLs"].get(str(int(method)))
        if e is not None:
            nd, v1, v2 = int(e.get("ND", 3)), float(e.get("V1", 0.0)), float(e.get("V2", 0.0))
    k6 = model.get("PARAM_K6ROT", 100.0)
    k6rot = float(k6) if isinstance(k6, (int, float)) else 100.0
    asp = model.get("PARAM_AUTOSPC", True)
    if isinstance(asp, str):
        autospc = asp.strip().upper() not in ("", "NO", "N", "FALSE", "F", "OFF", "0")
    elif isinstance(asp, (int, float)) and not isinstance(asp, bool):
        autospc = abs(float(asp)) > 1e-12
    else:
        autospc = bool(asp)
    snorm = model.get("PARAM_SNORM", 0.0)
    if isinstance(snorm, (int, float)) and snorm > 0:
        raise UnsupportedModel("PARAM,SNORM > 0 is not supported")
    q4a = np.asarray(q4, dtype=np.int32).reshape(-1, 4)
    if len(q4a):
        dev, L = _flat_dev(xyz[q4a])
        warped = np.flatnonzero(dev >= 1e-6 * np.maximum(L, 1e-12))
        if len(warped):
            raise UnsupportedModel(f"element {q4eid[warped[0]]}: warped CQUAD4 (curved-shell gates) not supported")
    t3a = np.asarray(t3, dtype=np.int32).reshape(-1, 3)
    bara = np.asarray(bars, dtype=np.int32).reshape(-1, 2)
    has_shell = np.zeros(n, dtype=np.uint8)
    has_bar = np.zeros(n, dtype=np.uint8)
    if q4a.size:
        has_shell[q4a.ravel()] = 1
    if t3a.size:
        has_shell[t3a.ravel()] = 1
    if bara.size:
        has_bar[bara.ravel()] = 1
    return Sol105Arrays(
        node_ids=np.asarray(ids, dtype=np.int64), xyz=xyz,
        q4=q4a, q4_eid=np.asarray(q4eid, dtype=np.int64), q4_pid=np.asarray(q4pid, dtype=np.int64),
        q4prop=np.asarray(q4prop, dtype=float).reshape(-1, 6),
        t3=t3a, t3_eid=np.asarray(t3eid, dtype=np.int64), t3_pid=np.asarray(t3pid, dtype=np.int64),
        t3prop=np.asarray(t3prop, dtype=float).reshape(-1, 6),
        bar=bara, bar_eid=np.asarray(bareid, dtype=np.int64),
        bar_vref=np.asarray(vref, dtype=float).reshape(-1, 3),
        barprop=np.asarray(barprop, dtype=float).reshape(-1, 8),
        spc_mask=spc_mask, F=F, nd=nd, v1=v1, v2=v2, k6rot=k6rot, autospc=autospc,
        node_has_shell=has_shell, node_has_bar=has_bar,
        buckling_subcase=int(buck_sid), static_subcase=int(stat_sid),
        shell_cards=len(model["CSHELLs"]),
        bar_points=np.asarray(barpts, dtype=float).reshape(-1, 8),
        rbe2=np.asarray(rbe2, dtype=np.int32).reshape(-1, 3), rbe3=np.asarray(rbe3, dtype=np.int32).reshape(-1, 4),
        rbe3_grp=np.asarray(rbe3_grp, dtype=np.int32).reshape(-1, 3), rbe3_wt=np.asarray(rbe3_wt, dtype=float),
        rbe3_node=np.asarray(rbe3_node, dtype=np.int32),
        mpc_eq=np.asarray(mpc_eq, dtype=np.int32).reshape(-1, 2),
        mpc_term=np.asarray(mpc_term, dtype=np.int32).reshape(-1, 2), mpc_coef=np.asarray(mpc_coef, dtype=float),
        rspline=np.asarray(rspline, dtype=np.int32).reshape(-1, 4),
        rspline_indep=np.asarray(rspline_indep, dtype=np.int32),
        rspline_dep=np.asarray(rspline_dep, dtype=np.int32).reshape(-1, 2),
        rbe1=np.asarray(rbe1, dtype=np.int32).reshape(-1, 4),
        rbe1_indep=np.asarray(rbe1_indep, dtype=np.int32).reshape(-1, 2),
        rbe1_dep=np.asarray(rbe1_dep, dtype=np.int32).reshape(-1, 2),
    )
