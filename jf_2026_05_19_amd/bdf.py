"""Minimal Nastran bulk-data reader for the SOL 105 card subset.

Builds the same model dictionary shape as the reference's ModelBuilder
(``JFEM/src/ModelBuilder.jl:421-823``) for the cards this path supports:
GRID, CQUAD4, CTRIA3, CBAR, CBEAM, PSHELL, PBARL, MAT1, SPC1, SPC, FORCE,
EIGRL, PARAM, RBE1, RBE2, RBE3, RBAR, MPC, RSPLINE and the case-control keys SOL / SUBCASE / SPC / LOAD / METHOD /
STATSUB.  Field parsing follows ``JFEM/src/parsing/utilities.jl`` (small and
large fixed fields, free-field commas, ``+``/``*`` continuations, Nastran
exponent shorthand such as ``1.-5``).

Property and material extraction restate
``JFEM/src/parsing/extract_properties.jl:3-31`` (PSHELL, blank MID2 means
membrane only), ``:40-190`` (PBARL BAR section constants), ``:540-560``
(Cowper shear factor for a solid rectangle with nu = 0.3),
``JFEM/src/parsing/extract_materials.jl:3-27`` (MAT1 completion) and
``:465-480`` (EIGRL, ND defaults to 3).  RBE2, RBE3 and RBAR follow
``JFEM/src/parsing/extract_constraints.jl:46-141`` (RBE3 weight groups: a
field with '.', an exponent or an inner sign starts a new group; continuation
markers are skipped), RBAR becomes an RBE2 as ``ModelBuilder.jl:708-725``
converts it, MPC follows ``extract_constraints.jl:251-271``, RSPLINE
``:202-250`` and RBE1 ``:147-200``.
"""
from __future__ import annotations

import math
import re

__all__ = ["parse_number", "read_bdf_text", "read_bdf", "cards_to_model"]

_EXP_SHORTHAND = re.compile(r"^([+-]?[0-9.]+)([+-][0-9]+)$")


def parse_number(tok, default=None):
    """Nastran number: ints, floats, ``1.5+3`` / ``1.-5`` exponent shorthand."""
    if tok is None:
        return default
    s = str(tok).strip()
    if not s:
        return default
    try:
        if re.fullmatch(r"[+-]?\d+", s):
            return int(s)
        return float(s.replace("D", "E").replace("d", "e"))
    except ValueError:
        m = _EXP_SHORTHAND.match(s)
        if m:
            return float(m.group(1) + "e" + m.group(2))
        return default


def _split_fixed(line: str, large: bool):
    """Split one fixed-format line into its data fields (name excluded)."""
    width = 16 if large else 8
    name = line[:8]
    body = line[8:72] if not large else line[8:72]
    fields = [body[i:i + width] for i in range(0, len(body), width)]
    return name, fields


def _logical_cards(lines):
    """Yield (name, [fields...]) with continuations merged."""
    current = None
    for raw in lines:
        line = raw.rstrip("\n").rstrip("\r")
        if "$" in line:
            line = line[: line.index("$")]
        if not line.strip():
            continue
        if "," in line:
            parts = [p.strip() for p in line.split(",")]
            name = parts[0]
            fields = parts[1:]
            if name.startswith("+") or name.startswith("*") or name == "":
                if current is not None:
                    current[1].extend(fields)
                continue
            if current is not None:
                yield current
            current = (name.rstrip("*").upper(), fields)
            continue
        head = line[:8]
        large = head.strip().endswith("*") or head.startswith("*")
        if head.startswith("+") or head.startswith("*") or (head.strip() == "" and current is not None):
            _, fields = _split_fixed(line.ljust(80), large)
            if current is not None:
                current[1].extend(fields)
            continue
        if current is not None:
            yield current
        name, fields = _split_fixed(line.ljust(80), large)
        current = (name.strip().rstrip("*").upper(), fields)
    if current is not None:
        yield current


def read_bdf_text(text: str):
    """Return (cards, case_control) for a deck given as text."""
    lines = text.splitlines()
    case_lines, bulk_lines = [], []
    section = "exec"
    for ln in lines:
        u = ln.strip().upper()
        if section == "exec":
            if u.startswith("CEND"):
                section = "case"
                continue
            if u.startswith("SOL"):
                case_lines.append(ln)
            continue
        if section == "case":
            if u.startswith("BEGIN") and "BULK" in u:
                section = "bulk"
                continue
            case_lines.append(ln)
            continue
        if u.startswith("ENDDATA"):
            break
        bulk_lines.append(ln)
    if section == "exec":  # bulk-only fragment
        bulk_lines = lines
    cards = {}
    for name, fields in _logical_cards(bulk_lines):
        cards.setdefault(name, []).append([f.strip() for f in fields])
    cc = _parse_case_control(case_lines)
    return cards, cc


def read_bdf(path: str):
    with open(path, "r") as fh:
        return read_bdf_text(fh.read())


def _parse_case_control(lines):
    cc = {"SOL": 101, "SUBCASES": {}, "GLOBAL": {}}
    current = cc["GLOBAL"]
    for ln in lines:
        s = ln.split("$")[0].strip()
        if not s:
            continue
        u = s.upper()
        if u.startswith("SOL"):
            tok = u.split()
            if len(tok) > 1:
                cc["SOL"] = parse_number(tok[1], 101)
            continue
        if u.startswith("SUBCASE"):
            sid = parse_number(u.split()[1], 1)
            current = cc["SUBCASES"].setdefault(sid, {})
            continue
        if "=" in u:
            key, val = [p.strip() for p in u.split("=", 1)]
            key = key.split("(")[0].strip()
            num = parse_number(val.split()[0] if val else "", None)
            current[key] = num if num is not None else val
    if not cc["SUBCASES"]:
        cc["SUBCASES"][1] = {}
    for sub in cc["SUBCASES"].values():
        for k, v in cc["GLOBAL"].items():
            sub.setdefault(k, v)
    return cc


def _f(card, i, default=None):
    return parse_number(card[i], default) if i < len(card) else default


def _to_id(v):
    """to_id(parse_nastran_number(v)): integers as they are, reals rounded,
    anything else 0 (JFEM/src/parsing/utilities.jl:8-27)."""
    x = parse_number(v, None)
    if x is None:
        return 0
    if isinstance(x, float):
        if abs(x) >= 0.5 and abs(x - round(x)) < 1e-8:
            return int(round(x))
        return int(round(x))
    return int(x)


def _is_cont(sv):
    sv = sv.strip()
    return bool(sv) and sv[0] in "+*"


def _is_real_field(sv):
    sv = sv.strip()
    if not sv or _is_cont(sv):
        return False
    return "." in sv or re.search(r"[EeDd]", sv) is not None or (len(sv) > 1 and re.search(r"[+-]", sv[1:]) is not None)


def _rbe3_groups(c):
    """Weight groups of an RBE3 card from its fifth field on
    (extract_constraints.jl:88-122)."""
    groups, pos, n = [], 4, len(c)
    while pos < n:
        wt_str = str(c[pos]).strip()
        if not wt_str or _is_cont(wt_str):
            pos += 1
            continue
        wt = parse_number(wt_str, None)
        if wt is None:
            break
        pos += 1
        while pos < n:
            cs = str(c[pos]).strip()
            if cs and not _is_cont(cs):
                break
            pos += 1
        if pos >= n:
            break
        comps = _to_id(c[pos])
        pos += 1
        grids = []
        while pos < n:
            g = str(c[pos]).strip()
            if not g or _is_cont(g):
                pos += 1
                continue
            if _is_real_field(g):
                break
            gid = parse_number(g, None)
            if gid is None:
                break
            grids.append(int(gid))
            pos += 1
        if grids:
            groups.append({"wt": float(wt), "comps": comps, "grids": grids})
    return groups


def _pbarl_bar(dims, nu=0.3):
    """PBARL TYPE=BAR section constants (extract_properties.jl:109-121, 554-557)."""
    w, h = dims[0], dims[1]
    A = w * h
    I1 = w * h ** 3 / 12
    I2 = h * w ** 3 / 12
    if w >= h and w > 0:
        J = w * h ** 3 / 3 * (1 - 0.63 * h / w)
    elif h > 0:
        J = h * w ** 3 / 3 * (1 - 0.63 * w / h)
    else:
        J = I1 + I2
    K = 10.0 * (1.0 + nu) / (12.0 + 11.0 * nu)
    return dict(A=A, I=I1, I1=I1, I2=I2, J=J, K1=K, K2=K)


def cards_to_model(cards, cc):
    """Cards -> model dict with the reference's key layout."""
    model = {
        "GRIDs": {}, "CSHELLs": {}, "CBARs": {}, "CBEAMs": {}, "PSHELLs": {}, "PBARLs": {},
        "MATs": {}, "SPC1s": [], "SPCADDs": {}, "FORCEs": [], "MOMENTs": [], "EIGRLs": {},
        "CORDs": {}, "CASE_CONTROL": cc, "SOL": cc.get("SOL", 101),
    }
    for c in cards.get("GRID", []):
        gid = _f(c, 0)
        model["GRIDs"][str(gid)] = {
            "ID": gid, "CP": _f(c, 1, 0) or 0,
            "X": [float(_f(c, 2, 0.0)), float(_f(c, 3, 0.0)), float(_f(c, 4, 0.0))],
            "CD": _f(c, 5, 0) or 0, "PS": str(c[6]).strip() if len(c) > 6 else "",
        }
    for name, n in (("CQUAD4", 4), ("CTRIA3", 3)):
        for c in cards.get(name, []):
            eid = _f(c, 0)
            nodes = [_f(c, 2 + k) for k in range(n)]
            model["CSHELLs"][str(eid)] = {"ID": eid, "PID": _f(c, 1), "NODES": nodes,
                                         "THETA": 0.0 if len(c) <= 2 + n else float(_f(c, 2 + n, 0.0) or 0.0),
                                         "TYPE": name}
    for name, key in (("CBAR", "CBARs"), ("CBEAM", "CBEAMs")):
        for c in cards.get(name, []):
            eid = _f(c, 0)
            x1 = c[4] if len(c) > 4 else ""
            x1s = str(x1).strip()
            is_g0 = bool(x1s) and re.fullmatch(r"[+-]?\d+", x1s) is not None and not (
                len(c) > 5 and str(c[5]).strip())
            entry = {"ID": eid, "PID": _f(c, 1), "GA": _f(c, 2), "GB": _f(c, 3),
                     "PA": _f(c, 8, 0) or 0, "PB": _f(c, 9, 0) or 0,
                     "WA": [float(_f(c, 10, 0.0) or 0.0), float(_f(c, 11, 0.0) or 0.0), float(_f(c, 12, 0.0) or 0.0)],
                     "WB": [float(_f(c, 13, 0.0) or 0.0), float(_f(c, 14, 0.0) or 0.0), float(_f(c, 15, 0.0) or 0.0)]}
            if is_g0:
                entry["G0"] = int(x1s)
                entry["V"] = [0.0, 0.0, 0.0]
            else:
                entry["G0"] = 0
                entry["V"] = [float(_f(c, 4, 0.0) or 0.0), float(_f(c, 5, 0.0) or 0.0), float(_f(c, 6, 0.0) or 0.0)]
            model[key][str(eid)] = entry
    for c in cards.get("PSHELL", []):
        pid = _f(c, 0)
        mid2 = _f(c, 3, None)
        blank = mid2 is None or mid2 == 0
        mid3 = _f(c, 5, None)
        model["PSHELLs"][str(pid)] = {
            "PID": pid, "MID": _f(c, 1), "T": float(_f(c, 2, 0.0) or 0.0),
            "MID2": 0 if blank else mid2,
            "BEND_RATIO": 0.0 if blank else float(_f(c, 4, 1.0) if _f(c, 4, None) is not None else 1.0),
            "MID3": 0 if (mid3 is None or mid3 == 0) else mid3,
            "TS_T": float(_f(c, 6, 5.0 / 6.0) if _f(c, 6, None) is not None else 5.0 / 6.0),
            "NSM": float(_f(c, 7, 0.0) or 0.0),
        }
    for c in cards.get("PBARL", []):
        pid = _f(c, 0)
        typ = "ROD"
        dim_start = None
        for k in range(2, min(10, len(c))):
            v = str(c[k]).strip()
            if v and re.match(r"^[A-Za-z]", v):
                typ = v.upper()
                dim_start = k + 1
                break
        nums = [parse_number(v) for v in c[dim_start:]] if dim_start is not None else []
        nums = [float(v) for v in nums if isinstance(v, (int, float))]
        entry = {"PID": pid, "MID": _f(c, 1), "TYPE": typ, "DIMS": nums[:2]}
        if typ == "BAR" and len(nums) >= 2:
            entry.update(_pbarl_bar(nums[:2]))
        else:
            entry["UNSUPPORTED"] = True
        model["PBARLs"][str(pid)] = entry
    for c in cards.get("MAT1", []):
        mid = _f(c, 0)
        E = _f(c, 1, None); G = _f(c, 2, None); nu = _f(c, 3, None)
        E = 0.0 if E is None else float(E)
        G = 0.0 if G is None else float(G)
        nu = -1.0 if nu is None else float(nu)
        if E > 0 and G > 0 and nu < 0:
            nu = E / (2 * G) - 1.0
        elif E > 0 and nu >= 0 and G <= 0:
            G = E / (2 * (1 + nu))
        elif G > 0 and nu >= 0 and E <= 0:
            E = 2 * G * (1 + nu)
        if nu < 0:
            nu = 0.3
        model["MATs"][str(mid)] = {"MID": mid, "E": E, "G": G, "NU": nu, "TYPE": "MAT1"}
    for c in cards.get("SPC1", []):
        sid = _f(c, 0)
        comps = str(c[1]).strip()
        nodes = []
        rest = [v for v in c[2:] if str(v).strip()]
        if len(rest) >= 3 and str(rest[1]).strip().upper() == "THRU":
            a, b = int(parse_number(rest[0])), int(parse_number(rest[2]))
            nodes = list(range(a, b + 1))
        else:
            nodes = [int(parse_number(v)) for v in rest]
        model["SPC1s"].append({"SID": sid, "C": sorted(set(comps)), "NODES": nodes, "D": 0.0})
    for c in cards.get("SPC", []):
        sid = _f(c, 0)
        for k in range(1, len(c) - 1, 3):
            g = _f(c, k, None)
            if g is None:
                continue
            comps = str(c[k + 1]).strip()
            d = float(_f(c, k + 2, 0.0) or 0.0)
            model["SPC1s"].append({"SID": sid, "C": sorted(set(comps)), "NODES": [g], "D": d})
    for c in cards.get("FORCE", []):
        model["FORCEs"].append({"TYPE": "FORCE", "SID": _f(c, 0), "GID": _f(c, 1), "CID": _f(c, 2, 0) or 0,
                                "Mag": float(_f(c, 3, 0.0) or 0.0),
                                "Dir": [float(_f(c, 4, 0.0) or 0.0), float(_f(c, 5, 0.0) or 0.0),
                                        float(_f(c, 6, 0.0) or 0.0)]})
    for c in cards.get("EIGRL", []):
        sid = _f(c, 0)
        nd = _f(c, 3, 0) or 0
        model["EIGRLs"][str(sid)] = {"SID": sid, "V1": float(_f(c, 1, 0.0) or 0.0),
                                     "V2": float(_f(c, 2, 0.0) or 0.0), "ND": int(nd) if nd else 3}
    for c in cards.get("PARAM", []):
        key = str(c[0]).strip().upper()
        val = str(c[1]).strip() if len(c) > 1 else ""
        num = parse_number(val, None)
        model["PARAM_" + key] = num if num is not None else val
    model["RBE2s"], model["RBE3s"] = {}, {}
    for c in cards.get("RBE2", []):
        eid, gn, cm = _to_id(c[0] if c else ""), _to_id(c[1] if len(c) > 1 else ""), _to_id(c[2] if len(c) > 2 else "")
        slaves = [g for g in (_to_id(v) for v in c[3:]) if g > 0]
        if eid > 0 and gn > 0:
            model["RBE2s"][str(eid)] = {"ID": eid, "GN": gn, "CM": cm, "GM": slaves}
    for c in cards.get("RBAR", []):
        f = [_to_id(c[k]) if k < len(c) else 0 for k in range(7)]
        eid, ga, gb, cna, cnb = f[0], f[1], f[2], f[3], f[4]
        if eid > 0 and ga > 0 and gb > 0:
            if cnb > 0 and cna == 0:
                master, slave, cm = gb, ga, cnb
            elif cna > 0 and cnb == 0:
                master, slave, cm = ga, gb, cna
            else:
                master, slave, cm = ga, gb, 123456
            model["RBE2s"][str(eid)] = {"ID": eid, "GN": master, "CM": cm, "GM": [slave]}
    for c in cards.get("RBE3", []):
        eid = _to_id(c[0] if c else "")
        model["RBE3s"][str(eid)] = {"ID": eid, "REFGRID": _to_id(c[2] if len(c) > 2 else ""),
                                    "REFC": _to_id(c[3] if len(c) > 3 else ""), "WT_GROUPS": _rbe3_groups(c)}
    # MPC: SID, then (G, C, A) triples while a whole triple remains; terms
    # with G > 0 and C > 0 kept, equations with two or more terms kept
    # (extract_constraints.jl:251-271); every MPC card applies (constraints.jl:291)
    model["MPCs"] = []
    for c in cards.get("MPC", []):
        terms = []
        k = 1
        while k + 2 < len(c):
            gid, comp = _to_id(c[k]), _to_id(c[k + 1])
            coeff = float(parse_number(c[k + 2], 0.0) or 0.0)
            if gid > 0 and comp > 0:
                terms.append({"G": gid, "C": comp, "A": coeff})
            k += 3
        if len(terms) >= 2:
            model["MPCs"].append({"SID": _to_id(c[0] if c else ""), "TERMS": terms})
    # RSPLINE: EID, D/L, then GRIDs in polyline order; a GRID followed by a
    # component code (digits 0-6, at most 123456) is dependent in each
    # non-zero digit, any other GRID independent (extract_constraints.jl:202-250)
    model["RSPLINEs"] = {}
    for c in cards.get("RSPLINE", []):
        eid = _to_id(c[0] if c else "")
        g1 = _to_id(c[2] if len(c) > 2 else "")
        if eid <= 0 or g1 <= 0:
            continue
        indep, dep = [g1], []
        k = 3
        while k < len(c):
            g = _to_id(c[k])
            if g <= 0:
                k += 1
                continue
            if k + 1 < len(c):
                nxt = _to_id(c[k + 1]) if parse_number(c[k + 1], None) is not None else -1
                if 0 < nxt <= 123456 and all(ch in "0123456" for ch in str(nxt)):
                    dep.extend((g, int(ch)) for ch in str(nxt) if ch != "0")
                    k += 2
                    continue
            indep.append(g)
            k += 1
        if dep and len(indep) >= 2:
            model["RSPLINEs"][str(eid)] = {"ID": eid, "INDEP_GRIDS": indep, "DEP": dep}
    # RBE1: EID, (GRID, components) pairs up to the "UM" field, then the
    # dependent pairs; every digit of a component field is one DOF
    # (extract_constraints.jl:147-200)
    model["RBE1s"] = {}
    for c in cards.get("RBE1", []):
        eid = _to_id(c[0] if c else "")
        um = next((k for k in range(len(c)) if str(c[k]).strip().upper() == "UM"), -1)
        if eid <= 0 or um < 0:
            continue

        def pairs(k, stop):
            out = []
            while k + 1 < len(c) and k < stop:
                g, cn = _to_id(c[k]), _to_id(c[k + 1])
                if g > 0 and cn > 0:
                    out.extend((g, int(ch)) for ch in str(cn) if ch.isdigit())
                k += 2
            return out
        indep, dep = pairs(1, um), pairs(um + 1, len(c))
        if indep and dep:
            model["RBE1s"][str(eid)] = {"ID": eid, "INDEP": indep, "DEP": dep}
    unsupported = sorted(set(cards) - {"GRID", "CQUAD4", "CTRIA3", "CBAR", "CBEAM", "PSHELL", "PBARL",
                                       "MAT1", "SPC1", "SPC", "FORCE", "EIGRL", "PARAM", "RBE2", "RBE3", "RBAR",
                                       "MPC", "RSPLINE", "RBE1"})
    model["_UNSUPPORTED_CARDS"] = unsupported
    return model


def bdf_to_model_text(text: str):
    cards, cc = read_bdf_text(text)
    return cards_to_model(cards, cc)
