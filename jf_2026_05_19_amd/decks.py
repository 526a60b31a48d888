"""Seeded synthetic SOL 105 decks (free-field BDF text) for the workloads the
path is measured on.  The decks use only the card subset of
``jf_2026_05_19_amd.bdf`` so the same text can be fed to the reference's
parser (``JFEM/src/parsing/card_processor.jl``).

* ``plate_deck``  -- flat square plate, CQUAD4 mesh, simply supported on all
  edges, uniform compressive edge load on x = a reacted at x = 0, consistent
  nodal lumping (half load at the corners).
* ``tri_frac`` (plate and panel) -- a seeded fraction of the quads is split
  into two CTRIA3 along an alternating diagonal (mixed CQUAD4/CTRIA3 mesh).
  A region meshed only with CTRIA3 has two zero-energy drilling patterns
  (theta_z alternating over a 3-colouring of its nodes: the one-point
  drilling penalty of each triangle sees only the mean of its three
  theta_z); ``spc_drill`` constrains theta_z everywhere for such decks.
* ``panel_deck``  -- flat panel with CBAR stringers (PBARL BAR) on every k-th
  longitudinal mesh line, clamped loaded edges, axial compression.
* ``variant_sweep`` -- seeded design variants of the plate (mesh size,
  thickness, modulus).
* ``served_families`` -- seeded stiffened-panel meshes for a request stream
  (DOF log-uniform, 10 % of the quads split into CTRIA3, CBAR stringers).
* ``cylinder_deck`` -- axially compressed cylindrical shell of flat CQUAD4
  facets with a circumferential band of CTRIA3 (each quad of the band split
  in two), clamped at z = 0, axial ring load at z = L.
* ``glued_plate_deck`` -- the plate meshed as two halves with duplicated
  seam GRIDs tied by RBE2 (123456, coincident nodes): the same structure as
  the one-piece plate.
* ``eccentric_panel_deck`` -- a plate with CBAR stringers on their own GRIDs
  at an eccentricity above the skin, each tied to the skin node beneath by an
  RBE2 (123456).
* ``rbe3_load_plate_deck`` -- the plate with its edge load applied at one
  reference GRID and spread over the loaded edge by an RBE3.
"""
from __future__ import annotations

import numpy as np

__all__ = ["plate_deck", "panel_deck", "variant_sweep", "cylinder_deck", "served_families", "glued_plate_deck",
           "eccentric_panel_deck", "rbe3_load_plate_deck"]


def _fmt(v):
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, str):
        return v
    return repr(float(v))


def _card(*fields):
    return ",".join(_fmt(f) for f in fields)


def _header(nd, sol=105):
    return [
        f"SOL {sol}",
        "CEND",
        "SUBCASE 1",
        "  SPC = 1",
        "  LOAD = 10",
        "SUBCASE 2",
        "  SPC = 1",
        "  METHOD = 100",
        "  STATSUB = 1",
        "BEGIN BULK",
        _card("EIGRL", 100, "", "", nd),
    ]


def _shells(L, nx, ny, gid, tri_frac, tri_seed):
    """CQUAD4 per mesh cell, or two CTRIA3 for a seeded fraction of cells."""
    rng = np.random.default_rng(tri_seed)
    split = rng.random(nx * ny) < tri_frac
    eid = 1
    for j in range(ny):
        for i in range(nx):
            c = (gid(i, j), gid(i + 1, j), gid(i + 1, j + 1), gid(i, j + 1))
            if split[j * nx + i]:
                if (i + j) % 2 == 0:
                    L.append(_card("CTRIA3", eid, 1, c[0], c[1], c[2]))
                    L.append(_card("CTRIA3", 5000000 + eid, 1, c[0], c[2], c[3]))
                else:
                    L.append(_card("CTRIA3", eid, 1, c[0], c[1], c[3]))
                    L.append(_card("CTRIA3", 5000000 + eid, 1, c[1], c[2], c[3]))
            else:
                L.append(_card("CQUAD4", eid, 1, *c))
            eid += 1


def plate_deck(nx=20, ny=20, a=500.0, b=500.0, t=2.0, E=70e3, nu=0.33, line_load=1.0, nd=5,
               k6rot=None, tri_frac=0.0, tri_seed=20260520, spc_drill=False):
    """Simply supported plate under uniaxial edge compression (N/mm)."""
    L = _header(nd)
    if k6rot is not None:
        L.append(_card("PARAM", "K6ROT", float(k6rot)))
    L.append(_card("MAT1", 1, float(E), "", float(nu)))
    L.append(_card("PSHELL", 1, 1, float(t), 1, "", ""))
    gid = lambda i, j: j * (nx + 1) + i + 1
    for j in range(ny + 1):
        for i in range(nx + 1):
            L.append(_card("GRID", gid(i, j), "", a * i / nx, b * j / ny, 0.0))
    _shells(L, nx, ny, gid, tri_frac, tri_seed)
    edge = sorted({gid(i, j) for j in range(ny + 1) for i in range(nx + 1)
                   if i in (0, nx) or j in (0, ny)})
    for g in edge:
        L.append(_card("SPC1", 1, "3", g))
    for j in range(ny + 1):
        L.append(_card("SPC1", 1, "1", gid(0, j)))
    L.append(_card("SPC1", 1, "2", gid(0, 0)))
    if spc_drill:
        for j in range(ny + 1):
            for i in range(nx + 1):
                L.append(_card("SPC1", 1, "6", gid(i, j)))
    dy = b / ny
    for j in range(ny + 1):
        w = 0.5 if j in (0, ny) else 1.0
        L.append(_card("FORCE", 10, gid(nx, j), 0, line_load * dy * w, -1.0, 0.0, 0.0))
    L.append("ENDDATA")
    return "\n".join(L) + "\n"


def panel_deck(nx=200, ny=100, a=1000.0, b=500.0, t=1.5, E=70e3, nu=0.33, stringer_every=10,
               bar_dims=(3.0, 20.0), line_load=1.0, nd=10, tri_frac=0.0, tri_seed=20260520):
    """Stringer-stiffened panel: clamped loaded edges, axial compression."""
    L = _header(nd)
    L.append(_card("MAT1", 1, float(E), "", float(nu)))
    L.append(_card("PSHELL", 1, 1, float(t), 1, "", ""))
    L.append(_card("PBARL", 2, 1, "", "BAR", float(bar_dims[0]), float(bar_dims[1])))
    gid = lambda i, j: j * (nx + 1) + i + 1
    for j in range(ny + 1):
        for i in range(nx + 1):
            L.append(_card("GRID", gid(i, j), "", a * i / nx, b * j / ny, 0.0))
    _shells(L, nx, ny, gid, tri_frac, tri_seed)
    bid = 1000001
    for j in range(0, ny + 1, stringer_every):
        for i in range(nx):
            L.append(_card("CBAR", bid, 2, gid(i, j), gid(i + 1, j), 0.0, 0.0, 1.0))
            bid += 1
    for j in range(ny + 1):
        L.append(_card("SPC1", 1, "123456", gid(0, j)))
        L.append(_card("SPC1", 1, "23456", gid(nx, j)))
    dy = b / ny
    for j in range(ny + 1):
        w = 0.5 if j in (0, ny) else 1.0
        L.append(_card("FORCE", 10, gid(nx, j), 0, line_load * dy * w, -1.0, 0.0, 0.0))
    L.append("ENDDATA")
    return "\n".join(L) + "\n"


def variant_sweep(n=256, seed=20260519, n_min=10, n_max=80):
    """Seeded plate design variants: (deck text, description) pairs."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        m = int(rng.integers(n_min, n_max + 1))
        t = float(rng.uniform(1.0, 4.0))
        steel = bool(rng.integers(0, 2))
        E, nu = (210e3, 0.30) if steel else (70e3, 0.33)
        out.append((plate_deck(nx=m, ny=m, t=t, E=E, nu=nu), dict(mesh=m, t=t, E=E, nu=nu)))
    return out


def served_families(n=10, seed=20260520, dof_min=2e3, dof_max=5e5, nd=5):
    """Seeded panel meshes of a request stream: DOF log-uniform in
    [dof_min, dof_max] (nx = 2 ny, 5 mm cells), 10 % of the quads split into
    two CTRIA3, CBAR stringers on about every tenth longitudinal line.
    Returns (deck text, description) pairs."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        dof = float(np.exp(rng.uniform(np.log(dof_min), np.log(dof_max))))
        ny = max(4, int(round(np.sqrt(dof / 12.0))))
        nx = 2 * ny
        every = max(2, ny // 10)
        txt = panel_deck(nx=nx, ny=ny, a=5.0 * nx, b=5.0 * ny, stringer_every=every, nd=nd, tri_frac=0.10,
                         tri_seed=seed + k)
        out.append((txt, dict(nx=nx, ny=ny, ndof=6 * (nx + 1) * (ny + 1), stringer_every=every)))
    return out


def cylinder_deck(n_circ=416, n_axial=400, R=500.0, L=1000.0, t=1.0, E=70e3, nu=0.33, line_load=1.0, nd=10,
                  tri_band=0.10):
    """Cylindrical shell under axial compression (N/mm of circumference).

    n_circ x n_axial GRIDs on the radius-R surface (theta_i = 2 pi i / n_circ,
    z_j = L j / (n_axial - 1)); one flat CQUAD4 per grid cell (a chord of the
    circle extruded along z, so every facet is planar).  The element rows
    whose centres lie in the middle tri_band fraction of the length are split
    into two CTRIA3 each (alternating diagonals), i.e. tri_band of the area.
    End z = 0 clamped (123456); end z = L: components 1, 2, 4, 5 fixed and a
    compressive axial FORCE ring, line_load times the node's share of the
    circumference.
    """
    L_ = _header(nd)
    L_.append(_card("MAT1", 1, float(E), "", float(nu)))
    L_.append(_card("PSHELL", 1, 1, float(t), 1, "", ""))
    gid = lambda i, j: j * n_circ + (i % n_circ) + 1
    for j in range(n_axial):
        z = L * j / (n_axial - 1)
        for i in range(n_circ):
            th = 2.0 * np.pi * i / n_circ
            L_.append(_card("GRID", gid(i, j), "", R * np.cos(th), R * np.sin(th), z))
    n_rows = n_axial - 1
    lo, hi = 0.5 * (1.0 - tri_band), 0.5 * (1.0 + tri_band)
    eid = 1
    for j in range(n_rows):
        zc = (j + 0.5) / n_rows
        band = tri_band > 0 and lo <= zc < hi
        for i in range(n_circ):
            c = (gid(i, j), gid(i + 1, j), gid(i + 1, j + 1), gid(i, j + 1))
            if band:
                if (i + j) % 2 == 0:
                    L_.append(_card("CTRIA3", eid, 1, c[0], c[1], c[2]))
                    L_.append(_card("CTRIA3", 5000000 + eid, 1, c[0], c[2], c[3]))
                else:
                    L_.append(_card("CTRIA3", eid, 1, c[0], c[1], c[3]))
                    L_.append(_card("CTRIA3", 5000000 + eid, 1, c[1], c[2], c[3]))
            else:
                L_.append(_card("CQUAD4", eid, 1, *c))
            eid += 1
    for i in range(n_circ):
        L_.append(_card("SPC1", 1, "123456", gid(i, 0)))
        L_.append(_card("SPC1", 1, "1245", gid(i, n_axial - 1)))
    f = line_load * 2.0 * np.pi * R / n_circ
    for i in range(n_circ):
        L_.append(_card("FORCE", 10, gid(i, n_axial - 1), 0, f, 0.0, 0.0, -1.0))
    L_.append("ENDDATA")
    return "\n".join(L_) + "\n"


def _plate_bc_and_load(L, nx, ny, b, gid_at, line_load, load_edge=True):
    """plate_deck's supports and edge load on the GRID ids gid_at(i, j)."""
    edge = sorted({gid_at(i, j) for j in range(ny + 1) for i in range(nx + 1) if i in (0, nx) or j in (0, ny)})
    for g in edge:
        L.append(_card("SPC1", 1, "3", g))
    for j in range(ny + 1):
        L.append(_card("SPC1", 1, "1", gid_at(0, j)))
    L.append(_card("SPC1", 1, "2", gid_at(0, 0)))
    if load_edge:
        dy = b / ny
        for j in range(ny + 1):
            w = 0.5 if j in (0, ny) else 1.0
            L.append(_card("FORCE", 10, gid_at(nx, j), 0, line_load * dy * w, -1.0, 0.0, 0.0))


def glued_plate_deck(nx=12, ny=8, a=400.0, b=300.0, t=1.5, E=70e3, nu=0.33, line_load=1.0, nd=4, tie="rbe2"):
    """plate_deck's plate (nx even) as two half meshes: the cells right of
    x = a/2 use their own GRIDs (ids + 100000) and each seam GRID of the
    right half is the RBE2 (123456) slave of the coincident left one
    (tie="rbe2"), tied to it by six MPC equations u_r - u_l = 0
    (tie="mpc"), or the dependent GRID of an RBE1 whose six independent
    DOFs are the left GRID's (tie="rbe1", the same rigid link)."""
    assert nx % 2 == 0
    L = _header(nd)
    L.append(_card("MAT1", 1, float(E), "", float(nu)))
    L.append(_card("PSHELL", 1, 1, float(t), 1, "", ""))
    h = nx // 2
    gl = lambda i, j: j * (nx + 1) + i + 1
    gr = lambda i, j: 100000 + j * (nx + 1) + i + 1
    gid_at = lambda i, j: gl(i, j) if i <= h else gr(i, j)
    for j in range(ny + 1):
        for i in range(nx + 1):
            if i <= h:
                L.append(_card("GRID", gl(i, j), "", a * i / nx, b * j / ny, 0.0))
            if i >= h:
                L.append(_card("GRID", gr(i, j), "", a * i / nx, b * j / ny, 0.0))
    eid = 1
    for j in range(ny):
        for i in range(nx):
            g = gl if i < h else gr
            L.append(_card("CQUAD4", eid, 1, g(i, j), g(i + 1, j), g(i + 1, j + 1), g(i, j + 1)))
            eid += 1
    for j in range(ny + 1):
        if tie == "rbe1":
            L.append(_card("RBE1", 900000 + j, gl(h, j), 123456, "UM", gr(h, j), 123456))
        elif tie == "mpc":
            for c in range(1, 7):
                L.append(_card("MPC", 1, gr(h, j), c, 1.0, gl(h, j), c, -1.0))
        else:
            L.append(_card("RBE2", 900000 + j, gl(h, j), 123456, gr(h, j)))
    _plate_bc_and_load(L, nx, ny, b, gid_at, line_load)
    L.append("ENDDATA")
    return "\n".join(L) + "\n"


def eccentric_panel_deck(nx=16, ny=8, a=400.0, b=200.0, t=1.5, E=70e3, nu=0.33, stringer_every=4, ecc=12.0,
                         dims=(3.0, 20.0), line_load=1.0, nd=4):
    """A simply supported plate with CBAR stringers (PBARL BAR) on every
    stringer_every-th longitudinal mesh line, on their own GRIDs at z = ecc
    (ids + 200000), each tied to the skin GRID beneath by an RBE2 (123456,
    the skin node independent)."""
    L = _header(nd)
    L.append(_card("MAT1", 1, float(E), "", float(nu)))
    L.append(_card("PSHELL", 1, 1, float(t), 1, "", ""))
    L.append(_card("PBARL", 2, 1, "", "BAR", float(dims[0]), float(dims[1])))
    gid = lambda i, j: j * (nx + 1) + i + 1
    for j in range(ny + 1):
        for i in range(nx + 1):
            L.append(_card("GRID", gid(i, j), "", a * i / nx, b * j / ny, 0.0))
    _shells(L, nx, ny, gid, 0.0, 0)
    beid = 100000
    for j in range(stringer_every, ny, stringer_every):
        for i in range(nx + 1):
            L.append(_card("GRID", 200000 + gid(i, j), "", a * i / nx, b * j / ny, float(ecc)))
            L.append(_card("RBE2", 300000 + gid(i, j), gid(i, j), 123456, 200000 + gid(i, j)))
        for i in range(nx):
            L.append(_card("CBAR", beid, 2, 200000 + gid(i, j), 200000 + gid(i + 1, j), 0.0, 0.0, 1.0))
            beid += 1
    _plate_bc_and_load(L, nx, ny, b, gid, line_load)
    L.append("ENDDATA")
    return "\n".join(L) + "\n"


def rbe3_load_plate_deck(nx=10, ny=8, a=400.0, b=300.0, t=1.5, E=70e3, nu=0.33, line_load=1.0, nd=4,
                         offset=10.0):
    """plate_deck's plate whose edge load is one FORCE at a reference GRID
    (id 999999) beside the middle of the loaded edge, spread over the edge
    GRIDs by an RBE3 (REFC 1; one group: weight 1.0, component 1)."""
    L = _header(nd)
    L.append(_card("MAT1", 1, float(E), "", float(nu)))
    L.append(_card("PSHELL", 1, 1, float(t), 1, "", ""))
    gid = lambda i, j: j * (nx + 1) + i + 1
    for j in range(ny + 1):
        for i in range(nx + 1):
            L.append(_card("GRID", gid(i, j), "", a * i / nx, b * j / ny, 0.0))
    L.append(_card("GRID", 999999, "", a + float(offset), b / 2, 0.0))
    _shells(L, nx, ny, gid, 0.0, 0)
    L.append(_card("RBE3", 999000, "", 999999, 1, 1.0, 1, *[gid(nx, j) for j in range(ny + 1)]))
    _plate_bc_and_load(L, nx, ny, b, gid, line_load, load_edge=False)
    L.append(_card("FORCE", 10, 999999, 0, line_load * b, -1.0, 0.0, 0.0))
    L.append("ENDDATA")
    return "\n".join(L) + "\n"


def rspline_load_plate_deck(nx=10, ny=8, a=400.0, b=300.0, t=1.5, E=70e3, nu=0.33, line_load=1.0, nd=4,
                            offset=10.0, frac=0.25, P=-150.0):
    """plate_deck's plate with an extra x force P at a GRID (id 999999)
    beside the loaded edge, tied in x by an RSPLINE to the last three edge
    GRIDs (fraction frac along the last segment): the force reaches those
    GRIDs with weights 1 - frac and frac.  With nx = 10 the edge GRIDs of
    the spline are 77, 88 and 99, ids that the RSPLINE field rule cannot
    take for component codes."""
    base = plate_deck(nx=nx, ny=ny, a=a, b=b, t=t, E=E, nu=nu, line_load=line_load, nd=nd)
    gid = lambda i, j: j * (nx + 1) + i + 1
    g = [gid(nx, ny - 2), gid(nx, ny - 1), gid(nx, ny)]
    y = b * (ny - 1) / ny + frac * b / ny
    extra = [_card("GRID", 999999, "", a + float(offset), y, 0.0),
             _card("RSPLINE", 1, 0.1, g[0], g[1], g[2], 999999, 1),
             _card("FORCE", 10, 999999, 0, float(P), 1.0, 0.0, 0.0)]
    return base.replace("ENDDATA", "\n".join(extra) + "\nENDDATA")
