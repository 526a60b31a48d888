"""RBE1 / RBE2 / RBE3 / RSPLINE / MPC constraints on the CPU: pins of the oracle's restatement of
the reference's dependent-DOF elimination (JFEM/src/solver/constraints.jl:
37-364, the load redistribution of solve_case.jl:40-55 and the recovery of
boundary_conditions.jl:978-988, solve_case.jl:3153-3162), the library's host
constraint code (constraints.cpp, built with g++) against the oracle, and the
native deck reader against the host builder on constraint decks.

Pins (expectations from mechanics, not from the restatement):
* two half meshes glued by RBE2 (123456) pairs of coincident GRIDs are the
  one-piece plate: K and F on the independent DOFs, u, lambda and the modes;
* a cantilever CBAR with an RBE2 arm of length e at its tip, loaded at the
  arm's end: tip deflection P L^3/(3 E I) + P L/(k G A), twist e P L/(G J),
  and the arm end moves rigidly with the tip (w + e theta_x);
* two half meshes tied by MPC equations u_r - u_l = 0 are the one-piece
  plate; a load at a GRID defined by an MPC average of two GRIDs acts as
  half the load at each of them; an MPC replaces the RBE2 entries of its
  DOF, and a later MPC an earlier one;
* RBE3: a rigid motion of the independent grids gives the reference the
  same rigid motion (the least-squares fit is exact for it), and a load at
  the reference spreads into grid forces with the same resultant force and
  moment about the reference.
"""
import ctypes
import os
import subprocess
import types

import numpy as np
import pytest

from jf_2026_05_19_amd import bdf, decks, model as M
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native_host", "constraint_probe.cpp")
LIBDIR = os.path.join(HERE, "native_host", "_build")
LIB = os.path.join(LIBDIR, "libconstraint_probe.so")


def _arr(txt):
    return M.build_arrays(bdf.bdf_to_model_text(txt))


def _static(arr):
    """Oracle static solve with the constraints (dependents recovered)."""
    cmap = O.constraint_map(arr)
    K = O.assemble_K(arr)
    ac = O.constrained_arrays(arr, cmap)
    u, free = O.static_solve(ac, K)
    O.recover_dependent(cmap, u)
    return u, K, ac, cmap


def test_glued_halves_equal_one_piece_plate():
    g = _arr(decks.glued_plate_deck(nx=12, ny=8))
    p = _arr(decks.plate_deck(nx=12, ny=8, a=400.0, b=300.0, t=1.5, nd=4))
    rg, rp = O.solve(g), O.solve(p)
    assert np.allclose(rg["eigenvalues"], rp["eigenvalues"], rtol=1e-10, atol=0)
    key = {(round(x[0], 6), round(x[1], 6)): i for i, x in enumerate(p.xyz)}
    for i, x in enumerate(g.xyz):
        j = key[(round(x[0], 6), round(x[1], 6))]
        assert np.allclose(rg["u"][6 * i:6 * i + 6], rp["u"][6 * j:6 * j + 6], rtol=0,
                           atol=1e-10 * np.abs(rp["u"]).max())
        for m in range(3):
            assert abs(abs(rg["modes"][6 * i + 2, m]) - abs(rp["modes"][6 * j + 2, m])) <= 1e-8


def _cantilever_deck(L=300.0, e=40.0, s=10.0, P=50.0, E=70e3, nu=0.33):
    lines = ["SOL 101", "CEND", "SUBCASE 1", "  SPC = 1", "  LOAD = 10", "BEGIN BULK",
             "MAT1,1,%r,,%r" % (E, nu), "PBARL,2,1,,BAR,%r,%r" % (s, s),
             "GRID,1,,0.0,0.0,0.0", "GRID,2,,%r,0.0,0.0" % L, "GRID,3,,%r,%r,0.0" % (L, e),
             "CBAR,7,2,1,2,0.0,0.0,1.0", "RBE2,8,2,123456,3", "SPC1,1,123456,1",
             "FORCE,10,3,0,%r,0.0,0.0,1.0" % P, "ENDDATA"]
    return "\n".join(lines) + "\n"


def test_rbe2_arm_on_cantilever_closed_form():
    L, e, s, P, E, nu = 300.0, 40.0, 10.0, 50.0, 70e3, 0.33
    arr = _arr(_cantilever_deck(L, e, s, P, E, nu))
    u, K, ac, cmap = _static(arr)
    G = E / (2 * (1 + nu))
    A, I = s * s, s ** 4 / 12
    J = s * s ** 3 / 3 * (1 - 0.63)
    kap = 10 * (1 + 0.3) / (12 + 11 * 0.3)
    w2 = P * L ** 3 / (3 * E * I) + P * L / (kap * G * A)
    th = e * P * L / (G * J)
    assert np.isclose(u[6 + 2], w2, rtol=1e-10)          # tip deflection (GRID 2, uz)
    assert np.isclose(u[6 + 3], th, rtol=1e-10)          # tip twist (thx)
    assert np.isclose(u[12 + 2], w2 + e * th, rtol=1e-10)  # arm end: uz + e thx
    assert np.isclose(u[12 + 3], th, rtol=1e-12)
    # the force moved to the tip as a force and a torque
    assert np.isclose(ac.F[6 + 2], P) and np.isclose(ac.F[6 + 3], e * P) and not ac.F[12:].any()


def _rbe3_ns(xyz, ref, refc, groups):
    """An arrays stand-in holding only what constraint_map reads."""
    grp, wt, nodes, = [], [], []
    for w, comps, g in groups:
        grp.append([comps, len(nodes), len(g)])
        wt.append(w)
        nodes.extend(g)
    return types.SimpleNamespace(xyz=np.asarray(xyz, float), rbe2=np.zeros((0, 3), np.int32),
                                 rbe3=np.array([[ref, refc, 0, len(grp)]], np.int32),
                                 rbe3_grp=np.array(grp, np.int32), rbe3_wt=np.array(wt, float),
                                 rbe3_node=np.array(nodes, np.int32))


def _rigid(xyz, t, th):
    u = np.zeros(6 * len(xyz))
    for i, x in enumerate(xyz):
        u[6 * i:6 * i + 3] = t + np.cross(th, x)
        u[6 * i + 3:6 * i + 6] = th
    return u


def test_rbe3_reproduces_rigid_motion_and_balances_load():
    rng = np.random.default_rng(4)
    xyz = np.vstack([rng.normal(size=(6, 3)) * 50.0, [[3.0, -4.0, 2.0]]])   # grids 0..5, reference 6
    ns = _rbe3_ns(xyz, 6, 123456, [(1.0, 123, [0, 1, 2]), (2.5, 123, [3, 4, 5])])
    cmap = O.constraint_map(ns)
    assert sorted(cmap) == [36 + c for c in range(6)]
    t, th = rng.normal(size=3), rng.normal(size=3) * 1e-2
    u = _rigid(xyz, t, th)
    ref = u.copy()
    O.recover_dependent(cmap, ref)
    assert np.allclose(ref[36:42], u[36:42], rtol=0, atol=1e-12 * np.abs(u).max())
    # a load (force and moment) at the reference
    Fr = rng.normal(size=6)
    F = np.zeros(42)
    F[36:42] = Fr
    for dep, pairs in cmap.items():
        for ind, c in pairs:
            F[ind] += F[dep] * c
    grid = F[:36].reshape(6, 6)
    assert np.allclose(grid[:, :3].sum(axis=0), Fr[:3], rtol=1e-12, atol=1e-12)
    mom = sum(np.cross(xyz[i] - xyz[6], grid[i, :3]) for i in range(6))
    assert np.allclose(mom, Fr[3:], rtol=1e-10, atol=1e-10)


def test_mpc_glued_halves_equal_one_piece_plate():
    g = _arr(decks.glued_plate_deck(nx=12, ny=8, tie="mpc"))
    p = _arr(decks.plate_deck(nx=12, ny=8, a=400.0, b=300.0, t=1.5, nd=4))
    assert len(g.mpc_eq) == 6 * 9 and len(g.rbe2) == 0
    rg, rp = O.solve(g), O.solve(p)
    assert np.allclose(rg["eigenvalues"], rp["eigenvalues"], rtol=1e-10, atol=0)


def test_mpc_equation_moves_a_load_by_its_coefficients():
    """u_R1 = 0.5 u_A1 + 0.5 u_B1 (MPC 2 u_R1 - u_A1 - u_B1 = 0): a force P
    at the otherwise free GRID R acts as P/2 at A and at B (the transposed
    equation), so the model equals the plate loaded at A and B directly,
    and R moves with the average of A and B."""
    nx, ny, P = 10, 8, -150.0
    base = decks.plate_deck(nx=nx, ny=ny, a=400.0, b=300.0, t=1.5, nd=4)
    A, B = 3 * (nx + 1) + nx + 1, 5 * (nx + 1) + nx + 1     # GRIDs on the loaded edge
    viampc = base.replace("ENDDATA", "GRID,999999,,410.0,150.0,0.0\nMPC,1,999999,1,2.0,%d,1,-1.0,%d,1,-1.0\n"
                          "FORCE,10,999999,0,%r,1.0,0.0,0.0\nENDDATA" % (A, B, P))
    direct = base.replace("ENDDATA", "FORCE,10,%d,0,%r,1.0,0.0,0.0\nFORCE,10,%d,0,%r,1.0,0.0,0.0\nENDDATA"
                          % (A, P / 2, B, P / 2))
    m, d = _arr(viampc), _arr(direct)
    rm, rd = O.solve(m), O.solve(d)
    assert np.allclose(rm["eigenvalues"], rd["eigenvalues"], rtol=1e-11, atol=0)
    nd = 6 * d.n_nodes                                       # GRID 999999 is the last rank
    assert np.allclose(rm["u"][:nd], rd["u"], rtol=0, atol=1e-12 * np.abs(rd["u"]).max())
    ia, ib = 6 * (A - 1), 6 * (B - 1)
    assert np.isclose(rm["u"][nd], 0.5 * (rm["u"][ia] + rm["u"][ib]), rtol=1e-14)


def test_mpc_replaces_rbe2_entries_and_last_card_wins():
    txt = decks.glued_plate_deck(nx=4, ny=2)
    h = 2
    gl, gr = h + 1, 100000 + h + 1                            # seam GRIDs of row 0
    extra = "MPC,7,%d,3,1.0,%d,3,-0.5\nMPC,8,%d,3,1.0,%d,3,-0.25\nENDDATA" % (gr, gl, gr, gl)
    arr = _arr(txt.replace("ENDDATA", extra))
    cmap = O.constraint_map(arr)
    rank = {int(g): i for i, g in enumerate(arr.node_ids)}
    dep = 6 * rank[gr] + 2
    assert cmap[dep] == [(6 * rank[gl] + 2, 0.25)]           # the later MPC, not the RBE2 row
    assert cmap[6 * rank[gr] + 3] == [(6 * rank[gl] + 3, 1.0)]  # other components keep the RBE2


def test_rspline_projects_and_moves_a_load_by_its_weights():
    """RSPLINE over edge GRIDs 77, 88, 99 with a dependent GRID beside the
    last segment at a quarter of its length: its x motion is 0.75 u(88) +
    0.25 u(99) (the nearest segment, not the first), so a force P there
    equals 0.75 P at 88 and 0.25 P at 99 applied directly."""
    P = -150.0
    s = _arr(decks.rspline_load_plate_deck(nx=10, ny=8, frac=0.25, P=P))
    assert s.rspline.tolist() == [[0, 3, 0, 1]] and s.rspline_dep.tolist() == [[s.n_nodes - 1, 1]]
    base = decks.plate_deck(nx=10, ny=8, a=400.0, b=300.0, t=1.5, nd=4)
    direct = _arr(base.replace("ENDDATA", "FORCE,10,88,0,%r,1.0,0.0,0.0\nFORCE,10,99,0,%r,1.0,0.0,0.0\nENDDATA"
                               % (0.75 * P, 0.25 * P)))
    cmap = O.constraint_map(s)
    assert cmap == {6 * (s.n_nodes - 1): [(6 * 87, 0.75), (6 * 98, 0.25)]}
    rs, rd = O.solve(s), O.solve(direct)
    assert np.allclose(rs["eigenvalues"], rd["eigenvalues"], rtol=1e-11, atol=0)
    nd = 6 * direct.n_nodes
    assert np.allclose(rs["u"][:nd], rd["u"], rtol=0, atol=1e-12 * np.abs(rd["u"]).max())
    assert np.isclose(rs["u"][nd], 0.75 * rs["u"][6 * 87] + 0.25 * rs["u"][6 * 98], rtol=1e-14)


def test_rbe1_on_one_grid_is_the_rbe2_link():
    """An RBE1 whose six independent DOFs are one GRID's is the rigid link
    of an RBE2 from that GRID: the same map and the one-piece plate."""
    e1 = _arr(decks.glued_plate_deck(nx=8, ny=4, tie="rbe1"))
    e2 = _arr(decks.glued_plate_deck(nx=8, ny=4))
    m1, m2 = O.constraint_map(e1), O.constraint_map(e2)
    assert sorted(m1) == sorted(m2) and len(e1.rbe1) == 5
    for k in m2:
        a = dict(m1[k]); b = dict(m2[k])
        assert sorted(a) == sorted(b) and all(abs(a[i] - b[i]) <= 1e-13 for i in a)
    p = _arr(decks.plate_deck(nx=8, ny=4, a=400.0, b=300.0, t=1.5, nd=4))
    assert np.allclose(O.solve(e1)["eigenvalues"], O.solve(p)["eigenvalues"], rtol=1e-10, atol=0)


def test_rbe1_reproduces_rigid_motion_and_balances_load():
    """3-2-1 independent DOFs on three GRIDs: a rigid motion of the grids
    gives the dependent GRID the same rigid motion, and a load there moves
    onto the independent DOFs with the same resultant force and moment."""
    rng = np.random.default_rng(7)
    xyz = rng.normal(size=(4, 3)) * 40.0
    ns = types.SimpleNamespace(xyz=xyz, rbe2=np.zeros((0, 3), np.int32), rbe3=np.zeros((0, 4), np.int32),
                               rbe1=np.array([[0, 6, 0, 6]], np.int32),
                               rbe1_indep=np.array([[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]], np.int32),
                               rbe1_dep=np.array([[3, c] for c in range(1, 7)], np.int32))
    cmap = O.constraint_map(ns)
    assert sorted(cmap) == [18 + c for c in range(6)]
    t, th = rng.normal(size=3), rng.normal(size=3) * 1e-2
    u = _rigid(xyz, t, th)
    ref = u.copy()
    ref[18:] = 0.0
    O.recover_dependent(cmap, ref)
    assert np.allclose(ref[18:24], u[18:24], rtol=0, atol=1e-12 * np.abs(u).max())
    Fr = rng.normal(size=6)
    F = np.zeros(24)
    F[18:24] = Fr
    for dep, pairs in cmap.items():
        for ind, c in pairs:
            F[ind] += F[dep] * c
    grid = F[:18].reshape(3, 6)
    assert np.allclose(grid[:, :3].sum(axis=0), Fr[:3], rtol=1e-10, atol=1e-10)
    mom = sum(np.cross(xyz[i] - xyz[3], grid[i, :3]) for i in range(3))
    assert np.allclose(mom, Fr[3:], rtol=1e-9, atol=1e-9)


def test_rspline_field_rule_takes_small_digit_ids_for_components():
    """The reference's RSPLINE reader calls a GRID dependent when the next
    field reads as a component code (extract_constraints.jl:222-237): in
    11, 22, 33, 44, 1 the 33 makes 22 dependent in 3 (twice) and the 1
    makes 44 dependent in 1, which leaves one independent GRID and drops
    the element; in 11, 22, 77, 33, 1 only 77 is dependent and the trailing
    1 is an independent GRID."""
    base = decks.plate_deck(nx=10, ny=8, nd=4)
    m = bdf.bdf_to_model_text(base.replace("ENDDATA", "RSPLINE,5,0.1,11,22,33,44,1\nENDDATA"))
    assert m["RSPLINEs"] == {}
    m = bdf.bdf_to_model_text(base.replace("ENDDATA", "RSPLINE,5,0.1,11,22,77,33,1\nENDDATA"))
    assert m["RSPLINEs"]["5"]["INDEP_GRIDS"] == [11, 22, 1] and m["RSPLINEs"]["5"]["DEP"] == [(77, 3), (77, 3)]


# --------------------------------------------------------------------------
# the library's constraint code (host build) against the oracle
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def probe():
    os.makedirs(LIBDIR, exist_ok=True)
    csrc = os.path.join(HERE, "..", "jf_2026_05_19_amd", "csrc")
    tmp = LIB + ".tmp%d" % os.getpid()
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", SRC, os.path.join(csrc, "constraints.cpp"),
                           os.path.join(csrc, "symbolic.cpp"), "-o", tmp])
    os.replace(tmp, LIB)
    return ctypes.CDLL(LIB)


def _desc(arr):
    from jf_2026_05_19_amd import _native as N
    return N._model_desc(arr)


DECKS = [("glued", lambda: decks.glued_plate_deck(nx=8, ny=4)),
         ("eccentric", lambda: decks.eccentric_panel_deck(nx=8, ny=8, stringer_every=4)),
         ("rbe3_load", lambda: decks.rbe3_load_plate_deck(nx=6, ny=4)),
         ("glued_mpc", lambda: decks.glued_plate_deck(nx=8, ny=4, tie="mpc")),
         ("rspline_load", lambda: decks.rspline_load_plate_deck(nx=10, ny=8)),
         ("glued_rbe1", lambda: decks.glued_plate_deck(nx=8, ny=4, tie="rbe1")),
         ("mpc_over_rbe2", lambda: decks.glued_plate_deck(nx=4, ny=2).replace(
             "ENDDATA", "MPC,7,100003,3,1.0,3,3,-0.5,5,3,0.25\nMPC,8,100003,3,2.0,3,3,-1.0\nENDDATA"))]


@pytest.mark.parametrize("name,deck", DECKS, ids=[d[0] for d in DECKS])
def test_library_map_and_redistribution_match_oracle(probe, name, deck):
    arr = _arr(deck())
    d, keep = _desc(arr)
    sizes = np.zeros(2, np.int64)
    err = ctypes.create_string_buffer(256)
    assert probe.probe_map_sizes(ctypes.byref(d), sizes.ctypes.data_as(ctypes.c_void_p), err, 256) == 0, err.value
    nd, npair = int(sizes[0]), int(sizes[1])
    dep = np.zeros(nd, np.int32); ptr = np.zeros(nd + 1, np.int32)
    ind = np.zeros(max(npair, 1), np.int32); coef = np.zeros(max(npair, 1))
    F = np.zeros(arr.ndof)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert probe.probe_map(ctypes.byref(d), vp(dep), vp(ptr), vp(ind), vp(coef), vp(F)) == 0
    cmap = O.constraint_map(arr)
    assert sorted(cmap) == dep.tolist()
    for k, dd in enumerate(dep):
        got = list(zip(ind[ptr[k]:ptr[k + 1]].tolist(), coef[ptr[k]:ptr[k + 1]].tolist()))
        ref = cmap[int(dd)]
        assert [a for a, _ in got] == [a for a, _ in ref]
        assert np.allclose([c for _, c in got], [c for _, c in ref], rtol=1e-12, atol=1e-14)
    assert np.allclose(F, O.constrained_arrays(arr, cmap).F, rtol=1e-13, atol=1e-12)
    # K' through the block pattern and the redistribution plan
    import copy
    plain = copy.copy(arr)
    plain.rbe2 = np.zeros((0, 3), np.int32)
    plain.rbe3 = np.zeros((0, 4), np.int32)
    plain.mpc_eq = np.zeros((0, 2), np.int32)
    plain.rspline = np.zeros((0, 4), np.int32)
    plain.rbe1 = np.zeros((0, 4), np.int32)
    K0 = O.assemble_K(plain).toarray()
    K1 = O.assemble_K(arr).toarray()
    out = np.zeros_like(K0)
    assert probe.probe_redistribute(ctypes.byref(d), vp(np.ascontiguousarray(K0)), vp(out)) == 0
    assert np.abs(out - K1).max() <= 1e-12 * np.abs(K1).max()


def test_library_refuses_ambiguous_constraints(probe):
    arr = _arr(decks.glued_plate_deck(nx=4, ny=2))
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    err = ctypes.create_string_buffer(256)
    sizes = np.zeros(2, np.int64)
    import copy
    dup = copy.copy(arr)                      # one slave claimed by two RBE2 pairs
    dup.rbe2 = np.vstack([arr.rbe2, arr.rbe2[:1]])
    d, keep = _desc(dup)
    assert probe.probe_map_sizes(ctypes.byref(d), vp(sizes), err, 256) != 0
    assert b"two constraints" in err.value
    chain = copy.copy(arr)                    # a master that is itself a slave
    r = arr.rbe2[0]
    chain.rbe2 = np.vstack([arr.rbe2, [[0, 123456, r[0]]]]).astype(np.int32)
    d, keep = _desc(chain)
    probe.probe_map_sizes(ctypes.byref(d), vp(sizes), err, 256)
    assert b"chain" in err.value


@pytest.mark.parametrize("name,deck", DECKS, ids=[d[0] for d in DECKS])
def test_native_reader_matches_host_builder(name, deck):
    from jf_2026_05_19_amd import _native as N
    txt = deck()
    a, b = _arr(txt), N.read_deck(txt)
    for f in ("rbe2", "rbe3", "rbe3_grp", "rbe3_wt", "rbe3_node", "mpc_eq", "mpc_term", "mpc_coef", "rspline",
              "rspline_indep", "rspline_dep", "rbe1", "rbe1_indep", "rbe1_dep", "xyz", "F",
              "spc_mask", "q4", "bar"):
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape and np.array_equal(x, y), (name, f)
