"""The SOL 101 results layout (jf_2026_05_19_amd/results.py) on the CPU, fed
with the oracle's recovery rows: the reference's keys and row shapes
(JFEM/src/solver/solve_case.jl:126-186, JFEM/src/solver/stress_recovery.jl:
419-430, 301-310, 485-511), the reaction-row test, strain rows as membrane
strain and curvature, von Mises of the fibre stresses, and bars shorter than
1e-9 left out."""
import math

import numpy as np

from jf_2026_05_19_amd import bdf, decks, model as M
from jf_2026_05_19_amd import results as R
from oracle import oracle as O
from oracle import recovery as RC


def _solved(txt):
    arr = M.build_arrays(bdf.bdf_to_model_text(txt))
    K = O.assemble_K(arr)
    u, free = O.static_solve(arr, K)
    return arr, u, RC.recover(arr, u, K=K, bar_points=arr.bar_points)


def test_layout_keys_and_rows():
    arr, u, (res, q4, t3, bar) = _solved(decks.panel_deck(nx=8, ny=4, stringer_every=2, nd=3, tri_frac=0.3))
    u_out, scal, js = R.static_results(arr, u, res, q4, t3, bar, {"backend": "x"})
    assert set(js) == {"displacements", "spc_forces", "forces", "forces_bilin", "stresses", "strains",
                       "solver_diagnostics"}
    assert set(js["forces"]) == {"cbar", "quad4", "tria3", "crod", "conrod", "celas1"}
    assert set(js["stresses"]) == {"cbar", "quad4", "tria3", "crod", "conrod", "celas1", "ctetra", "chexa", "cpenta"}
    assert set(js["forces_bilin"]) == {"quad4", "tria3"}
    d0 = js["displacements"][0]
    assert set(d0) == {"grid_id", "t1", "t2", "t3", "r1", "r2", "r3"}
    assert np.array_equal(u_out, u)
    f = js["forces"]["quad4"][0]
    assert list(f) == ["eid", "fx", "fy", "fxy", "mx", "my", "mxy", "qx", "qy"]
    assert f["eid"] == int(arr.q4_eid[0]) and f["qy"] == q4[0, 7]
    s = js["stresses"]["quad4"][0]
    t = float(arr.q4prop[0, 2])
    assert s["z1"]["fiber_dist"] == -t / 2 and s["z2"]["fiber_dist"] == t / 2
    sx, sy, txy = q4[0, 8:11]
    assert math.isclose(s["z1"]["von_mises"], math.sqrt(sx * sx - sx * sy + sy * sy + 3 * txy * txy))
    assert s["z1"]["major"] == 0.0 and s["z1"]["minor"] == 0.0
    e = js["strains"]["quad4"][0]
    np.testing.assert_allclose([e["z1"]["normal_x"], e["z2"]["normal_x"]],
                               [(q4[0, 14] + q4[0, 17]) / 2, (q4[0, 14] - q4[0, 17]) / t])
    assert e["z1"]["fiber_dist"] == 0.0 and e["z2"]["fiber_dist"] == -1.0
    bil = js["forces_bilin"]["quad4"][:5]
    assert bil[0]["grid_id"] == "CEN/4" and [b["grid_id"] for b in bil[1:]] == [int(arr.node_ids[n]) for n in arr.q4[0]]
    assert bil[1]["fx"] == q4[0, 20] and bil[4]["my"] == q4[0, 32 + 9 + 1] and bil[2]["qx"] == q4[0, 6]
    cb = js["stresses"]["cbar"][0]
    assert set(cb) == {"eid", "end_a", "end_b", "axial"} and set(cb["end_a"]) == {"p1", "p2", "p3", "p4"}
    assert list(js["forces"]["cbar"][0]) == ["eid", "axial", "shear_1", "shear_2", "torque", "moment_a1",
                                             "moment_a2", "moment_b1", "moment_b2"]
    assert scal[int(arr.bar_eid[0])] == abs(bar[0, 8])
    sxx, syy, sxy = q4[0, 8:11]
    assert math.isclose(scal[int(arr.q4_eid[0])], (sxx + syy) / 2 + math.hypot((sxx - syy) / 2, sxy))
    # reaction rows: norm > 1e-20, sums in the diagnostics
    rows = [i for i in range(arr.n_nodes) if np.linalg.norm(res[6 * i:6 * i + 6]) > 1e-20]
    assert [r["grid_id"] for r in js["spc_forces"]] == [int(arr.node_ids[i]) for i in rows]
    eq = js["solver_diagnostics"]["equilibrium"]
    assert math.isclose(eq["spc_reaction_sum"]["fx"], sum(res[6 * i] for i in rows), rel_tol=1e-12, abs_tol=1e-12)
    assert js["solver_diagnostics"]["backend"] == "x"


def test_short_bar_is_left_out():
    arr, u, (res, q4, t3, bar) = _solved(decks.panel_deck(nx=8, ny=4, stringer_every=2, nd=3))
    bar = bar.copy()
    bar[0, 17] = 0.0
    _, scal, js = R.static_results(arr, u, res, q4, t3, bar)
    assert len(js["forces"]["cbar"]) == arr.n_bar - 1
    assert int(arr.bar_eid[0]) not in scal
