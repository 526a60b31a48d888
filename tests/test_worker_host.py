"""The JSONL worker's protocol, manifest schema and result files on the CPU,
with the oracle standing in for the device solver (injected here; the
worker's own solver is the device request server).  Covers the commands of
JFEM/tools/jfem_worker_jsonl.jl:58-146, the manifest rules of
JFEM/tools/manifest_batch_core.jl:100-505 and the .BUCKLING.JSON layout of
JFEM/src/Export.jl:2267-2330."""
import csv
import io
import json
import os

import numpy as np
import pytest

from jf_2026_05_19_amd import bdf, decks, model as M, worker as W
from oracle import oracle as O


class OracleSolver:
    """Same call shape as worker.DeviceSolver, answered by the oracle."""
    workers = 1

    def __init__(self):
        self.calls = []

    def solve(self, arrs, want_modes):
        self.calls.append((len(arrs), want_modes))
        out = []
        for a in arrs:
            r = O.solve(a)
            out.append((r["eigenvalues"], r["modes"] if want_modes else None, None, 0.01))
        return out


def _deck(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def _decks(tmp_path):
    plate = _deck(tmp_path, "plate_a.bdf", decks.plate_deck(nx=6, ny=5, a=300.0, b=250.0, t=1.5, nd=2))
    plate_b = _deck(tmp_path, "plate_b.bdf", decks.plate_deck(nx=6, ny=5, a=300.0, b=250.0, t=1.8, nd=2))
    panel = _deck(tmp_path, "panel.dat", decks.panel_deck(nx=10, ny=6, a=200.0, b=120.0, stringer_every=3, nd=3))
    bad = _deck(tmp_path, "bad.bdf", decks.plate_deck(nx=3, ny=3).replace("ENDDATA", "CONM2,900,1,,5.0\nENDDATA"))
    return plate, plate_b, panel, bad


def _serve(worker, lines):
    inp = io.StringIO("".join(json.dumps(x) + "\n" if not isinstance(x, str) else x + "\n" for x in lines))
    out = io.StringIO()
    worker.serve(inp, out, err=io.StringIO())
    return [json.loads(l) for l in out.getvalue().splitlines()]


def test_protocol_commands(tmp_path):
    plate, plate_b, panel, bad = _decks(tmp_path)
    w = W.Worker(OracleSolver(), {"JFEM_FOO": "1"})
    man = {"output_root": str(tmp_path / "out"), "batch_id": "b1",
           "defaults": {"output_options": {"json": True, "eigenvalues_only": True, "report": False}},
           "cases": [{"input": plate}]}
    resp = _serve(w, [{"command": "status", "id": "s"}, {"cmd": "ping"}, {"command": "gc"},
                      "not json", [1, 2], {"command": "bogus"},
                      {"command": "run_batch", "id": 7, "manifest_data": man},
                      {"command": "run_batch"},
                      {"command": "quit"}, {"command": "status"}])
    assert resp[0]["status"] == "ready" and resp[0]["pid"] == os.getpid() and "worker_session_id" in resp[0]
    assert resp[1]["status"] == "ok" and resp[1]["id"] == "s" and resp[1]["commands"] == 1
    assert resp[2]["status"] == "ok" and resp[2]["id"] == 2          # no id: the command counter
    assert resp[3]["status"] == "ok"
    assert resp[4]["status"] == "error" and resp[4]["id"] is None
    assert resp[5]["status"] == "error" and "object" in resp[5]["error"]
    assert resp[6]["status"] == "error" and "unknown worker command" in resp[6]["error"]
    r = resp[7]
    assert r["status"] == "ok" and r["id"] == 7 and r["completed"] == 1 and r["failed"] == 0 and r["total"] == 1
    assert r["batch_id"] == "b1" and os.path.isfile(r["summary_json"]) and os.path.isfile(r["summary_csv"])
    assert resp[8]["status"] == "error" and "manifest" in resp[8]["error"]
    assert resp[9]["status"] == "stopping"
    assert len(resp) == 10                                              # nothing read after quit
    st = w.handle({"command": "status"})
    assert st["completed_batches"] == 1 and st["failed_batches"] == 0


def test_manifest_outputs_against_oracle(tmp_path):
    plate, plate_b, panel, bad = _decks(tmp_path)
    mpath = tmp_path / "m" / "cases.json"
    mpath.parent.mkdir()
    man = {"output_root": "runs", "batch_id": "it1",
           "defaults": {"output_options": {"json": True, "binary": False, "report": True}},
           "cases": [{"input": plate, "case_id": "p/1"},
                     {"input": plate_b, "case_id": "p/1"},
                     {"input": panel, "output_dir": "custom_dir",
                      "output_options": {"eigenvectors": True, "report": False}},
                     {"input": bad},
                     {"input": plate, "output_options": {"eigenvalues_only": True}}]}
    mpath.write_text(json.dumps(man))
    solver = OracleSolver()
    w = W.Worker(solver)
    r = w.handle({"command": "run_batch", "manifest": str(mpath)})
    assert r["status"] == "failed" and r["completed"] == 4 and r["failed"] == 1 and r["total"] == 5
    root = str(mpath.parent / "runs")
    assert r["output_root"] == root
    summ = json.load(open(r["summary_json"]))
    rows = summ["cases"]
    # slugs: duplicate case ids get _2; output_dir relative to the manifest
    assert rows[0]["output_dir"] == os.path.join(root, "p_1") and rows[1]["output_dir"] == os.path.join(root, "p_1_2")
    assert rows[2]["output_dir"] == str(mpath.parent / "custom_dir")
    # the unsupported deck fails with its error, the batch goes on
    assert rows[3]["status"] == "failed" and "CONM2" in rows[3]["error"]
    # eigenvalues against the oracle, summary columns
    for k, deck in ((0, plate), (1, plate_b), (2, panel), (4, plate)):
        ref = O.solve(M.build_arrays(bdf.bdf_to_model_text(open(deck).read())))
        row = rows[k]
        assert row["status"] == "ok" and row["sol_type"] == 105
        assert np.allclose(row["eigenvalues"], ref["eigenvalues"], rtol=0, atol=0)
        assert row["eigenvalue_count"] == len(ref["eigenvalues"]) and row["first_eigenvalue"] == row["eigenvalues"][0]
        assert row["eigenvalues_csv"] == ";".join("%.17g" % v for v in ref["eigenvalues"])
    # json + modes: .BUCKLING.JSON with the mode shapes, GRID ids ascending
    bj = json.load(open(rows[2]["result_json"]))
    assert rows[2]["result_json"].endswith("panel.dat.BUCKLING.JSON")    # only a .bdf suffix is dropped
    arr = M.build_arrays(bdf.bdf_to_model_text(open(panel).read()))
    ref = O.solve(arr)
    assert bj["analysis_type"] == "SOL105_BUCKLING" and bj["grid_id_order"] == sorted(bj["grid_id_order"])
    assert len(bj["modes"]) == len(ref["eigenvalues"]) and rows[2]["mode_shapes_available"]
    m0 = bj["modes"][0]
    assert m0["mode_number"] == 1 and m0["eigenvalue"] == ref["eigenvalues"][0]
    shape = np.array([[d[c] for c in ("t1", "t2", "t3", "r1", "r2", "r3")] for d in m0["mode_shape"]]).ravel()
    assert np.array_equal(shape, ref["modes"][:, 0])
    assert rows[2]["report"] == ""                                     # report switched off for that case
    assert os.path.isfile(rows[0]["report"]) and rows[0]["report"].endswith("plate_a.REPORT.md")
    # eigenvalues only: the JSON holds the loads and no mode shapes
    bj4 = json.load(open(rows[4]["result_json"]))
    assert bj4["modes"] == [] and bj4["eigenvalues"] == rows[4]["eigenvalues"]
    assert rows[4]["mode_shape_count"] == 0 and not rows[4]["mode_shapes_available"]
    # the CSV has the reference's columns, one row per case
    with open(r["summary_csv"]) as fh:
        table = list(csv.reader(fh))
    assert tuple(table[0]) == W.CSV_COLUMNS and len(table) == 6
    assert table[4][2] == "failed" and table[1][6] == rows[0]["eigenvalues_csv"]
    # every case has its run manifest and log
    for row in rows:
        rm = json.load(open(os.path.join(row["output_dir"], "run_manifest.json")))
        assert rm["bdf"]["path"] == row["input"] and rm["applied_flags"]["JFEM_EXPORT_BINARY"] == "false"
        assert os.path.isfile(row["log"])
    # the parsed cases went to the solver together: the three JSON exports
    # with their mode shapes, the eigenvalues-only case without
    assert sorted(solver.calls) == [(1, False), (3, True)]


def test_stop_on_error_and_schema_errors(tmp_path):
    plate, _, _, bad = _decks(tmp_path)
    w = W.Worker(OracleSolver())
    man = {"output_root": str(tmp_path / "o"), "defaults": {"stop_on_error": True},
           "cases": [{"input": bad}, {"input": plate}]}
    with pytest.raises(RuntimeError):
        w.handle({"command": "run_batch", "manifest_data": man})
    for bad_man, msg in (({"cases": [{"input": plate}]}, "output_root"),
                         ({"output_root": "x", "cases": []}, "empty"),
                         ({"output_root": "x", "cases": [{"case_id": "a"}]}, "requires input"),
                         ({"output_root": "x", "cases": [{"input": str(tmp_path / "missing.bdf")}]}, "not found"),
                         ({"output_root": "x", "cases": "abc"}, "cases array")):
        with pytest.raises(W.ManifestError, match=msg):
            w.handle({"command": "run_batch", "manifest_data": bad_man})


def test_flags_and_options_rules():
    assert W.string_dict("A=1;B=x=y") == {"A": "1", "B": "x=y"} and W.string_dict({"A": 2}) == {"A": "2"}
    with pytest.raises(W.ManifestError):
        W.string_dict("NOEQUALS")
    eo = W.export_options({"eigenvectors": True, "eigenvalues_only": True}, {})
    assert eo["export_json"] and not eo["eigenvalues_only"] and eo["export_jfem_binary"]
    eo = W.export_options({"values_only": "yes", "vtk": True}, {"JFEM_EXPORT_BINARY": "false"})
    assert eo["eigenvalues_only"] and not eo["export_vtk"] and not eo["export_json"] and eo["export_report"]
    assert W.case_slug(" a/b c?.bdf ") == "a_b_c_.bdf" and W.case_slug("///") == "case"
    assert W.csv_escape('a,"b"') == '"a,""b"""' and W.float_cell([1.0, 0.1]) == "1;0.10000000000000001"


def test_mesh_fingerprint_ignores_thickness_only():
    a = M.build_arrays(bdf.bdf_to_model_text(decks.plate_deck(nx=5, ny=4, t=1.5)))
    b = M.build_arrays(bdf.bdf_to_model_text(decks.plate_deck(nx=5, ny=4, t=2.5)))
    c = M.build_arrays(bdf.bdf_to_model_text(decks.plate_deck(nx=5, ny=4, t=1.5, line_load=2.0)))
    assert W.mesh_fingerprint(a) == W.mesh_fingerprint(b) != W.mesh_fingerprint(c)


def test_mesh_fingerprint_sees_constraints_and_range():
    import copy
    g = M.build_arrays(bdf.bdf_to_model_text(decks.glued_plate_deck(nx=4, ny=2)))
    fewer = copy.copy(g)
    fewer.rbe2 = g.rbe2[:-1]
    ranged = copy.copy(g)
    ranged.v1, ranged.v2 = 0.5, 50.0
    thick = g.with_shell_thickness(g.shell_thickness() * 1.1)
    prints = [W.mesh_fingerprint(x) for x in (g, fewer, ranged)]
    assert len(set(prints)) == 3 and W.mesh_fingerprint(thick) == prints[0]
    r3 = M.build_arrays(bdf.bdf_to_model_text(decks.rbe3_load_plate_deck(nx=4, ny=2)))
    heavier = copy.copy(r3)
    heavier.rbe3_wt = r3.rbe3_wt * 2.0
    assert W.mesh_fingerprint(r3) != W.mesh_fingerprint(heavier)
