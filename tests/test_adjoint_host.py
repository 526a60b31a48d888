"""Per-PSHELL buckling sensitivities (jf_2026_05_19_amd/adjoint.py) on the
CPU: the configuration checks of parse_adjoint_config
(JFEM/src/solver/adjoint.jl:16-38), the design-variable values
(adjoint.jl:789-800), and the group formula of
JFEM/src/solver/adjoint_buckling.jl:2611 on hand-built element terms, with a
device-model stand-in that returns those terms."""
import json

import numpy as np
import pytest

from jf_2026_05_19_amd import adjoint as A, bdf, decks, model as M

CFG = {"responses": [{"id": "buck", "type": "buckling"}],
       "design_variables": [{"id": "t", "type": "shell_thickness", "pids": [1, 2, 7]}]}


def test_config_checks(tmp_path):
    p = tmp_path / "adj.json"
    p.write_text(json.dumps(CFG))
    assert A.parse_adjoint_config(str(p))["design_variables"][0]["pids"] == [1, 2, 7]
    with pytest.raises(A.AdjointConfigError, match="not found"):
        A.parse_adjoint_config(str(tmp_path / "missing.json"))
    for bad, msg in (({"design_variables": CFG["design_variables"]}, "response"),
                     ({"responses": CFG["responses"]}, "design variable"),
                     ({"responses": [{"type": "x"}], "design_variables": CFG["design_variables"]}, "'id'"),
                     ({"responses": CFG["responses"], "design_variables": [{"id": "a"}]}, "'type'")):
        with pytest.raises(A.AdjointConfigError, match=msg):
            A.parse_adjoint_config(bad)


def _two_pid_model():
    m = bdf.bdf_to_model_text(decks.plate_deck(nx=6, ny=5, a=300.0, b=250.0, nd=3))
    m["PSHELLs"]["2"] = dict(m["PSHELLs"]["1"], PID=2, T=1.6)
    for k in sorted(m["CSHELLs"], key=int)[::2]:
        m["CSHELLs"][k]["PID"] = 2
    return m


def test_design_variable_values_and_unsupported_types():
    m = _two_pid_model()
    assert A.design_variable_values(CFG["design_variables"][0], m) == {"PID_1": 2.0, "PID_2": 1.6}
    with pytest.raises(M.UnsupportedModel):
        A.design_variable_values({"id": "e", "type": "material_E", "mids": [1]}, m)


class _Ctx:
    OPT_SENS_MODE = 5

    def __init__(self):
        self.modes = []

    def set_option(self, opt, value):
        assert opt == self.OPT_SENS_MODE
        self.modes.append(value)


class _Stub:
    """Returns fixed per-shell terms and phi' KG phi for every mode."""

    def __init__(self, arr, terms, pkp):
        self.arr, self.terms, self.pkp = arr, terms, pkp
        self.ctx = _Ctx()

    def sensitivity_terms(self, mode):
        return self.terms[mode], None, self.pkp[mode]


def test_group_formula_and_layout():
    m = _two_pid_model()
    arr = M.build_arrays(m)
    assert set(np.unique(arr.q4_pid)) == {1, 2}
    rng = np.random.default_rng(3)
    terms = [rng.normal(size=(arr.n_shell, 3)) for _ in range(3)]
    pkp = [-2.5, 1e-31, -0.75]
    lam = np.array([11.0, 12.0, 13.5])
    res = A.solve_adjoint_buckling(_Stub(arr, terms, pkp), lam, m, CFG)
    assert res["eigenvalue_values"] == {"mode_1": 11.0, "mode_2": 12.0, "mode_3": 13.5}
    assert res["design_variable_values"] == {"t": {"PID_1": 2.0, "PID_2": 1.6}}
    assert res["sensitivities"]["mode_2"] == {}                       # |phi' KG phi| < 1e-30
    for i in (0, 2):
        s = res["sensitivities"]["mode_%d" % (i + 1)]["t"]
        assert set(s) == {"PID_1", "PID_2", "PID_7"} and s["PID_7"] == 0.0
        per_elem = -(terms[i][:, 0] + lam[i] * terms[i][:, 1] - terms[i][:, 2]) / pkp[i]
        for p in (1, 2):
            assert abs(s["PID_%d" % p] - per_elem[arr.q4_pid == p].sum()) <= 1e-12 * np.abs(per_elem).sum()
        # the groups partition the shells: their sum is the all-shell derivative
        assert abs(s["PID_1"] + s["PID_2"] - per_elem.sum()) <= 1e-12 * np.abs(per_elem).sum()


def test_reference_mode_whole_kg_shortcut():
    """Reference mode: a design variable that is the only PSHELL of a model
    without bars takes phi' KG phi / T for its dKG term (the reference's
    _dKg_dx_phi_thickness shortcut); two PSHELLs keep the element sums.  The
    context's sensitivity mode is set for the terms and restored after."""
    m = bdf.bdf_to_model_text(decks.plate_deck(nx=4, ny=3, a=200.0, b=150.0, t=1.25, nd=2))
    arr = M.build_arrays(m)
    rng = np.random.default_rng(5)
    terms = [rng.normal(size=(arr.n_shell, 3))]
    pkp = [-3.0]
    lam = np.array([7.0])
    cfg1 = {"responses": [{"id": "b", "type": "buckling"}],
            "design_variables": [{"id": "t", "type": "shell_thickness", "pids": [1]}]}
    stub = _Stub(arr, terms, pkp)
    res = A.solve_adjoint_buckling(stub, lam, m, cfg1, mode="reference")
    assert stub.ctx.modes == [1, 0]
    t = terms[0].sum(axis=0)
    expect = -(t[0] + lam[0] * pkp[0] / 1.25 - t[2]) / pkp[0]
    assert abs(res["sensitivities"]["mode_1"]["t"]["PID_1"] - expect) <= 1e-12 * abs(expect)
    exact = A.solve_adjoint_buckling(stub, lam, m, cfg1)
    assert abs(exact["sensitivities"]["mode_1"]["t"]["PID_1"] - (-(t[0] + lam[0] * t[1] - t[2]) / pkp[0])) <= \
        1e-12 * abs(expect)
    with pytest.raises(ValueError):
        A.solve_adjoint_buckling(stub, lam, m, cfg1, mode="legacy")
