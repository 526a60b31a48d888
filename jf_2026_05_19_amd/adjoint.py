"""Buckling sensitivities per design variable, in the reference's result
layout, from the device's element sensitivities.

``solve_adjoint_buckling`` mirrors ``Solver.solve_adjoint_buckling``
(JFEM/src/solver/adjoint_buckling.jl:2406-2725) for ``shell_thickness``
design variables: for every output mode ``mode_i`` and design variable
``id`` it returns ``sensitivities["mode_i"][id]["PID_p"]``, the derivative
of lambda_i with respect to the thickness T of PSHELL p, i.e. with all shells
of that property moving together:

    d lambda / d T_p = -(sum_e phi' dKe/dt phi + lambda sum_e phi' KGe/t phi
                         - sum_e psi' dKe/dt u) / phi' KG phi,

the sums over the shells e of property p (the group sums of
adjoint_buckling.jl:2611), with the device's per-shell terms
(``sol105_sensitivity_terms``).  A mode with |phi' KG phi| < 1e-30 gets an
empty dictionary (adjoint_buckling.jl:2479-2484).  The configuration follows
``parse_adjoint_config`` (JFEM/src/solver/adjoint.jl:16-38) and the design
variable values ``get_design_variable_values`` (adjoint.jl:789-800).

``mode="exact"`` (default) takes the per-shell terms of the exact derivative
of the forward path.  ``mode="reference"`` takes the reference's own analytic
path for CQUAD4 (SOL105_OPT_SENS_MODE = 1: dKe/dt of the legacy MITC4 element
stiffness_quad4_generic, the unit-basis adjoint load, the context KG with six
condensed membrane modes when the model has no bars) and, like
``_dKg_dx_phi_thickness`` (adjoint_buckling.jl:2059-2089), replaces the
group's sum of phi' KGe phi / t by phi' KG phi / T when the design variable
is one PSHELL that owns every shell and the model has no bars.

Other design-variable types (material E and NU, bar area, node coordinates,
topology density) have no device path and raise ``UnsupportedModel``.
"""
from __future__ import annotations

import json
import os

import numpy as np

from .model import UnsupportedModel

__all__ = ["AdjointConfigError", "parse_adjoint_config", "design_variable_values", "solve_adjoint_buckling"]


class AdjointConfigError(ValueError):
    pass


def parse_adjoint_config(config):
    """A configuration from a JSON file path or an already parsed dict, with
    the reference's checks: at least one response and one design variable,
    each with an ``id`` and a ``type``."""
    if isinstance(config, (str, os.PathLike)):
        if not os.path.isfile(config):
            raise AdjointConfigError("[ADJOINT] Config file not found: %s" % config)
        with open(config) as fh:
            config = json.load(fh)
    if not isinstance(config, dict):
        raise AdjointConfigError("[ADJOINT] Config must be a JSON object")
    if not config.get("responses"):
        raise AdjointConfigError("[ADJOINT] Config must contain at least one response")
    if not config.get("design_variables"):
        raise AdjointConfigError("[ADJOINT] Config must contain at least one design variable")
    for resp in config["responses"]:
        if "id" not in resp:
            raise AdjointConfigError("[ADJOINT] Each response must have an 'id'")
        if "type" not in resp:
            raise AdjointConfigError("[ADJOINT] Response '%s' must have a 'type'" % resp["id"])
    for dv in config["design_variables"]:
        if "id" not in dv:
            raise AdjointConfigError("[ADJOINT] Each design variable must have an 'id'")
        if "type" not in dv:
            raise AdjointConfigError("[ADJOINT] Design variable '%s' must have a 'type'" % dv["id"])
    return config


def _check_supported(dv):
    if dv["type"] != "shell_thickness":
        raise UnsupportedModel("design variable type %r has no device path" % dv["type"])


def design_variable_values(dv, model):
    """Current values of a design variable: ``{"PID_p": T}`` for the
    PSHELLs listed that exist in the model."""
    _check_supported(dv)
    pshells = model["PSHELLs"]
    out = {}
    for pid in dv["pids"]:
        key = str(int(pid))
        if key in pshells:
            out["PID_" + key] = float(pshells[key]["T"])
    return out


def group_sensitivities(terms, lam, phi_kg_phi, shell_pids, pids, whole_kg_over_t=None):
    """d lambda / d T_p for each listed PSHELL id p from the per-shell terms
    ``[phi' dKe/dt phi, phi' KGe phi / t, psi' dKe/dt u]`` (rows in shell
    order, ``shell_pids`` the property of each row).  A PSHELL without shells
    has derivative zero.  ``whole_kg_over_t`` (phi' KG phi / T of the
    assembled KG) replaces the group's second term when given."""
    terms = np.asarray(terms, dtype=float)
    shell_pids = np.asarray(shell_pids, dtype=np.int64)
    out = {}
    for pid in pids:
        rows = shell_pids == int(pid)
        t = terms[rows].sum(axis=0) if rows.any() else np.zeros(3)
        t2 = t[1] if whole_kg_over_t is None else whole_kg_over_t
        out["PID_%d" % int(pid)] = float(-(t[0] + lam * t2 - t[2]) / phi_kg_phi) if rows.any() else 0.0
    return out


def _whole_kg_shortcut(dv, shell_pids, n_bar):
    """The reference's single-PSHELL shortcut of the dKG/dT term applies."""
    pids = [int(p) for p in dv["pids"]]
    return len(pids) == 1 and set(np.asarray(shell_pids).tolist()) <= {pids[0]} and n_bar == 0


def solve_adjoint_buckling(dm, lam, model, config, mode="exact"):
    """The reference's result dictionary for the solved device model ``dm``
    (its last buckling solve gave the eigenvalues ``lam``): sensitivities,
    eigenvalue_values and design_variable_values.  ``mode``: "exact" or
    "reference" (see the module docstring)."""
    if mode not in ("exact", "reference"):
        raise ValueError("mode is 'exact' or 'reference'")
    config = parse_adjoint_config(config)
    dvs = config["design_variables"]
    for dv in dvs:
        _check_supported(dv)
    arr = dm.arr
    shell_pids = np.concatenate([np.asarray(arr.q4_pid, np.int64), np.asarray(arr.t3_pid, np.int64)])
    thick = np.concatenate([np.asarray(arr.q4prop, float).reshape(-1, 6)[:, 2],
                            np.asarray(arr.t3prop, float).reshape(-1, 6)[:, 2]])
    ref = mode == "reference"
    sens, eig = {}, {}
    dm.ctx.set_option(dm.ctx.OPT_SENS_MODE, 1 if ref else 0)
    try:
        for i, li in enumerate(np.asarray(lam, dtype=float)):
            mode_id = "mode_%d" % (i + 1)
            eig[mode_id] = float(li)
            terms, _, pkp = dm.sensitivity_terms(i)
            if abs(pkp) < 1e-30:
                sens[mode_id] = {}
                continue
            sens[mode_id] = {}
            for dv in dvs:
                whole = None
                if ref and len(thick) and _whole_kg_shortcut(dv, shell_pids, arr.n_bar):
                    whole = pkp / float(thick[0])
                sens[mode_id][dv["id"]] = group_sensitivities(terms, li, pkp, shell_pids, dv["pids"], whole)
    finally:
        dm.ctx.set_option(dm.ctx.OPT_SENS_MODE, 0)
    return {
        "sensitivities": sens,
        "eigenvalue_values": eig,
        "design_variable_values": {dv["id"]: design_variable_values(dv, model) for dv in dvs},
    }
