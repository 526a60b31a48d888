"""GPU: the static solve's failure handling and the eigensolver's error path,
through the C ABI, against the oracle.

* All-CTRIA3 plates without theta_z constraints have two zero-energy drilling
  patterns.  Which DOFs the refix (or, after a failed Cholesky, the
  zero-pivot step) fixes depends on the elimination order (the device's
  nested dissection, the oracle's natural order), but the translations of
  u and the buckling loads do not (the patterns carry no load and no
  geometric stiffness).  Checked: same number of DOFs fixed, translations
  within 1e-10, lambda within 1e-8, translational MAC of the modes.
* A model with every DOF constrained returns SOL105_E_SINGULAR.
* An eigensolver capped from the host (32 basis vectors, no restart) returns
  SOL105_E_NOCONV, and the same context solves again once the caps are
  lifted.
"""
import numpy as np
import pytest

from jf_2026_05_19_amd import bdf, decks, model as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from jf_2026_05_19_amd import _native as N
    c = N.Context(0)
    yield c
    c.close()


def _arr(**kw):
    return M.build_arrays(bdf.bdf_to_model_text(decks.plate_deck(**kw)))


def _mac(a, b):
    return float((a @ b) ** 2 / ((a @ a) * (b @ b)))


@pytest.mark.parametrize("nx,ny,a,b,t", [(8, 6, 400.0, 300.0, 1.5), (12, 10, 480.0, 400.0, 1.4)])
def test_tri_plate_mechanisms_parity(ctx, nx, ny, a, b, t):
    from jf_2026_05_19_amd import _native as N
    from oracle import oracle as O
    arr = _arr(nx=nx, ny=ny, a=a, b=b, t=t, nd=3, tri_frac=1.0)
    print("tri plate", nx, ny, "ndof", arr.ndof, flush=True)
    ref = O.solve(arr)
    assert ref["static_info"]["path"] in ("refix", "zero_pivots")
    dm = N.DeviceModel(ctx, arr)
    u, free = dm.static_solve()
    assert int(free.sum()) == len(ref["free"])
    tr = np.arange(arr.ndof) % 6 < 3
    assert np.abs(u[tr] - ref["u"][tr]).max() <= 1e-10 * np.abs(ref["u"][tr]).max()
    lam, modes = dm.solve()
    lr = ref["eigenvalues"]
    assert len(lam) == len(lr)
    assert np.max(np.abs(lam - lr) / np.abs(lr)) <= 1e-8
    for i in range(len(lam)):
        gap = min([abs(lr[i] - lr[j]) / abs(lr[i]) for j in range(len(lr)) if j != i] + [1.0])
        if gap < 1e-6:
            continue
        assert _mac(modes[tr, i], ref["modes"][tr, i]) >= 1 - 1e-8
    dm.free()


def test_all_dofs_constrained_is_singular(ctx):
    from jf_2026_05_19_amd import _native as N
    arr = _arr(nx=4, ny=3, a=200.0, b=150.0, nd=1)
    arr.spc_mask = np.ones(arr.ndof, np.uint8)
    dm = N.DeviceModel(ctx, arr)
    with pytest.raises(N.Sol105Error) as ei:
        dm.static_solve()
    assert ei.value.code == -4                    # SOL105_E_SINGULAR
    dm.free()


def test_capped_eigensolver_reports_noconv(ctx):
    from jf_2026_05_19_amd import _native as N
    arr = _arr(nx=12, ny=12, a=480.0, b=480.0, nd=5)
    dm = N.DeviceModel(ctx, arr)
    lam_ok, _ = dm.solve(want_modes=False)
    try:
        ctx.set_option(N.Context.OPT_LANCZOS_MAX_VECTORS, 32)
        ctx.set_option(N.Context.OPT_LANCZOS_MAX_RESTARTS, 0)
        with pytest.raises(N.Sol105Error) as ei:
            dm.solve(want_modes=False)
        assert ei.value.code == -5                # SOL105_E_NOCONV
    finally:
        ctx.set_option(N.Context.OPT_LANCZOS_MAX_VECTORS, 1024)
        ctx.set_option(N.Context.OPT_LANCZOS_MAX_RESTARTS, 40)
    lam, _ = dm.solve(want_modes=False)
    assert len(lam) == len(lam_ok) and np.max(np.abs(lam - lam_ok) / np.abs(lam_ok)) <= 1e-12
    with pytest.raises(N.Sol105Error):
        ctx.set_option(N.Context.OPT_LANCZOS_MAX_VECTORS, 16)
    with pytest.raises(N.Sol105Error):
        ctx.set_option(99, 1)
    dm.free()
