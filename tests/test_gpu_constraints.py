"""GPU: RBE1 / RBE2 / RBE3 / RSPLINE / MPC constraints (dependent-DOF elimination)
through the C ABI against the oracle (constraints.jl:37-364 restated in
oracle.py):

* glued half meshes (RBE2 123456 pairs of coincident GRIDs, six MPC
  equations per pair, or one RBE1 per pair), a plate with eccentric CBAR stringers on their own
  GRIDs tied by RBE2, a plate loaded through an RBE3 at one reference GRID,
  and one with a force at a GRID tied to the loaded edge by an RSPLINE;
* K and KG (redistributed, dependent rows and columns cleared) within 1e-12
  of their largest entry, bit-identical run to run; the free set, u within
  1e-10 (dependent DOFs recovered); lambda within 1e-8 and modes at MAC >=
  1 - 1e-8 (pairs with a relative gap below 1e-6 skipped);
* the glued plate gives the one-piece plate's eigenvalues;
* batches and the no-allocation K assembly refuse constrained models.
"""
import numpy as np
import pytest

from jf_2026_05_19_amd import bdf, decks, model as M

pytestmark = pytest.mark.gpu

CASES = {
    "glued": lambda: decks.glued_plate_deck(nx=12, ny=8),
    "eccentric": lambda: decks.eccentric_panel_deck(nx=16, ny=8, stringer_every=4),
    "rbe3_load": lambda: decks.rbe3_load_plate_deck(nx=10, ny=8),
    "glued_mpc": lambda: decks.glued_plate_deck(nx=12, ny=8, tie="mpc"),
    "rspline_load": lambda: decks.rspline_load_plate_deck(nx=10, ny=8),
    "glued_rbe1": lambda: decks.glued_plate_deck(nx=12, ny=8, tie="rbe1"),
}


def _arr(txt):
    return M.build_arrays(bdf.bdf_to_model_text(txt))


@pytest.fixture(scope="module")
def ctx():
    from jf_2026_05_19_amd import _native as N
    c = N.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def solved(ctx):
    from jf_2026_05_19_amd import _native as N
    from oracle import oracle as O
    out = {}
    for name, mk in CASES.items():
        arr = _arr(mk())
        print("case", name, "nodes", arr.n_nodes, "rbe2", len(arr.rbe2), "rbe3", len(arr.rbe3), flush=True)
        out[name] = (arr, N.DeviceModel(ctx, arr), O.solve(arr))
    yield out
    for arr, dm, ref in out.values():
        dm.free()


@pytest.mark.parametrize("name", list(CASES))
def test_K_KG_static_parity(solved, name):
    arr, dm, ref = solved[name]
    print("K", name, flush=True)
    v1, v2 = dm.assemble_K(), dm.assemble_K()
    assert v1.tobytes() == v2.tobytes()
    K = dm.bsr_to_csr(v1)
    assert abs(K - ref["K"]).max() <= 1e-12 * abs(ref["K"]).max()
    u, free = dm.static_solve()
    assert np.array_equal(np.flatnonzero(free), ref["free"])
    assert np.abs(u - ref["u"]).max() <= 1e-10 * np.abs(ref["u"]).max()
    KG = dm.bsr_to_csr(dm.assemble_KG(u=ref["u"]))
    assert abs(KG - ref["KG"]).max() <= 1e-12 * abs(ref["KG"]).max()


@pytest.mark.parametrize("name", list(CASES))
def test_buckling_parity(solved, name):
    from oracle import oracle as O
    arr, dm, ref = solved[name]
    print("buckling", name, flush=True)
    lam, modes = dm.solve()
    lr = np.asarray(ref["eigenvalues"])
    assert len(lam) == len(lr)
    assert np.max(np.abs(lam - lr) / np.abs(lr)) <= 1e-8, (lam, lr)
    for i in range(len(lam)):
        gap = min([abs(lr[i] - lr[j]) / abs(lr[i]) for j in range(len(lr)) if j != i] + [1.0])
        if gap < 1e-6:
            continue
        assert O.mac(modes[:, i], ref["modes"][:, i]) >= 1 - 1e-8, i
        assert abs(np.abs(modes[:, i]).max() - 1.0) <= 1e-14


def test_glued_equals_one_piece(solved, ctx):
    from jf_2026_05_19_amd import _native as N
    dp = N.DeviceModel(ctx, _arr(decks.plate_deck(nx=12, ny=8, a=400.0, b=300.0, t=1.5, nd=4)))
    lam_p, _ = dp.solve()
    dp.free()
    for name in ("glued", "glued_mpc", "glued_rbe1"):
        arr, dm, ref = solved[name]
        lam_g, _ = dm.solve()
        assert np.allclose(lam_g, lam_p, rtol=1e-9, atol=0), (name, lam_g, lam_p)


def test_batch_and_noalloc_refuse_constraints(solved, ctx):
    from jf_2026_05_19_amd import _native as N
    arr, dm, ref = solved["eccentric"]
    with pytest.raises(N.Sol105Error):
        N.Batch(ctx, [arr, _arr(decks.plate_deck(nx=6, ny=4, nd=3))])
    with pytest.raises(N.Sol105Error):
        ctx.check(ctx.L.sol105_assemble_K_noalloc(ctx.h, dm.h, None))
