"""HIP-graph replay of the block Lanczos steps (SOL105_OPT_STEP_GRAPHS).

A model's first solve launches every kernel directly and arms its step-graph
cache; the second captures each expansion step as a graph when it reaches
it; later solves replay the graphs.  Replayed steps run the same kernels on
the same addresses, so every solve must give bitwise the eigenvalues and
modes of direct launches in a context with the graphs off: for a single
model on the Krylov path, after a thickness change (a new numeric factor in
the same buffers, the same graphs replayed), and for a batch.
"""
import numpy as np
import pytest

from jf_2026_05_19_amd import bdf, decks, model as M

pytestmark = pytest.mark.gpu


def _arr(txt):
    return M.build_arrays(bdf.bdf_to_model_text(txt))


@pytest.fixture(scope="module")
def ctxs():
    from jf_2026_05_19_amd import _native as N
    g = N.Context(0)
    d = N.Context(0)
    d.set_option(N.Context.OPT_STEP_GRAPHS, 0)
    yield g, d
    g.close()
    d.close()


def test_option_values(ctxs):
    from jf_2026_05_19_amd import _native as N
    g, _ = ctxs
    with pytest.raises(N.Sol105Error):
        g.set_option(N.Context.OPT_STEP_GRAPHS, 2)
    g.set_option(N.Context.OPT_STEP_GRAPHS, 1)


def test_single_model_replay_is_bitwise(ctxs):
    from jf_2026_05_19_amd import _native as N
    g, d = ctxs
    arr = _arr(decks.panel_deck(nx=40, ny=20, a=400.0, b=200.0, stringer_every=5, nd=5, tri_frac=0.1))
    mg = N.DeviceModel(g, arr)
    md = N.DeviceModel(d, arr)
    try:
        ld, vd = md.solve()
        for rep in range(4):   # direct, capture, replay, replay
            lg, vg = mg.solve()
            print("solve", rep, lg[:3], flush=True)
            assert np.array_equal(lg, ld), (rep, lg, ld)
            assert np.array_equal(vg, vd), rep
        # a new design: new K, factor and KG in the same buffers
        t = np.asarray(arr.q4prop)[:, 2].copy()
        rng = np.random.default_rng(7)
        ts = np.concatenate([t * rng.uniform(0.9, 1.1, t.size),
                             np.asarray(arr.t3prop)[:, 2] * rng.uniform(0.9, 1.1, arr.n_t3)])
        mg.set_shell_thickness(ts)
        md.set_shell_thickness(ts)
        ld2, vd2 = md.solve()
        assert not np.array_equal(ld2, ld)
        for rep in range(2):
            lg2, vg2 = mg.solve()
            assert np.array_equal(lg2, ld2), (rep, lg2, ld2)
            assert np.array_equal(vg2, vd2), rep
    finally:
        mg.free()
        md.free()


def test_batch_replay_is_bitwise(ctxs):
    from jf_2026_05_19_amd import _native as N
    g, d = ctxs
    arrs = [_arr(txt) for txt, _ in decks.variant_sweep(n=5, n_min=10, n_max=40)]
    arrs.append(_arr(decks.panel_deck(nx=30, ny=16, a=300.0, b=160.0, stringer_every=4, nd=5)))
    print("batch", [a.ndof for a in arrs], flush=True)
    Bg = N.Batch(g, arrs)
    Bd = N.Batch(d, arrs)
    try:
        ld, sd = Bd.solve(16)
        md = [Bd.modes(i) for i in range(len(arrs))]
        assert all(s == 0 for s in sd), sd
        for rep in range(3):
            lg, sg = Bg.solve(16)
            print("batch solve", rep, [x[:1] for x in lg], flush=True)
            assert list(sg) == list(sd)
            for i in range(len(arrs)):
                assert np.array_equal(lg[i], ld[i]), (rep, i, lg[i], ld[i])
                assert np.array_equal(Bg.modes(i), md[i]), (rep, i)
    finally:
        Bg.free()
        Bd.free()
