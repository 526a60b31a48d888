"""GPU: launch shapes of the triangular-solve products and the autotuner
(sol105_solve_cfg_* / sol105_solve_autotune / sol105_batch_autotune), and the
persisted table (jf_2026_05_19_amd/tuning.py).

Every shape computes the same solve up to the summation order of k-slices:
eigenvalues of a stiffened panel within 1e-12 relative of the default
shapes' (and of the oracle within the parity bound 1e-8) for forced splits
(kmax = ks = 128), wide slices (384), each chunk length and each count of
k-groups per workgroup (their parts summed in a fixed order); a tuned model and
a tuned batch give the same eigenvalues as before tuning; a table written
from one model's tuning sets the same shapes on a fresh model of the mesh.
"""
import os

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


def _panel():
    return M.build_arrays(bdf.bdf_to_model_text(decks.panel_deck(nx=40, ny=20, stringer_every=4, nd=4,
                                                                 tri_frac=0.1)))


def test_shapes_give_the_same_solve(ctx):
    from jf_2026_05_19_amd import _native as N
    from oracle import oracle as O
    arr = _panel()
    print("shapes on panel ndof", arr.ndof, flush=True)
    ref = O.solve(arr)["eigenvalues"]
    dm = N.DeviceModel(ctx, arr)
    try:
        lam0, _ = dm.solve(want_modes=False)
        assert np.max(np.abs(lam0 - ref) / np.abs(ref)) <= 1e-8
        nh = len(dm.solve_heights())
        for shape in [(128, 128, 8, 1), (384, 384, 16, 1), (256, 128, 32, 1), (384, 256, 8, 1),
                      (384, 128, 8, 2), (384, 384, 16, 4), (128, 128, 0, 4), (256, 256, 16, 2), (0, 0, 0, 0)]:
            cfg = np.tile(np.array(shape, np.int32), (nh, 2, 1))
            dm.set_solve_cfg(cfg)
            got = dm.solve_cfg()
            if shape[0]:
                assert (got == cfg).all()
            lam, _ = dm.solve(want_modes=False)
            assert np.max(np.abs(lam - lam0) / np.abs(lam0)) <= 1e-12, shape
        for bad in [(384, 128, 12, 1), (384, 128, 8, 3), (384, 128, 32, 2)]:   # pch, kg, pch 32 with kg > 1
            with pytest.raises(N.Sol105Error):
                dm.set_solve_cfg(np.tile(np.array(bad, np.int32), (nh, 2, 1)))
    finally:
        dm.free()


def test_autotune_and_table(ctx, tmp_path):
    from jf_2026_05_19_amd import _native as N, tuning
    arr = _panel()
    dm = N.DeviceModel(ctx, arr)
    try:
        lam0, _ = dm.solve(want_modes=False)
        table, rep = tuning.tune(dm, {}, reps=3)
        before, after, timed, heights = rep
        print("autotune", rep, flush=True)
        assert before > 0 and after > 0 and timed >= heights > 0
        lam, _ = dm.solve(want_modes=False)
        assert np.max(np.abs(lam - lam0) / np.abs(lam0)) <= 1e-12
        tuned = dm.solve_cfg()
        path = os.path.join(str(tmp_path), "shapes.json")
        tuning.save(path, table)
    finally:
        dm.free()
    dm2 = N.DeviceModel(ctx, arr)
    try:
        dm2.solve(want_modes=False)
        hit = tuning.apply(dm2, tuning.load(path))
        assert hit == 2 * len(tuned)
        assert np.array_equal(dm2.solve_cfg(), tuned)
        lam2, _ = dm2.solve(want_modes=False)
        assert np.max(np.abs(lam2 - lam0) / np.abs(lam0)) <= 1e-12
    finally:
        dm2.free()


def test_batch_autotune_keeps_results(ctx):
    from jf_2026_05_19_amd import _native as N
    arrs = [M.build_arrays(bdf.bdf_to_model_text(decks.plate_deck(nx=n, ny=n, nd=3))) for n in (10, 14, 18, 22)]
    B = N.Batch(ctx, arrs)
    try:
        lam0, st0 = B.solve(8)
        rep = B.autotune(3)
        print("batch autotune", rep, flush=True)
        assert rep[0] > 0 and rep[1] > 0
        assert B.solve_heights().shape[0] == rep[3]
        lam1, st1 = B.solve(8)
        assert np.array_equal(st0, st1)
        for a, b in zip(lam0, lam1):
            assert len(a) == len(b) and np.max(np.abs(a - b) / np.abs(a)) <= 1e-12
    finally:
        B.free()
