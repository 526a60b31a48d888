"""sol105_assemble_K_noalloc: K assembly with the element-block staging in the
library's static device scratch (no allocation in the call).

Checked through the C ABI against sol105_assemble_K on two calls in a row,
the second with a smaller model: the values must be equal bit for bit (same
element kernels, same fixed-order gather), the model's device K must be the
one later phases use (the static solve after it equals the one after
sol105_assemble_K), and a model whose staging exceeds the scratch must be
refused before any launch.
"""
import ctypes
import os

import numpy as np
import pytest

from jf_2026_05_19_amd import bdf, decks, model as M

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 * 1024 * 1024   # SOL105_NOALLOC_SCRATCH_BYTES (include/sol105.h)


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            H = ctypes.CDLL(name)
            break
        except OSError:
            H = None
    assert H is not None, "HIP runtime library not found"
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


def _arr(txt):
    return M.build_arrays(bdf.bdf_to_model_text(txt))


def _staging_bytes(arr):
    return (16 * arr.n_q4 + 9 * arr.n_t3 + 4 * arr.n_bar) * 36 * 8


@pytest.fixture(scope="module")
def ctx():
    from jf_2026_05_19_amd import _native as N
    c = N.Context(0)
    yield c
    c.close()


def _noalloc_K(ctx, dm, H):
    nbytes = dm.n_blocks * 36 * 8
    d = ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(d), nbytes) == 0
    try:
        ctx.check(ctx.L.sol105_assemble_K_noalloc(ctx.h, dm.h, d))
        out = np.zeros(dm.n_blocks * 36)
        assert H.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), d, nbytes, 2) == 0   # hipMemcpyDeviceToHost
    finally:
        H.hipFree(d)
    return out


def test_noalloc_K_equals_assemble_K_two_calls(ctx):
    from jf_2026_05_19_amd import _native as N
    H = _hip()
    # first a mixed CQUAD4/CTRIA3 panel with CBAR stringers, then a smaller plate
    cases = [("mixed_panel_36x18", decks.panel_deck(nx=36, ny=18, a=360.0, b=180.0, stringer_every=6, nd=4,
                                                    tri_frac=0.2)),
             ("plate_8x6", decks.plate_deck(nx=8, ny=6, a=400.0, b=300.0, t=1.5, nd=3))]
    sizes = []
    for name, txt in cases:
        arr = _arr(txt)
        assert _staging_bytes(arr) <= SCRATCH_BYTES
        sizes.append(arr.ndof)
        print("no-alloc K:", name, "ndof", arr.ndof, "q4", arr.n_q4, "t3", arr.n_t3, "bars", arr.n_bar, flush=True)
        dm = N.DeviceModel(ctx, arr)
        ref = dm.assemble_K()
        u_ref, _ = dm.static_solve()
        got = _noalloc_K(ctx, dm, H)
        assert got.tobytes() == ref.tobytes(), "no-allocation K differs from sol105_assemble_K"
        # K stays in the model's device buffer: the static solve runs on it
        u, _ = dm.static_solve()
        assert u.tobytes() == u_ref.tobytes()
        # NULL output: K only on the device
        ctx.check(ctx.L.sol105_assemble_K_noalloc(ctx.h, dm.h, None))
        u2, _ = dm.static_solve()
        assert u2.tobytes() == u_ref.tobytes()
        dm.free()
    assert sizes[1] < sizes[0]


def test_noalloc_K_refuses_oversize_model(ctx):
    from jf_2026_05_19_amd import _native as N
    arr = _arr(decks.plate_deck(nx=125, ny=120, a=1250.0, b=1200.0, t=1.5, nd=3))
    assert _staging_bytes(arr) > SCRATCH_BYTES
    dm = N.DeviceModel(ctx, arr)
    rc = ctx.L.sol105_assemble_K_noalloc(ctx.h, dm.h, None)
    assert rc == -1, rc   # SOL105_E_INVALID, refused before any launch
    msg = ctx.L.sol105_last_error(ctx.h).decode()
    assert "no-allocation scratch" in msg, msg
    dm.free()
