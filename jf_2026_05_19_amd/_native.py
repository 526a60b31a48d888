"""ctypes binding of libsol105_hip.so (the C ABI in include/sol105.h).

The library is required: importing succeeds without it (so CPU-only tooling
can inspect the package), but every call raises ``NativeLibraryMissing``
when the built library is absent.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsol105_hip.so")

SOL105_OK = 0
ERRORS = {-1: "SOL105_E_INVALID", -2: "SOL105_E_UNSUPPORTED", -3: "SOL105_E_DEVICE", -4: "SOL105_E_SINGULAR",
          -5: "SOL105_E_NOCONV", -6: "SOL105_E_STATE", -7: "SOL105_E_PENDING"}
SOL105_E_UNSUPPORTED = -2
SOL105_E_PENDING = -7
SERVER_MAX_MODES = 64   # a server request keeps at most this many modes

EXPORTS = [
    "sol105_ctx_create", "sol105_ctx_destroy", "sol105_last_error", "sol105_ctx_set_option", "sol105_model_upload", "sol105_model_free",
    "sol105_model_set_q4_thickness", "sol105_model_set_shell_thickness", "sol105_pattern", "sol105_assemble_K",
    "sol105_assemble_K_noalloc", "sol105_static_solve",
    "sol105_assemble_KG", "sol105_buckling", "sol105_buckling_info", "sol105_solve", "sol105_sensitivity_thickness",
    "sol105_sensitivity_terms", "sol105_profile_enable", "sol105_profile_read", "sol105_symeig_top", "sol105_last_timings",
    "sol105_last_stats", "sol105_batch_create", "sol105_batch_free", "sol105_batch_solve", "sol105_batch_modes",
    "sol105_batch_stats", "sol105_server_create", "sol105_server_destroy", "sol105_server_last_error",
    "sol105_server_add_model", "sol105_server_submit", "sol105_server_submit_ex", "sol105_server_poll",
    "sol105_server_wait", "sol105_server_wait_modes",
    "sol105_server_stats", "sol105_deck_read", "sol105_deck_free", "sol105_deck_sizes", "sol105_deck_desc",
    "sol105_deck_ids", "sol105_deck_bar_points", "sol105_static_recover", "sol105_solve_cfg_get",
    "sol105_solve_cfg_set", "sol105_solve_autotune", "sol105_solve_heights", "sol105_batch_autotune",
    "sol105_batch_model",
]


class NativeLibraryMissing(RuntimeError):
    pass


class Sol105Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, str(code)), msg))
        self.code = code


class ModelDesc(ctypes.Structure):
    _fields_ = [
        ("n_nodes", ctypes.c_int32), ("xyz", ctypes.POINTER(ctypes.c_double)),
        ("n_q4", ctypes.c_int32), ("q4_nodes", ctypes.POINTER(ctypes.c_int32)),
        ("q4_prop", ctypes.POINTER(ctypes.c_double)),
        ("n_t3", ctypes.c_int32), ("t3_nodes", ctypes.POINTER(ctypes.c_int32)),
        ("t3_prop", ctypes.POINTER(ctypes.c_double)),
        ("n_bar", ctypes.c_int32), ("bar_nodes", ctypes.POINTER(ctypes.c_int32)),
        ("bar_vref", ctypes.POINTER(ctypes.c_double)), ("bar_prop", ctypes.POINTER(ctypes.c_double)),
        ("spc_mask", ctypes.POINTER(ctypes.c_uint8)), ("force", ctypes.POINTER(ctypes.c_double)),
        ("node_has_shell", ctypes.POINTER(ctypes.c_uint8)), ("node_has_bar", ctypes.POINTER(ctypes.c_uint8)),
        ("k6rot", ctypes.c_double), ("autospc", ctypes.c_int32), ("nd", ctypes.c_int32),
        ("eigrl_v1", ctypes.c_double), ("eigrl_v2", ctypes.c_double), ("n_shell_cards", ctypes.c_int32),
        ("n_rbe2", ctypes.c_int32), ("rbe2", ctypes.POINTER(ctypes.c_int32)),
        ("n_rbe3", ctypes.c_int32), ("rbe3", ctypes.POINTER(ctypes.c_int32)),
        ("n_rbe3_grp", ctypes.c_int32), ("rbe3_grp", ctypes.POINTER(ctypes.c_int32)),
        ("rbe3_wt", ctypes.POINTER(ctypes.c_double)),
        ("n_rbe3_node", ctypes.c_int32), ("rbe3_node", ctypes.POINTER(ctypes.c_int32)),
        ("n_mpc", ctypes.c_int32), ("mpc_eq", ctypes.POINTER(ctypes.c_int32)),
        ("n_mpc_term", ctypes.c_int32), ("mpc_term", ctypes.POINTER(ctypes.c_int32)),
        ("mpc_coef", ctypes.POINTER(ctypes.c_double)),
        ("n_rspline", ctypes.c_int32), ("rspline", ctypes.POINTER(ctypes.c_int32)),
        ("n_rspline_indep", ctypes.c_int32), ("rspline_indep", ctypes.POINTER(ctypes.c_int32)),
        ("n_rspline_dep", ctypes.c_int32), ("rspline_dep", ctypes.POINTER(ctypes.c_int32)),
        ("n_rbe1", ctypes.c_int32), ("rbe1", ctypes.POINTER(ctypes.c_int32)),
        ("n_rbe1_indep", ctypes.c_int32), ("rbe1_indep", ctypes.POINTER(ctypes.c_int32)),
        ("n_rbe1_dep", ctypes.c_int32), ("rbe1_dep", ctypes.POINTER(ctypes.c_int32)),
    ]


_lib = None


def lib():
    """Load the native library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                "libsol105_hip.so is not built: run `python -m jf_2026_05_19_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.sol105_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.sol105_ctx_destroy.argtypes = [vp]
        L.sol105_last_error.argtypes = [vp]
        L.sol105_last_error.restype = ctypes.c_char_p
        L.sol105_ctx_set_option.argtypes = [vp, ctypes.c_int32, ctypes.c_double]
        L.sol105_model_upload.argtypes = [vp, ctypes.POINTER(ModelDesc), ctypes.POINTER(vp)]
        L.sol105_model_free.argtypes = [vp, vp]
        L.sol105_model_set_q4_thickness.argtypes = [vp, vp, dp]
        L.sol105_model_set_shell_thickness.argtypes = [vp, vp, dp]
        L.sol105_pattern.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_int32)]
        L.sol105_assemble_K.argtypes = [vp, vp, dp]
        L.sol105_assemble_K_noalloc.argtypes = [vp, vp, ctypes.c_void_p]
        L.sol105_static_solve.argtypes = [vp, vp, dp, ctypes.POINTER(ctypes.c_uint8)]
        L.sol105_assemble_KG.argtypes = [vp, vp, dp, dp]
        L.sol105_buckling.argtypes = [vp, vp, ctypes.c_int32, dp, dp, ctypes.POINTER(ctypes.c_int32)]
        L.sol105_solve.argtypes = [vp, vp, ctypes.c_int32, dp, dp, ctypes.POINTER(ctypes.c_int32)]
        L.sol105_buckling_info.argtypes = [vp, vp, dp]
        L.sol105_sensitivity_thickness.argtypes = [vp, vp, ctypes.c_int32, dp]
        L.sol105_sensitivity_terms.argtypes = [vp, vp, ctypes.c_int32, dp, dp, dp]
        L.sol105_profile_enable.argtypes = [vp, ctypes.c_int32]
        L.sol105_profile_read.argtypes = [vp, dp]
        L.sol105_symeig_top.argtypes = [vp, ctypes.c_int32, dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, dp, dp]
        L.sol105_last_timings.argtypes = [vp, dp]
        L.sol105_last_stats.argtypes = [vp, dp]
        ip = ctypes.POINTER(ctypes.c_int32)
        L.sol105_batch_create.argtypes = [vp, ctypes.POINTER(ModelDesc), ctypes.c_int32, ctypes.POINTER(vp)]
        L.sol105_batch_free.argtypes = [vp, vp]
        L.sol105_batch_solve.argtypes = [vp, vp, ctypes.c_int32, dp, ip, ip]
        L.sol105_batch_modes.argtypes = [vp, vp, ctypes.c_int32, dp]
        L.sol105_batch_stats.argtypes = [vp, vp, dp]
        L.sol105_server_create.argtypes = [ctypes.c_int, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                           ctypes.POINTER(vp)]
        L.sol105_server_destroy.argtypes = [vp]
        L.sol105_server_last_error.argtypes = [vp]
        L.sol105_server_last_error.restype = ctypes.c_char_p
        L.sol105_server_add_model.argtypes = [vp, ctypes.POINTER(ModelDesc), ip]
        L.sol105_server_submit.argtypes = [vp, ctypes.c_int32, dp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]
        L.sol105_server_poll.argtypes = [vp, ctypes.c_int64, dp, ip, ip, dp]
        L.sol105_server_wait.argtypes = [vp, ctypes.c_int64, dp, ip, ip, dp]
        L.sol105_server_submit_ex.argtypes = [vp, ctypes.c_int32, dp, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.POINTER(ctypes.c_int64)]
        L.sol105_server_wait_modes.argtypes = [vp, ctypes.c_int64, dp, ip, ip, dp, dp]
        L.sol105_server_stats.argtypes = [vp, dp]
        lp = ctypes.POINTER(ctypes.c_int64)
        L.sol105_deck_read.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(vp),
                                       ctypes.c_char_p, ctypes.c_int32]
        L.sol105_deck_free.argtypes = [vp]
        L.sol105_deck_sizes.argtypes = [vp, lp]
        L.sol105_deck_desc.argtypes = [vp, ctypes.POINTER(ModelDesc)]
        L.sol105_deck_ids.argtypes = [vp, lp, lp, lp, lp, lp, lp]
        L.sol105_deck_bar_points.argtypes = [vp, dp]
        L.sol105_static_recover.argtypes = [vp, vp, dp, dp, dp, dp, dp, dp]
        L.sol105_solve_cfg_get.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        L.sol105_solve_cfg_set.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
        L.sol105_solve_autotune.argtypes = [vp, vp, ctypes.c_int32, dp]
        L.sol105_batch_autotune.argtypes = [vp, vp, ctypes.c_int32, dp]
        L.sol105_batch_model.argtypes = [vp]
        L.sol105_batch_model.restype = vp
        L.sol105_solve_heights.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def read_deck(text, buckling_index=0, with_sol=False):
    """Sol105Arrays of a Nastran deck (text or bytes) read by the native
    reader (sol105_deck_read), with the deck's SOL as (arrays, sol) when
    with_sol; raises model.UnsupportedModel for decks outside the supported
    branch, ValueError for malformed ones."""
    from .model import Sol105Arrays, UnsupportedModel
    L = lib()
    data = text.encode() if isinstance(text, str) else bytes(text)
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(1024)
    rc = L.sol105_deck_read(data, len(data), int(buckling_index), ctypes.byref(h), err, len(err))
    if rc == SOL105_E_UNSUPPORTED:
        raise UnsupportedModel(err.value.decode())
    if rc != 0:
        raise ValueError(err.value.decode())
    try:
        sz = np.zeros(9, np.int64)
        lp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        L.sol105_deck_sizes(h, lp(sz))
        n, nq, nt, nb = (int(x) for x in sz[:4])
        d = ModelDesc()
        L.sol105_deck_desc(h, ctypes.byref(d))
        ids = [np.zeros(k, np.int64) for k in (n, nq, nq, nt, nt, nb)]
        L.sol105_deck_ids(h, *[lp(a) for a in ids])
        bar_points = np.zeros((nb, 8))
        L.sol105_deck_bar_points(h, bar_points.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

        def arr(ptr, count, dtype, shape):
            out = np.empty(count, dtype)
            if count:
                ctypes.memmove(out.ctypes.data, ptr, count * out.itemsize)
            return out.reshape(shape)
        dbl, i32, u8 = np.float64, np.int32, np.uint8
        arrays = Sol105Arrays(
            node_ids=ids[0], xyz=arr(d.xyz, 3 * n, dbl, (n, 3)),
            q4=arr(d.q4_nodes, 4 * nq, i32, (nq, 4)), q4_eid=ids[1], q4_pid=ids[2],
            q4prop=arr(d.q4_prop, 6 * nq, dbl, (nq, 6)),
            t3=arr(d.t3_nodes, 3 * nt, i32, (nt, 3)), t3_eid=ids[3], t3_pid=ids[4],
            t3prop=arr(d.t3_prop, 6 * nt, dbl, (nt, 6)),
            bar=arr(d.bar_nodes, 2 * nb, i32, (nb, 2)), bar_eid=ids[5],
            bar_vref=arr(d.bar_vref, 3 * nb, dbl, (nb, 3)), barprop=arr(d.bar_prop, 8 * nb, dbl, (nb, 8)),
            spc_mask=arr(d.spc_mask, 6 * n, u8, (6 * n,)), F=arr(d.force, 6 * n, dbl, (6 * n,)),
            nd=int(d.nd), v1=float(d.eigrl_v1), v2=float(d.eigrl_v2), k6rot=float(d.k6rot),
            autospc=bool(d.autospc), node_has_shell=arr(d.node_has_shell, n, u8, (n,)),
            node_has_bar=arr(d.node_has_bar, n, u8, (n,)),
            buckling_subcase=int(sz[5]), static_subcase=int(sz[6]), shell_cards=int(sz[7]),
            bar_points=bar_points,
            rbe2=arr(d.rbe2, 3 * d.n_rbe2, i32, (d.n_rbe2, 3)), rbe3=arr(d.rbe3, 4 * d.n_rbe3, i32, (d.n_rbe3, 4)),
            rbe3_grp=arr(d.rbe3_grp, 3 * d.n_rbe3_grp, i32, (d.n_rbe3_grp, 3)),
            rbe3_wt=arr(d.rbe3_wt, d.n_rbe3_grp, dbl, (d.n_rbe3_grp,)),
            rbe3_node=arr(d.rbe3_node, d.n_rbe3_node, i32, (d.n_rbe3_node,)),
            mpc_eq=arr(d.mpc_eq, 2 * d.n_mpc, i32, (d.n_mpc, 2)),
            mpc_term=arr(d.mpc_term, 2 * d.n_mpc_term, i32, (d.n_mpc_term, 2)),
            mpc_coef=arr(d.mpc_coef, d.n_mpc_term, dbl, (d.n_mpc_term,)),
            rspline=arr(d.rspline, 4 * d.n_rspline, i32, (d.n_rspline, 4)),
            rspline_indep=arr(d.rspline_indep, d.n_rspline_indep, i32, (d.n_rspline_indep,)),
            rspline_dep=arr(d.rspline_dep, 2 * d.n_rspline_dep, i32, (d.n_rspline_dep, 2)),
            rbe1=arr(d.rbe1, 4 * d.n_rbe1, i32, (d.n_rbe1, 4)),
            rbe1_indep=arr(d.rbe1_indep, 2 * d.n_rbe1_indep, i32, (d.n_rbe1_indep, 2)),
            rbe1_dep=arr(d.rbe1_dep, 2 * d.n_rbe1_dep, i32, (d.n_rbe1_dep, 2)))
        return (arrays, int(sz[8])) if with_sol else arrays
    finally:
        L.sol105_deck_free(h)


class Context:
    """One HIP device context (stream + error state)."""

    def __init__(self, device=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.sol105_ctx_create(int(device), ctypes.byref(h))
        if rc != SOL105_OK:
            raise Sol105Error(rc, "could not create a HIP context on device %d" % device)
        self.h = h
        self.L = L

    def check(self, rc):
        if rc != SOL105_OK:
            raise Sol105Error(rc, self.L.sol105_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.sol105_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    OPT_LANCZOS_MAX_VECTORS = 1
    OPT_LANCZOS_MAX_RESTARTS = 2
    OPT_WARM_START = 3
    OPT_STEP_GRAPHS = 4
    OPT_SENS_MODE = 5

    def set_option(self, option, value):
        """Context options (SOL105_OPT_*): basis size before a thick restart,
        thick restarts before SOL105_E_NOCONV, warm starts, step graphs."""
        self.check(self.L.sol105_ctx_set_option(self.h, int(option), float(value)))

    def symeig_top(self, A, k, bw=0, algebraic=False):
        """(theta (k,), Z (m, k)) of the projected eigensolver on a symmetric
        matrix A (diagnostics, sol105_symeig_top)."""
        A = np.asfortranarray(np.asarray(A, dtype=np.float64))
        m = A.shape[0]
        th = np.zeros(k)
        Z = np.zeros((k, m))
        self.check(self.L.sol105_symeig_top(self.h, int(m), A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(k),
                                            int(bw), 1 if algebraic else 0, _dp(th), _dp(Z)))
        return th, Z.T.copy()

    def profile_enable(self, on=True):
        self.check(self.L.sol105_profile_enable(self.h, 1 if on else 0))

    def profile_read(self):
        """Timed product launches: dict of launches, ms and algorithmic bytes
        over all heights ("all"), the leaf height ("leaf") and the heights
        above it ("upper")."""
        out = np.zeros(8)
        self.check(self.L.sol105_profile_read(self.h, _dp(out)))
        return {"all": (int(out[0]), float(out[1]), float(out[2])),
                "leaf": (int(out[3]), float(out[4]), float(out[5])),
                "upper": (int(out[0] - out[3]), float(out[6]), float(out[7]))}

    def timings(self):
        out = np.zeros(8)
        self.check(self.L.sol105_last_timings(self.h, _dp(out)))
        return out

    def stats(self):
        out = np.zeros(8)
        self.check(self.L.sol105_last_stats(self.h, _dp(out)))
        return out


def _model_desc(arr):
    """(ModelDesc, arrays it points into) for a Sol105Arrays."""
    keep = dict(
        xyz=np.ascontiguousarray(arr.xyz, dtype=np.float64),
        q4=np.ascontiguousarray(arr.q4, dtype=np.int32),
        q4prop=np.ascontiguousarray(arr.q4prop, dtype=np.float64),
        t3=np.ascontiguousarray(arr.t3, dtype=np.int32),
        t3prop=np.ascontiguousarray(arr.t3prop, dtype=np.float64),
        bar=np.ascontiguousarray(arr.bar, dtype=np.int32),
        vref=np.ascontiguousarray(arr.bar_vref, dtype=np.float64),
        barprop=np.ascontiguousarray(arr.barprop, dtype=np.float64),
        spc=np.ascontiguousarray(arr.spc_mask, dtype=np.uint8),
        F=np.ascontiguousarray(arr.F, dtype=np.float64),
        hs=np.ascontiguousarray(arr.node_has_shell, dtype=np.uint8),
        hb=np.ascontiguousarray(arr.node_has_bar, dtype=np.uint8),
        rbe2=np.ascontiguousarray(arr.rbe2, dtype=np.int32).reshape(-1),
        rbe3=np.ascontiguousarray(arr.rbe3, dtype=np.int32).reshape(-1),
        rbe3_grp=np.ascontiguousarray(arr.rbe3_grp, dtype=np.int32).reshape(-1),
        rbe3_wt=np.ascontiguousarray(arr.rbe3_wt, dtype=np.float64),
        rbe3_node=np.ascontiguousarray(arr.rbe3_node, dtype=np.int32),
        mpc_eq=np.ascontiguousarray(arr.mpc_eq, dtype=np.int32).reshape(-1),
        mpc_term=np.ascontiguousarray(arr.mpc_term, dtype=np.int32).reshape(-1),
        mpc_coef=np.ascontiguousarray(arr.mpc_coef, dtype=np.float64),
        rspline=np.ascontiguousarray(arr.rspline, dtype=np.int32).reshape(-1),
        rspline_indep=np.ascontiguousarray(arr.rspline_indep, dtype=np.int32),
        rspline_dep=np.ascontiguousarray(arr.rspline_dep, dtype=np.int32).reshape(-1),
        rbe1=np.ascontiguousarray(arr.rbe1, dtype=np.int32).reshape(-1),
        rbe1_indep=np.ascontiguousarray(arr.rbe1_indep, dtype=np.int32).reshape(-1),
        rbe1_dep=np.ascontiguousarray(arr.rbe1_dep, dtype=np.int32).reshape(-1),
    )
    k = keep
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    up = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    d = ModelDesc(arr.n_nodes, _dp(k["xyz"]), arr.n_q4, ip(k["q4"]), _dp(k["q4prop"]),
                  arr.n_t3, ip(k["t3"]), _dp(k["t3prop"]), arr.n_bar, ip(k["bar"]), _dp(k["vref"]), _dp(k["barprop"]),
                  up(k["spc"]), _dp(k["F"]), up(k["hs"]), up(k["hb"]),
                  float(arr.k6rot), int(arr.autospc), int(arr.nd),
                  float(arr.v1), float(arr.v2), int(arr.n_shell_cards),
                  len(arr.rbe2), ip(k["rbe2"]), len(arr.rbe3), ip(k["rbe3"]), len(arr.rbe3_grp), ip(k["rbe3_grp"]),
                  _dp(k["rbe3_wt"]), len(arr.rbe3_node), ip(k["rbe3_node"]),
                  len(arr.mpc_eq), ip(k["mpc_eq"]), len(arr.mpc_term), ip(k["mpc_term"]), _dp(k["mpc_coef"]),
                  len(arr.rspline), ip(k["rspline"]), len(arr.rspline_indep), ip(k["rspline_indep"]),
                  len(arr.rspline_dep), ip(k["rspline_dep"]),
                  len(arr.rbe1), ip(k["rbe1"]), len(arr.rbe1_indep), ip(k["rbe1_indep"]),
                  len(arr.rbe1_dep), ip(k["rbe1_dep"]))
    return d, keep


class _ModelRef:
    """A model handle owned elsewhere (a batch's stacked model): the
    solve-shape calls only."""

    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, ctypes.c_void_p(h)

    def solve_heights(self):
        return DeviceModel.solve_heights(self)

    def solve_cfg(self):
        return DeviceModel.solve_cfg(self)

    def set_solve_cfg(self, cfg):
        DeviceModel.set_solve_cfg(self, cfg)


class DeviceModel:
    """A model uploaded to the device (pattern, maps and ordering built)."""

    def __init__(self, ctx: Context, arr):
        self.ctx = ctx
        self.arr = arr
        L = ctx.L
        d, self._keep = _model_desc(arr)
        h = ctypes.c_void_p()
        ctx.check(L.sol105_model_upload(ctx.h, ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        nb = ctypes.c_int64()
        ctx.check(L.sol105_pattern(ctx.h, h, ctypes.byref(nb), None, None))
        self.n_blocks = int(nb.value)
        self.row_ptr = np.zeros(arr.n_nodes + 1, np.int32)
        self.col = np.zeros(self.n_blocks, np.int32)
        ctx.check(L.sol105_pattern(ctx.h, h, None, self.row_ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                   self.col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))

    def free(self):
        if self.h:
            self.ctx.L.sol105_model_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # --- phases -----------------------------------------------------------
    def set_q4_thickness(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        self.ctx.check(self.ctx.L.sol105_model_set_q4_thickness(self.ctx.h, self.h, _dp(t)))

    def set_shell_thickness(self, t):
        """Thickness of every shell (CQUAD4 first, then CTRIA3)."""
        t = np.ascontiguousarray(t, dtype=np.float64)
        assert t.shape == (self.arr.n_shell,)
        self.ctx.check(self.ctx.L.sol105_model_set_shell_thickness(self.ctx.h, self.h, _dp(t)))

    def assemble_K(self, want=True):
        vals = np.zeros(self.n_blocks * 36) if want else None
        self.ctx.check(self.ctx.L.sol105_assemble_K(self.ctx.h, self.h, _dp(vals) if want else None))
        return vals

    def static_solve(self):
        u = np.zeros(self.arr.ndof)
        fm = np.zeros(self.arr.ndof, np.uint8)
        self.ctx.check(self.ctx.L.sol105_static_solve(self.ctx.h, self.h, _dp(u),
                                                      fm.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return u, fm.astype(bool)

    def static_recover(self, bar_points=None, timings=False):
        """SOL 101 result recovery on the device-resident static state:
        ``(residual K u - F, CQUAD4 rows (n_q4, 44), CTRIA3 rows (n_t3, 20),
        CBAR rows (n_bar, 18))`` in the column layout of
        sol105_static_recover, plus the four kernel times in ms when
        ``timings``."""
        a = self.arr
        pts = a.bar_points if bar_points is None else bar_points
        pts = None if pts is None or a.n_bar == 0 else np.ascontiguousarray(pts, dtype=np.float64).reshape(a.n_bar, 8)
        res = np.zeros(a.ndof)
        q4 = np.zeros((44, a.n_q4)); t3 = np.zeros((20, a.n_t3)); bar = np.zeros((18, a.n_bar))
        ms = np.zeros(4)
        self.ctx.check(self.ctx.L.sol105_static_recover(
            self.ctx.h, self.h, _dp(pts) if pts is not None else None, _dp(res), _dp(q4), _dp(t3), _dp(bar),
            _dp(ms) if timings else None))
        out = (res, q4.T.copy(), t3.T.copy(), bar.T.copy())
        return out + (ms,) if timings else out

    def solve_cfg(self):
        """Launch shapes of the triangular-solve products per tree height:
        array (heights, 2, 4) of (kmax, ks, pch, kg), forward then backward."""
        L, c = self.ctx.L, self.ctx
        n = ctypes.c_int32(0)
        c.check(L.sol105_solve_cfg_get(c.h, self.h, None, ctypes.byref(n)))
        cfg = np.zeros((n.value, 2, 4), np.int32)
        c.check(L.sol105_solve_cfg_get(c.h, self.h, cfg.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                       ctypes.byref(n)))
        return cfg

    def solve_heights(self):
        """Per tree height: (fronts, largest m, largest ns, nonzero factor
        entries)."""
        L, c = self.ctx.L, self.ctx
        n = ctypes.c_int32(0)
        c.check(L.sol105_solve_heights(c.h, self.h, None, ctypes.byref(n)))
        sh = np.zeros((n.value, 4), np.int64)
        c.check(L.sol105_solve_heights(c.h, self.h, sh.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       ctypes.byref(n)))
        return sh

    def set_solve_cfg(self, cfg):
        cfg = np.ascontiguousarray(cfg, dtype=np.int32).reshape(-1, 2, 4)
        self.ctx.check(self.ctx.L.sol105_solve_cfg_set(self.ctx.h, self.h,
                                                       cfg.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                       len(cfg)))

    def autotune(self, reps=7):
        """Time every height's product launch for candidate shapes on this
        model's factor and keep the fastest; returns (block solve ms before,
        after, launches timed, heights)."""
        rep = np.zeros(4)
        self.ctx.check(self.ctx.L.sol105_solve_autotune(self.ctx.h, self.h, int(reps), _dp(rep)))
        return float(rep[0]), float(rep[1]), int(rep[2]), int(rep[3])

    def assemble_KG(self, u=None, want=True):
        vals = np.zeros(self.n_blocks * 36) if want else None
        uu = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        self.ctx.check(self.ctx.L.sol105_assemble_KG(self.ctx.h, self.h, _dp(uu) if uu is not None else None,
                                                     _dp(vals) if want else None))
        return vals

    def buckling(self, max_modes=64, want_modes=True):
        lam = np.zeros(max_modes)
        modes = np.zeros((max_modes, self.arr.ndof)) if want_modes else None
        n = ctypes.c_int32()
        self.ctx.check(self.ctx.L.sol105_buckling(self.ctx.h, self.h, max_modes, _dp(lam),
                                                  _dp(modes) if want_modes else None, ctypes.byref(n)))
        k = n.value
        return lam[:k].copy(), (modes[:k].T.copy() if want_modes else None)

    def buckling_info(self):
        """Selection record of the last eigensolve: eigenpairs computed,
        positive candidates in the EIGRL range, modes the cluster filter
        skipped, and whether the computed pairs decide the output as the
        complete in-range spectrum would."""
        out = np.zeros(4)
        self.ctx.check(self.ctx.L.sol105_buckling_info(self.ctx.h, self.h, _dp(out)))
        return dict(pairs=int(out[0]), in_range=int(out[1]), cluster_skip=int(out[2]), decided=bool(out[3]))

    def solve(self, max_modes=64, want_modes=True):
        lam = np.zeros(max_modes)
        modes = np.zeros((max_modes, self.arr.ndof)) if want_modes else None
        n = ctypes.c_int32()
        self.ctx.check(self.ctx.L.sol105_solve(self.ctx.h, self.h, max_modes, _dp(lam),
                                               _dp(modes) if want_modes else None, ctypes.byref(n)))
        k = n.value
        return lam[:k].copy(), (modes[:k].T.copy() if want_modes else None)

    def sensitivity_thickness(self, mode=0):
        """d(lambda_mode)/dt of every shell (CQUAD4 first, then CTRIA3)."""
        out = np.zeros(self.arr.n_q4 + self.arr.n_t3)
        self.ctx.check(self.ctx.L.sol105_sensitivity_thickness(self.ctx.h, self.h, int(mode), _dp(out)))
        return out

    def sensitivity_terms(self, mode=0):
        """(terms (n_q4 + n_t3) x 3, adjoint load ndof, phi' KG phi) of d(lambda)/dt."""
        terms = np.zeros((self.arr.n_q4 + self.arr.n_t3, 3))
        g = np.zeros(self.arr.ndof)
        pkp = np.zeros(1)
        self.ctx.check(self.ctx.L.sol105_sensitivity_terms(self.ctx.h, self.h, int(mode), _dp(terms), _dp(g),
                                                           _dp(pkp)))
        return terms, g, float(pkp[0])

    def bsr_to_csr(self, vals):
        """BSR values (n_blocks*36) -> scipy CSR over all DOFs."""
        import scipy.sparse as sp
        nb = self.n_blocks
        brow = np.repeat(np.arange(self.arr.n_nodes), np.diff(self.row_ptr))
        ci = np.arange(36) // 6
        cj = np.arange(36) % 6
        I = (6 * brow[:, None] + ci[None, :]).ravel()
        J = (6 * self.col[:, None].astype(np.int64) + cj[None, :]).ravel()
        return sp.csr_matrix((np.asarray(vals).reshape(nb * 36), (I, J)), shape=(self.arr.ndof, self.arr.ndof))


class Batch:
    """Independent models solved together as one block-diagonal system
    (sol105_batch_*): results per member, as a DeviceModel.solve of each
    member alone would give them."""

    def __init__(self, ctx: Context, arrs):
        self.ctx = ctx
        self.arrs = list(arrs)
        pairs = [_model_desc(a) for a in self.arrs]
        self._keep = [k for _, k in pairs]
        descs = (ModelDesc * len(pairs))(*[d for d, _ in pairs])
        h = ctypes.c_void_p()
        ctx.check(ctx.L.sol105_batch_create(ctx.h, descs, len(pairs), ctypes.byref(h)))
        self.h = h
        self.n_out = None

    def free(self):
        if self.h:
            self.ctx.L.sol105_batch_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def solve(self, max_modes=16):
        """(list of eigenvalue arrays, status codes); a member's own error is
        reported in its status, not raised."""
        n = len(self.arrs)
        lam = np.zeros((n, max_modes))
        nout = np.zeros(n, np.int32)
        st = np.zeros(n, np.int32)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        self.ctx.check(self.ctx.L.sol105_batch_solve(self.ctx.h, self.h, int(max_modes), _dp(lam), ip(nout), ip(st)))
        self.n_out = nout
        return [lam[i, :nout[i]].copy() for i in range(n)], st

    def modes(self, member):
        k = int(self.n_out[member])
        out = np.zeros((k, self.arrs[member].ndof))
        if k:
            self.ctx.check(self.ctx.L.sol105_batch_modes(self.ctx.h, self.h, int(member), _dp(out)))
        return out.T.copy()

    def stats(self):
        """dict of the last batch solve: device ms, Lanczos size, steps,
        members re-solved alone, stacked free DOFs, front bytes, factor flops,
        tree height."""
        out = np.zeros(8)
        self.ctx.check(self.ctx.L.sol105_batch_stats(self.ctx.h, self.h, _dp(out)))
        keys = ("ms", "lanczos_vectors", "block_steps", "fallbacks", "n_free", "front_bytes", "factor_flops",
                "tree_height")
        return dict(zip(keys, out.tolist()))

    def _shapes(self):
        m = _ModelRef(self.ctx, self.ctx.L.sol105_batch_model(self.h))
        return m

    def solve_heights(self):
        return self._shapes().solve_heights()

    def solve_cfg(self):
        return self._shapes().solve_cfg()

    def set_solve_cfg(self, cfg):
        self._shapes().set_solve_cfg(cfg)

    def autotune(self, reps=7):
        """The solve-product autotuner on the stacked model (after a solve):
        (block solve ms before, after, launches timed, heights)."""
        rep = np.zeros(4)
        self.ctx.check(self.ctx.L.sol105_batch_autotune(self.ctx.h, self.h, int(reps), _dp(rep)))
        return float(rep[0]), float(rep[1]), int(rep[2]), int(rep[3])


class Server:
    """Device-resident request server (sol105_server_*): meshes registered
    once, designs submitted as shell thicknesses, requests served by worker
    streams with stacking of small pending requests and, on request,
    warm-started eigensolves."""

    def __init__(self, device=0, workers=4, hbm_budget=0.0, batching=True, warm_start=False):
        self.L = lib()
        h = ctypes.c_void_p()
        flags = (1 if batching else 0) | (2 if warm_start else 0)   # SOL105_SERVER_STACKING, _WARM_START
        rc = self.L.sol105_server_create(int(device), int(workers), float(hbm_budget), flags, ctypes.byref(h))
        if rc != SOL105_OK:
            raise Sol105Error(rc, "could not create the server")
        self.h = h
        self.arrs = []
        self._tickets = {}

    def _check(self, rc):
        if rc != SOL105_OK:
            raise Sol105Error(rc, self.L.sol105_server_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.sol105_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_model(self, arr):
        d, keep = _model_desc(arr)
        mid = ctypes.c_int32()
        self._check(self.L.sol105_server_add_model(self.h, ctypes.byref(d), ctypes.byref(mid)))
        self.arrs.append(arr)
        return int(mid.value)

    def submit(self, model_id, thickness=None, max_modes=16, modes=False):
        """Ticket of a new request; with modes=True the request keeps its mode
        shapes for wait_modes()."""
        t = None if thickness is None else np.ascontiguousarray(thickness, dtype=np.float64)
        if t is not None:
            assert t.shape == (self.arrs[model_id].n_shell,)
        tk = ctypes.c_int64()
        self._check(self.L.sol105_server_submit_ex(self.h, int(model_id), _dp(t) if t is not None else None,
                                                   int(max_modes), 1 if modes else 0, ctypes.byref(tk)))
        self._tickets[int(tk.value)] = (int(model_id), min(int(max_modes), SERVER_MAX_MODES))
        return int(tk.value)

    def wait_modes(self, ticket):
        """(eigenvalues, status, info, modes 6 n_nodes x n) of a request
        submitted with modes=True, blocking until it is done."""
        mid, mm = self._tickets.pop(ticket, (0, SERVER_MAX_MODES))
        ndof = self.arrs[mid].ndof
        lam = np.zeros(SERVER_MAX_MODES)
        modes = np.zeros((max(mm, 1), ndof))
        n, st, info = ctypes.c_int32(), ctypes.c_int32(), np.zeros(8)
        self._check(self.L.sol105_server_wait_modes(self.h, ctypes.c_int64(ticket), _dp(lam), ctypes.byref(n),
                                                    ctypes.byref(st), _dp(info), _dp(modes)))
        keys = ("latency_ms", "queue_ms", "service_ms", "members", "start_ms", "finish_ms", "lanczos_vectors",
                "cached")
        k = n.value
        return lam[:k].copy(), int(st.value), dict(zip(keys, info.tolist())), modes[:k].T.copy()

    def _result(self, fn, ticket):
        lam = np.zeros(SERVER_MAX_MODES)
        n = ctypes.c_int32()
        st = ctypes.c_int32()
        info = np.zeros(8)
        rc = fn(self.h, ctypes.c_int64(ticket), _dp(lam), ctypes.byref(n), ctypes.byref(st), _dp(info))
        if rc == SOL105_E_PENDING:
            return None
        self._tickets.pop(ticket, None)
        self._check(rc)
        keys = ("latency_ms", "queue_ms", "service_ms", "members", "start_ms", "finish_ms", "lanczos_vectors",
                "cached")
        return lam[:n.value].copy(), int(st.value), dict(zip(keys, info.tolist()))

    def poll(self, ticket):
        """(eigenvalues, status, info) when done, else None."""
        return self._result(self.L.sol105_server_poll, ticket)

    def wait(self, ticket):
        return self._result(self.L.sol105_server_wait, ticket)

    def stats(self):
        out = np.zeros(10)
        self._check(self.L.sol105_server_stats(self.h, _dp(out)))
        keys = ("done", "stacked_solves", "stacked_requests", "queued", "inflight_bytes", "budget_bytes",
                "workers", "meshes", "cached", "peak_inflight_bytes")
        return dict(zip(keys, out.tolist()))
