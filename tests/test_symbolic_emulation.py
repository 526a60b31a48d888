"""CPU checks of the device algorithm's host-side structures.

The block pattern, the deterministic element->block gather map and the
nested-dissection multifrontal structure built by ``csrc/symbolic.cpp`` are
compiled into a host-only probe (g++) and exercised here with a NumPy
emulation of exactly what the device kernels do (block gather, K scatter into
fronts, child extend-add, partial dense Cholesky, front-vector triangular
solves).  Element matrices come from the oracle.  This pins the layout
conventions and maps without a GPU.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from jf_2026_05_19_amd import bdf, decks, model as M
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native_host", "symbolic_probe.cpp")
LIBDIR = os.path.join(HERE, "native_host", "_build")
LIB = os.path.join(LIBDIR, "libsymprobe.so")


def _probe():
    os.makedirs(LIBDIR, exist_ok=True)
    csrc = os.path.join(HERE, "..", "jf_2026_05_19_amd", "csrc")
    tmp = LIB + ".tmp%d" % os.getpid()
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", SRC, os.path.join(csrc, "symbolic.cpp"),
                           os.path.join(csrc, "solve_plan.cpp"), "-o", tmp])
    os.replace(tmp, LIB)
    L = ctypes.CDLL(LIB)
    vp = ctypes.c_void_p
    L.probe_analyze.restype = vp
    L.probe_analyze.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, vp,
                                ctypes.c_int]
    L.probe_free.argtypes = [vp]
    L.probe_sizes.argtypes = [vp, vp]
    L.probe_pattern.argtypes = [vp, vp, vp, vp, vp]
    L.probe_fronts.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.probe_plans.restype = vp
    L.probe_plans.argtypes = [vp]
    L.probe_plans_free.argtypes = [vp]
    L.probe_plan_sizes.argtypes = [vp, vp]
    L.probe_plan.argtypes = [vp] * 10
    L.probe_pull_sizes.argtypes = [vp, vp]
    L.probe_pull.argtypes = [vp, vp, vp]
    return L


@pytest.fixture(scope="module")
def probe():
    return _probe()


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def analyze(L, arr, free_mask, leaf=8):
    q4 = np.ascontiguousarray(arr.q4, np.int32)
    t3 = np.ascontiguousarray(arr.t3, np.int32)
    bar = np.ascontiguousarray(arr.bar, np.int32)
    xyz = np.ascontiguousarray(arr.xyz, np.float64)
    fm = np.ascontiguousarray(free_mask, np.uint8)
    h = L.probe_analyze(arr.n_nodes, arr.n_q4, _ptr(q4), arr.n_t3, _ptr(t3), arr.n_bar, _ptr(bar), _ptr(xyz),
                        _ptr(fm), leaf)
    sz = np.zeros(9, np.int64)
    L.probe_sizes(h, _ptr(sz))
    nb, nc, nfree, nf, nrows, nrel, nk, fs, height = [int(x) for x in sz]
    rp = np.zeros(arr.n_nodes + 1, np.int32); col = np.zeros(nb, np.int32)
    cptr = np.zeros(nb + 1, np.int32); cidx = np.zeros(nc, np.int32)
    L.probe_pattern(h, _ptr(rp), _ptr(col), _ptr(cptr), _ptr(cidx))
    info = np.zeros((nf, 9), np.int64); rows = np.zeros(nrows, np.int32); rel = np.zeros(nrel, np.int32)
    eod = np.zeros(arr.ndof, np.int32); ksrc = np.zeros(nk, np.int32); kdst = np.zeros(nk, np.int64)
    L.probe_fronts(h, _ptr(info), _ptr(rows), _ptr(rel), _ptr(eod), _ptr(ksrc), _ptr(kdst))
    q = L.probe_plans(h)
    ps = np.zeros(5, np.int64)
    L.probe_plan_sizes(q, _ptr(ps))
    nft, nbt, H, mht, pst = [int(x) for x in ps]
    plan = dict(ft=np.zeros(3 * nft, np.int32), bt=np.zeros(3 * nbt, np.int32), ftb=np.zeros(nf, np.int32),
                btb=np.zeros(nf, np.int32), fnk=np.zeros(nf, np.int32), bnk=np.zeros(nf, np.int32),
                flist=np.zeros(nf, np.int32), hinfo=np.zeros((H + 1, 8), np.int64), poff=np.zeros(nf, np.int64),
                max_height_tiles=mht, p_storage=pst)
    L.probe_plan(q, *[_ptr(plan[k]) for k in ("ft", "bt", "ftb", "btb", "fnk", "bnk", "flist", "hinfo", "poff")])
    psz = np.zeros(2, np.int64)
    L.probe_pull_sizes(q, _ptr(psz))
    plan["pull_ptr"] = np.zeros(int(psz[0]), np.int32)
    plan["pull_src"] = np.zeros(2 * int(psz[1]), np.int32)
    L.probe_pull(q, _ptr(plan["pull_ptr"]), _ptr(plan["pull_src"]))
    L.probe_plans_free(q)
    L.probe_free(h)
    return dict(rp=rp, col=col, cptr=cptr, cidx=cidx, nfree=nfree, info=info, rows=rows, rel=rel, eod=eod,
                ksrc=ksrc, kdst=kdst, fs=fs, height=height, plan=plan)


def element_blocks(arr):
    """Oracle element matrices laid out exactly like the device buffer."""
    eb = np.zeros((16 * arr.n_q4 + 9 * arr.n_t3 + 4 * arr.n_bar, 6, 6))
    for e in range(arr.n_q4):
        p = arr.xyz[arr.q4[e]].ravel()
        Ke, _, _ = O.q4_ke_global(p, arr.q4prop[e], arr.k6rot)
        for a in range(4):
            for b in range(4):
                eb[e * 16 + a * 4 + b] = Ke[6 * a:6 * a + 6, 6 * b:6 * b + 6]
    off = 16 * arr.n_q4
    for e in range(arr.n_t3):
        Ke, _, _ = O.t3_ke_global(arr.xyz[arr.t3[e]].ravel(), arr.t3prop[e], arr.k6rot)
        for a in range(3):
            for b in range(3):
                eb[off + e * 9 + a * 3 + b] = Ke[6 * a:6 * a + 6, 6 * b:6 * b + 6]
    off += 9 * arr.n_t3
    for i in range(arr.n_bar):
        n1, n2 = arr.bar[i]
        Kb = O.bar_ke_global(arr.xyz[n1], arr.xyz[n2], arr.bar_vref[i], arr.barprop[i])
        for a in range(2):
            for c in range(2):
                eb[off + i * 4 + a * 2 + c] = Kb[6 * a:6 * a + 6, 6 * c:6 * c + 6]
    return eb


def bsr_csr(S, vals, ndof):
    nb = len(S["col"])
    brow = np.repeat(np.arange(len(S["rp"]) - 1), np.diff(S["rp"]))
    I = (6 * brow[:, None] + np.arange(36)[None, :] // 6).ravel()
    J = (6 * S["col"][:, None].astype(np.int64) + np.arange(36)[None, :] % 6).ravel()
    return sp.csr_matrix((vals.reshape(nb * 36), (I, J)), shape=(ndof, ndof))


def emulate_factor(S, kvals):
    F = np.zeros(S["fs"])
    F[S["kdst"]] = kvals[S["ksrc"]]
    info = S["info"]
    nf = len(info)
    children = [[] for _ in range(nf)]
    for f in range(nf):
        if info[f, 4] >= 0:
            children[info[f, 4]].append(f)
    order = sorted(range(nf), key=lambda f: (info[f, 5], f))
    for f in order:
        ns, nu, m, piv0, par, h, off, roff, reloff = info[f]
        Fp = F[off:off + m * m].reshape(m, m, order="F")  # view, column-major
        for c in children[f]:
            nsc, nuc, mc, offc, relc = info[c, 0], info[c, 1], info[c, 2], info[c, 6], info[c, 8]
            Fc = F[offc:offc + mc * mc].reshape(mc, mc, order="F")
            rel = S["rel"][relc:relc + nuc]
            U = np.tril(Fc[nsc:, nsc:])
            for j in range(nuc):
                Fp[rel[j:], rel[j]] += U[j:, j]
        A11 = np.tril(Fp[:ns, :ns])
        A11 = A11 + np.tril(A11, -1).T
        L11 = np.linalg.cholesky(A11)
        Fp[:ns, :ns] = np.tril(L11)
        if nu:
            L21 = np.linalg.solve(L11, Fp[ns:, :ns].T).T
            Fp[ns:, :ns] = L21
            Fp[ns:, ns:] -= np.tril(L21 @ L21.T)
    return F


def assemble_L(S, F):
    n = S["nfree"]
    Ld = np.zeros((n, n))
    for f in range(len(S["info"])):
        ns, nu, m, piv0, par, h, off, roff, reloff = S["info"][f]
        Fp = F[off:off + m * m].reshape(m, m, order="F")
        rows = S["rows"][roff:roff + m]
        Ld[np.ix_(rows, np.arange(piv0, piv0 + ns))] = np.tril(Fp[:, :ns], 0) if ns == m else np.vstack(
            [np.tril(Fp[:ns, :ns]), Fp[ns:, :ns]])
    return Ld


def emulate_solve(S, F, b):
    """front-vector forward/backward solve as the device kernels do it."""
    info = S["info"]
    nf = len(info)
    children = [[] for _ in range(nf)]
    for f in range(nf):
        if info[f, 4] >= 0:
            children[info[f, 4]].append(f)
    x = b.copy()
    Vs = {}
    for f in sorted(range(nf), key=lambda f: (info[f, 5], f)):
        ns, nu, m, piv0, par, h, off, roff, reloff = info[f]
        Fp = F[off:off + m * m].reshape(m, m, order="F")
        V = np.zeros(m)
        V[:ns] = x[piv0:piv0 + ns]
        for c in children[f]:
            nsc, nuc, relc = info[c, 0], info[c, 1], info[c, 8]
            V[S["rel"][relc:relc + nuc]] += Vs[c][nsc:]
        y = np.linalg.solve(np.tril(Fp[:ns, :ns]), V[:ns])
        V[:ns] = y
        V[ns:] -= Fp[ns:, :ns] @ y
        Vs[f] = V
        x[piv0:piv0 + ns] = y
    for f in sorted(range(nf), key=lambda f: (-info[f, 5], f)):
        ns, nu, m, piv0, par, h, off, roff, reloff = info[f]
        Fp = F[off:off + m * m].reshape(m, m, order="F")
        rows = S["rows"][roff:roff + m]
        z = x[piv0:piv0 + ns] - Fp[ns:, :ns].T @ x[rows[ns:]]
        x[piv0:piv0 + ns] = np.linalg.solve(np.tril(Fp[:ns, :ns]).T, z)
    return x


@pytest.mark.parametrize("deck", ["plate", "panel", "mixed_panel"])
def test_gather_and_multifrontal_emulation(probe, deck):
    if deck == "plate":
        txt = decks.plate_deck(nx=7, ny=5, a=350.0, b=250.0)
    elif deck == "mixed_panel":
        txt = decks.panel_deck(nx=12, ny=8, a=240.0, b=160.0, stringer_every=4, nd=4, tri_frac=0.35)
    else:
        txt = decks.panel_deck(nx=12, ny=8, a=240.0, b=160.0, stringer_every=4, nd=4)
    arr = M.build_arrays(bdf.bdf_to_model_text(txt))
    K_or = O.assemble_K(arr)
    free_mask = ~O.autospc(arr, K_or.diagonal())
    S = analyze(probe, arr, free_mask, leaf=6)
    # 1. deterministic block gather reproduces the oracle's COO sum
    eb = element_blocks(arr)
    kv = np.zeros((len(S["col"]), 6, 6))
    for blk in range(len(S["col"])):
        for k in range(S["cptr"][blk], S["cptr"][blk + 1]):
            kv[blk] += eb[S["cidx"][k]]
    K = bsr_csr(S, kv, arr.ndof)
    scale = abs(K_or).max()
    assert abs(K - K_or).max() <= 1e-13 * scale
    # every element block is used exactly once
    assert np.array_equal(np.sort(S["cidx"]), np.arange(len(eb)))
    # 2. elimination numbering covers the free set
    eod = S["eod"]
    free = np.flatnonzero(free_mask)
    assert np.array_equal(np.sort(eod[free]), np.arange(S["nfree"]))
    assert np.all(eod[~free_mask] == -1)
    # 3. multifrontal emulation: L L^T equals the permuted K_ff
    F = emulate_factor(S, kv.reshape(-1))
    Ld = assemble_L(S, F)
    perm = np.empty(S["nfree"], np.int64)
    perm[eod[free]] = free
    Kp = K_or[perm][:, perm].toarray()
    assert np.abs(Ld @ Ld.T - Kp).max() <= 1e-10 * scale
    # 4. front-vector solves reproduce K_ff^-1 F
    b = arr.F[perm]
    x = emulate_solve(S, F, b)
    u_ref, _ = O.static_solve(arr, K_or)
    assert np.abs(x - u_ref[perm]).max() <= 1e-9 * np.abs(u_ref).max()
    # 5. k-map destinations are inside their fronts' lower triangles
    info = S["info"]
    for f in range(len(info)):
        ns, nu, m, off = info[f, 0], info[f, 1], info[f, 2], info[f, 6]
        sel = (S["kdst"] >= off) & (S["kdst"] < off + m * m)
        loc = S["kdst"][sel] - off
        r, c = loc % m, loc // m
        assert np.all(r >= c) and np.all(c < ns)


RT, KS = 64, 128  # product tile rows / k-slice (plan.h SOLVE_RT / SOLVE_KS)


def emulate_tiled_solve(S, F, b):
    """The GEMM-form solve exactly as the device tiles it: P = [L11^-1; L21 L11^-1],
    per-height tile lists, k-slice partial sums and the fixed-order finish."""
    info, plan = S["info"], S["plan"]
    nf = len(info)
    Pf, PTf = {}, {}
    for f in range(nf):
        ns, nu, m, off = info[f, 0], info[f, 1], info[f, 2], info[f, 6]
        Fp = F[off:off + m * m].reshape(m, m, order="F")
        Linv = np.tril(np.linalg.inv(np.tril(Fp[:ns, :ns])))
        P = np.vstack([Linv, Fp[ns:, :ns] @ Linv])
        Pf[f], PTf[f] = P, P.T.copy()
    children = [[] for _ in range(nf)]
    for f in range(nf):
        if info[f, 4] >= 0:
            children[info[f, 4]].append(f)
    rel_of = lambda c: S["rel"][info[c, 8]:info[c, 8] + info[c, 1]]
    x = b.copy()
    V = {}
    H = plan["hinfo"]
    for h in range(len(H)):
        lo, n, fo, bo, nfw, nbw = H[h, :6]
        fl = plan["flist"][lo:lo + n]
        for f in fl:
            ns, m, piv0 = info[f, 0], info[f, 2], info[f, 3]
            v = np.zeros(m)
            v[:ns] = x[piv0:piv0 + ns]
            for c in children[f]:
                v[rel_of(c)] += V[c][info[c, 0]:]
            V[f] = v
        part = np.zeros((nfw, RT))
        for tt in range(nfw):
            f, rt, ks = plan["ft"][3 * (fo + tt):3 * (fo + tt) + 3]
            ns, m = info[f, 0], info[f, 2]
            r0, r1 = rt * RT, min(m, rt * RT + RT)
            k0, k1 = ks * KS, min(ns, ks * KS + KS)
            part[tt, :r1 - r0] = Pf[f][r0:r1, k0:k1] @ V[f][k0:k1]
        for f in fl:
            ns, m, piv0 = info[f, 0], info[f, 2], info[f, 3]
            nk, base = plan["fnk"][f], plan["ftb"][f]
            for i in range(m):
                s = sum(part[base + (i // RT) * nk + k, i % RT] for k in range(nk))
                if i < ns:
                    x[piv0 + i] = s
                else:
                    V[f][i] -= s
    for h in reversed(range(len(H))):
        lo, n, fo, bo, nfw, nbw = H[h, :6]
        fl = plan["flist"][lo:lo + n]
        U = {}
        for f in fl:
            ns, m, roff = info[f, 0], info[f, 2], info[f, 7]
            rows = S["rows"][roff:roff + m]
            u = x[rows].copy()
            u[ns:] *= -1.0
            U[f] = u
        part = np.zeros((nbw, RT))
        for tt in range(nbw):
            f, rt, ks = plan["bt"][3 * (bo + tt):3 * (bo + tt) + 3]
            ns, m = info[f, 0], info[f, 2]
            r0, r1 = rt * RT, min(ns, rt * RT + RT)
            k0, k1 = ks * KS, min(m, ks * KS + KS)
            part[tt, :r1 - r0] = PTf[f][r0:r1, k0:k1] @ U[f][k0:k1]
        for f in fl:
            ns, piv0 = info[f, 0], info[f, 3]
            nk, base = plan["bnk"][f], plan["btb"][f]
            for i in range(ns):
                x[piv0 + i] = sum(part[base + (i // RT) * nk + k, i % RT] for k in range(nk))
    return x


def test_tiled_gemm_solve_emulation(probe):
    txt = decks.panel_deck(nx=14, ny=10, a=280.0, b=200.0, stringer_every=5, nd=4)
    arr = M.build_arrays(bdf.bdf_to_model_text(txt))
    K_or = O.assemble_K(arr)
    free_mask = ~O.autospc(arr, K_or.diagonal())
    S = analyze(probe, arr, free_mask, leaf=4)
    plan = S["plan"]
    # tile bookkeeping: every (front, row tile, k slice) appears exactly once
    info = S["info"]
    want_f = sum(-(-info[f, 2] // RT) * -(-info[f, 0] // KS) for f in range(len(info)))
    want_b = sum(-(-info[f, 0] // RT) * -(-info[f, 2] // KS) for f in range(len(info)))
    assert len(plan["ft"]) // 3 == want_f and len(plan["bt"]) // 3 == want_b
    assert plan["max_height_tiles"] == max(plan["hinfo"][:, 4].max(), plan["hinfo"][:, 5].max())
    eb = element_blocks(arr)
    kv = np.zeros((len(S["col"]), 6, 6))
    for blk in range(len(S["col"])):
        for k in range(S["cptr"][blk], S["cptr"][blk + 1]):
            kv[blk] += eb[S["cidx"][k]]
    F = emulate_factor(S, kv.reshape(-1))
    eod = S["eod"]
    free = np.flatnonzero(free_mask)
    perm = np.empty(S["nfree"], np.int64)
    perm[eod[free]] = free
    x = emulate_tiled_solve(S, F, arr.F[perm])
    u_ref, _ = O.static_solve(arr, K_or)
    assert np.abs(x - u_ref[perm]).max() <= 1e-9 * np.abs(u_ref).max()


@pytest.mark.parametrize("deck", ["plate", "panel"])
def test_forward_pull_map_equals_extend_add(probe, deck):
    """The forward solve assembles each front vector by pulling child update
    rows through a precomputed map; it must equal the push form
    V_f[rel_c[j]] += w_c[j] with children in their stored order, bit for bit."""
    if deck == "plate":
        arr = M.build_arrays(bdf.bdf_to_model_text(decks.plate_deck(nx=12, ny=10, nd=2)))
    else:
        arr = M.build_arrays(bdf.bdf_to_model_text(decks.panel_deck(nx=24, ny=12, stringer_every=4, nd=2)))
    free = (~arr.spc_mask.astype(bool)).astype(np.uint8)
    S = analyze(probe, arr, free, leaf=8)
    info = S["info"]
    ns, nu, m, parent = info[:, 0], info[:, 1], info[:, 2], info[:, 4]
    vec_off = np.concatenate([[0], np.cumsum(m)])[:-1]
    nslot = int(m.sum())
    pp, ps = S["plan"]["pull_ptr"], S["plan"]["pull_src"]
    assert len(pp) == nslot + 1 and pp[-1] == nu[parent >= 0].sum()
    rng = np.random.default_rng(5)
    W = rng.standard_normal(nslot)
    push = np.zeros(nslot)
    children = [[] for _ in range(len(m))]
    for c in range(len(m)):
        if parent[c] >= 0:
            children[parent[c]].append(c)
    for f in range(len(m)):
        for c in children[f]:
            rel = S["rel"][info[c, 8]:info[c, 8] + nu[c]]
            for j in range(nu[c]):
                push[vec_off[f] + rel[j]] += W[vec_off[c] + ns[c] + j]
    pull = np.zeros(nslot)
    for t in range(nslot):
        v = 0.0
        for q in range(pp[t], pp[t + 1]):
            c, row = ps[2 * q], ps[2 * q + 1]
            v += W[vec_off[c] + row]
        pull[t] = v
    assert np.array_equal(push, pull)
