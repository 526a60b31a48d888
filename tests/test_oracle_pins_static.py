"""Pins of the oracle's static-solve path (boundary_conditions.jl:734-881):
the singular-pivot refix and the factorisation AUTOSPC, on matrices whose
Cholesky pivots are known in closed form, and on a real mesh with
mechanisms, through invariants that do not depend on which DOFs are fixed.

A coupled pair [[k, -k(1-d)], [-k(1-d), k]] (large diagonals, so the
diagonal AUTOSPC keeps both DOFs) has second pivot k - k(1-d)^2 = k d (2 - d):
pivot ratio d (2 - d), so d = 1e-8 is refixed and d = 1e-7 is not
(ratio 2e-7).  With d = 0 the pivot is exactly zero, the Cholesky fails and
the shifted probe K + s I gives L^2 = (2ks + s^2)/(k + s), L ~ sqrt(2 s),
below 3 sqrt(s): the mechanism DOF is fixed.
"""
import types

import numpy as np
import pytest
import scipy.sparse as sp

from jf_2026_05_19_amd import bdf, decks, model as M
from oracle import oracle as O


def _case(pairs_d, k=1e6, extra=()):
    """Two-DOF blocks along the diagonal, 6 DOFs per node; every other DOF
    stiff and uncoupled.  pairs_d: {first_dof: d}."""
    n = 12
    K = np.diag(np.full(n, 5e5))
    for i, d in pairs_d.items():
        K[i, i] = K[i + 1, i + 1] = k
        K[i, i + 1] = K[i + 1, i] = -k * (1 - d)
    for blk in extra:
        K[np.ix_(blk, blk)] = k
    F = np.zeros(n)
    F[0] = 1.0
    arr = types.SimpleNamespace(spc_mask=np.zeros(n, np.uint8), autospc=True, ndof=n,
                                node_has_shell=np.ones(2, np.uint8), node_has_bar=np.zeros(2, np.uint8), F=F)
    return arr, sp.csr_matrix(K)


def test_refix_threshold_on_pivot_ratio():
    arr, K = _case({1: 1e-8, 7: 1e-7})
    info = {}
    u, free = O.static_solve_cholesky(arr, K, 1e6, info)
    assert info["path"] == "refix" and info["refix"] == [2]      # ratio 2e-8 fixed, 2e-7 kept
    assert 2 not in free and 8 in free
    Kff = K.toarray()[np.ix_(free, free)]
    assert np.abs(Kff @ u[free] - arr.F[free]).max() <= 1e-12 * np.abs(arr.F).max() * np.abs(Kff).max()


def test_failed_cholesky_takes_the_shifted_probes():
    arr, K = _case({3: 0.0})
    info = {}
    u, free = O.static_solve_cholesky(arr, K, 1e6, info)
    assert info["path"] == "factorization_autospc" and info["mechanism"] == [4]
    assert 4 not in free and np.all(np.isfinite(u))


def test_spread_mechanism_missed_by_the_probes_fixed_by_its_zero_pivot():
    """A 12-DOF chain k (path Laplacian) has the null vector ones(12): in
    natural order its last pivot is exactly zero, so the Cholesky fails; the
    probe K + sI gives that pivot s ||v||^2 / v_last^2 = 12 s, L = 3.46
    sqrt(s) > 3 sqrt(s), so the probe flags nothing; the zero pivot of the
    failed factorisation is then fixed."""
    n, k = 12, 1e6
    K = np.zeros((n, n))
    for i in range(n - 1):
        K[np.ix_([i, i + 1], [i, i + 1])] += k * np.array([[1.0, -1.0], [-1.0, 1.0]])
    F = np.zeros(n); F[0], F[5] = 1.0, -1.0
    arr = types.SimpleNamespace(spc_mask=np.zeros(n, np.uint8), autospc=True, ndof=n,
                                node_has_shell=np.ones(2, np.uint8), node_has_bar=np.zeros(2, np.uint8), F=F)
    info = {}
    u, free = O.static_solve_cholesky(arr, sp.csr_matrix(K), k, info)
    assert info["path"] == "zero_pivots" and info["zero_pivots"] == [11] and info["mechanism"] == []
    assert np.abs(K[np.ix_(free, free)] @ u[free] - F[free]).max() <= 1e-12 * k * np.abs(u).max() + 1e-12


def test_factorization_autospc_skips_when_most_dofs_flagged():
    """Rank-one 3x3 blocks k ones(3, 3): two of three pivots per block are
    exactly zero, so the probe flags two thirds of the free set, more than
    half, and fixes nothing; the zero pivots of the failed factorisation are
    then fixed (DOFs 1, 2 of each block), leaving k on the rest."""
    arr, K = _case({}, extra=[[0, 1, 2], [3, 4, 5], [6, 7, 8], [9, 10, 11]])
    info = {}
    u, free = O.static_solve_cholesky(arr, K, 1e6, info)
    assert info["path"] == "zero_pivots" and info["mechanism"] == []
    assert list(free) == [0, 3, 6, 9] and abs(u[0] - 1e-6) <= 1e-18


def test_indefinite_stiffness_is_singular():
    """[[k, 2k], [2k, k]] has eigenvalues 3k and -k with positive diagonals:
    AUTOSPC keeps both DOFs, every probe fails (the negative eigenvalue is far
    beyond the shifts), and the -3k pivot is not a zero pivot."""
    arr, K = _case({})
    K = K.toarray()
    K[np.ix_([1, 2], [1, 2])] = 1e6 * np.array([[1.0, 2.0], [2.0, 1.0]])
    with pytest.raises(O.SingularModel):
        O.static_solve_cholesky(arr, sp.csr_matrix(K), 1e6)


# which branch a mesh takes depends on the rounding of its mechanism pivots;
# these two decks take one each
@pytest.mark.parametrize("nx,ny,a,b,t,path", [(8, 6, 400.0, 300.0, 1.5, "refix"),
                                              (12, 10, 480.0, 400.0, 1.4, "zero_pivots")])
def test_tri_mesh_drilling_mechanisms_fixed_by_the_static_path(nx, ny, a, b, t, path):
    """An all-CTRIA3 plate has two zero-energy theta_z patterns.  The static
    path must fix exactly as many DOFs as K_ff's nullity (counted here from
    its spectrum), leave K_ff positive definite, and give the translations
    and buckling loads of any other choice of DOFs that removes the
    mechanisms (they are pure theta_z, unloaded, and carry no geometric
    stiffness)."""
    arr = M.build_arrays(bdf.bdf_to_model_text(decks.plate_deck(nx=nx, ny=ny, a=a, b=b, t=t, nd=3, tri_frac=1.0)))
    K, mes = O.assemble_K(arr, with_max_elem_stiff=True)
    fixed0 = O.autospc(arr, K.diagonal())
    free0 = np.flatnonzero(~fixed0)
    Kff = K.toarray()[np.ix_(free0, free0)]
    ev, vec = np.linalg.eigh(0.5 * (Kff + Kff.T))
    null = ev < 1e-10 * ev[-1]
    assert null.sum() == 2
    info = {}
    u, free = O.static_solve_cholesky(arr, K, mes, info)
    assert info["path"] == path and len(free0) - len(free) == 2, info
    Kf = K.toarray()[np.ix_(free, free)]
    assert np.linalg.eigvalsh(0.5 * (Kf + Kf.T))[0] > 0
    # another removal of the mechanisms: the two theta_z DOFs on which the
    # null vectors are most independent, away from the ones the path fixed
    N = vec[:, null]
    drop = set(range(len(free0))) - {int(np.flatnonzero(free0 == d)[0]) for d in set(free0) - set(free)}
    best, pick = -1.0, None
    cand = [i for i in drop if free0[i] % 6 == 5]
    for a in cand[:40]:
        for b in cand[-40:]:
            if a < b:
                s = abs(np.linalg.det(N[[a, b], :]))
                if s > best:
                    best, pick = s, (a, b)
    free_alt = np.delete(free0, list(pick))
    Ka = K.toarray()[np.ix_(free_alt, free_alt)]
    u_alt = np.zeros(arr.ndof)
    u_alt[free_alt] = np.linalg.solve(Ka, arr.F[free_alt])
    tr = np.arange(arr.ndof) % 6 < 3
    assert np.abs(u[tr] - u_alt[tr]).max() <= 1e-9 * np.abs(u[tr]).max()
    KG = O.assemble_KG(arr, u)
    lam = O.buckling(arr, K, KG, free)["eigenvalues"]
    lam_alt = O.buckling(arr, K, O.assemble_KG(arr, u_alt), free_alt)["eigenvalues"]
    assert len(lam) == len(lam_alt) and np.abs(lam - lam_alt).max() <= 1e-9 * np.abs(lam).max()
