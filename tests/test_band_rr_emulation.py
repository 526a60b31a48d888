"""CPU emulation of the banded Rayleigh-Ritz path (csrc/rr_eig.hip,
symeig_band_kernel): band-to-tridiagonal bulge chasing with the kernel's task
structure, its pipelined schedule (sweep i runs step ts - 3 i at time step
ts) and its reverse-sweep back-transformation.

What the mathematics fixes, on random banded symmetric matrices:
the reduced matrix is exactly tridiagonal, its eigenvalues equal numpy's
eigenvalues of the original (an orthogonal similarity), and the
back-transformed tridiagonal eigenvectors are orthonormal eigenvectors of the
original matrix.  The schedule's same-step tasks touch disjoint band
entries; a planted mistake in it (sweeps one step apart, run concurrently)
must break the similarity.
"""
import numpy as np
import pytest


def house(x):
    """H = I - tau v v^T, v[0] = 1, H x = mu e_0 (Golub & Van Loan Alg. 5.1.1)."""
    a = x[0]
    sig = float(np.dot(x[1:], x[1:]))
    v = np.zeros_like(x)
    v[0] = 1.0
    if sig == 0.0:
        return v, 0.0, a
    mu = np.sqrt(a * a + sig)
    v1 = a - mu if a <= 0 else -sig / (a + mu)
    tau = 2 * v1 * v1 / (sig + v1 * v1)
    v[1:] = x[1:] / v1
    return v, tau, mu


def nsteps(m, bw, i):
    return -(-(m - 1 - i) // bw)


def task_rows(m, bw, i, s):
    """Index range [lo, hi] of the rows and columns a task reads or writes."""
    lo = i if s == 0 else i + 1 + (s - 1) * bw
    return lo, min(i + (s + 1) * bw, m - 1)


def task_entries(m, bw, i, s):
    """Band entries (r, c), r >= c, that task (i, s) reads or writes: column i
    below the diagonal (s == 0) or the block below the previous reflector's
    rows, and the lower triangle of the diagonal block."""
    r0 = i + 1 + s * bw
    L = min(bw, m - r0)
    if s == 0:
        cols = [(r0 + p, i) for p in range(L)]
    else:
        cs = i + 1 + (s - 1) * bw
        cols = [(r0 + p, cs + q) for p in range(L) for q in range(min(bw, m - cs))]
    return set(cols) | {(r0 + p, r0 + q) for p in range(L) for q in range(p + 1)}


def band_reduce(T, bw, stagger=2, clashes=None, entry_clashes=None):
    """The kernel's schedule.  Tasks of one time step run concurrently on the
    device; here they run in decreasing sweep order (the reverse of the
    sequential order), so a schedule that lets them interact gives a wrong
    result.  clashes (a list) collects overlapping same-step task pairs."""
    A = np.tril(T).copy()
    m = len(A)
    refl, prev = {}, {}
    if m < 3:
        return np.diag(A).copy(), np.diag(A, -1).copy(), refl, A
    tend = stagger * (m - 3) + nsteps(m, bw, m - 3)
    for ts in range(tend):
        active = [i for i in range(m - 2) if 0 <= ts - stagger * i < nsteps(m, bw, i)]
        if clashes is not None:
            rng = [task_rows(m, bw, i, ts - stagger * i) for i in active]
            clashes += [(ts, a, b) for a in range(len(rng)) for b in range(a + 1, len(rng))
                        if rng[a][0] <= rng[b][1] and rng[b][0] <= rng[a][1]]
        if entry_clashes is not None:
            ent = [task_entries(m, bw, i, ts - stagger * i) for i in active]
            entry_clashes += [(ts, a, b) for a in range(len(ent)) for b in range(a + 1, len(ent))
                              if ent[a] & ent[b]]
        for i in reversed(active):
            s = ts - stagger * i
            if s == 0:
                J = np.arange(i + 1, min(i + bw, m - 1) + 1)
                v, tau, mu = house(A[J, i].copy())
                A[J, i] = 0.0
                A[i + 1, i] = mu
            else:
                vp, tp, Jp = prev[i]
                J = np.arange(i + 1 + s * bw, min(i + (s + 1) * bw, m - 1) + 1)
                C = A[np.ix_(J, Jp)]
                C = C - tp * np.outer(C @ vp, vp)
                v, tau, mu = house(C[:, 0].copy())
                C[:, 0] = 0.0
                C[0, 0] = mu
                C[:, 1:] = C[:, 1:] - tau * np.outer(v, v @ C[:, 1:])
                A[np.ix_(J, Jp)] = C
            S = A[np.ix_(J, J)]
            S = np.tril(S) + np.tril(S, -1).T
            p = tau * S @ v
            w = p - 0.5 * tau * (p @ v) * v
            A[np.ix_(J, J)] = np.tril(S - np.outer(v, w) - np.outer(w, v))
            prev[i] = (v, tau, J)
            refl[(i, s)] = (v, tau, J)
    return np.diag(A).copy(), np.diag(A, -1).copy(), refl, A


def back_transform(Z, refl, m, bw):
    Z = Z.copy()
    for i in range(m - 3, -1, -1):
        for s in range(nsteps(m, bw, i)):
            v, tau, J = refl[(i, s)]
            Z[J] -= tau * np.outer(v, v @ Z[J])
    return Z


def banded(m, bw, seed):
    X = np.random.default_rng(seed).standard_normal((m, m))
    T = np.tril(X) - np.tril(X, -bw - 1)
    return T + np.tril(T, -1).T


@pytest.mark.parametrize("m,bw", [(3, 4), (5, 4), (17, 16), (40, 4), (64, 8), (97, 16), (160, 16), (130, 8)])
def test_band_reduction_is_an_orthogonal_similarity(m, bw):
    T = banded(m, bw, m * 31 + bw)
    clashes = []
    d, e, refl, A = band_reduce(T, bw, entry_clashes=clashes)
    assert clashes == []                                            # same-step tasks touch disjoint entries
    assert np.abs(np.tril(A, -2)).max(initial=0.0) == 0.0          # exactly tridiagonal
    Tt = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w_ref = np.linalg.eigvalsh(T)
    w, U = np.linalg.eigh(Tt)
    scale = np.abs(w_ref).max()
    assert np.abs(w - w_ref).max() <= 1e-13 * scale
    Z = back_transform(U, refl, m, bw)
    assert np.abs(T @ Z - Z * w).max() <= 1e-13 * scale
    assert np.abs(Z.T @ Z - np.eye(m)).max() <= 1e-13
    # the number of reflector records matches the kernel's workspace formula
    F = lambda n: bw * (n // bw) * (n // bw + 1) // 2 + (n % bw) * (n // bw + 1)
    assert len(refl) == (F(m - 1) - F(1) if m >= 3 else 0)
    assert len(refl) * (bw + 1) <= m + (5 * m * m) // 8 + 17 * m + 64


def test_overlapping_schedule_is_detected():
    """Sweeps one step apart share band entries within a time step: the
    overlap is reported and, run in the reverse order, the result is no longer
    a tridiagonal similarity.  Two steps apart (the kernel's schedule) the
    row ranges still touch at one index, but the entries are disjoint (the
    shared index is a column of one task and the diagonal of the other), and
    the result is bitwise that of three steps apart."""
    m, bw = 64, 8
    T = banded(m, bw, 7)
    ent = []
    d, e, refl, A = band_reduce(T, bw, stagger=1, entry_clashes=ent)
    assert ent
    Tt = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    off = np.abs(np.tril(A, -2)).max()
    err = np.abs(np.linalg.eigvalsh(Tt) - np.linalg.eigvalsh(T)).max()
    assert off > 1e-8 or err > 1e-8
    rng, ent2 = [], []
    d2, e2, _, _ = band_reduce(T, bw, stagger=2, clashes=rng, entry_clashes=ent2)
    assert rng and not ent2
    d3, e3, _, _ = band_reduce(T, bw, stagger=3)
    assert np.array_equal(d2, d3) and np.array_equal(e2, e3)
