"""The buckling output selection: EIGRL range, cluster filter, mode count.

* The oracle's cluster filter (oracle.cluster_filter_skip, restating
  JFEM/src/solver/solve_case.jl:2995-3139) is pinned on hand-built spectra
  whose answers follow from the rule's statement: which jump qualifies, the
  singleton threshold at the first mode, the upper jump bound, the
  pre-jump spread bound, the "last" rule and every guard.
* The oracle's whole selection from Ritz values (oracle.select_modes) is
  pinned on spectra with a known answer (range cuts, the dense path's whole
  spectrum under a range, negative roots ahead of the positive branch).
* The device library's selection (jf_2026_05_19_amd/csrc/select.cpp,
  compiled for the host with g++) must choose the same pairs as the oracle
  on random spectra, and its decidedness rule must hold: whenever it says a
  truncated spectrum decides the output, appending any larger eigenvalues
  leaves the output unchanged.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native_host", "select_probe.cpp")
LIBDIR = os.path.join(HERE, "native_host", "_build")
LIB = os.path.join(LIBDIR, "libselectprobe.so")

skip = O.cluster_filter_skip


# --------------------------------------------------------------------------
# cluster filter pins
# --------------------------------------------------------------------------

def test_cluster_singleton_jump_needs_three():
    # first mode alone below a tight cluster: a jump of 3.5 >= 3 skips it
    assert skip([1.0, 3.5, 3.6, 3.7, 3.8], True, 1e9, 500) == 1
    # a jump of 2.9 at the first mode is below the singleton threshold, and
    # no later jump reaches 1.25
    assert skip([1.0, 2.9, 3.0, 3.1, 3.2], True, 1e9, 500) == 0
    # a jump of 2.0 after the first mode is enough from the second mode on
    assert skip([1.0, 2.0, 2.1, 2.2, 2.3], True, 1e9, 500) == 0
    assert skip([1.0, 1.1, 2.2, 2.3, 2.4], True, 1e9, 500) == 2


def test_cluster_multi_mode_gap():
    # three low modes, a 1.33x gap, then a cluster within 6 %
    assert skip([1.0, 1.1, 1.2, 1.6, 1.65, 1.7, 1.75], True, 1e9, 500) == 3
    # the same gap with the post-gap modes spread 40 %: no cluster
    assert skip([1.0, 1.1, 1.2, 1.6, 1.9, 2.24, 2.5], True, 1e9, 500) == 0
    # a gap just below 1.25: nothing
    assert skip([1.0, 1.1, 1.2, 1.49, 1.5, 1.51, 1.52], True, 1e9, 500) == 0


def test_cluster_last_and_first_rule():
    e = [1.0, 1.05, 1.4, 1.45, 1.5, 2.0, 2.05, 2.1, 2.15]
    assert skip(e, True, 1e9, 500) == 5                 # default: the last qualifying gap
    assert skip(e, True, 1e9, 500, last=False) == 2


def test_cluster_jump_upper_bound():
    assert skip([1.0, 9.0, 9.1, 9.2, 9.3], True, 1e9, 500) == 0     # 9x > 8x
    assert skip([1.0, 7.9, 8.0, 8.1, 8.2], True, 1e9, 500) == 1     # 7.9x is allowed


def test_cluster_pre_jump_spread_bound():
    # the gap 6.0 -> 7.9 follows modes spanning 6x (> 5x): rejected; the gap
    # 1.5 -> 6.0 has a cluster spread of 33 %: rejected too
    e = [1.0, 1.5, 6.0, 7.9, 8.0, 8.1]
    assert skip(e, True, 1e9, 500) == 0
    assert skip(e, True, 1e9, 500, max_pre_spread=10.0) == 3


def test_cluster_guards():
    e = [1.0, 3.5, 3.6, 3.7, 3.8]
    assert skip(e, False, 0.0, 500) == 0                 # no EIGRL range
    assert skip(e, True, 0.99, 500) == 0                 # V2 below 1
    assert skip(e, True, 1.0, 500) == 1
    assert skip(e, True, 1e9, 99) == 0                   # fewer than 100 shell cards
    assert skip(e, True, 1e9, 100) == 1
    assert skip(e[:4], True, 1e9, 500) == 0              # fewer than dense_n + 2 modes
    assert skip([-1.0] + e, True, 1e9, 500) == 0          # a non-positive mode


# --------------------------------------------------------------------------
# whole selection pins (oracle)
# --------------------------------------------------------------------------

def _lams(theta, idx):
    return [1.0 / theta[i] for i in idx]


def test_selection_range_cut_and_count():
    # lambda = 1 .. 12 (theta = 1/lambda), dense, ND = 3, range [2.5, 9]
    lam = np.arange(1.0, 13.0)
    th = 1.0 / lam
    idx, rec = O.select_modes(th, 3, 500, 2.5, 9.0, 10)
    assert _lams(th, idx) == [3.0, 4.0, 5.0] and rec["in_range"] == 7
    # no range: 3 ND modes by |lambda| (dense), ascending
    idx, _ = O.select_modes(th, 3, 500, 0.0, 0.0, 10)
    assert _lams(th, idx) == list(lam[:9])
    # the Krylov path keeps ND without a range
    idx, _ = O.select_modes(th, 3, 5000, 0.0, 0.0, 10)
    assert _lams(th, idx) == [1.0, 2.0, 3.0]


def test_selection_dense_range_keeps_whole_spectrum():
    # twenty load-reversal roots (negative lambda) of smaller |lambda| than
    # the positive branch: the dense path under a range still finds the
    # positive modes, the Krylov path's requested count (ceil(3 ND)) does not
    lam = np.concatenate([-np.linspace(0.5, 0.9, 20), [1.0, 1.2, 1.4, 1.6]])
    th = 1.0 / lam
    idx, _ = O.select_modes(th, 3, 500, 0.0, 1e9, 10)
    assert _lams(th, idx) == [1.0, 1.2, 1.4]
    idx, _ = O.select_modes(th, 3, 5000, 0.0, 1e9, 10)
    assert idx.size == 0
    # without a range the dense path keeps 3 ND by |lambda|: all negative
    idx, _ = O.select_modes(th, 3, 500, 0.0, 0.0, 10)
    assert idx.size == 0


def test_selection_range_lower_bound_clamped_and_localization():
    lam = np.array([1e-12, 0.5, 1.0, 2.0, 3.0, 4.0])
    th = 1.0 / lam
    # a negative V1 does not re-admit anything at or below 1e-10
    idx, _ = O.select_modes(th, 3, 500, -1e-4, 10.0, 10)
    assert _lams(th, idx) == [0.5, 1.0, 2.0]
    # 100 shell cards: a localized candidate (share 0.2) is dropped
    shares = np.zeros(len(th))
    shares[2] = 0.2
    idx, rec = O.select_modes(th, 3, 500, -1e-4, 10.0, 100, shares)
    assert _lams(th, idx) == [0.5, 2.0, 3.0] and rec["localization_dropped"] == 1
    # 99 shell cards: the filter does not run
    idx, _ = O.select_modes(th, 3, 500, -1e-4, 10.0, 99, shares)
    assert _lams(th, idx) == [0.5, 1.0, 2.0]


def test_request_count():
    assert O.request_count(3, 500) == 9 and O.request_count(3, 5000) == 3
    assert O.request_count(3, 5000, 0.0, 1e9) == 9 and O.request_count(3, 500, 0.0, 1e9) == 9
    assert O.request_count(3, 6) == 4                     # n_free - 2
    assert O.request_count(3, 500, 5.0, 1.0) == 9         # V2 <= V1: no range


# --------------------------------------------------------------------------
# device selection (host build of select.cpp) against the oracle
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def sel():
    os.makedirs(LIBDIR, exist_ok=True)
    csrc = os.path.join(HERE, "..", "jf_2026_05_19_amd", "csrc")
    tmp = LIB + ".tmp%d" % os.getpid()
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", SRC, os.path.join(csrc, "select.cpp"),
                           "-o", tmp])
    os.replace(tmp, LIB)
    L = ctypes.CDLL(LIB)
    vp = ctypes.c_void_p
    L.probe_select.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                               ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_double, vp, vp]
    L.probe_requested.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    return L


def _device_select(L, theta, nd, n_free, v1, v2, n_cards, shares, cover=np.inf, all_pairs=False):
    th = np.ascontiguousarray(theta, np.float64)
    sh = np.ascontiguousarray(shares, np.float64)
    out = np.zeros(max(len(th), 1), np.int32)
    info = np.zeros(4, np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    n = L.probe_select(nd, n_free, v1, v2, n_cards, int(all_pairs), p(th), len(th), p(sh), cover, p(out), p(info))
    return out[:n].copy(), info


def test_request_count_matches_oracle(sel):
    for nd in (1, 3, 7, 20):
        for nf in (3, 10, 64, 500, 4000, 4001, 90000):
            for v1, v2 in ((0.0, 0.0), (0.0, 1e9), (-1e-4, 50.0), (5.0, 1.0)):
                assert sel.probe_requested(nd, nf, v1, v2) == O.request_count(nd, nf, v1, v2), (nd, nf, v1, v2)


def _random_case(rng):
    """A spectrum as an eigensolve returns it: the pairs of largest |theta|,
    some negative, a few clustered, with a cluster gap now and then."""
    n_free = int(rng.choice([40, 300, 3000, 6000, 20000]))
    nd = int(rng.integers(1, 8))
    k = int(rng.integers(3, 40))
    base = np.sort(rng.uniform(0.5, 30.0, size=k))
    if rng.random() < 0.4:                                 # a spurious low band and a gap
        m = int(rng.integers(1, max(2, k // 3)))
        base[:m] = np.sort(rng.uniform(0.2, 0.4, size=m))
    sign = np.where(rng.random(k) < 0.25, -1.0, 1.0)
    lam = base * sign
    if rng.random() < 0.1:
        lam[rng.integers(k)] = 1e-8                       # below the positive tolerance
    theta = 1.0 / lam
    rng.shuffle(theta)
    r = rng.random()
    if r < 0.3:
        v1, v2 = 0.0, 0.0
    elif r < 0.6:
        v1, v2 = 0.0, 1e9
    else:
        v1 = float(rng.choice([-1e-4, 0.0, rng.uniform(0.3, 5.0)]))
        v2 = float(rng.uniform(v1 + 0.5, 40.0)) if rng.random() < 0.8 else 0.9
    n_cards = int(rng.choice([50, 100, 400]))
    shares = np.where(rng.random(k) < 0.15, 0.3, 0.02)
    return theta, nd, n_free, v1, v2, n_cards, shares


def test_device_selection_matches_oracle(sel):
    rng = np.random.default_rng(2024)
    n_range = n_skip = 0
    for case in range(3000):
        theta, nd, n_free, v1, v2, n_cards, shares = _random_case(rng)
        ref, rec = O.select_modes(theta, nd, n_free, v1, v2, n_cards, shares)
        got, info = _device_select(sel, theta, nd, n_free, v1, v2, n_cards, shares)
        assert list(got) == list(ref), (case, theta, nd, n_free, v1, v2, n_cards)
        assert info[2] == rec["cluster_skip"], case
        n_range += (v1 != 0.0 or v2 != 0.0) and v2 > v1
        n_skip += rec["cluster_skip"] > 0
    assert n_range > 1000 and n_skip > 50        # the random cases reach the range and cluster branches


def test_decided_truncations_keep_the_output(sel):
    """A spectrum cut at |lambda| < cover: whenever the device rule calls the
    cut decided, the full spectrum's selection (any eigenvalues above the
    cut appended) equals the cut's."""
    rng = np.random.default_rng(77)
    n_decided = n_open = 0
    for case in range(1500):
        theta, nd, _, v1, v2, n_cards, shares = _random_case(rng)
        n_free = 300                                       # dense path: the whole spectrum counts
        lam = 1.0 / theta
        order = np.argsort(np.abs(lam), kind="stable")
        cut = int(rng.integers(1, len(theta) + 1))
        if not ((v1 != 0.0 or v2 != 0.0) and v2 > v1):
            # without a range the eigensolve always returns the requested count
            cut = max(cut, min(O.request_count(nd, n_free), len(theta)))
        kept = order[:cut]
        cover = np.abs(lam[order[cut]]) if cut < len(theta) else np.inf
        got, info = _device_select(sel, theta[kept], nd, n_free, v1, v2, n_cards, shares[kept], cover)
        if not info[3]:
            n_open += 1
            continue
        n_decided += 1
        full, _ = O.select_modes(theta, nd, n_free, v1, v2, n_cards, shares)
        assert list(lam[kept][got]) == list(lam[full]), (case, cut)
    assert n_decided > 300 and n_open > 100


CLUSTER_SPECTRA = [
    [1.0, 3.5, 3.6, 3.7, 3.8], [1.0, 2.9, 3.0, 3.1, 3.2], [1.0, 2.0, 2.1, 2.2, 2.3], [1.0, 1.1, 2.2, 2.3, 2.4],
    [1.0, 1.1, 1.2, 1.6, 1.65, 1.7, 1.75], [1.0, 1.1, 1.2, 1.6, 1.9, 2.24, 2.5],
    [1.0, 1.1, 1.2, 1.49, 1.5, 1.51, 1.52], [1.0, 1.05, 1.4, 1.45, 1.5, 2.0, 2.05, 2.1, 2.15],
    [1.0, 9.0, 9.1, 9.2, 9.3], [1.0, 8.2, 8.3, 8.4, 8.5], [1.0, 7.9, 8.0, 8.1, 8.2], [1.0, 1.5, 6.0, 7.9, 8.0, 8.1],
]


@pytest.mark.parametrize("e", CLUSTER_SPECTRA)
def test_device_cluster_filter_on_pinned_spectra(sel, e):
    """The hand-built spectra of the cluster pins through the device rule
    (ND large enough to show every mode after the skip)."""
    th = 1.0 / np.array(e)
    got, info = _device_select(sel, th, 20, 300, 0.0, 1e9, 500, np.zeros(len(e)))
    n_skip = skip(e, True, 1e9, 500)
    assert info[2] == n_skip
    assert list(got) == list(range(n_skip, len(e)))
