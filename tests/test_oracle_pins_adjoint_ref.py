"""Pins of oracle/adjoint_ref.py, the restatement of the reference's default
analytic thickness sensitivity (the reference-exact device mode is checked
against it on the GPU).

Each expectation comes from the mathematics of the element or from an
independent implementation, not from the restatement's own formulas:

* Legacy CQUAD4 (stiffness_quad4_generic, FEMKernels.jl:1473-1658): symmetric
  with exactly the six rigid-body motions in its null space; a constant
  membrane strain has energy A eps' Cm eps; the constant-curvature Kirchhoff
  field on a rectangle has energy A k' Cb k (the MITC4 tying shears vanish
  for it); a linear w with zero rotations has the MITC4 shear energy
  phi^2 (5/6) G h |grad w|^2 A on a parallelogram, with
  phi^2 = min(1, 10 h^2 / A) (A = 4 |det J_c|), both below and at the
  saturation.
* d(Ke)/dh: central differences of the legacy element in h on both sides of
  the phi^2 saturation, and invariance of the global dKe under a rotation of
  the element frame in its plane.
* Membrane modes: the four-mode map and the principal-transverse KG equal
  the C oracle's independent routines (or_q4_condensation_map,
  or_q4_kg_local); with six modes a constant strain on a parallelogram
  needs no mode, and the r s mode relaxes a bilinear ux = c x y entirely
  (amplitude -c a b / 4 on a rectangle, zero stress) where the four modes
  keep pure bending exact; more modes never stiffen:
  u'(Km - Kc6)u <= u'(Km - Kc4)u <= u'Km u.
* Adjoint loads: the CQUAD4 unit-basis load equals phi' KG(sigma(e_j)) phi
  through the C oracle's KG; it is quadratic in phi; the CTRIA3 and CBAR
  loads are linear in u, so u . g = lambda phi' KG(u) phi.
"""
import numpy as np
import pytest

from oracle import adjoint_ref as R
from oracle import oracle as O

E, NU = 70000.0, 0.33
G = E / (2 * (1 + NU))


def _rect(a=40.0, b=25.0):
    return np.array([[-a / 2, -b / 2], [a / 2, -b / 2], [a / 2, b / 2], [-a / 2, b / 2]])


def _para(a=40.0, b=25.0, sk=9.0):
    return np.array([[-a / 2 - sk / 2, -b / 2], [a / 2 - sk / 2, -b / 2], [a / 2 + sk / 2, b / 2],
                     [-a / 2 + sk / 2, b / 2]])


def _quad():
    return np.array([[-21.0, -12.5], [19.0, -13.5], [23.0, 14.0], [-17.5, 11.0]])


def _area(lc):
    x, y = lc[:, 0], lc[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def _field(lc, f):
    """24-vector of nodal values f(x, y) -> (u, v, w, thx, thy, thz)."""
    u = np.zeros(24)
    for k in range(4):
        u[6 * k:6 * k + 6] = f(lc[k, 0], lc[k, 1])
    return u


def _rbm(lc):
    modes = []
    for t in range(3):
        modes.append(_field(lc, lambda x, y, t=t: np.eye(6)[t]))
    modes.append(_field(lc, lambda x, y: np.array([0, 0, y, 1.0, 0, 0])))     # rotation about x
    modes.append(_field(lc, lambda x, y: np.array([0, 0, -x, 0, 1.0, 0])))    # rotation about y
    modes.append(_field(lc, lambda x, y: np.array([-y, x, 0, 0, 0, 1.0])))    # rotation about z
    return np.array(modes).T


@pytest.mark.parametrize("lc", [_rect(), _para(), _quad()])
def test_legacy_symmetric_with_rigid_null_space(lc):
    K = R.legacy_ke(lc, E, NU, 1.5)
    assert np.allclose(K, K.T, rtol=0, atol=1e-10 * np.abs(K).max())
    assert np.abs(K @ _rbm(lc)).max() <= 1e-9 * np.abs(K).max()
    s = np.linalg.svd(K, compute_uv=False)
    assert s[17] > 1e-8 * s[0] and s[18] <= 1e-12 * s[0]


@pytest.mark.parametrize("lc", [_rect(), _para()])
def test_legacy_membrane_patch(lc):
    h = 1.5
    ex, ey, gxy = 1e-3, -4e-4, 7e-4
    u = _field(lc, lambda x, y: np.array([ex * x + 0.5 * gxy * y, ey * y + 0.5 * gxy * x, 0, 0, 0, 0]))
    Cm = E * h / (1 - NU ** 2) * R.iso_D(NU)
    eps = np.array([ex, ey, gxy])
    assert np.isclose(u @ R.legacy_ke(lc, E, NU, h) @ u, _area(lc) * eps @ Cm @ eps, rtol=1e-11)


def test_legacy_constant_curvature_on_rectangle():
    lc, h = _rect(), 1.5
    kx, ky, kxy = 2e-4, -1e-4, 3e-4
    # w = -(kx x^2 + ky y^2 + kxy x y)/2, thy = -dw/dx... with kx = d thy/dx,
    # ky = -d thx/dy, kxy = d thy/dy - d thx/dx (Bb rows) and gamma = dw + th = 0
    def f(x, y):
        thy = kx * x + 0.5 * kxy * y
        thx = -(ky * y + 0.5 * kxy * x)
        w = -(0.5 * kx * x * x + 0.5 * ky * y * y + 0.5 * kxy * x * y)
        return np.array([0, 0, w, thx, thy, 0])
    u = _field(lc, f)
    Cb = E * h ** 3 / (12 * (1 - NU ** 2)) * R.iso_D(NU)
    k = np.array([kx, ky, kxy])
    assert np.isclose(u @ R.legacy_ke(lc, E, NU, h) @ u, _area(lc) * k @ Cb @ k, rtol=1e-11)


@pytest.mark.parametrize("h", [1.0, 12.0])    # phi^2 < 1 and saturated
@pytest.mark.parametrize("lc", [_rect(), _para()])
def test_legacy_mitc4_shear_with_phi2(lc, h):
    a, b = 3e-3, -2e-3
    u = _field(lc, lambda x, y: np.array([0, 0, a * x + b * y, 0, 0, 0]))
    A = _area(lc)
    phi2 = min(1.0, 10.0 * h * h / A)
    expect = phi2 * (5.0 / 6.0) * G * h * (a * a + b * b) * A
    assert np.isclose(u @ R.legacy_ke(lc, E, NU, h) @ u, expect, rtol=1e-11)
    assert (phi2 < 1.0) == (h == 1.0)


@pytest.mark.parametrize("h", [1.0, 1.5, 12.0])
@pytest.mark.parametrize("lc", [_rect(), _para(), _quad()])
def test_legacy_dke_dh_is_the_derivative(lc, h):
    d = 1e-5 * h
    fd = (R.legacy_ke(lc, E, NU, h + d) - R.legacy_ke(lc, E, NU, h - d)) / (2 * d)
    an = R.legacy_dke_dh(lc, E, NU, h)
    assert np.abs(fd - an).max() <= 1e-7 * np.abs(an).max()


def test_legacy_dke_frame_invariant():
    lc, h = _quad(), 1.5
    beta = 0.37
    c, s = np.cos(beta), np.sin(beta)
    Q = np.array([[c, s], [-s, c]])       # new in-plane axes
    lc2 = lc @ Q.T
    R3 = np.eye(3); R3[:2, :2] = Q
    T = np.zeros((24, 24))
    for k in range(8):
        T[3 * k:3 * k + 3, 3 * k:3 * k + 3] = R3
    K1 = R.legacy_dke_dh(lc, E, NU, h)
    K2 = T.T @ R.legacy_dke_dh(lc2, E, NU, h) @ T
    assert np.abs(K1 - K2).max() <= 1e-10 * np.abs(K1).max()


@pytest.mark.parametrize("lc", [_rect(), _para(), _quad()])
def test_four_mode_map_and_kg_match_the_c_oracle(lc):
    h = 1.5
    Cm = E * h / (1 - NU ** 2) * R.iso_D(NU)
    A4 = R.membrane_map(lc, Cm, 4)
    Ac = O.q4_condensation_map(lc, Cm)
    assert np.abs(A4 - Ac).max() <= 1e-12 * np.abs(Ac).max()
    rng = np.random.default_rng(3)
    for _ in range(3):
        sig = rng.normal(size=(4, 3)) * 10.0
        K1 = R.kg_local(lc, sig, h, A4)
        K2 = O.q4_kg_local(lc, sig, h, Ac)
        assert np.abs(K1 - K2).max() <= 1e-11 * np.abs(K2).max()


def test_six_modes_constant_strain_on_parallelogram():
    lc, h = _para(), 1.5
    Cm = E * h / (1 - NU ** 2) * R.iso_D(NU)
    eps = np.array([1e-3, -4e-4, 7e-4])
    u = _field(lc, lambda x, y: np.array([eps[0] * x + eps[2] * y, eps[1] * y, 0, 0, 0, 0]))
    A6 = R.membrane_map(lc, Cm, 6)
    assert np.abs(A6 @ u).max() <= 1e-12 * np.abs(A6).max() * np.abs(u).max()
    sig = R.sigma_gp(lc, u, E, NU, h, 6)
    assert np.allclose(sig, np.tile(Cm @ eps / h, (4, 1)), rtol=1e-12, atol=1e-9)


def test_pure_bending_four_modes_exact_six_modes_relax_it():
    """On a rectangle the four Wilson modes reproduce in-plane pure bending
    (sigma = [E k y, 0, 0]); the r s mode of the six-mode set spans the
    bilinear ux = c x y = c (a b / 4) r s, so it relaxes that field entirely:
    amplitude -c a b / 4 and zero stress at every Gauss point."""
    a, b, h, k = 40.0, 25.0, 1.5, 1e-5
    lc = _rect(a, b)
    u = _field(lc, lambda x, y: np.array([k * x * y, -0.5 * k * (x * x + NU * y * y), 0, 0, 0, 0]))
    sig4 = R.sigma_gp(lc, u, E, NU, h, 4)
    for g, (r, s) in enumerate(R.GPS):
        assert np.allclose(sig4[g], [E * k * s * b / 2, 0.0, 0.0], rtol=1e-10, atol=1e-10 * E * k * b)
    c = 3e-6
    u = _field(lc, lambda x, y: np.array([c * x * y, 0, 0, 0, 0, 0]))
    Cm = E * h / (1 - NU ** 2) * R.iso_D(NU)
    amp = R.membrane_map(lc, Cm, 6) @ u
    assert np.allclose(amp, [0, 0, 0, 0, -c * a * b / 4, 0], rtol=1e-10, atol=1e-12 * c * a * b)
    assert np.abs(R.sigma_gp(lc, u, E, NU, h, 6)).max() <= 1e-9 * E * c * b


@pytest.mark.parametrize("lc", [_rect(), _para(), _quad()])
def test_more_modes_never_stiffen(lc):
    h = 1.5
    Cm = E * h / (1 - NU ** 2) * R.iso_D(NU)
    Km = np.zeros((24, 24))
    Kab = {4: np.zeros((24, 4)), 6: np.zeros((24, 6))}
    for r, s in R.GPS:
        det, iJ, dx, dy, _ = R._gp(lc, r, s)
        adj = abs(det)
        Bm = R.Bm_mat(dx, dy)
        Km += Bm.T @ Cm @ Bm * adj
        for n in (4, 6):
            Kab[n] += Bm.T @ Cm @ R.mode_B(r, s, iJ, n) * adj
    rng = np.random.default_rng(5)
    for _ in range(20):
        u = rng.normal(size=24)
        e0 = u @ Km @ u
        e4 = e0 + u @ Kab[4] @ (R.membrane_map(lc, Cm, 4) @ u)
        e6 = e0 + u @ Kab[6] @ (R.membrane_map(lc, Cm, 6) @ u)
        assert e6 <= e4 * (1 + 1e-12) and e4 <= e0 * (1 + 1e-12)
        assert e6 >= -1e-9 * e0


def test_unit_basis_load_matches_c_oracle_kg():
    lc, h = _quad(), 1.5
    Cm = E * h / (1 - NU ** 2) * R.iso_D(NU)
    Ac = O.q4_condensation_map(lc, Cm)
    rng = np.random.default_rng(11)
    phi = rng.normal(size=24)
    f = R.q4_rhs_unit_basis(lc, E, NU, h, phi, 4)
    for j in range(24):
        e = np.zeros(24); e[j] = 1.0
        sig = O.q4_membrane_force_field(lc, e, Cm) / h
        ref = phi @ O.q4_kg_local(lc, sig, h, Ac) @ phi
        assert np.isclose(f[j], ref, rtol=1e-11, atol=1e-11 * np.abs(f).max())
    assert np.allclose(R.q4_rhs_unit_basis(lc, E, NU, h, 2 * phi, 6), 4 * R.q4_rhs_unit_basis(lc, E, NU, h, phi, 6),
                       rtol=1e-12, atol=1e-12 * np.abs(f).max())


def test_triangle_and_bar_loads_are_linear_in_u():
    from jf_2026_05_19_amd import bdf, decks, model as M
    arr = M.build_arrays(bdf.bdf_to_model_text(
        decks.panel_deck(nx=6, ny=4, a=120.0, b=80.0, stringer_every=2, nd=3, tri_frac=1.0)))
    assert arr.n_q4 == 0 and arr.n_t3 > 0 and arr.n_bar > 0
    rng = np.random.default_rng(2)
    phi = rng.normal(size=arr.ndof)
    u = rng.normal(size=arr.ndof)
    g = R.adjoint_load(arr, 1.0, phi, 4)
    KG = O.assemble_KG(arr, u)
    assert np.isclose(u @ g, phi @ (KG @ phi), rtol=1e-10)
