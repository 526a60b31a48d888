"""Closed-form pins of the CQUAD4 oracle routines that the patch tests of
test_oracle_pins.py leave open: the MacNeal transverse-shear block, the
:diag element frame, the condensed Wilson modes (membrane recovery and the
condensation map), the principal-transverse geometric stiffness with the
augmented gradients, and the drilling penalty.

Each expectation is written from the mathematics of the element, not from
the oracle's code:

* MacNeal shear (FEMKernels.jl:4715-4978).  On a parallelogram every shear
  sample point has the same |J| = A/4, so on the uniform-shear subspace
  (gamma_xz, gamma_yz constant) the 4x4 compliance Z = Zs + Zb acts as
  2 [G^-1 / A + diag(zb_x, zb_y)] and the strain energy is
      u'Ku = A g' (G^-1 + diag(0.65 Dx^2/(12 D11), 0.65 Dy^2/(12 D22)))^-1 g,
  MacNeal's residual bending flexibility added to the shear compliance
  (Dx, Dy the projected side lengths in the element frame).  The
  differential (linearly varying) shear of w = c x y on a rectangle sees
  the anisotropy factors a = eps/(eps + (1-eps) Dx^2/Dy^2), b likewise:
      u'Ku = A px^2/(1/Gxx + 0.65 a Dx^2/(12 D11)) + A py^2/(1/Gyy + ...),
  px = c Ly/(2 sqrt 3), py = c Lx/(2 sqrt 3).  A wrong zb scale, eps,
  swapped Dx/Dy, a/b, flex_x/flex_y, Gxx/Gyy or a wrong cross-coupling
  factor fails one of these.
* :diag frame (helpers.jl:1134-1193): x along d13/|d13| - d24/|d24|,
  computed here by hand for a skewed parallelogram; the same quad's shear
  energy is then compared with the closed form in that frame, which
  differs from the edge-frame value, so the frame reaches Ke through Dx, Dy.
* Wilson modes (FEMKernels.jl:6341-6539, 6708-6785): on a parallelogram a
  constant strain needs no incompatible mode (A u = 0, N_gp = Cm eps); on a
  rectangle in-plane pure bending (ux = k x y, uy = -k (x^2 + nu y^2)/2) is
  reproduced exactly, so A u = [0, k Lx^2/8, 0, nu k Ly^2/8] (the parabolas
  left over by the bilinear interpolation) and N_gp = [E h k y_gp, 0, 0].
* Principal-transverse KG (FEMKernels.jl:6891-6972, 7554-7985): for
  uniform sigma and linear fields u = G x, w = g.x,
      u'KGu = h A [g' sigma g + sum_k lambda_k (n_k_perp' G n_k)^2],
  (lambda_k, n_k) the principal pairs (numpy eigh, no angle formula),
  n_perp = (s, -c); for equal biaxial stress the branch takes the x and y
  axes: h A s [G21^2 + G12^2 + |g|^2].  With the condensation map, the
  pure-bending field under sigma_xx gives h sigma_xx k^2 Lx^3 Ly / 12 from
  the augmented d(uy)/dx = -k x (zero without the map).
* Drilling (FEMKernels.jl:3557-3559, 3683-3686, 3780): a theta_z-only field
  has u'Ku = (k6rot/1e5) G h int theta_z^2 dA.
"""
import numpy as np
import pytest

from oracle import oracle as O

E, NU, T = 70000.0, 0.33, 2.0
G = E / (2 * (1 + NU))
TST = 5.0 / 6.0
PROP = np.array([E, NU, T, 1.0, TST, G])
ZB, EPS = 0.65, 0.04
PT = 1.0 / np.sqrt(3.0)


def _dmat(nu=NU):
    return np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])


def _iso(h=T):
    Cm = E * h / (1 - NU ** 2) * _dmat()
    Cb = E * h ** 3 / (12 * (1 - NU ** 2)) * _dmat()
    Cs = TST * h * G * np.eye(2)
    return Cm, Cb, Cs


def _shear_block(lc, Cm, Cb, Cs, h=T):
    """Ke(Cs) - Ke(Cs = 0): the MacNeal block alone (it is the only term
    that depends on Cs)."""
    return O.q4_ke_local(lc, Cm, Cb, Cs, h, E) - O.q4_ke_local(lc, Cm, Cb, np.zeros((2, 2)), h, E)


def _rect(Lx, Ly):
    return np.array([[-Lx / 2, -Ly / 2], [Lx / 2, -Ly / 2], [Lx / 2, Ly / 2], [-Lx / 2, Ly / 2]])


def _uniform_shear_energy(A, Gs, D11, D22, Dx, Dy, g):
    M = np.linalg.inv(Gs) + np.diag([ZB * Dx ** 2 / (12 * D11), ZB * Dy ** 2 / (12 * D22)])
    return A * g @ np.linalg.solve(M, g)


# ---------------------------------------------------------------- MacNeal shear

ANISO_CB = np.array([[3.1e5, 0.6e5, 0.0], [0.6e5, 1.2e5, 0.0], [0.0, 0.0, 0.5e5]])
ANISO_CS = np.array([[4200.0, 900.0], [900.0, 2600.0]])


@pytest.mark.parametrize("g", [(1e-3, 0.0), (0.0, 1e-3), (0.7e-3, -1.3e-3)])
def test_macneal_uniform_shear_closed_form(g):
    Lx, Ly = 60.0, 25.0
    lc = _rect(Lx, Ly)
    Cm, _, _ = _iso()
    Ks = _shear_block(lc, Cm, ANISO_CB, ANISO_CS)
    g = np.array(g)
    ref = _uniform_shear_energy(Lx * Ly, ANISO_CS, ANISO_CB[0, 0], ANISO_CB[1, 1], Lx, Ly, g)
    # the same uniform shear three ways: w linear, theta_y = gamma_x,
    # theta_x = -gamma_y (gamma_xz = w,x + theta_y, gamma_yz = w,y - theta_x)
    for kind in ("w", "theta", "mixed"):
        u = np.zeros((4, 6))
        for k, (x, y) in enumerate(lc):
            if kind == "w":
                u[k, 2] = g[0] * x + g[1] * y
            elif kind == "theta":
                u[k, 4] = g[0]
                u[k, 3] = -g[1]
            else:
                u[k, 2] = g[0] * x
                u[k, 3] = -g[1]
        U = u.ravel() @ Ks @ u.ravel()
        assert abs(U - ref) <= 1e-10 * ref, (kind, U, ref)


def test_macneal_differential_shear_closed_form():
    Lx, Ly, c = 60.0, 25.0, 2e-5
    lc = _rect(Lx, Ly)
    Cm, _, _ = _iso()
    Ks = _shear_block(lc, Cm, ANISO_CB, ANISO_CS)
    u = np.zeros((4, 6))
    for k, (x, y) in enumerate(lc):
        u[k, 2] = c * x * y
    U = u.ravel() @ Ks @ u.ravel()
    A = Lx * Ly
    a = EPS / (EPS + (1 - EPS) * Lx ** 2 / Ly ** 2)
    b = EPS / (EPS + (1 - EPS) * Ly ** 2 / Lx ** 2)
    px, py = c * Ly / (2 * np.sqrt(3)), c * Lx / (2 * np.sqrt(3))
    ref = (A * px ** 2 / (1 / ANISO_CS[0, 0] + ZB * a * Lx ** 2 / (12 * ANISO_CB[0, 0])) +
           A * py ** 2 / (1 / ANISO_CS[1, 1] + ZB * b * Ly ** 2 / (12 * ANISO_CB[1, 1])))
    assert abs(U - ref) <= 1e-10 * ref, (U, ref)
    # uniform and differential shear are Z-orthogonal: energies add
    g = np.array([0.4e-3, -0.9e-3])
    u2 = u.copy()
    for k, (x, y) in enumerate(lc):
        u2[k, 2] += g[0] * x + g[1] * y
    U2 = u2.ravel() @ Ks @ u2.ravel()
    ref2 = ref + _uniform_shear_energy(A, ANISO_CS, ANISO_CB[0, 0], ANISO_CB[1, 1], Lx, Ly, g)
    assert abs(U2 - ref2) <= 1e-10 * ref2


def test_macneal_block_rank_and_placement():
    """K_plate = D' Z^-1 D has rank 4 (four shear samples) and lives on the
    w, theta_x, theta_y DOFs only."""
    lc = np.array([[0.0, 0.0], [50.0, 8.0], [57.0, 41.0], [4.0, 30.0]])
    lc = lc - lc.mean(axis=0)
    Cm, Cb, Cs = _iso()
    Ks = _shear_block(lc, Cm, Cb, Cs)
    mask = np.zeros(24, bool)
    for k in range(4):
        mask[6 * k + 2:6 * k + 5] = True
    scale = np.abs(Ks).max()
    assert np.abs(Ks[~mask]).max() <= 1e-14 * scale and np.abs(Ks[:, ~mask]).max() <= 1e-14 * scale
    ev = np.linalg.eigvalsh(0.5 * (Ks + Ks.T))
    assert np.all(np.abs(ev[:-4]) <= 1e-10 * ev[-1]) and ev[-4] > 1e-6 * ev[-1]


# ---------------------------------------------------------------- :diag frame

PARA = np.array([[0.0, 0.0], [50.0, 8.0], [62.0, 38.0], [12.0, 30.0]])   # p1 + p3 = p2 + p4


def _diag_frame_by_hand(p):
    d13, d24 = p[2] - p[0], p[3] - p[1]
    x = d13 / np.linalg.norm(d13) - d24 / np.linalg.norm(d24)
    gxi = 0.25 * (-p[0] + p[1] + p[2] - p[3])
    geta = 0.25 * (-p[0] - p[1] + p[2] + p[3])
    n = np.cross(gxi, geta)
    n /= np.linalg.norm(n)
    x = x - (x @ n) * n
    v1 = x / np.linalg.norm(x)
    return v1, np.cross(n, v1), n


def test_diag_frame_hand_computed():
    p = np.c_[PARA, np.zeros(4)]
    v1, v2, v3 = _diag_frame_by_hand(p)
    assert np.allclose(v3, [0, 0, 1])
    # in the plane the axis bisects d13 and -d24: here, by hand,
    # d13 = (62, 38), d24 = (-38, 22)
    a = np.array([62.0, 38.0]) / np.hypot(62, 38) + np.array([38.0, -22.0]) / np.hypot(38, 22)
    assert np.allclose(v1[:2], a / np.linalg.norm(a), atol=1e-15)
    _, R, lc = O.q4_ke_global(p.ravel(), PROP)
    assert np.abs(R - np.array([v1, v2, v3])).max() <= 1e-14
    c = p.mean(axis=0)
    assert np.abs(lc - np.c_[(p - c) @ v1, (p - c) @ v2]).max() <= 1e-12
    # a rigid 3-D placement carries the frame along
    Q = np.linalg.qr(np.array([[0.3, -0.5, 0.8], [0.9, 0.1, -0.2], [0.1, 0.7, 0.4]]))[0]
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    pq = p @ Q.T + np.array([5.0, -3.0, 8.0])
    _, Rq, lcq = O.q4_ke_global(pq.ravel(), PROP)
    assert np.abs(Rq - R @ Q.T).max() <= 1e-13
    assert np.abs(lcq - lc).max() <= 1e-11


def test_diag_frame_reaches_ke_through_macneal_lengths():
    p = np.c_[PARA, np.zeros(4)]
    v1, v2, _ = _diag_frame_by_hand(p)
    c = p.mean(axis=0)
    lc = np.c_[(p - c) @ v1, (p - c) @ v2]
    Dx = 0.5 * (lc[1, 0] + lc[2, 0] - lc[0, 0] - lc[3, 0])
    Dy = 0.5 * (lc[2, 1] + lc[3, 1] - lc[0, 1] - lc[1, 1])
    A = 0.5 * abs(np.cross(p[2] - p[0], p[3] - p[1])[2])
    Dplate = E * T ** 3 / (12 * (1 - NU ** 2))
    Gs = TST * T * G * np.eye(2)
    Ks = O.q4_ke_global(p.ravel(), PROP)[0] - O.q4_ke_global(p.ravel(), np.r_[PROP[:4], 0.0, PROP[5]])[0]
    g_glob = np.array([1.1e-3, -0.3e-3])            # uniform shear (dw/dX, dw/dY)
    u = np.zeros((4, 6))
    u[:, 2] = (p[:, :2] - c[:2]) @ g_glob
    U = u.ravel() @ Ks @ u.ravel()
    g_loc = np.array([g_glob @ v1[:2], g_glob @ v2[:2]])
    ref = _uniform_shear_energy(A, Gs, Dplate, Dplate, Dx, Dy, g_loc)
    assert abs(U - ref) <= 1e-10 * ref, (U, ref)
    # the edge frame (x along p2 - p1) would give a different energy
    e1 = (p[1] - p[0]) / np.linalg.norm(p[1] - p[0])
    e2 = np.cross([0, 0, 1.0], e1)
    lce = np.c_[(p - c) @ e1, (p - c) @ e2]
    Dxe = 0.5 * (lce[1, 0] + lce[2, 0] - lce[0, 0] - lce[3, 0])
    Dye = 0.5 * (lce[2, 1] + lce[3, 1] - lce[0, 1] - lce[1, 1])
    alt = _uniform_shear_energy(A, Gs, Dplate, Dplate, Dxe, Dye, np.array([g_glob @ e1[:2], g_glob @ e2[:2]]))
    assert abs(alt - ref) > 1e-4 * ref


# ---------------------------------------------------------------- Wilson modes

def test_membrane_constant_strain_needs_no_incompatible_mode():
    p = np.c_[PARA, np.zeros(4)]
    _, _, lc = O.q4_ke_global(p.ravel(), PROP)
    Cm, _, _ = _iso()
    ex, ey, gxy = 1.1e-3, -0.6e-3, 0.8e-3
    u = np.zeros((4, 6))
    for k, (x, y) in enumerate(lc):
        u[k, 0] = ex * x + 0.5 * gxy * y
        u[k, 1] = 0.5 * gxy * x + ey * y
    A = O.q4_condensation_map(lc, Cm)
    assert np.abs(A @ u.ravel()).max() <= 1e-12 * np.abs(u).max()
    N = O.q4_membrane_force_field(lc, u.ravel(), Cm)
    ref = Cm @ np.array([ex, ey, gxy])
    assert np.abs(N - ref).max() <= 1e-10 * np.abs(ref).max()


def _pure_bending(lc, kap, axis):
    u = np.zeros((4, 6))
    for k, (x, y) in enumerate(lc):
        if axis == 0:      # bending about z with sigma_xx = E k y
            u[k, 0] = kap * x * y
            u[k, 1] = -0.5 * kap * (x * x + NU * y * y)
        else:              # sigma_yy = E k x
            u[k, 1] = kap * x * y
            u[k, 0] = -0.5 * kap * (y * y + NU * x * x)
    return u.ravel()


def test_wilson_modes_reproduce_inplane_pure_bending():
    Lx, Ly, kap = 48.0, 20.0, 3e-5
    lc = _rect(Lx, Ly)
    Cm, _, _ = _iso()
    A = O.q4_condensation_map(lc, Cm)
    gp = np.array([[-PT, -PT], [PT, -PT], [PT, PT], [-PT, PT]]) * np.array([Lx / 2, Ly / 2])
    # sigma_xx = E k y
    u = _pure_bending(lc, kap, 0)
    alpha = A @ u
    ref = np.array([0.0, kap * Lx ** 2 / 8, 0.0, NU * kap * Ly ** 2 / 8])
    assert np.abs(alpha - ref).max() <= 1e-12 * np.abs(ref).max(), alpha
    N = O.q4_membrane_force_field(lc, u, Cm)
    Nref = np.c_[E * T * kap * gp[:, 1], np.zeros(4), np.zeros(4)]
    assert np.abs(N - Nref).max() <= 1e-10 * np.abs(Nref).max(), N
    # sigma_yy = E k x
    u = _pure_bending(lc, kap, 1)
    alpha = A @ u
    ref = np.array([NU * kap * Lx ** 2 / 8, 0.0, kap * Ly ** 2 / 8, 0.0])
    assert np.abs(alpha - ref).max() <= 1e-12 * np.abs(ref).max(), alpha
    N = O.q4_membrane_force_field(lc, u, Cm)
    Nref = np.c_[np.zeros(4), E * T * kap * gp[:, 0], np.zeros(4)]
    assert np.abs(N - Nref).max() <= 1e-10 * np.abs(Nref).max(), N


def test_wilson_modes_sheared_parallelogram_bending():
    """x = a xi, y = b xi + c eta (vertical xi-edges, d eta/dx != 0).  Under
    sigma_yy = E k x the edges xi = +-1 are traction free and the exact field
    ux = -k (y^2 + nu x^2)/2, uy = k x y is quadratic in (xi, eta), so the
    condensed modes reproduce it: with f = c4 xi^2 + c5 eta^2 + bilinear the
    mode amplitudes are (-c4, -c5), i.e. A u = [k (b^2 + nu a^2)/2, -k a b,
    k c^2/2, 0]; N_gp = [0, E h k x_gp, 0]; and the augmented gradient gives
    u'KGu = h s k^2 int y^2 dA = h s k^2 (4 a c / 3)(b^2 + c^2)."""
    a, b, c, kap = 24.0, 9.0, 13.0, 3e-5
    lc = np.array([[-a, -b - c], [a, b - c], [a, b + c], [-a, -b + c]])
    Cm, _, _ = _iso()
    A = O.q4_condensation_map(lc, Cm)
    u = _pure_bending(lc, kap, 1)
    ref = np.array([kap * (b * b + NU * a * a) / 2, -kap * a * b, kap * c * c / 2, 0.0])
    assert np.abs(A @ u - ref).max() <= 1e-12 * np.abs(ref).max(), A @ u
    gp = np.array([[-PT, -PT], [PT, -PT], [PT, PT], [-PT, PT]])
    x_gp = a * gp[:, 0]
    N = O.q4_membrane_force_field(lc, u, Cm)
    Nref = np.c_[np.zeros(4), E * T * kap * x_gp, np.zeros(4)]
    assert np.abs(N - Nref).max() <= 1e-10 * np.abs(Nref).max(), N
    s0 = -2.5
    KG = O.q4_kg_local(lc, np.tile([0.0, s0, 0.0], (4, 1)), T, A)
    kg_ref = T * s0 * kap ** 2 * (4 * a * c / 3) * (b * b + c * c)
    assert abs(u @ KG @ u - kg_ref) <= 1e-10 * abs(kg_ref)


# ---------------------------------------------------------------- KG

def test_kg_augmented_gradients_pure_bending():
    Lx, Ly, kap, s0 = 48.0, 20.0, 3e-5, -3.0
    lc = _rect(Lx, Ly)
    Cm, _, _ = _iso()
    A = O.q4_condensation_map(lc, Cm)
    u = _pure_bending(lc, kap, 0)
    KG = O.q4_kg_local(lc, np.tile([s0, 0.0, 0.0], (4, 1)), T, A)
    ref = T * s0 * kap ** 2 * Lx ** 3 * Ly / 12.0          # h sigma int (d uy / dx)^2, d uy / dx = -k x
    assert abs(u @ KG @ u - ref) <= 1e-10 * abs(ref)
    KG0 = O.q4_kg_local(lc, np.tile([s0, 0.0, 0.0], (4, 1)), T, None)
    assert abs(u @ KG0 @ u) <= 1e-12 * abs(ref)          # bilinear uy is constant on this field
    u = _pure_bending(lc, kap, 1)
    KG = O.q4_kg_local(lc, np.tile([0.0, s0, 0.0], (4, 1)), T, A)
    ref = T * s0 * kap ** 2 * Lx * Ly ** 3 / 12.0
    assert abs(u @ KG @ u - ref) <= 1e-10 * abs(ref)


def _linear_field(lc, Gm, g):
    u = np.zeros((4, 6))
    for k, xy in enumerate(lc):
        u[k, :2] = Gm @ xy
        u[k, 2] = g @ xy
    return u.ravel()


@pytest.mark.parametrize("sig", [(-2.0, 0.7, 1.3), (1.5, -0.4, -2.2), (-1.0, -3.0, 0.25), (0.0, 0.0, 1.7)])
def test_kg_principal_decomposition_closed_form(sig):
    p = np.c_[PARA, np.zeros(4)]
    _, _, lc = O.q4_ke_global(p.ravel(), PROP)
    Cm, _, _ = _iso()
    A = O.q4_condensation_map(lc, Cm)
    area = 0.5 * abs(np.cross(p[2] - p[0], p[3] - p[1])[2])
    S = np.array([[sig[0], sig[2]], [sig[2], sig[1]]])
    lam, vec = np.linalg.eigh(S)
    KG = O.q4_kg_local(lc, np.tile(sig, (4, 1)), T, A)
    rng = np.random.default_rng(7)
    for _ in range(4):
        Gm = rng.standard_normal((2, 2)) * 1e-3
        g = rng.standard_normal(2) * 1e-3
        u = _linear_field(lc, Gm, g)
        assert np.abs(A @ u).max() <= 1e-12 * np.abs(u).max()  # linear fields: no incompatible mode
        inplane = sum(l * (np.array([n[1], -n[0]]) @ Gm @ n) ** 2 for l, n in zip(lam, vec.T))
        ref = T * area * (g @ S @ g + inplane)
        got = u @ KG @ u
        assert abs(got - ref) <= 1e-10 * (T * area * np.abs(lam).max() * (Gm ** 2).sum() + abs(ref))
    # out-of-plane part alone recombines to g' sigma g; a rigid in-plane
    # rotation w gets h A w^2 tr(sigma)
    u = _linear_field(lc, np.zeros((2, 2)), np.array([1e-3, 2e-3]))
    assert abs(u @ KG @ u - T * area * np.array([1e-3, 2e-3]) @ S @ np.array([1e-3, 2e-3])) <= 1e-10 * T * area * 1e-5
    w = 1e-3
    u = _linear_field(lc, np.array([[0.0, -w], [w, 0.0]]), np.zeros(2))
    assert abs(u @ KG @ u - T * area * w * w * np.trace(S)) <= 1e-10 * T * area * w * w * np.abs(lam).max()


def test_kg_equal_biaxial_takes_the_element_axes():
    p = np.c_[PARA, np.zeros(4)]
    _, _, lc = O.q4_ke_global(p.ravel(), PROP)
    area = 0.5 * abs(np.cross(p[2] - p[0], p[3] - p[1])[2])
    s0 = -1.7
    KG = O.q4_kg_local(lc, np.tile([s0, s0, 0.0], (4, 1)), T, None)
    Gm = np.array([[0.3e-3, -1.1e-3], [0.8e-3, 0.5e-3]])
    g = np.array([0.6e-3, -0.2e-3])
    u = _linear_field(lc, Gm, g)
    ref = T * area * s0 * (Gm[1, 0] ** 2 + Gm[0, 1] ** 2 + g @ g)
    assert abs(u @ KG @ u - ref) <= 1e-10 * abs(ref)


# ---------------------------------------------------------------- drilling

@pytest.mark.parametrize("k6rot", [100.0, 250.0])
def test_drilling_energy_closed_form(k6rot):
    pts2 = np.array([[0.0, 0.0], [52.0, 3.0], [47.0, 41.0], [-4.0, 36.0]])   # general flat quad
    p = np.c_[pts2, np.zeros(4)]
    Ke, _, _ = O.q4_ke_global(p.ravel(), PROP, k6rot=k6rot)
    th = np.array([0.7e-3, -0.2e-3, 1.1e-3, 0.4e-3])
    u = np.zeros((4, 6)); u[:, 5] = th
    U = u.ravel() @ Ke @ u.ravel()
    # int theta_z^2 dA of the bilinear field by 6x6 Gauss
    g, wts = np.polynomial.legendre.leggauss(6)
    xi_n = np.array([-1, 1, 1, -1.0]); eta_n = np.array([-1, -1, 1, 1.0])
    integ = 0.0
    for xi, wx in zip(g, wts):
        for eta, wy in zip(g, wts):
            N = 0.25 * (1 + xi * xi_n) * (1 + eta * eta_n)
            dxi = 0.25 * xi_n * (1 + eta * eta_n)
            deta = 0.25 * eta_n * (1 + xi * xi_n)
            J = np.array([[dxi @ pts2[:, 0], dxi @ pts2[:, 1]], [deta @ pts2[:, 0], deta @ pts2[:, 1]]])
            integ += wx * wy * abs(np.linalg.det(J)) * (N @ th) ** 2
    ref = k6rot / 1e5 * G * T * integ
    assert abs(U - ref) <= 1e-10 * ref, (U, ref)
