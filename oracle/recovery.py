"""CPU restatement of the reference's SOL 101 result recovery -- TEST
INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this module, and only as the checker; the product path never
imports it.  Plain numpy, one element at a time, in the order the reference
states the arithmetic (default flags: ``:diag`` CQUAD4 frame, no SNORM, no
curvature-membrane term, no center-Jacobian incompatible modes, isotropic
PSHELL, PBARL TYPE=BAR, no pressure loads):

* residual and reaction rows: ``R = K u - F``, rows whose 6-vector norm
  exceeds 1e-20, and the reaction sums (``JFEM/src/solver/solve_case.jl:
  866``, ``:126-176``);
* CQUAD4 centroid forces and fibre stresses/strains
  (``stress_strain_quad4``, ``JFEM/src/FEMKernels.jl:6003-6195``): Wilson
  membrane modes condensed with ``alpha = -Kbb^-1 Kab' u`` at |det J|-weighted
  2x2 Gauss points, membrane strain and curvature averaged over the Gauss
  points with the same weights, ``N = Cm eps``, ``M = -br D kappa h^3/12``;
* CQUAD4 transverse shear: the MITC4 centre resultant with
  ``phi2 = min(1, 10 h^2 / (4 |det Jc|))``, ``G = E/2(1+nu)``, ``ts/t = 5/6``
  (``quad4_mitc4_center_shear_resultant``, ``FEMKernels.jl:5927-6001``);
* CQUAD4 corner forces by 2x2 Gauss extrapolation
  (``quad4_bilinear_corner_forces``, ``FEMKernels.jl:6197-6338``,
  ``interp_2x2_gauss_sigma``, ``:5210-5227``), the equilibrium shear
  ``(dMx/dx, dMy/dy)`` from the corner moments replacing each shear component
  it undercuts in magnitude, and the twisting moment held constant over the
  corners (``JFEM/src/solver/stress_recovery.jl:89-133``, ``:263-311``);
* CTRIA3 constant-strain membrane and curvature, the macro-element shear (the
  three virtual CQUAD4 plate blocks with ``Cb = D h^3/12``,
  ``Cs = 5/6 G h``, interior DOFs from the condensation, area-weighted MITC4
  shears) and the macro moment blend with weight
  ``clamp((shared edges - 1)/2, 0, 1)`` (``FEMKernels.jl:7213-7311``,
  ``:7391-7432``; ``stress_recovery.jl:20-87``, ``:315-370``);
* CBAR end forces ``k u`` in the bar frame and the stresses at the four
  recovery points C, D, E, F with the BAR section's sign -1
  (``stress_recovery.jl:432-513``, ``FEMKernels.jl:215-229``).

Outputs are per-element rows in the device layout (``Q4_COLS``,
``T3_COLS``, ``BAR_COLS``) so the parity tests compare arrays.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

PT = 1.0 / np.sqrt(3.0)
GAUSS = ((-PT, -PT), (PT, -PT), (PT, PT), (-PT, PT))
CORNERS = ((-1.0, -1.0), (1.0, -1.0), (1.0, 1.0), (-1.0, 1.0))

# row layouts: N(3) M(3) Q(2) sigma_z1(3) sigma_z2(3) eps_z1(3) eps_z2(3)
# [+ corner N (4x3) + corner M (4x3) for CQUAD4]
Q4_COLS = 44
T3_COLS = 20
# axial shear_1 shear_2 torque moment_a1 moment_a2 moment_b1 moment_b2,
# axial stress, end A stresses C D E F, end B stresses C D E F, valid
BAR_COLS = 18


def _dmat(E, nu):
    return (E / (1.0 - nu * nu)) * np.array([[1.0, nu, 0.0], [nu, 1.0, 0.0], [0.0, 0.0, (1.0 - nu) / 2.0]])


def _shape_derivs(r, s):
    dNr = np.array([-0.25 * (1 - s), 0.25 * (1 - s), 0.25 * (1 + s), -0.25 * (1 + s)])
    dNs = np.array([-0.25 * (1 - r), -0.25 * (1 + r), 0.25 * (1 + r), 0.25 * (1 - r)])
    return dNr, dNs


def _jac(lc, r, s):
    dNr, dNs = _shape_derivs(r, s)
    return np.array([dNr, dNs]) @ lc, dNr, dNs       # [[J11, J12], [J21, J22]]


def _bm(dxy):
    B = np.zeros((3, 24))
    for k in range(4):
        i = 6 * k
        B[0, i] = dxy[0, k]; B[1, i + 1] = dxy[1, k]
        B[2, i] = dxy[1, k]; B[2, i + 1] = dxy[0, k]
    return B


def _bb(dxy):
    B = np.zeros((3, 24))
    for k in range(4):
        i = 6 * k
        B[0, i + 4] = dxy[0, k]
        B[1, i + 3] = -dxy[1, k]
        B[2, i + 4] = dxy[1, k]
        B[2, i + 3] = -dxy[0, k]
    return B


def _bi(iJ, r, s):
    """Wilson membrane modes (1 - xi^2), (1 - eta^2) on ux, uy; Gauss-point
    Jacobian.  FEMKernels.jl:1831-1860."""
    p1x, p1y = iJ[0, 0] * (-2 * r), iJ[1, 0] * (-2 * r)
    p2x, p2y = iJ[0, 1] * (-2 * s), iJ[1, 1] * (-2 * s)
    return np.array([[p1x, 0.0, p2x, 0.0], [0.0, p1y, 0.0, p2y], [p1y, p1x, p2y, p2x]])


def q4_mitc4_shear(lc, ul, G, h, ts_t=5.0 / 6.0, phi2_alpha=10.0):
    """MITC4 centre shear resultant [Qx, Qy].  FEMKernels.jl:5927-6001."""
    tying = ((0.0, -1.0), (0.0, 1.0), (-1.0, 0.0), (1.0, 0.0))
    Bs = np.zeros((4, 24))
    for t, (xi, eta) in enumerate(tying):
        J, dNr, dNs = _jac(lc, xi, eta)
        N = 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])
        for k in range(4):
            i = 6 * k
            if t < 2:
                Bs[t, i + 2] = dNr[k]; Bs[t, i + 3] = -J[0, 1] * N[k]; Bs[t, i + 4] = J[0, 0] * N[k]
            else:
                Bs[t, i + 2] = dNs[k]; Bs[t, i + 3] = -J[1, 1] * N[k]; Bs[t, i + 4] = J[1, 0] * N[k]
    Jc, _, _ = _jac(lc, 0.0, 0.0)
    det = Jc[0, 0] * Jc[1, 1] - Jc[0, 1] * Jc[1, 0]
    adet = max(abs(det), 1e-30)
    i11, i12, i21, i22 = Jc[1, 1] / det, -Jc[0, 1] / det, -Jc[1, 0] / det, Jc[0, 0] / det
    gcov = np.array([0.5 * (Bs[0] + Bs[1]), 0.5 * (Bs[2] + Bs[3])]) @ ul
    phi2 = 1.0
    if phi2_alpha > 0.0:
        phi2 = min(1.0, phi2_alpha * h * h / max(4.0 * adet, 1e-30))
    k = ts_t * G * h
    t11, t12, t21, t22 = k * i11, k * i12, k * i21, k * i22
    C = np.array([[i11 * t11 + i21 * t21, i11 * t12 + i21 * t22], [i12 * t11 + i22 * t21, i12 * t12 + i22 * t22]])
    Qc = phi2 * C @ gcov
    return np.array([Jc[0, 0] * Qc[0] + Jc[1, 0] * Qc[1], Jc[0, 1] * Qc[0] + Jc[1, 1] * Qc[1]])


def _q4_gauss_fields(lc, ul, Cm):
    """Per Gauss point: (|det J| clamped, membrane strain with the condensed
    Wilson modes, curvature)."""
    geo = []
    Kab = np.zeros((24, 4)); Kbb = np.zeros((4, 4))
    for r, s in GAUSS:
        J, dNr, dNs = _jac(lc, r, s)
        adj = max(abs(np.linalg.det(J)), 1e-12)
        iJ = np.linalg.inv(J)
        dxy = iJ @ np.array([dNr, dNs])
        Bm, Bi = _bm(dxy), _bi(iJ, r, s)
        Kab += (Bm.T @ Cm @ Bi) * adj
        Kbb += (Bi.T @ Cm @ Bi) * adj
        geo.append((adj, Bm, _bb(dxy), Bi))
    alpha = -np.linalg.solve(Kbb, Kab.T @ ul)
    return [(adj, Bm @ ul + Bi @ alpha, Bb @ ul) for adj, Bm, Bb, Bi in geo]


def q4_stress_strain(lc, ul, E, nu, h, br):
    """(N, M, Q, s_z1, s_z2, e_z1, e_z2) at the centroid.
    FEMKernels.jl:6003-6195 (stress_strain_quad4, default keywords)."""
    D = _dmat(E, nu)
    Cm = D * h
    gp = _q4_gauss_fields(lc, ul, Cm)
    area = sum(a for a, _, _ in gp)
    eps = sum(a * e for a, e, _ in gp) / area
    kap = sum(a * k for a, _, k in gp) / area
    N = Cm @ eps
    M = -br * (D @ kap) * (h ** 3 / 12.0)
    G = E / (2 * (1 + nu))
    Q = np.zeros(2) if br <= 1e-12 else q4_mitc4_shear(lc, ul, G, h)
    e1, e2 = eps + (-h / 2.0) * kap, eps + (h / 2.0) * kap
    return N, M, Q, D @ e1, D @ e2, e1, e2


def q4_corner_forces(lc, ul, E, nu, h, br):
    """Corner N and M (4 x 3 each) extrapolated from the Gauss points.
    FEMKernels.jl:6197-6338, 5210-5227."""
    D = _dmat(E, nu)
    Cm, Cb = D * h, br * D * (h ** 3 / 12.0)
    gp = _q4_gauss_fields(lc, ul, Cm)
    Ngp = np.array([Cm @ e for _, e, _ in gp])
    Mgp = np.array([-(Cb @ k) for _, _, k in gp])
    Nc, Mc = np.zeros((4, 3)), np.zeros((4, 3))
    for i, (xi, eta) in enumerate(CORNERS):
        lx1, lx2 = (PT - xi) / (2 * PT), (xi + PT) / (2 * PT)
        ly1, ly2 = (PT - eta) / (2 * PT), (eta + PT) / (2 * PT)
        w = np.array([lx1 * ly1, lx2 * ly1, lx2 * ly2, lx1 * ly2])
        Nc[i], Mc[i] = w @ Ngp, w @ Mgp
    return Nc, Mc


def q4_equilibrium_shear(lc, Mc):
    """(dMx/dx, dMy/dy) from the corner moments, or None on a degenerate
    centre Jacobian.  stress_recovery.jl:89-119."""
    J, dNr, dNs = _jac(lc, 0.0, 0.0)
    det = J[0, 0] * J[1, 1] - J[0, 1] * J[1, 0]
    if abs(det) < 1e-12:
        return None
    iJ = np.array([[J[1, 1], -J[0, 1]], [-J[1, 0], J[0, 0]]]) / det
    gx = iJ @ np.array([dNr @ Mc[:, 0], dNs @ Mc[:, 0]])
    gy = iJ @ np.array([dNr @ Mc[:, 1], dNs @ Mc[:, 1]])
    return np.array([gx[0], gy[1]])


def q4_recover(lc, ul, E, nu, h, br):
    """One CQUAD4 row (Q4_COLS).  stress_recovery.jl:242-311."""
    N, M, Q, s1, s2, e1, e2 = q4_stress_strain(lc, ul, E, nu, h, br)
    Nc, Mc = q4_corner_forces(lc, ul, E, nu, h, br)
    Qo = Q.copy()
    if abs(br) > 1e-12:
        eq = q4_equilibrium_shear(lc, Mc)
        if eq is not None:
            for i in range(2):
                if np.isfinite(eq[i]) and abs(eq[i]) < abs(Qo[i]):
                    Qo[i] = eq[i]
    Mc[:, 2] = M[2]
    return np.concatenate([N, M, Qo, s1, s2, e1, e2, Nc.ravel(), Mc.ravel()])


# --------------------------------------------------------------------------
# CTRIA3
# --------------------------------------------------------------------------

MACRO_QUADS = ((0, 3, 6, 5), (1, 4, 6, 3), (2, 5, 6, 4))
PLATE_IDX = (2, 3, 4, 8, 9, 10, 14, 15, 16, 20, 21, 22)


def _macro_points(lc):
    p = np.zeros((7, 2))
    p[:3] = lc
    p[3] = (lc[0] + lc[1]) / 2; p[4] = (lc[1] + lc[2]) / 2; p[5] = (lc[2] + lc[0]) / 2
    p[6] = (lc[0] + lc[1] + lc[2]) / 3
    return p


def _virtual_quad_area(qc):
    a1 = (qc[1, 0] - qc[0, 0]) * (qc[3, 1] - qc[0, 1]) - (qc[3, 0] - qc[0, 0]) * (qc[1, 1] - qc[0, 1])
    a2 = (qc[2, 0] - qc[1, 0]) * (qc[3, 1] - qc[1, 1]) - (qc[3, 0] - qc[1, 0]) * (qc[2, 1] - qc[1, 1])
    return (abs(a1) + abs(a2)) / 2


def t3_macro_matrix(lc, E, nu, h):
    """21 x 21 macro plate matrix with the recovery constitutive set
    (Cb = D h^3/12, Cs = 5/6 G h, G = E/2(1+nu)).  FEMKernels.jl:7156-7209."""
    D = _dmat(E, nu)
    G = E / (2 * (1 + nu))
    gs = 5.0 / 6.0 * G * h
    _, K21 = O.t3_plate_macro(lc, D * h, D * (h ** 3 / 12.0), np.array([gs, 0.0, 0.0, gs]), h, G, 100.0)
    return K21


def t3_macro_field(lc, u_plate, E, nu, h):
    """(w, thx, thy) at the 7 macro points: the corners' values and the
    interior ones from the condensation -Kii^-1 Kie u."""
    K21 = t3_macro_matrix(lc, E, nu, h)
    Aint = -np.linalg.solve(K21[9:, 9:], K21[9:, :9])
    return np.concatenate([u_plate, Aint @ u_plate])


def t3_macro_average(lc, u_all, E, nu, h, br):
    """Area-weighted (Q, M) of the three virtual quads carrying the 7-point
    plate field u_all.  FEMKernels.jl:7244-7265, 7289-7310."""
    pts = _macro_points(lc)
    Qs, Ms, As = np.zeros(2), np.zeros(3), 0.0
    for quad in MACRO_QUADS:
        qc = pts[list(quad)]
        uq = np.zeros(24)
        for a, nid in enumerate(quad):
            uq[list(PLATE_IDX[3 * a:3 * a + 3])] = u_all[3 * nid:3 * nid + 3]
        _, Mq, Qq, _, _, _, _ = q4_stress_strain(qc, uq, E, nu, h, br)
        area = _virtual_quad_area(qc)
        Qs += area * Qq; Ms += area * Mq; As += area
    if As <= 1e-12:
        return np.zeros(2), np.zeros(3)
    return Qs / As, Ms / As


def t3_macro_resultants(lc, u_plate, E, nu, h, br):
    """Macro-element (Q, M) of a CTRIA3.  FEMKernels.jl:7213-7311."""
    if br <= 1e-12:
        return np.zeros(2), np.zeros(3)
    return t3_macro_average(lc, t3_macro_field(lc, u_plate, E, nu, h), E, nu, h, br)


def t3_recover(lc, ul, E, nu, h, br, weight):
    """One CTRIA3 row (T3_COLS).  FEMKernels.jl:7391-7432,
    stress_recovery.jl:52-87, 356-366."""
    x, y = lc[:, 0], lc[:, 1]
    A = 0.5 * abs(x[0] * (y[1] - y[2]) + x[1] * (y[2] - y[0]) + x[2] * (y[0] - y[1]))
    if A < 1e-12:
        return np.zeros(T3_COLS)
    b = np.array([y[1] - y[2], y[2] - y[0], y[0] - y[1]]) / (2 * A)
    c = np.array([x[2] - x[1], x[0] - x[2], x[1] - x[0]]) / (2 * A)
    D = _dmat(E, nu)
    u = ul[0::6]; v = ul[1::6]; tx = ul[3::6]; ty = ul[4::6]
    eps = np.array([b @ u, c @ v, c @ u + b @ v])
    kap = np.array([b @ ty, -(c @ tx), c @ ty - b @ tx])
    N = (D @ eps) * h
    M = -br * (D @ kap) * (h ** 3 / 12.0)
    u_plate = np.concatenate([ul[6 * k + 2:6 * k + 5] for k in range(3)])
    Q, Mmac = t3_macro_resultants(lc, u_plate, E, nu, h, br)
    e1, e2 = eps + (-h / 2.0) * kap, eps + (h / 2.0) * kap
    s1, s2 = D @ e1, D @ e2
    if weight > 0.0 and abs(br) > 1e-12 and np.all(np.isfinite(Mmac)):
        M = (1.0 - weight) * M + weight * Mmac
        kap = -(12.0 / (br * h ** 3)) * np.linalg.solve(D, M)
        em = (e1 + e2) / 2.0
        e1, e2 = em + (-h / 2.0) * kap, em + (h / 2.0) * kap
        s1, s2 = D @ e1, D @ e2
    return np.concatenate([N, M, Q, s1, s2, e1, e2])


def t3_macro_weights(q4, t3):
    """clamp((edges shared with another shell - 1) / 2, 0, 1) per CTRIA3.
    stress_recovery.jl:20-51."""
    count = {}
    for el in list(np.asarray(q4)) + list(np.asarray(t3)):
        n = len(el)
        for k in range(n):
            a, b = int(el[k]), int(el[(k + 1) % n])
            key = (min(a, b), max(a, b))
            count[key] = count.get(key, 0) + 1
    out = np.zeros(len(t3))
    for e, el in enumerate(np.asarray(t3)):
        shared = sum(count[(min(int(el[k]), int(el[(k + 1) % 3])), max(int(el[k]), int(el[(k + 1) % 3])))] > 1
                     for k in range(3))
        out[e] = min(max((shared - 1) / 2.0, 0.0), 1.0)
    return out


# --------------------------------------------------------------------------
# CBAR
# --------------------------------------------------------------------------

def bar_recover(p1, p2, vref, bp, points, ug1, ug2):
    """One CBAR row (BAR_COLS).  stress_recovery.jl:432-513 (no offsets,
    I12 = 0, sign -1 of TYPE=BAR), FEMKernels.jl:215-229."""
    row = np.zeros(BAR_COLS)
    R, L = O.bar_frame(p1, p2, vref)
    if L < 1e-9:
        return row
    A, I1, I2, J, K1, K2, E, G = (float(v) for v in bp)
    Iy, Iz = I2, I1
    if Iy == 0.0:
        Iy = Iz
    if Iz == 0.0:
        Iz = Iy
    ul = np.concatenate([R @ ug1[:3], R @ ug1[3:], R @ ug2[:3], R @ ug2[3:]])
    k = O.frame3d_ke(L, A, Iy, Iz, J, E, G, K1 * A if K1 > 0 else np.inf, K2 * A if K2 > 0 else np.inf)
    f = k @ ul
    forces = [f[6], -f[1], -f[2], -f[3], -f[5], f[4], f[11], -f[10]]
    row[:8] = forces
    row[8] = f[6] / A if abs(A) > 1e-30 else 0.0

    def surf(m1, m2, yj, zj):
        den = Iz * Iy
        s = (m1 * Iy * yj + m2 * Iz * zj) / den if abs(den) > 1e-30 else \
            m1 * yj / max(Iz, 1e-30) + m2 * zj / max(Iy, 1e-30)
        return -s
    pts = np.asarray(points, float).reshape(4, 2)
    for j in range(4):
        row[9 + j] = surf(forces[4], forces[5], pts[j, 0], pts[j, 1])
        row[13 + j] = surf(forces[6], forces[7], pts[j, 0], pts[j, 1])
    row[17] = 1.0
    return row


# --------------------------------------------------------------------------
# whole model
# --------------------------------------------------------------------------

def _local_u(R, u, nodes):
    return np.concatenate([np.concatenate([R @ u[6 * n:6 * n + 3], R @ u[6 * n + 3:6 * n + 6]]) for n in nodes])


def recover(arr, u, K=None, bar_points=None):
    """(residual R = K u - F, Q4 rows, T3 rows, bar rows) for the model
    arrays and the analysis-frame displacement u."""
    u = np.asarray(u, float)
    if K is None:
        K = O.assemble_K(arr)
    res = K @ u - np.asarray(arr.F, float)
    xyz = np.asarray(arr.xyz, float)
    q4rows = np.zeros((arr.n_q4, Q4_COLS))
    for e in range(arr.n_q4):
        nodes = [int(n) for n in arr.q4[e]]
        R, lc = O.q4_frame(xyz[nodes].ravel())
        E, nu, h, br = (float(v) for v in arr.q4prop[e][:4])
        q4rows[e] = q4_recover(lc, _local_u(R, u, nodes), E, nu, h, br)
    w = t3_macro_weights(arr.q4, arr.t3)
    t3rows = np.zeros((arr.n_t3, T3_COLS))
    for e in range(arr.n_t3):
        nodes = [int(n) for n in arr.t3[e]]
        R, lc = O.t3_frame(xyz[nodes].ravel())
        E, nu, h, br = (float(v) for v in arr.t3prop[e][:4])
        t3rows[e] = t3_recover(lc, _local_u(R, u, nodes), E, nu, h, br, w[e])
    pts = np.zeros((arr.n_bar, 8)) if bar_points is None else np.asarray(bar_points, float)
    barrows = np.zeros((arr.n_bar, BAR_COLS))
    for b in range(arr.n_bar):
        n1, n2 = (int(n) for n in arr.bar[b])
        barrows[b] = bar_recover(xyz[n1], xyz[n2], arr.bar_vref[b], arr.barprop[b], pts[b],
                                 u[6 * n1:6 * n1 + 6], u[6 * n2:6 * n2 + 6])
    return res, q4rows, t3rows, barrows
