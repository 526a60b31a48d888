"""Oracle (test infrastructure only): the reference's own analytic SOL 105
thickness sensitivity for flat isotropic shells, restated in numpy.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may use this
module, as the checker.

For a ``shell_thickness`` design variable the reference computes, per mode
and PSHELL group p (JFEM/src/solver/adjoint_buckling.jl:2406-2637),

    d lambda / d T_p = -(phi' dK_p phi + lambda phi' dKG_p phi - psi' dK_p u) / phi' KG phi,
    K psi = lambda d(phi' KG phi)/du,

with these ingredients by default (no ENV overrides):

* dK_p: ForwardDiff in h of the legacy element ``stiffness_quad4_generic``
  (dKdx.jl:219-247, FEMKernels.jl:1473-1658): MITC4 transverse shear with
  the phi^2 = min(1, 10 h^2 / (4 |det J_c|)) factor, standard bending B,
  Wilson membrane modes, drilling (k6rot 100) -- not the MacNeal element of
  the forward K.  Every part of it scales with a power of h (membrane,
  drilling and the condensation with h, bending with h^3, shear with
  phi^2(h) h), so dK/dh is exact in closed form; at phi^2 = 1 the branch is
  the saturated one (derivative of the constant 1).
* The adjoint load of a flat isotropic CQUAD4 (``_flat_iso_quad4_rhs_local``,
  adjoint_buckling.jl:1024-1042): f_j = phi_l' KG(sigma(e_j)) phi_l for
  the 24 unit nodal displacements e_j -- a unit-basis evaluation of the
  principal-transverse KG, not its derivative.  The membrane stresses come
  from ``quad4_membrane_force_field`` at the Gauss points, and both the
  stresses and the KG gradients carry condensed membrane modes: the four
  Wilson modes when the model has line elements, else the six "enhanced"
  modes (Wilson plus the r s mode; ``use_enhanced_modes`` is set by the
  default JFEM_SOL105_EIG_FLAT_ISO_DKMQ=true when no CBAR/CBEAM/CROD is
  present, adjoint_buckling.jl:930-1005).
* dKG_p at fixed u: phi' KG phi / h of the forward KG when the design
  variable is one PSHELL that owns every shell and there are no bars
  (adjoint_buckling.jl:2059-2089); otherwise the sum over the group's
  elements of phi' KG_ctx phi / h with the context's KG (the modes above).
* CTRIA3 (the generic unit-stress branch, adjoint_buckling.jl:1887-1906) and
  CBAR adjoint loads are linear in u, so d/du = the unit-basis evaluation;
  they are computed here by unit vectors through the C oracle's element KG.

Local frames: the legacy dK is frame-invariant in the element's plane (the
pins check it), so the forward ``:diag`` frame is used throughout.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as O

PT = 1.0 / np.sqrt(3.0)
GPS = ((-PT, -PT), (PT, -PT), (PT, PT), (-PT, PT))
TYING = ((0.0, -1.0), (0.0, 1.0), (-1.0, 0.0), (1.0, 0.0))
PHI2_ALPHA = 10.0   # FEMKernels.jl:10


def shape_derivs(r, s):
    dNr = 0.25 * np.array([-(1 - s), (1 - s), (1 + s), -(1 + s)])
    dNs = 0.25 * np.array([-(1 - r), -(1 + r), (1 + r), (1 - r)])
    return dNr, dNs


def shape_vals(r, s):
    return 0.25 * np.array([(1 - r) * (1 - s), (1 + r) * (1 - s), (1 + r) * (1 + s), (1 - r) * (1 + s)])


def _jac(lc, r, s):
    dNr, dNs = shape_derivs(r, s)
    J11, J12 = dNr @ lc[:, 0], dNr @ lc[:, 1]
    J21, J22 = dNs @ lc[:, 0], dNs @ lc[:, 1]
    return dNr, dNs, J11, J12, J21, J22


def _gp(lc, r, s):
    """(det J, inverse Jacobian entries, dN/dx, dN/dy, N) at (r, s)."""
    dNr, dNs, J11, J12, J21, J22 = _jac(lc, r, s)
    det = J11 * J22 - J12 * J21
    iJ = (J22 / det, -J12 / det, -J21 / det, J11 / det)
    dx = iJ[0] * dNr + iJ[1] * dNs
    dy = iJ[2] * dNr + iJ[3] * dNs
    return det, iJ, dx, dy, shape_vals(r, s)


def iso_D(nu):
    return np.array([[1.0, nu, 0.0], [nu, 1.0, 0.0], [0.0, 0.0, (1.0 - nu) / 2.0]])


def Bm_mat(dx, dy):
    B = np.zeros((3, 24))
    for k in range(4):
        B[0, 6 * k] = dx[k]
        B[1, 6 * k + 1] = dy[k]
        B[2, 6 * k] = dy[k]
        B[2, 6 * k + 1] = dx[k]
    return B


def Bb_mat(dx, dy):
    """Standard bending B (kx = d thy/dx, ky = -d thx/dy, kxy = d thy/dy -
    d thx/dx), FEMKernels.jl:1597-1604."""
    B = np.zeros((3, 24))
    for k in range(4):
        B[0, 6 * k + 4] = dx[k]
        B[1, 6 * k + 3] = -dy[k]
        B[2, 6 * k + 4] = dy[k]
        B[2, 6 * k + 3] = -dx[k]
    return B


def Bd_vec(dx, dy, N):
    b = np.zeros(24)
    for k in range(4):
        b[6 * k] = 0.5 * dy[k]
        b[6 * k + 1] = -0.5 * dx[k]
        b[6 * k + 5] = N[k]
    return b


def mode_B(r, s, iJ, n_modes):
    """Membrane mode B (3 x n_modes): Wilson modes (1 - r^2), (1 - s^2) on ux
    and uy, plus r s on ux and uy for the six enhanced modes, with the Gauss
    point's own inverse Jacobian (FEMKernels.jl:1831-1896)."""
    iJ11, iJ12, iJ21, iJ22 = iJ
    p1x, p1y = iJ11 * (-2.0 * r), iJ21 * (-2.0 * r)
    p2x, p2y = iJ12 * (-2.0 * s), iJ22 * (-2.0 * s)
    B = np.zeros((3, n_modes))
    B[0, 0] = p1x; B[2, 0] = p1y
    B[1, 1] = p1y; B[2, 1] = p1x
    B[0, 2] = p2x; B[2, 2] = p2y
    B[1, 3] = p2y; B[2, 3] = p2x
    if n_modes == 6:
        qx, qy = iJ11 * s + iJ12 * r, iJ21 * s + iJ22 * r
        B[0, 4] = qx; B[2, 4] = qy
        B[1, 5] = qy; B[2, 5] = qx
    return B


# --------------------------------------------------------------------------
# legacy element (stiffness_quad4_generic) and its thickness derivative
# --------------------------------------------------------------------------

def mitc4_tying_B(lc):
    """Covariant transverse shear rows at the four tying points (0,-1),
    (0,1), (-1,0), (1,0): [gamma_r at the r-edges; gamma_s at the s-edges]
    (FEMKernels.jl:1514-1542)."""
    B = np.zeros((4, 24))
    for t, (r, s) in enumerate(TYING):
        dNr, dNs, J11, J12, J21, J22 = _jac(lc, r, s)
        N = shape_vals(r, s)
        for k in range(4):
            if t < 2:
                B[t, 6 * k + 2] = dNr[k]
                B[t, 6 * k + 3] = -J12 * N[k]
                B[t, 6 * k + 4] = J11 * N[k]
            else:
                B[t, 6 * k + 2] = dNs[k]
                B[t, 6 * k + 3] = -J22 * N[k]
                B[t, 6 * k + 4] = J21 * N[k]
    return B


def legacy_parts(lc, E, nu, h, k6rot=100.0, ts_t=5.0 / 6.0, bend_ratio=1.0):
    """Pieces of the legacy CQUAD4 stiffness at thickness h: membrane Km,
    bending Kb, drilling Kd, MITC4 shear without phi^2 (Ks), the Wilson-mode
    condensation Kc, phi^2 and whether it saturated."""
    D = iso_D(nu)
    Cm = E * h / (1 - nu * nu) * D
    Cb = bend_ratio * E * h ** 3 / (12 * (1 - nu * nu)) * D
    G = E / (2 * (1 + nu))
    Cs = ts_t * G * h * np.eye(2)
    Btp = mitc4_tying_B(lc)
    _, _, J11c, J12c, J21c, J22c = _jac(lc, 0.0, 0.0)
    detc = abs(J11c * J22c - J12c * J21c)
    trial = PHI2_ALPHA * h * h / max(4.0 * detc, 1e-30)
    phi2 = trial if trial < 1.0 else 1.0
    alpha = (k6rot / 1e5) * G * h
    Km = np.zeros((24, 24)); Kb = np.zeros((24, 24)); Kd = np.zeros((24, 24)); Ks = np.zeros((24, 24))
    Kab = np.zeros((24, 4)); Kbb = np.zeros((4, 4))
    for r, s in GPS:
        det, iJ, dx, dy, N = _gp(lc, r, s)
        adj = max(abs(det), 1e-12)
        Bm, Bb, bd = Bm_mat(dx, dy), Bb_mat(dx, dy), Bd_vec(dx, dy, N)
        Bi = mode_B(r, s, iJ, 4)
        Km += Bm.T @ Cm @ Bm * adj
        Kb += Bb.T @ Cb @ Bb * adj
        Kd += np.outer(bd, bd) * (adj * alpha)
        Bs = np.vstack([0.5 * (1 - s) * Btp[0] + 0.5 * (1 + s) * Btp[1],
                        0.5 * (1 - r) * Btp[2] + 0.5 * (1 + r) * Btp[3]])
        iJm = np.array([[iJ[0], iJ[1]], [iJ[2], iJ[3]]])
        Ks += Bs.T @ (iJm.T @ Cs @ iJm) @ Bs * adj
        Kab += Bm.T @ Cm @ Bi * adj
        Kbb += Bi.T @ Cm @ Bi * adj
    Kc = Kab @ np.linalg.solve(Kbb, Kab.T) if np.abs(Kbb).max() > 0 else np.zeros((24, 24))
    return dict(Km=Km, Kb=Kb, Kd=Kd, Ks=Ks, Kc=Kc, phi2=phi2, saturated=not (trial < 1.0))


def legacy_ke(lc, E, nu, h, **kw):
    P = legacy_parts(lc, E, nu, h, **kw)
    return P["Km"] + P["Kb"] + P["Kd"] + P["phi2"] * P["Ks"] - P["Kc"]


def legacy_dke_dh(lc, E, nu, h, **kw):
    """d/dh of ``legacy_ke``: (Km + Kd - Kc)/h + 3 Kb/h + d(phi^2 h)/dh Ks/h."""
    P = legacy_parts(lc, E, nu, h, **kw)
    dphi2h = 1.0 if P["saturated"] else 3.0 * P["phi2"]
    return (P["Km"] + P["Kd"] - P["Kc"]) / h + 3.0 * P["Kb"] / h + dphi2h * P["Ks"] / h


# --------------------------------------------------------------------------
# membrane modes, Gauss-point stresses and the principal-transverse KG
# --------------------------------------------------------------------------

def membrane_map(lc, Cm, n_modes):
    """A = -Kbb^-1 Kab^T (n_modes x 24), the condensed mode amplitudes per
    unit nodal displacement (FEMKernels.jl:6708-6785, 6787-6864)."""
    Kab = np.zeros((24, n_modes)); Kbb = np.zeros((n_modes, n_modes))
    for r, s in GPS:
        det, iJ, dx, dy, _ = _gp(lc, r, s)
        adj = max(abs(det), 1e-12)
        Bm, Bi = Bm_mat(dx, dy), mode_B(r, s, iJ, n_modes)
        Kab += Bm.T @ Cm @ Bi * adj
        Kbb += Bi.T @ Cm @ Bi * adj
    return -np.linalg.solve(Kbb, Kab.T)


def sigma_gp(lc, u, E, nu, h, n_modes):
    """Membrane stresses N/h (4 x 3) at the Gauss points with the condensed
    modes (quad4_membrane_force_field, FEMKernels.jl:6341-6539)."""
    Cm = E * h / (1 - nu * nu) * iso_D(nu)
    alpha = membrane_map(lc, Cm, n_modes) @ u
    out = np.zeros((4, 3))
    for g, (r, s) in enumerate(GPS):
        det, iJ, dx, dy, _ = _gp(lc, r, s)
        eps = Bm_mat(dx, dy) @ u + mode_B(r, s, iJ, n_modes) @ alpha
        out[g] = Cm @ eps / h
    return out


def principal2d(sxx, syy, sxy):
    mean, hd = 0.5 * (sxx + syy), 0.5 * (sxx - syy)
    rad = np.sqrt(hd * hd + sxy * sxy)
    if rad <= 1e-30:
        return ((mean, 1.0, 0.0), (mean, 0.0, 1.0))
    th = 0.5 * np.arctan2(2.0 * sxy, sxx - syy)
    c, s = np.cos(th), np.sin(th)
    return ((mean + rad, c, s), (mean - rad, -s, c))


def kg_grads(lc, r, s, A):
    """(dux/dx, dux/dy, duy/dx, duy/dy, duz/dx, duz/dy) per element DOF (6 x
    24) with the condensed modes' gradients added (FEMKernels.jl:7642-7704)."""
    det, iJ, dx, dy, _ = _gp(lc, r, s)
    G = np.zeros((6, 24))
    for k in range(4):
        G[0, 6 * k] = dx[k]; G[1, 6 * k] = dy[k]
        G[2, 6 * k + 1] = dx[k]; G[3, 6 * k + 1] = dy[k]
        G[4, 6 * k + 2] = dx[k]; G[5, 6 * k + 2] = dy[k]
    if A is not None:
        Bi = mode_B(r, s, iJ, A.shape[0])
        # mode column m of ux: Bi[0, m] = d/dx, Bi[2, m] = d/dy; of uy: Bi[2, m], Bi[1, m]
        for m in range(A.shape[0]):
            if m % 2 == 0:   # an ux mode
                G[0] += Bi[0, m] * A[m]; G[1] += Bi[2, m] * A[m]
            else:            # an uy mode
                G[2] += Bi[2, m] * A[m]; G[3] += Bi[1, m] * A[m]
    return det, G


def kg_local(lc, sig, h, A):
    """Principal-transverse KG (24 x 24, local) for Gauss-point stresses sig
    (4 x 3): h |J| sum_k l_k [(s_k gx - c_k gy)^2 + gz^2], g = c_k d/dx + s_k
    d/dy, directions with |l_k| <= 1e-30 dropped (FEMKernels.jl:6891-6972)."""
    KG = np.zeros((24, 24))
    for g, (r, s) in enumerate(GPS):
        det, Gr = kg_grads(lc, r, s, A)
        sc = h * max(abs(det), 1e-12)
        for lam, c, sn in principal2d(*sig[g]):
            if abs(lam) <= 1e-30:
                continue
            gx = c * Gr[0] + sn * Gr[1]
            gy = c * Gr[2] + sn * Gr[3]
            gz = c * Gr[4] + sn * Gr[5]
            va = sn * gx - c * gy
            KG += sc * lam * (np.outer(va, va) + np.outer(gz, gz))
    return KG


def q4_ctx_kg(lc, u_l, E, nu, h, n_modes):
    Cm = E * h / (1 - nu * nu) * iso_D(nu)
    return kg_local(lc, sigma_gp(lc, u_l, E, nu, h, n_modes), h, membrane_map(lc, Cm, n_modes))


def q4_rhs_unit_basis(lc, E, nu, h, phi_l, n_modes):
    """f_j = phi_l' KG(sigma(e_j)) phi_l, j = 0..23 (local frame)."""
    Cm = E * h / (1 - nu * nu) * iso_D(nu)
    A = membrane_map(lc, Cm, n_modes)
    f = np.zeros(24)
    for j in range(24):
        e = np.zeros(24); e[j] = 1.0
        f[j] = phi_l @ kg_local(lc, sigma_gp(lc, e, E, nu, h, n_modes), h, A) @ phi_l
    return f


# --------------------------------------------------------------------------
# model level
# --------------------------------------------------------------------------

def _block_T(R, n_nodes):
    T = np.zeros((6 * n_nodes, 6 * n_nodes))
    for k in range(2 * n_nodes):
        T[3 * k:3 * k + 3, 3 * k:3 * k + 3] = R
    return T


def _dofs(nodes):
    return np.array([6 * int(n) + c for n in nodes for c in range(6)])


def q4_element_terms(arr, e, u, phi, psi, n_modes):
    """[phi' dKe phi, phi' KG_ctx phi / h, psi' dKe u] of CQUAD4 e (global
    vectors), with the legacy dKe and the context KG."""
    nodes = [int(v) for v in arr.q4[e]]
    xyz = np.asarray(arr.xyz)
    R, lc = O.q4_frame(xyz[nodes].ravel())
    E, nu, h = (float(v) for v in arr.q4prop[e][:3])
    T = _block_T(R, 4)
    d = _dofs(nodes)
    ul, pl, sl = T @ u[d], T @ phi[d], T @ psi[d]
    dK = legacy_dke_dh(lc, E, nu, h)
    KGc = q4_ctx_kg(lc, ul, E, nu, h, n_modes)
    return np.array([pl @ dK @ pl, pl @ KGc @ pl / h, sl @ dK @ ul])


def adjoint_load(arr, lam, phi, n_modes):
    """lambda d(phi' KG phi)/du in the reference's default evaluation
    (global, all DOFs)."""
    xyz = np.asarray(arr.xyz)
    g = np.zeros(arr.ndof)
    for e in range(arr.n_q4):
        nodes = [int(v) for v in arr.q4[e]]
        R, lc = O.q4_frame(xyz[nodes].ravel())
        E, nu, h = (float(v) for v in arr.q4prop[e][:3])
        T = _block_T(R, 4)
        d = _dofs(nodes)
        g[d] += T.T @ q4_rhs_unit_basis(lc, E, nu, h, T @ phi[d], n_modes)
    for e in range(arr.n_t3):
        nodes = [int(v) for v in arr.t3[e]]
        p = xyz[nodes].ravel()
        d = _dofs(nodes)
        pe = phi[d]
        for j in range(18):
            ej = np.zeros(18); ej[j] = 1.0
            KG, _ = O.t3_kg_global(p, arr.t3prop[e], ej)
            g[d[j]] += pe @ KG @ pe
    for b in range(arr.n_bar):
        n1, n2 = (int(v) for v in arr.bar[b])
        d = _dofs([n1, n2])
        pe = phi[d]
        for j in range(12):
            ej = np.zeros(12); ej[j] = 1.0
            KG = np.zeros((12, 12))
            O.lib().or_bar_kg_global(O._p(O._c(xyz[n1], float)), O._p(O._c(xyz[n2], float)),
                                     O._p(O._c(arr.bar_vref[b], float)), O._p(O._c(arr.barprop[b], float)),
                                     O._p(ej), O._p(KG), None)
            g[d[j]] += pe @ KG @ pe
    return lam * g


def sensitivities(arr, res, mode, pids_groups, shell_pids):
    """d lambda_mode / d T_p for each group (a list of PSHELL id lists, one
    per design variable) from an oracle solve ``res`` (oracle.solve): the
    reference's default analytic path.  Returns (list of {"PID_p": value}
    per design variable, per-element Q4 terms, adjoint load)."""
    lam = float(res["eigenvalues"][mode])
    phi = np.asarray(res["modes"][:, mode], dtype=float)
    u = np.asarray(res["u"], dtype=float)
    K, KG, free = res["K"], res["KG"], np.asarray(res["free"])
    n_modes = 4 if arr.n_bar > 0 else 6
    g = adjoint_load(arr, lam, phi, n_modes)
    fidx = np.flatnonzero(free) if free.dtype == bool else free
    Kff = sp.csc_matrix(K[fidx][:, fidx])
    psi = np.zeros(arr.ndof)
    psi[fidx] = spla.spsolve(Kff, g[fidx])
    pkp = float(phi @ (KG @ phi))
    terms = np.array([q4_element_terms(arr, e, u, phi, psi, n_modes) for e in range(arr.n_q4)])
    # CTRIA3 rows: dK/dh by central differences of the C oracle's element
    # (the reference differentiates its own forward triangle), KG/h at u
    xyz = np.asarray(arr.xyz)
    t3rows = []
    for e in range(arr.n_t3):
        nodes = [int(v) for v in arr.t3[e]]
        p = xyz[nodes].ravel()
        d = _dofs(nodes)
        prop = np.array(arr.t3prop[e], dtype=float)
        h = prop[2]
        dh = 1e-6 * h
        pp, pm = prop.copy(), prop.copy()
        pp[2] += dh; pm[2] -= dh
        dK = (O.t3_ke_global(p, pp, arr.k6rot)[0] - O.t3_ke_global(p, pm, arr.k6rot)[0]) / (2 * dh)
        KGe, _ = O.t3_kg_global(p, prop, u[d])
        t3rows.append([phi[d] @ dK @ phi[d], phi[d] @ KGe @ phi[d] / h, psi[d] @ dK @ u[d]])
    all_terms = np.vstack([terms.reshape(-1, 3), np.array(t3rows).reshape(-1, 3)])
    shell_pids = np.asarray(shell_pids, dtype=np.int64)
    out = []
    for pids in pids_groups:
        single = len(pids) == 1 and set(shell_pids.tolist()) <= {int(pids[0])} and arr.n_bar == 0
        grp = {}
        for pid in pids:
            rows = shell_pids == int(pid)
            if not rows.any():
                grp["PID_%d" % int(pid)] = 0.0
                continue
            t = all_terms[rows].sum(axis=0)
            t2 = pkp / float(np.asarray(arr.q4prop)[:, 2][0] if arr.n_q4 else arr.t3prop[0][2]) if single else t[1]
            grp["PID_%d" % int(pid)] = float(-(t[0] + lam * t2 - t[2]) / pkp)
        out.append(grp)
    return out, all_terms, g
